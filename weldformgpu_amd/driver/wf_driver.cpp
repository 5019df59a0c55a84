// wf_driver.cpp — command-line host driver over libwfhip: reads a WeldFormGPU
// JSON deck, builds the domain through the C-ABI and runs the Fraser loop
// with the reference's output cadence and CSV snapshot format
// (src/WeldFormGPU.cu:142-926, src/cuda/SolverFraser.cu:30-409,
// src/cuda/Domain_d.cu:556-587).
//
//   wfdriver deck.json [--out DIR] [--max-steps N] [--dry-run]
//
// --dry-run builds the setup on the host only (no GPU) and prints it as JSON.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/wf_abi.h"
#include "deck.hpp"

namespace {

int check(int rc, const char* what) {
  if (rc != WF_OK) {
    std::fprintf(stderr, "wfdriver: %s failed (%d): %s\n", what, rc, wf_last_error());
    std::exit(1);
  }
  return rc;
}

template <class T>
std::vector<T> get(wf_domain* d, int field, int n, int comps) {
  std::vector<T> v((size_t)n * comps);
  check(wf_get_field(d, field, v.data(), v.size() * sizeof(T)), "wf_get_field");
  return v;
}

struct MeshRows {  // the rigid "particles" the reference appends after the solid
  std::vector<double> c0;  // initial centroids
  std::vector<double> radius;
  std::vector<int> el;  // connectivity, this mesh's node numbering
  int zone = 0;
};

// Domain_d::WriteCSV (Domain_d.cu:556-587): one row per particle, then one
// row per rigid triangle (the reference's rigid particles at the centroids)
void write_csv(const std::string& path, wf_domain* d, int n, const std::vector<MeshRows>& rows,
               double rho) {
  auto ID = get<int>(d, WF_F_ID, n, 1);
  auto x = get<double>(d, WF_F_X, n, 3);
  auto u = get<double>(d, WF_F_U, n, 3);
  auto v = get<double>(d, WF_F_V, n, 3);
  auto a = get<double>(d, WF_F_A, n, 3);
  auto r = get<double>(d, WF_F_RHO, n, 1);
  auto p = get<double>(d, WF_F_P, n, 1);
  auto seq = get<double>(d, WF_F_SIGMA_EQ, n, 1);
  auto ep = get<double>(d, WF_F_PL_STRAIN, n, 1);
  auto nb = get<int>(d, WF_F_NB_COUNT, n, 1);
  auto cnb = get<int>(d, WF_F_CONT_NB_COUNT, n, 1);
  auto cf = get<double>(d, WF_F_CONTFORCE, n, 3);
  auto nr = get<double>(d, WF_F_NORMAL, n, 3);
  auto h = get<double>(d, WF_F_H, n, 1);
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    std::fprintf(stderr, "wfdriver: cannot write %s\n", path.c_str());
    std::exit(1);
  }
  std::fprintf(f,
               "ID, X, Y, Z, posX,posY,posZ,Ux, Uy, Uz, Vx, Vy, Vz, Ax, Ay, Az, rho, p, SigmaEq, "
               "Pl_Strain, Nb, ContNb, CFx, CFy, CFz, Nx, Ny, Nz, h\n");
  const char* fmt =
      "%d, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, "
      "%.6e, %.6e, %.6e, %.6e, %.6e, %d, %d, %.6e, %.6e, %.6e, %.6e, %.6e, %.6e, %6e\n";
  for (int i = 0; i < n; i++) {
    const double* X = &x[3 * (size_t)i];
    std::fprintf(f, fmt, ID[i], X[0], X[1], X[2], X[0], X[1], X[2], u[3 * i], u[3 * i + 1],
                 u[3 * i + 2], v[3 * i], v[3 * i + 1], v[3 * i + 2], a[3 * i], a[3 * i + 1],
                 a[3 * i + 2], r[i], p[i], seq[i], ep[i], nb[i], cnb[i], cf[3 * i], cf[3 * i + 1],
                 cf[3 * i + 2], nr[3 * i], nr[3 * i + 1], nr[3 * i + 2], h[i]);
  }
  for (size_t m = 0; m < rows.size(); m++) {
    int nn = 0, ne = 0, zone = 0;
    check(wf_mesh_info(d, (int)m, &nn, &ne, &zone), "wf_mesh_info");
    std::vector<double> nodes(3 * (size_t)nn), nv(3 * (size_t)nn), c(3 * (size_t)ne),
        nrm(3 * (size_t)ne), pp(ne);
    check(wf_get_mesh(d, (int)m, nodes.data(), nv.data(), c.data(), nrm.data(), pp.data()),
          "wf_get_mesh");
    const std::vector<int>& el = rows[m].el;
    for (int e = 0; e < ne; e++) {
      const double* C = &c[3 * (size_t)e];
      const double* C0 = &rows[m].c0[3 * (size_t)e];
      double ve[3];  // UpdateContactParticles: mean of the node velocities
      for (int k = 0; k < 3; k++)
        ve[k] = (nv[3 * el[3 * e] + k] + nv[3 * el[3 * e + 1] + k] + nv[3 * el[3 * e + 2] + k]) / 3.;
      std::fprintf(f, fmt, zone, C[0], C[1], C[2], C[0], C[1], C[2], C[0] - C0[0], C[1] - C0[1],
                   C[2] - C0[2], ve[0], ve[1], ve[2], 0.0, 0.0, 0.0, rho, 0.0, 0.0, 0.0, 0, 0,
                   0.0, 0.0, 0.0, nrm[3 * e], nrm[3 * e + 1], nrm[3 * e + 2],
                   1.1 * rows[m].radius[e]);  // h = hfac * radius, hfac 1.1 (WeldFormGPU.cu:565)
    }
  }
  std::fclose(f);
}

void print_setup(const wfd::Deck& dk, const wfd::Setup& s) {
  const wf_config& c = s.cfg;
  std::printf("{\"n\": %d, \"dt\": %.17g, \"h\": %.17g, \"Cs\": %.17g, \"K\": %.17g, \"G\": %.17g,"
              " \"mass\": %.17g, \"sy0\": %.17g, \"T0\": %.17g, \"model\": %d, \"precision\": %d,"
              " \"return_map\": %d, \"thermal\": %d, \"contact\": %d, \"contact_law\": %d,"
              " \"fric_sta\": %.17g, \"fric_dyn\": %.17g, \"sim_time\": %.17g, \"out_time\": %.17g,",
              s.n, s.dt, s.h, s.Cs, s.K, s.G, s.mass, c.mat.sy0, s.T0, c.mat.model, c.precision,
              c.return_map, c.thermal, c.contact, c.contact_law, c.fric_sta, c.fric_dyn,
              dk.sim_time, dk.out_time);
  std::printf(" \"fric_heat\": %.17g, \"rebuild_policy\": %d, \"ns_skin\": %.17g,"
              " \"ep_hot\": %.17g, \"ns_global_every\": %d, \"axisymm\": %d,",
              c.fric_heat, c.rebuild_policy, c.ns_skin, c.ep_hot, c.ns_global_every, c.axisymm);
  std::printf(" \"mlp\": [%d, %d, %.17g, %d, %.17g],", c.mlp_train_interval, c.mlp_train_iters,
              c.mlp_lr, c.mlp_interval_max, c.mlp_refit_ratio);
  std::printf(" \"zones\": [");
  for (size_t z = 0; z < dk.zones.size(); z++)
    std::printf("%s[%d, %d]", z ? ", " : "", dk.zones[z].id, s.zone_count[z]);
  std::printf("], \"bcs\": [");
  for (size_t b = 0; b < s.bc_zone.size(); b++)
    std::printf("%s[%d, %.17g, %.17g, %.17g]", b ? ", " : "", s.bc_zone[b], s.bc_val[3 * b],
                s.bc_val[3 * b + 1], s.bc_val[3 * b + 2]);
  std::printf("], \"meshes\": [");
  for (size_t m = 0; m < s.meshes.size(); m++)
    std::printf("%s[%d, %zu, %zu, %.17g, %.17g, %.17g]", m ? ", " : "", s.meshes[m].zone,
                s.meshes[m].nodes.size() / 3, s.meshes[m].el.size() / 3, s.meshes[m].vel[0],
                s.meshes[m].vel[1], s.meshes[m].vel[2]);
  std::printf("], \"mesh_rotation\": [");  // angular velocity and rotation centre
  for (size_t m = 0; m < s.meshes.size(); m++)
    std::printf("%s[%.17g, %.17g, %.17g, %.17g, %.17g, %.17g]", m ? ", " : "",
                s.meshes[m].omega[0], s.meshes[m].omega[1], s.meshes[m].omega[2],
                s.meshes[m].center[0], s.meshes[m].center[1], s.meshes[m].center[2]);
  std::printf("]}\n");
}

std::string csv_name(const std::string& dir, double t) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "out_%.6f.csv", t);  // SolverFraser.cu:144,330
  return dir + "/" + buf;
}

}  // namespace

int main(int argc, char** argv) {
  std::string deck_path, out_dir = ".";
  bool dry = false;
  long max_steps_arg = -1;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "--dry-run") dry = true;
    else if (a == "--out" && i + 1 < argc) out_dir = argv[++i];
    else if (a == "--max-steps" && i + 1 < argc) max_steps_arg = std::atol(argv[++i]);
    else if (a[0] != '-') deck_path = a;
    else {
      std::fprintf(stderr, "wfdriver: unknown option %s\n", a.c_str());
      return 2;
    }
  }
  if (deck_path.empty()) {
    std::fprintf(stderr, "usage: wfdriver deck.json [--out DIR] [--max-steps N] [--dry-run]\n");
    return 2;
  }
  std::ifstream in(deck_path);
  if (!in) {
    std::fprintf(stderr, "wfdriver: cannot read %s\n", deck_path.c_str());
    return 2;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  wfd::Deck dk;
  wfd::Setup s;
  try {
    dk = wfd::parse_deck(wfj::parse(ss.str()));
    const size_t slash = deck_path.find_last_of('/');
    dk.dir = slash == std::string::npos ? "." : deck_path.substr(0, slash);
    s = wfd::build_setup(dk);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "wfdriver: %s\n", e.what());
    return 2;
  }
  if (dry) {
    print_setup(dk, s);
    return 0;
  }
  const long max_steps = max_steps_arg >= 0 ? max_steps_arg : dk.max_steps;
  const int n = s.n;
  wf_domain* d = nullptr;
  check(wf_create(&s.cfg, &d), "wf_create");
  check(wf_set_field(d, WF_F_X, s.x.data(), s.x.size() * sizeof(double)), "x");
  check(wf_fill_field(d, WF_F_RHO, dk.rho), "rho");    // SetDensity
  check(wf_fill_field(d, WF_F_RHO0, dk.rho), "rho0");
  check(wf_fill_field(d, WF_F_H, s.h), "h");            // Set_h
  check(wf_fill_field(d, WF_F_M, s.mass), "m");
  check(wf_fill_field(d, WF_F_CS, s.Cs), "Cs");         // SetCs
  check(wf_fill_field(d, WF_F_G, s.G), "G");            // SetShearModulus
  check(wf_fill_field(d, WF_F_SIGMA_Y, s.cfg.mat.sy0), "sigma_y");  // SetSigmay
  check(wf_fill_field(d, WF_F_T, s.T0), "T");
  if (dk.thermal) {
    check(wf_fill_field(d, WF_F_K_T, dk.k_T), "k_T");
    check(wf_fill_field(d, WF_F_CP_T, dk.cp_T), "cp_T");
  }
  check(wf_set_field(d, WF_F_ID, s.ID.data(), s.ID.size() * sizeof(int)), "ID");  // SetID
  check(wf_set_bcs(d, (int)s.bc_zone.size(), s.bc_zone.data(), s.bc_val.data()), "wf_set_bcs");
  std::vector<MeshRows> rows;
  for (const wfd::Mesh& me : s.meshes) {
    int mi = 0;
    const int nn = (int)(me.nodes.size() / 3), ne = (int)(me.el.size() / 3);
    check(wf_add_mesh(d, me.nodes.data(), nn, me.el.data(), ne, me.zone, &mi), "wf_add_mesh");
    check(wf_set_mesh_motion(d, mi, me.vel, me.omega, me.center), "wf_set_mesh_motion");
    MeshRows r;
    r.zone = me.zone;
    r.el = me.el;
    r.c0.resize(3 * (size_t)ne);
    r.radius.resize(ne);
    for (int e = 0; e < ne; e++) {
      double c[3];
      for (int k = 0; k < 3; k++)
        c[k] = (me.nodes[3 * me.el[3 * e] + k] + me.nodes[3 * me.el[3 * e + 1] + k] +
                me.nodes[3 * me.el[3 * e + 2] + k]) / 3.;
      double rmax = 0.0;
      for (int q = 0; q < 3; q++) {
        double d2 = 0.0;
        for (int k = 0; k < 3; k++) {
          const double t = me.nodes[3 * me.el[3 * e + q] + k] - c[k];
          d2 += t * t;
        }
        rmax = std::fmax(rmax, std::sqrt(d2));
      }
      for (int k = 0; k < 3; k++) r.c0[3 * e + k] = c[k];
      r.radius[e] = rmax;
    }
    rows.push_back(r);
  }
  check(wf_nsearch_build(d), "wf_nsearch_build");
  std::printf("wfdriver: %d particles, %zu rigid meshes, dt %.6e, h %.6e, Cs %.6e\n", n,
              s.meshes.size(), s.dt, s.h, s.Cs);
  write_csv(csv_name(out_dir, 0.0), d, n, rows, dk.rho);
  // MechFraserSolve: while (Time < tf) { step; Time += dt; output when Time >= t_out }
  double time = 0.0, t_out = dk.out_time;
  long steps = 0;
  while (time < dk.sim_time && (max_steps < 0 || steps < max_steps)) {
    // steps up to (and including) the next output time, with the reference's
    // accumulation of Time
    int k = 0;
    double t = time;
    bool out = false;
    while (t < dk.sim_time && (max_steps < 0 || steps + k < max_steps)) {
      t += s.dt;
      k++;
      if (t >= t_out) {
        out = true;
        break;
      }
    }
    if (k == 0) break;
    check(wf_step(d, k), "wf_step");
    steps += k;
    time = t;
    if (out) {
      check(wf_snapshot_positions(d), "wf_snapshot_positions");  // x_h = x
      write_csv(csv_name(out_dir, time), d, n, rows, dk.rho);
      t_out += dk.out_time;
      auto u = get<double>(d, WF_F_U, n, 3);
      auto ep = get<double>(d, WF_F_PL_STRAIN, n, 1);
      double mx[3] = {0, 0, 0}, mps = 0.0, cfs = 0.0;
      for (int i = 0; i < n; i++) {
        for (int c = 0; c < 3; c++) mx[c] = std::fmax(mx[c], u[3 * i + c]);
        mps = std::fmax(mps, ep[i]);
      }
      std::printf("Time %.6e, step %ld, max disp %.6e %.6e %.6e, max pl_strain %.6e", time, steps,
                  mx[0], mx[1], mx[2], mps);
      if (!s.meshes.empty()) {
        check(wf_contact_force_sum(d, &cfs), "wf_contact_force_sum");
        std::printf(", contact force sum %.6e", cfs);
      }
      std::printf("\n");
      std::fflush(stdout);
    }
  }
  write_csv(out_dir + "/test.csv", d, n, rows, dk.rho);
  wf_stats st{};
  check(wf_get_stats(d, &st), "wf_get_stats");
  std::printf("Total steps: %ld, time %.6e, neighbour builds %d\n", steps, time, st.ns_builds);
  wf_destroy(d);
  return 0;
}
