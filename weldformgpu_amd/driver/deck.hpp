// deck.hpp — the WeldFormGPU JSON input deck (Configuration, Materials,
// DomainBlocks, DomainZones, BoundaryConditions, RigidBodies, Contact) turned
// into a libwfhip configuration plus the initial particle arrays, the way
// src/WeldFormGPU.cu:142-906 builds its Domain_d.  Host only: the lattice,
// zone and plate builders are the library's host functions.
#pragma once

#include <cmath>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/wf_abi.h"
#include "json_min.hpp"

namespace wfd {

struct Zone {
  int id = 0;
  double start[3] = {0, 0, 0}, end[3] = {0, 0, 0};
};
struct Bc {
  int zone = 0, value_type = 0;
  double value[3] = {0, 0, 0}, value_ang[3] = {0, 0, 0};
};
struct Rigid {
  std::string type, file;
  double start[3] = {0, 0, 0}, dim[3] = {0, 0, 0}, move[3] = {0, 0, 0};
  // rotation centre of the body's angular velocity (an extension the
  // reference ignores; without it the body turns about the origin)
  double center[3] = {0, 0, 0};
  int dens = 10, zone = 0;
};

struct Deck {
  // Configuration (WeldFormGPU.cu:199-293)
  double r = 0.0, hfactor = 0.0, out_time = 0.0, sim_time = 0.0;
  double alpha = 1.0, beta = 0.0;
  // Materials[0] (:231-241, 753-792)
  std::string mat_type = "Bilinear";
  double rho = 0.0, E = 0.0, nu = 0.0, Fy = 0.0;
  std::vector<double> c;
  // DomainBlocks[0] (:313-400)
  std::string dom_type = "Box", grid_cs = "Cartesian";
  double slice_ang = 2.0000001 * M_PI;  // sliceAngle default (:311)
  double start[3] = {0, 0, 0}, L[3] = {0, 0, 0};
  std::vector<Zone> zones;      // DomainZones (:406-421)
  std::vector<Bc> bcs;          // BoundaryConditions (:457-485)
  std::vector<Rigid> rigid;     // RigidBodies (:425-434, 520-619)
  double fric_sta = 0.0, fric_dyn = 0.0;  // Contact[0] (:622-636)
  // "Extensions" (not read by the reference): options of this library
  int precision = WF_FP64, return_map = WF_RETURN_REFERENCE, surrogate = 0;
  int contact_law = WF_CONTACT_REFERENCE, quirks = WF_Q_ALL;
  double contact_kappa = 0.1;
  double fric_heat = 0.0;  // fraction of the sliding-friction power into the particle
  int rebuild_policy = WF_REBUILD_EVERY_N;
  double ns_skin = 0.0, ep_hot = 0.05;  // Extensions.rebuild (plastic-zone policy)
  int ns_global_every = 32;             // Extensions.rebuild.globalEvery (local re-lists)
  // Extensions.surrogateTraining: cadence and Adam settings of the online rounds
  int mlp_train_interval = 200, mlp_train_iters = 250, mlp_interval_max = 0;
  double mlp_lr = 2e-3, mlp_refit_ratio = 2.0;
  bool thermal = false;
  double T0 = -1.0, k_T = 0.0, cp_T = 0.0, taylor_quinney = 0.0, T_inf = 500.0, h_conv = 100.0;
  long max_steps = -1;
  std::string dir;  // the deck's directory, for relative mesh file names
};

inline Deck parse_deck(const wfj::Value& j) {
  using wfj::read;
  Deck d;
  const wfj::Value& cfg = j["Configuration"];
  read(cfg["particleRadius"], d.r);
  read(cfg["hFactor"], d.hfactor);
  read(cfg["outTime"], d.out_time);
  read(cfg["simTime"], d.sim_time);
  read(cfg["artifViscAlpha"], d.alpha);
  read(cfg["artifViscBeta"], d.beta);
  const wfj::Value& mat = j["Materials"][0];
  read(mat["type"], d.mat_type);
  read(mat["density0"], d.rho);
  read(mat["youngsModulus"], d.E);
  read(mat["poissonsRatio"], d.nu);
  read(mat["yieldStress0"], d.Fy);
  wfj::read_array(mat["const"], d.c);
  const wfj::Value& blk = j["DomainBlocks"][0];
  read(blk["type"], d.dom_type);
  read(blk["gridCoordSys"], d.grid_cs);
  read(blk["sliceAngle"], d.slice_ang);
  wfj::read3(blk["start"], d.start);
  wfj::read3(blk["dim"], d.L);
  for (const wfj::Value& z : j["DomainZones"].arr) {
    Zone zn;
    read(z["id"], zn.id);
    wfj::read3(z["start"], zn.start);
    wfj::read3(z["end"], zn.end);
    d.zones.push_back(zn);
  }
  for (const wfj::Value& b : j["BoundaryConditions"].arr) {
    Bc bc;
    read(b["zoneId"], bc.zone);
    read(b["valueType"], bc.value_type);  // decks spelling it "valuetype" get 0
    if (bc.value_type == 0) {
      wfj::read3(b["value"], bc.value);
      wfj::read3(b["valueAng"], bc.value_ang);
    }
    d.bcs.push_back(bc);
  }
  for (const wfj::Value& rb : j["RigidBodies"].arr) {
    Rigid r;
    if (!read(rb["type"], r.type)) continue;  // contact is on when a type is given
    wfj::read3(rb["start"], r.start);
    wfj::read3(rb["dim"], r.dim);
    read(rb["partSide"], r.dens);
    read(rb["zoneId"], r.zone);
    read(rb["fileName"], r.file);
    wfj::read3(rb["moveDir"], r.move);
    wfj::read3(rb["rotationCenter"], r.center);
    d.rigid.push_back(r);
  }
  const wfj::Value& ct = j["Contact"][0];
  read(ct["fricCoeffStatic"], d.fric_sta);
  read(ct["fricCoeffDynamic"], d.fric_dyn);
  const wfj::Value& ex = j["Extensions"];
  std::string s;
  if (read(ex["precision"], s)) d.precision = s == "mixed" ? WF_MIXED : WF_FP64;
  if (read(ex["returnMap"], s)) d.return_map = s == "newton" ? WF_RETURN_NEWTON : WF_RETURN_REFERENCE;
  read(ex["surrogate"], d.surrogate);
  if (read(ex["contactLaw"], s))
    d.contact_law = s == "stiffness" ? WF_CONTACT_STIFFNESS : WF_CONTACT_REFERENCE;
  read(ex["contactKappa"], d.contact_kappa);
  read(ex["frictionHeat"], d.fric_heat);
  const wfj::Value& rb = ex["rebuild"];
  if (!rb.is_null()) {
    if (read(rb["policy"], s)) {
      if (s == "plasticZone") d.rebuild_policy = WF_REBUILD_PLASTIC_ZONE;
      else if (s == "plasticZoneLocal") d.rebuild_policy = WF_REBUILD_PLASTIC_ZONE_LOCAL;
      else if (s != "everyN")
        throw std::runtime_error("Extensions.rebuild.policy " + s +
                                 " is not one of everyN, plasticZone, plasticZoneLocal");
    }
    read(rb["skin"], d.ns_skin);
    read(rb["epHot"], d.ep_hot);
    read(rb["globalEvery"], d.ns_global_every);
  }
  const wfj::Value& tr = ex["surrogateTraining"];
  if (!tr.is_null()) {
    read(tr["interval"], d.mlp_train_interval);
    read(tr["iterations"], d.mlp_train_iters);
    read(tr["learningRate"], d.mlp_lr);
    read(tr["intervalMax"], d.mlp_interval_max);
    read(tr["refitRatio"], d.mlp_refit_ratio);
  }
  read(ex["quirks"], d.quirks);
  double ms = -1.0;
  if (read(ex["maxSteps"], ms)) d.max_steps = (long)ms;
  const wfj::Value& th = ex["thermal"];
  if (!th.is_null()) {
    d.thermal = true;
    read(th["T0"], d.T0);
    read(th["k"], d.k_T);
    read(th["cp"], d.cp_T);
    read(th["taylorQuinney"], d.taylor_quinney);
    read(th["T_inf"], d.T_inf);
    read(th["h_conv"], d.h_conv);
  }
  if (!(d.r > 0.0)) throw std::runtime_error("Configuration.particleRadius must be positive");
  if (!(d.rho > 0.0 && d.E > 0.0)) throw std::runtime_error("Materials[0] needs density0 and youngsModulus");
  if (d.dom_type != "Box" && d.dom_type != "Cylinder")
    throw std::runtime_error("DomainBlocks[0].type " + d.dom_type + " is not supported (Box, Cylinder)");
  if (d.dom_type == "Cylinder" && d.grid_cs != "Cartesian" && d.grid_cs != "Cylindrical")
    throw std::runtime_error("DomainBlocks[0].gridCoordSys " + d.grid_cs +
                             " is not one of Cartesian, Cylindrical");
  for (const Rigid& r : d.rigid)
    if (r.type != "Plane" && r.type != "File")
      throw std::runtime_error("RigidBodies type " + r.type + " is not supported (Plane, File)");
  return d;
}

struct Mesh {
  std::vector<double> nodes;
  std::vector<int> el;
  int zone = 0;
  double vel[3] = {0, 0, 0}, omega[3] = {0, 0, 0}, center[3] = {0, 0, 0};
};

struct Setup {
  wf_config cfg{};
  int n = 0;
  double h = 0.0, K = 0.0, G = 0.0, Cs = 0.0, dt = 0.0, mass = 0.0, T0 = 0.0;
  std::vector<double> x;  // 3n
  std::vector<int> ID;
  std::vector<int> bc_zone;
  std::vector<double> bc_val;  // 3 per BC
  std::vector<int> zone_count;
  std::vector<Mesh> meshes;
};

// Johnson-Cook flow stress at eps_p = 0, rate 0 (taken as 1e-5), temperature
// T (Material.cuh:209-219) with the homologous temperature clamped to [0, 1]
inline double jc_initial_yield(const wf_material_params& m, double T) {
  double th = (T - m.Tt) / (m.Tm - m.Tt);
  th = th < 0.0 ? 0.0 : (th > 1.0 ? 1.0 : th);
  return m.A * (1.0 + m.C * std::log(1e-5 / m.eps0)) * (1.0 - std::pow(th, m.m));
}

inline Setup build_setup(const Deck& d) {
  Setup s;
  wf_config& c = s.cfg;
  wf_config_default(&c);
  const double dx = 2. * d.r;
  s.h = dx * d.hfactor;
  s.K = d.E / (3. * (1. - 2 * d.nu));
  s.G = d.E / (2. * (1. + d.nu));
  s.Cs = std::sqrt(s.K / d.rho);
  s.dt = 0.4 * s.h / (s.Cs + 10.0);  // VMAX = 10, WeldFormGPU.cu:62,906
  // lattice (Domain::AddBoxLength / AddCylinderLength / AddCylUniformLength,
  // chosen at WeldFormGPU.cu:341-375)
  int n = 0;
  bool axisymm = false;
  if (d.dom_type == "Box") {
    n = wf_gen_box(d.start, d.L[0], d.L[1], d.L[2], d.r, d.rho, nullptr, 0, nullptr);
    s.x.resize(3 * (size_t)n);
    wf_gen_box(d.start, d.L[0], d.L[1], d.L[2], d.r, d.rho, s.x.data(), n, &s.mass);
    s.ID.assign(n, 0);
  } else if (d.grid_cs == "Cartesian") {
    n = wf_gen_cylinder(d.start, d.L[0] / 2., d.L[2], d.r, d.rho, nullptr, 0, nullptr);
    s.x.resize(3 * (size_t)n);
    wf_gen_cylinder(d.start, d.L[0] / 2., d.L[2], d.r, d.rho, s.x.data(), n, &s.mass);
    s.ID.assign(n, 0);
  } else {
    // the whole circle only for a sliceAngle of exactly 2 pi; otherwise a
    // pi/16 slice of the outer half of the radius, edge rows marked 2 and 3,
    // and the axisymmetric acceleration clamp in the step loop
    const bool full = d.slice_ang == 2.0 * M_PI;
    const double ang = full ? 2.0 * M_PI : M_PI / 16.0, r_i = full ? 0.0 : d.L[0] / 4.0;
    axisymm = !full;
    n = wf_gen_cylinder_polar(0, d.L[0] / 2.0, d.L[2], d.r, d.rho, ang, 1, r_i, nullptr, nullptr,
                              0, nullptr);
    if (n > 0) {
      s.x.resize(3 * (size_t)n);
      s.ID.assign(n, 0);
      wf_gen_cylinder_polar(0, d.L[0] / 2.0, d.L[2], d.r, d.rho, ang, 1, r_i, s.x.data(),
                            s.ID.data(), n, &s.mass);
    }
  }
  if (n <= 0) throw std::runtime_error("the domain block holds no particles");
  s.n = n;
  for (const Zone& z : d.zones)
    s.zone_count.push_back(wf_assign_zone(s.x.data(), n, z.start, z.end, z.id, s.ID.data()));
  c.n_particles = n;
  c.solid_count = n;
  c.axisymm = axisymm ? 1 : 0;
  c.dt = s.dt;
  c.alpha = d.alpha;
  c.beta = d.beta;
  c.h_glob = s.h;
  c.precision = d.precision;
  c.quirks = d.quirks;
  c.surrogate = d.surrogate;
  c.return_map = d.return_map;
  wf_material_params& m = c.mat;
  m.E = d.E;
  m.nu = d.nu;
  m.sy0 = d.Fy;
  auto cc = [&](size_t k) { return k < d.c.size() ? d.c[k] : 0.0; };
  if (d.mat_type == "Bilinear") {
    m.model = WF_MAT_BILINEAR;
    m.Et = cc(0);
  } else if (d.mat_type == "Hollomon") {
    m.model = WF_MAT_HOLLOMON;
    m.K_h = cc(0);
    m.m_h = cc(1);
  } else if (d.mat_type == "JohnsonCook") {
    m.model = WF_MAT_JOHNSON_COOK;  // const order A, B, n, C, eps0, m, Tm, Tt
    m.A = cc(0);
    m.B = cc(1);
    m.n = cc(2);
    m.C = cc(3);
    m.eps0 = cc(4);
    m.m = cc(5);
    m.Tm = cc(6);
    m.Tt = cc(7);
  } else {
    throw std::runtime_error("unknown material type " + d.mat_type);
  }
  s.T0 = d.thermal && d.T0 >= 0.0 ? d.T0 : (m.model == WF_MAT_JOHNSON_COOK ? m.Tt : 0.0);
  if (m.model == WF_MAT_JOHNSON_COOK) m.sy0 = jc_initial_yield(m, s.T0);
  if (d.thermal) {
    c.thermal = 1;
    c.taylor_quinney = d.taylor_quinney;
    c.T_inf = d.T_inf;
    c.h_conv = d.h_conv;
  }
  for (const Bc& b : d.bcs) {
    if (b.value_type != 0) continue;  // amplitudes are parsed but unused by the reference
    s.bc_zone.push_back(b.zone);
    for (int k = 0; k < 3; k++) s.bc_val.push_back(b.value[k]);
  }
  if (s.bc_zone.size() > WF_MAX_BCS) throw std::runtime_error("too many boundary conditions");
  // rigid plates: AxisPlaneMesh(2, false, start, start + (dim.x, dim.y, 0), partSide);
  // a constant-velocity BC on the mesh's zone moves it (WeldFormGPU.cu:597-611)
  // "File": a Nastran triangle mesh translated by moveDir (WeldFormGPU.cu:546-558;
  // flipNormals is read there but not passed on, so the file's orientation holds)
  for (const Rigid& r : d.rigid) {
    Mesh me;
    me.zone = r.zone;
    for (int k = 0; k < 3; k++) me.center[k] = r.center[k];
    int nn = 0, ne = 0;
    if (r.type == "Plane") {
      const double p2[3] = {r.start[0] + r.dim[0], r.start[1] + r.dim[1], r.start[2]};
      ne = wf_gen_plane_mesh(r.start, p2, r.dens, nullptr, 0, nullptr, 0, &nn);
      if (ne <= 0) throw std::runtime_error("bad RigidBodies plate");
      me.nodes.resize(3 * (size_t)nn);
      me.el.resize(3 * (size_t)ne);
      wf_gen_plane_mesh(r.start, p2, r.dens, me.nodes.data(), nn, me.el.data(), ne, &nn);
    } else {
      std::string path = r.file;
      if (!std::ifstream(path) && !d.dir.empty()) path = d.dir + "/" + r.file;
      ne = wf_read_nastran(path.c_str(), 0, nullptr, 0, nullptr, 0, &nn);
      if (ne <= 0) throw std::runtime_error("cannot read the rigid mesh file " + r.file);
      me.nodes.resize(3 * (size_t)nn);
      me.el.resize(3 * (size_t)ne);
      wf_read_nastran(path.c_str(), 0, me.nodes.data(), nn, me.el.data(), ne, &nn);
      for (int q = 0; q < nn; q++)
        for (int k = 0; k < 3; k++) me.nodes[3 * q + k] += r.move[k];
    }
    for (const Bc& b : d.bcs)
      if (b.zone == r.zone && b.value_type == 0)
        for (int k = 0; k < 3; k++) {
          me.vel[k] = b.value[k];
          me.omega[k] = b.value_ang[k];
        }
    s.meshes.push_back(me);
  }
  c.rebuild_policy = d.rebuild_policy;
  c.ns_global_every = d.ns_global_every;
  c.mlp_train_interval = d.mlp_train_interval;
  c.mlp_train_iters = d.mlp_train_iters;
  c.mlp_lr = d.mlp_lr;
  c.mlp_interval_max = d.mlp_interval_max;
  c.mlp_refit_ratio = d.mlp_refit_ratio;
  c.ns_skin = d.ns_skin;
  c.ep_hot = d.ep_hot;
  if (!s.meshes.empty()) {
    c.contact = 1;
    c.contact_law = d.contact_law;
    c.contact_kappa = d.contact_kappa;
    c.contact_dx = dx;
    c.fric_sta = d.fric_sta;
    c.fric_dyn = d.fric_dyn;
    c.fric_heat = d.fric_heat;
  }
  return s;
}

}  // namespace wfd
