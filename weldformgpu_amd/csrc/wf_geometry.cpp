// wf_geometry.cpp — host-side lattice and plate-mesh builders of the C-ABI
// (no device code).
// The loops keep the reference's generation order and its floating-point
// steps (positions by formula for the box, by accumulation for the cylinder)
// so that counts and coordinates agree bit for bit with a reference run.
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/wf_abi.h"

namespace {

// number of particles on a half row at row offset xinc from the axis:
// calcHalfPartCount, Domain.cpp:473-488
int half_row(double r, double R, int xinc) {
  if (xinc <= 0) return -1;
  int cnt = 1;
  double yp = r;
  const double xp = r + (double)(xinc - 1) * 2. * r;
  double rad = std::sqrt(yp * yp + xp * xp);
  while (rad <= R - r) {
    yp += 2. * r;
    rad = std::sqrt(yp * yp + xp * xp);
    cnt++;
  }
  return cnt - 1;
}

inline void put(double* xyz, int cap, int k, double x, double y, double z) {
  if (xyz && k < cap) {
    xyz[3 * k] = x;
    xyz[3 * k + 1] = y;
    xyz[3 * k + 2] = z;
  }
}

// Nastran small-field real: 8 characters, exponent letter optional
// ("1.5-3" = 1.5e-3, "-2.5+2" = -250); a leading sign is the number's sign
double nastran_real(std::string t) {
  size_t b = t.find_first_not_of(' ');
  if (b == std::string::npos) return 0.0;
  t = t.substr(b);
  t = t.substr(0, t.find_last_not_of(' ') + 1);
  if (t.find_first_of("eEdD") == std::string::npos) {
    const size_t k = t.find_first_of("+-", 1);
    if (k != std::string::npos) t.insert(k, "E");
  } else {
    for (char& ch : t)
      if (ch == 'd' || ch == 'D') ch = 'E';
  }
  return std::strtod(t.c_str(), nullptr);
}

std::string field(const std::string& line, int k) {  // 8-character field k (0-based)
  const size_t p = 8 * (size_t)k;
  return p < line.size() ? line.substr(p, 8) : std::string();
}

}  // namespace

extern "C" {

int wf_gen_box(const double V[3], double Lx, double Ly, double Lz, double r, double density,
               double* xyz, int cap, double* mass_out) {
  int count = 0;
  double mx = V[0], my = V[1], mz = V[2];
  for (size_t k = 0;; k++) {
    // the first layer is always emitted: the loop test starts from V itself
    double zp = (k == 0) ? V[2] : V[2] + (2.0 * k + 1) * r;
    if (!(zp <= (V[2] + Lz - r))) break;
    for (size_t j = 0;; j++) {
      double yp = (j == 0) ? V[1] : V[1] + (2.0 * j + 1) * r;
      if (!(yp <= (V[1] + Ly - r))) break;
      for (size_t i = 0;; i++) {
        double xp = (i == 0) ? V[0] : V[0] + (2 * i + 1) * r;
        if (!(xp <= (V[0] + Lx - r))) break;
        double x = V[0] + (2.0 * i + 1) * r;
        double y = V[1] + (2.0 * j + 1) * r;
        double z = V[2] + (2.0 * k + 1) * r;
        put(xyz, cap, count, x, y, z);
        if (x > mx) mx = x;
        if (y > my) my = y;
        if (z > mz) mz = z;
        count++;
      }
    }
  }
  if (mass_out && count > 0) {
    mx += r;
    my += r;
    mz += r;
    *mass_out = (mx - V[0]) * (my - V[1]) * (mz - V[2]) * density / count;
  }
  return count;
}

int wf_gen_cylinder(const double V[3], double R, double Lz, double r, double density,
                    double* xyz, int cap, double* mass_out) {
  int count = 0;
  const int nxy = half_row(r, R, 1);
  for (double zp = V[2] + r; zp <= (V[2] + Lz - r); zp += 2. * r) {
    double yp = V[1] - r - (2. * r * (nxy - 1));
    int row = nxy, step = -1;
    for (int j = 0; j < 2 * nxy; j++) {
      const int nx = half_row(r, R, row);
      double xp = V[0] - r - (2. * r * (nx - 1));
      for (int i = 0; i < 2 * nx; i++) {
        put(xyz, cap, count, xp, yp, zp);
        count++;
        xp += 2. * r;
      }
      yp += 2. * r;
      row += step;
      if (row < 1) {  // crossed the axis: rows grow again
        row = 1;
        step = 1;
      }
    }
  }
  if (mass_out && count > 0) *mass_out = M_PI * R * R * Lz * density / count;
  return count;
}

int wf_gen_cylinder_polar(int tag, double R, double Lz, double r, double density, double ang,
                          int rows, double r_i, double* xyz, int* ID, int cap, double* mass_out) {
  if (!(r > 0.0) || !(ang > 0.0) || rows < 0) return WF_E_ARG;
  const bool slice = ang < 2.0 * M_PI;
  int count = 0;
  // the radius advances by accumulation and the layer height too, as the
  // reference's loops do, so ring and layer counts agree at the ends
  for (double zp = r; zp <= (r + Lz - r); zp += 2.0 * r) {
    for (double ri = r_i; ri < R; ri += 2. * r) {
      int ring = 1;
      double da = 0.0;
      if (ri != 0.) {
        ring = (int)std::ceil((ang * ri) / (2. * r));
        da = ang / ring;
      }
      for (int k = 0; k < ring; k++) {
        int id = tag;
        if (slice && k < rows) id = 2;
        else if (slice && k > ring - 1 - rows) id = 3;
        const double xp = ri * std::cos(k * da), yp = ri * std::sin(k * da);
        if (std::fabs(xp) < r / 10 && std::fabs(yp) < r / 10) id = 4;  // on the axis
        put(xyz, cap, count, xp, yp, zp);
        if (ID && count < cap) ID[count] = id;
        count++;
      }
    }
  }
  if (mass_out && count > 0) *mass_out = M_PI * R * R * Lz * density / count;
  return count;
}

int wf_assign_zone(const double* xyz, int n, const double start[3], const double end[3], int id,
                   int* ID) {
  if (!xyz || !ID || !start || !end || n < 0) return WF_E_ARG;
  int hits = 0;
  for (int a = 0; a < n; a++) {
    const double* p = xyz + 3 * a;
    bool in = true;
    for (int c = 0; c < 3; c++)
      if (p[c] < start[c] || p[c] > end[c]) in = false;
    if (in) {
      ID[a] = id;
      hits++;
    }
  }
  return hits;
}

// NastranReader::read + TriMesh_d::ReadFromNastran (NastranReader.cpp:5-153,
// Mesh_reader.cu:5-60): GRID cards give node id (field 2) and x, y, z
// (fields 4-6); CTRIA3 cards give three node ids (fields 4-6), mapped to
// 0-based node order.  flip reverses each triangle (normal -n).  Returns the
// triangle count, or a negative WF_E_* code.
int wf_read_nastran(const char* path, int flip, double* nodes, int node_cap, int* elnode,
                    int elem_cap, int* n_nodes) {
  if (!path) return WF_E_ARG;
  std::ifstream in(path);
  if (!in) return WF_E_ARG;
  std::vector<double> xyz;
  std::vector<int> ids;
  std::vector<int> tri;
  std::map<int, int> pos;
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.compare(0, 4, "GRID") == 0 && line.compare(0, 5, "GRID*") != 0) {
      const int id = std::atoi(field(line, 1).c_str());
      pos[id] = (int)ids.size();
      ids.push_back(id);
      for (int c = 0; c < 3; c++) xyz.push_back(nastran_real(field(line, 3 + c)));
    } else if (line.compare(0, 6, "CTRIA3") == 0) {
      for (int c = 0; c < 3; c++) tri.push_back(std::atoi(field(line, 3 + c).c_str()));
    }
  }
  const int nn = (int)ids.size(), ne = (int)(tri.size() / 3);
  if (n_nodes) *n_nodes = nn;
  for (int& t : tri) {
    auto it = pos.find(t);
    if (it == pos.end()) return WF_E_ARG;  // element names an unknown node
    t = it->second;
  }
  if (flip)
    for (int e = 0; e < ne; e++) std::swap(tri[3 * e + 1], tri[3 * e + 2]);
  if (nodes && node_cap >= nn)
    for (size_t k = 0; k < xyz.size(); k++) nodes[k] = xyz[k];
  if (elnode && elem_cap >= ne)
    for (size_t k = 0; k < tri.size(); k++) elnode[k] = tri[k];
  return ne;
}

// z-normal plate: nodes row by row (x fastest) with the spacing accumulated,
// two triangles per quad in the clockwise order that gives normal -z
// (TriMesh_d::AxisPlaneMesh(2, false, ...), Mesh.cu:48-194)
int wf_gen_plane_mesh(const double p1[3], const double p2[3], int dens, double* nodes,
                      int node_cap, int* elnode, int elem_cap, int* n_nodes) {
  if (!p1 || !p2 || dens <= 0 || dens > 4096) return WF_E_ARG;
  const int nn = (dens + 1) * (dens + 1), ne = 2 * dens * dens;
  if (n_nodes) *n_nodes = nn;
  const double dl = (p2[0] - p1[0]) / dens;
  if (nodes && node_cap >= nn) {
    int k = 0;
    double x2 = p1[1];
    for (int j = 0; j <= dens; j++) {
      double x1 = p1[0];
      for (int i = 0; i <= dens; i++) {
        nodes[3 * k] = x1;
        nodes[3 * k + 1] = x2;
        nodes[3 * k + 2] = p1[2];
        k++;
        x1 += dl;
      }
      x2 += dl;
    }
  }
  if (elnode && elem_cap >= ne) {
    int e = 0;
    for (int j = 0; j < dens; j++)
      for (int i = 0; i < dens; i++) {
        const int n0 = (dens + 1) * j + i, n1 = n0 + 1;
        const int n2 = (dens + 1) * (j + 1) + i, n3 = n2 + 1;
        const int tri[2][3] = {{n0, n2, n1}, {n1, n2, n3}};
        for (int t = 0; t < 2; t++, e++)
          for (int c = 0; c < 3; c++) elnode[3 * e + c] = tri[t][c];
      }
  }
  return ne;
}

}  // extern "C"
