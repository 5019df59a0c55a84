// wf_pair.hip — the pair loops of the step (gfx950), FP64 and mixed
// precision, on one kernel template (k_pair_gather).
//
// One workgroup per tile of kTileP Morton-consecutive particles, two threads
// per particle.  The tile's staged set (the particles that are a neighbour of
// at least one of the tile's particles, in cell order) is written to LDS as
// {tile-frame position, particle index} slots; each thread walks half of its
// particle's neighbour list (16-bit staged positions in ascending order,
// four per 8-byte word, every other word), reads the neighbour's position and
// index from LDS and gathers the neighbour's packed record from global
// memory.  Two threads per particle give a CU twice the wavefronts for the
// same LDS; the loops' registers then allow 4-6 (FP32) or 3-4 (FP64)
// wavefronts per SIMD.  The per-pair physics is the reference's
// (InteractionAlt.cu, Thermal.cu, Geometry.cu), with per-particle factors
// (m/rho, sigma/rho^2, 1/h) formed once instead of once per pair; the
// policies (DensD/DensG, RatesD/RatesG, AccelD/AccelG, CondD/CondG, SurfD,
// CondDens) say what a neighbour record holds and what a pair adds.
#include <climits>
#include <type_traits>

#include "wf_device.hpp"
#include "wf_tiles.hpp"

namespace wf {

// Neighbour-slot reader over the quad layout: four 16-bit slots per 8 bytes,
// the next quad prefetched one quad ahead.
struct SlotReader {
  const uint2* q;
  int64_t npad;
  int i, cnt, k;
  uint2 cur, nxt;
  __device__ __forceinline__ void init(const TileLists& tl, int i_, int cnt_) {
    q = tl.slotq;
    npad = tl.npad;
    i = i_;
    cnt = cnt_;
    k = 0;
    cur = cnt > 0 ? q[i] : make_uint2(0u, 0u);
    nxt = cnt > 4 ? q[npad + i] : make_uint2(0u, 0u);
  }
  __device__ __forceinline__ int slot() const {
    unsigned w = (k & 2) ? cur.y : cur.x;
    return (int)((w >> ((k & 1) * 16)) & 0xFFFFu);
  }
  __device__ __forceinline__ void advance() {
    k++;
    if ((k & 3) == 0) {
      cur = nxt;
      if (k + 4 < cnt) nxt = q[(int64_t)((k >> 2) + 1) * npad + i];
    }
  }
  // to the start of the next quad
  __device__ __forceinline__ void skip_quad() {
    k = (k | 3) + 1;
    cur = nxt;
    if (k + 4 < cnt) nxt = q[(int64_t)((k >> 2) + 1) * npad + i];
  }
};

// Workgroups are dispatched round-robin over the 8 XCDs (XCD = block % 8).
// Give each XCD one contiguous run of tiles so that tiles sharing staged
// cells share that XCD's L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int nt) {
  const int q = nt / kXcds, r = nt % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  return x * q + min(x, r) + k;
}

// FP64 policies need no per-tile context
struct NoCtx {};

// ---------------------------------------------------------------------------
// Per-pair cubic-spline gradient GK = (1/r) dW/dr (Functions.cu:76-88) from
// r^2: one refined reciprocal square root instead of sqrt + three IEEE
// divisions.  With hj == hi (the solver sets one h for all particles,
// Domain_d::Set_h) the thread's own 1/h and 1/(pi h^4) are reused, so the
// common case needs no division at all.  Differences from the reference's
// expression order are a few ulp.
struct HConst {
  double h, inv_h, C;
};
__device__ __forceinline__ HConst hconst(double h) {
  return HConst{h, 1.0 / h, 1.0 / (h * h * h * h * kPi)};
}
__device__ __forceinline__ double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = y * (1.5 - hx * y * y);
  y = y * (1.5 - hx * y * y);
  return y;
}
__device__ __forceinline__ double gk_pair(double r2, double hj, const HConst& hc, double& rinv,
                                          double& h_) {
  // coincident particles (never in a list built with d2 > 0, but possible
  // after motion) contribute nothing instead of 0 * inf
  rinv = r2 > 0.0 ? rsqrt_f64(r2) : 0.0;
  const double rij = r2 * rinv;
  double inv_h = hc.inv_h, C = hc.C;
  h_ = hc.h;
  if (hj != hc.h) {
    h_ = (hc.h + hj) / 2.0;
    inv_h = 1.0 / h_;
    C = 1.0 / (h_ * h_ * h_ * h_ * kPi);
  }
  const double q = rij * inv_h;
  const double t = 2.0 - q;
  const double poly = q < 1.0 ? (-3.0 * q + (9.0 / 4.0) * q * q)
                              : (q < 2.0 ? (-3.0 / 4.0) * t * t : 0.0);
  return C * rinv * poly;
}

// gather-kernel traits of a policy: the staged position type, positions per
// LDS pass, the per-tile context and a neighbour's staged position
// staged slots per LDS pass: 1024 x 32 B (FP64 position + index) = 32 KB;
// on a cubic lattice with h = 1.2 dx a tile's staged set holds ~740
// particles (at most ~910), so every tile takes the one-pass path
#ifndef WF_STAGE64
#define WF_STAGE64 1024
#endif
constexpr int kStage64 = WF_STAGE64;
// threads per particle in the pair loops (each takes every LANES-th slot
// quad); on a 759k-particle cylinder two threads take the FP32 step from
// 0.672 to 0.591 ms (four: 0.678 ms) and, with the 32 KB FP64 stage, the
// FP64 step from 1.00 to 0.90 ms
#ifndef WF_LANES64
#define WF_LANES64 2
#endif
#ifndef WF_LANES32
#define WF_LANES32 2
#endif
#define WF_GATHER64                                                                  \
  typedef double PT;                                                                 \
  static constexpr int STAGE = kStage64;                                             \
  static constexpr int LANES = WF_LANES64;                                           \
  typedef NoCtx Ctx;                                                                 \
  template <class P>                                                                 \
  __device__ static Ctx ctx(const Fields&, int, const P&) {                          \
    return Ctx{};                                                                    \
  }                                                                                  \
  __device__ static void pos(const Fields& f, int j, const Ctx&, double& x, double& y, \
                             double& z) {                                            \
    x = f.x[0][j];                                                                   \
    y = f.x[1][j];                                                                   \
    z = f.x[2][j];                                                                   \
  }

// ---------------------------------------------------------------------------
// FP64 pair loops (WF_FP64) on the gather kernel below: FP64 positions of the
// staged set in LDS, the neighbour fields from packed FP64 records.  Each
// thread sums its pairs in ascending staged order and the two threads of a
// particle add their partial sums, a fixed order, so results are
// deterministic.

// A: density rate drho_i = rho_i sum_j (m_j/rho_j) v_ij . (GK x_ij)
// (InteractionAlt.cu:166-246) [+ update rho_i + dt drho_i into rho_alt and
// the next {v, m/rho} record]
template <bool UH>
struct DensD {
  WF_GATHER64
  static constexpr int NBATCH = 4;
  struct Prm {
    int limit, n;
    double dt;
    int fuse;
    int mdj_alt;  // write m/rho into the next conduction record (fused with conduction)
  };
  struct Own {
    double x, y, z, vx, vy, vz, rho;
    HConst hc;
  };
  struct Nb {
    double4 v;  // v, m/rho
    double h;
  };
  struct Acc {
    double s;
  };
  __device__ static void merge(Acc& a, const Acc& b) {
    a.s += b.s;
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx&) {
    a.s = 0.0;
    if (!act) return;
    o.x = f.x[0][i]; o.y = f.x[1][i]; o.z = f.x[2][i];
    o.vx = f.v[0][i]; o.vy = f.v[1][i]; o.vz = f.v[2][i];
    o.rho = f.rho[i];
    o.hc = hconst(f.h[i]);
  }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    nb.v = f.kv[j];
    nb.h = UH ? 0.0 : f.h[j];
  }
  __device__ static void pair(double x, double y, double z, const Nb& nb, const Own& o, Acc& a,
                              const Prm&) {
    const double xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const double r2 = xx * xx + xy * xy + xz * xz;
    double rinv, h_;
    const double GK = gk_pair(r2, UH ? o.hc.h : nb.h, o.hc, rinv, h_);
    const double vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    a.s += nb.v.w * (GK * (vx * xx + vy * xy + vz * xz));
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own& o, const Acc& a,
                                const Prm& p) {
    if (act) {
      const double drho = o.rho * a.s;
      f.drho[i] = drho;
      if (p.fuse) {
        const double rn = o.rho + p.dt * drho;
        f.rho_alt[i] = rn;
        const double mdj = f.m[i] / rn;
        f.kv_alt[i] = make_double4(o.vx, o.vy, o.vz, mdj);
        (p.mdj_alt ? f.tk64_alt : f.tk64)[i].z = mdj;  // m/rho for the conduction loop
      }
    } else if (p.fuse && i < p.n) {
      f.rho_alt[i] = f.rho[i];
      f.kv_alt[i] = f.kv[i];
    }
  }
};

// B: strain rate -1/2 sum (m_j/rho_j) GK (v_ij (x) x_ij + x_ij (x) v_ij) and
// rotation rate -1/2 sum (m_j/rho_j) GK (v_ij (x) x_ij - x_ij (x) v_ij),
// upper triangle, xz optionally negated (InteractionAlt.cu:262-383)
template <bool UH>
struct RatesD {
  WF_GATHER64
  static constexpr int NBATCH = 4;
  struct Prm {
    int limit;
    double xz_sign;
  };
  struct Own {
    double x, y, z, vx, vy, vz;
    HConst hc;
  };
  typedef typename DensD<UH>::Nb Nb;
  struct Acc {
    double dxx, dyy, dzz, dxy, dyz, dxz, wxy, wyz, wxz;
  };
  __device__ static void merge(Acc& a, const Acc& b) {
    a.dxx += b.dxx; a.dyy += b.dyy; a.dzz += b.dzz;
    a.dxy += b.dxy; a.dyz += b.dyz; a.dxz += b.dxz;
    a.wxy += b.wxy; a.wyz += b.wyz; a.wxz += b.wxz;
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx&) {
    a = Acc{0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!act) return;
    o.x = f.x[0][i]; o.y = f.x[1][i]; o.z = f.x[2][i];
    o.vx = f.v[0][i]; o.vy = f.v[1][i]; o.vz = f.v[2][i];
    o.hc = hconst(f.h[i]);
  }
  __device__ static void load(const Fields& f, int j, Nb& nb) { DensD<UH>::load(f, j, nb); }
  __device__ static void pair(double x, double y, double z, const Nb& nb, const Own& o, Acc& a,
                              const Prm&) {
    const double xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const double r2 = xx * xx + xy * xy + xz * xz;
    double rinv, h_;
    const double w = gk_pair(r2, UH ? o.hc.h : nb.h, o.hc, rinv, h_) * nb.v.w;
    const double vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    const double ax = w * vx, ay = w * vy, az = w * vz;
    a.dxx += ax * xx;
    a.dyy += ay * xy;
    a.dzz += az * xz;
    a.dxy += ax * xy + ay * xx;
    a.dyz += ay * xz + az * xy;
    a.dxz += ax * xz + az * xx;
    a.wxy += ax * xy - ay * xx;
    a.wyz += ay * xz - az * xy;
    a.wxz += ax * xz - az * xx;
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own&, const Acc& a,
                                const Prm& p) {
    if (!act) return;
    f.strrate[0][i] = -a.dxx; f.strrate[1][i] = -a.dyy; f.strrate[2][i] = -a.dzz;
    f.strrate[3][i] = -0.5 * a.dxy; f.strrate[4][i] = -0.5 * a.dyz;
    f.strrate[5][i] = -0.5 * a.dxz;
    f.rotrate[3][i] = -0.5 * a.wxy; f.rotrate[4][i] = -0.5 * a.wyz;  // diagonal stays 0
    f.rotrate[5][i] = -0.5 * p.xz_sign * a.wxz;
  }
};

// D: a_i = sum_j m_j (sigma_i/rho_i^2 + sigma_j/rho_j^2 + Pi_ij I) . (GK x_ij),
// Pi_ij = (alpha C mu + beta mu^2)/rho_bar for approaching pairs,
// mu = h v.x/(r^2 + 0.01 h^2) (InteractionAlt.cu:18-150).  sigma_j/rho_j^2
// comes formed in the neighbour's record.
// neighbour records (96 B each) a lane holds in flight in the FP64 accel loop
#ifndef WF_ACCEL64_NBATCH
#define WF_ACCEL64_NBATCH 2
#endif
template <bool UH>
struct AccelD {
  WF_GATHER64
  static constexpr int NBATCH = WF_ACCEL64_NBATCH;
  struct Prm {
    int limit;
    double alpha, beta;
  };
  struct Own {
    double x, y, z, vx, vy, vz, di, Ci, s0, s1, s2, s3, s4, s5;
    HConst hc;
  };
  struct Nb {
    double4 v, sa, sb;  // {v, m}, sigma/rho^2 xx yy zz xy, {yz, xz, rho, Cs}
    double h;
  };
  struct Acc {
    double ax, ay, az;
  };
  __device__ static void merge(Acc& a, const Acc& b) {
    a.ax += b.ax; a.ay += b.ay; a.az += b.az;
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx&) {
    a = Acc{0, 0, 0};
    if (!act) return;
    o.x = f.x[0][i]; o.y = f.x[1][i]; o.z = f.x[2][i];
    o.vx = f.v[0][i]; o.vy = f.v[1][i]; o.vz = f.v[2][i];
    o.di = f.rho[i]; o.Ci = f.Cs[i];
    o.hc = hconst(f.h[i]);
    const double idi = 1.0 / (o.di * o.di);
    o.s0 = f.sigma[0][i] * idi; o.s1 = f.sigma[1][i] * idi; o.s2 = f.sigma[2][i] * idi;
    o.s3 = f.sigma[3][i] * idi; o.s4 = f.sigma[4][i] * idi; o.s5 = f.sigma[5][i] * idi;
  }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    nb.v = f.arec64[rec_at(f, j, 0)];
    nb.sa = f.arec64[rec_at(f, j, 1)];
    nb.sb = f.arec64[rec_at(f, j, 2)];
    nb.h = UH ? 0.0 : f.h[j];
  }
  __device__ static void pair(double x, double y, double z, const Nb& nb, const Own& o, Acc& a,
                              const Prm& p) {
    const double xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const double r2 = xx * xx + xy * xy + xz * xz;
    double rinv, h_;
    const double GK = gk_pair(r2, UH ? o.hc.h : nb.h, o.hc, rinv, h_);
    const double vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    const double vdx = vx * xx + vy * xy + vz * xz;
    double PI = 0.0;
    if (vdx < 0.0 && (p.alpha != 0.0 || p.beta != 0.0)) {
      // (alpha C mu + beta mu^2)/rho_bar with mu = h vdx/e, e = r^2 + 0.01 h^2,
      // as 2 h vdx (alpha C + beta mu)/(e (rho_i + rho_j)): one division
      // where the reference's form takes two (the beta term keeps its own)
      const double e = r2 + 0.01 * h_ * h_;
      const double Cij = 0.5 * (o.Ci + nb.sb.w);
      double coef = p.alpha * Cij;
      if (p.beta != 0.0) coef += p.beta * (h_ * vdx / e);
      PI = 2.0 * h_ * vdx * coef / (e * (o.di + nb.sb.z));
    }
    const double g = GK * nb.v.w;
    const double gx = g * xx, gy = g * xy, gz = g * xz;
    const double mxx = o.s0 + nb.sa.x + PI, myy = o.s1 + nb.sa.y + PI;
    const double mzz = o.s2 + nb.sa.z + PI;
    const double mxy = o.s3 + nb.sa.w, myz = o.s4 + nb.sb.x, mxz = o.s5 + nb.sb.y;
    a.ax += mxx * gx + mxy * gy + mxz * gz;
    a.ay += mxy * gx + myy * gy + myz * gz;
    a.az += mxz * gx + myz * gy + mzz * gz;
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own&, const Acc& a,
                                const Prm&) {
    if (!act) return;
    f.a[0][i] = a.ax;
    f.a[1][i] = a.ay;
    f.a[2][i] = a.az;
  }
};

// G: conduction dT_i/dt = 1/(rho_i cp_i) sum_j (m_j/rho_j) 4 k_i k_j/(k_i+k_j)
// (T_i - T_j) GK  (Thermal.cu:101-142, where x_ij.(GK x_ij)/r^2 = GK), then
// convection (Thermal.cu:36-64), plastic-work source and FIXED BC; optional
// fused leapfrog (Thermal.cu:16-31) writing T into T_alt and the next record.
template <bool UH>
struct CondD {
  WF_GATHER64
  static constexpr int NBATCH = 4;
  struct Prm {
    int limit;
    double T_inf, h_conv, dt;
    int leap, first;
  };
  struct Own {
    double x, y, z, k, T, mdj;
    HConst hc;
  };
  struct Nb {
    double4 tk;  // T, k, m/rho
    double h;
  };
  struct Acc {
    double q;
  };
  __device__ static void merge(Acc& a, const Acc& b) {
    a.q += b.q;
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx&) {
    a.q = 0.0;
    if (!act) return;
    o.x = f.x[0][i]; o.y = f.x[1][i]; o.z = f.x[2][i];
    o.k = f.k[i]; o.T = f.T[i];
    o.mdj = f.tk64[i].z;
    o.hc = hconst(f.h[i]);
  }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    nb.tk = f.tk64[j];
    nb.h = UH ? 0.0 : f.h[j];
  }
  __device__ static void pair(double x, double y, double z, const Nb& nb, const Own& o, Acc& a,
                              const Prm&) {
    const double xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const double r2 = xx * xx + xy * xy + xz * xz;
    double rinv, h_;
    const double GK = gk_pair(r2, UH ? o.hc.h : nb.h, o.hc, rinv, h_);
    const double kj = nb.tk.y;
    const double kh = (kj == o.k) ? 2.0 * o.k : 4.0 * o.k * kj / (o.k + kj);
    a.q += nb.tk.z * kh * (o.T - nb.tk.x) * GK;
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own& o, const Acc& a,
                                const Prm& p) {
    if (!act) return;
    const double rhoi = f.rho[i], cpi = f.cp[i];
    double acc = a.q * (1. / (rhoi * cpi));
    const int bct = f.BC_T[i];
    if (bct == WF_BCT_CONVECTION) {
      double mi = f.m[i];
      double dS2 = pow(mi / rhoi, 0.666666666);
      double q_conv = p.h_conv * dS2 * (p.T_inf - o.T) / mi;
      acc += 1. / cpi * q_conv;
    }
    double qp = f.qpl[i];
    if (qp != 0.0) acc += qp * (1. / (rhoi * cpi));
    const double qf = f.qfric[i];
    if (qf != 0.0) acc += qf / (f.m[i] * cpi);
    if (bct == WF_BCT_FIXED) acc = 0.0;
    f.dTdt[i] = acc;
    if (p.leap) {
      double Tn;
      if (p.first) {
        f.Ta[i] = o.T - p.dt / 2.0 * acc;
        Tn = o.T;
      } else {
        double ta = f.Ta[i];
        f.Tb[i] = ta;
        double tan = ta + acc * p.dt;
        f.Ta[i] = tan;
        Tn = (tan + ta) / 2.;
      }
      f.T_alt[i] = Tn;
      f.tk64_alt[i] = make_double4(Tn, o.k, o.mdj, 0.0);
    }
  }
};

// ===========================================================================
// Mixed precision (WF_MIXED): the same pair physics in FP32, on FP32 mirrors,
// with coordinates relative to the tile's base cell (x = xr + (c - c_tile) cs,
// c an integer grid cell), FP32 pair terms summed in FP64, and FP64 results.
// Per-particle terms stay FP64.
//
// The FP32 loops hold only the staged set's tile-frame positions in LDS
// (kStage32 particles, normally the whole set) and gather the other
// neighbour fields from packed per-particle float4/float2 mirrors, which stay
// in L1/L2 because the staged set is spatially compact.  A thread therefore
// walks its whole neighbour list without waiting at chunk boundaries for the
// busiest lane of its wavefront (with chunks of a few hundred staged
// particles a wavefront keeps only ~40-55 % of its lanes busy).

// sum of FP32 pair terms kept in FP64: with packed FP32 off, an FP64 add
// issues at the FP32 rate on gfx950, so widening each term costs two
// operations where compensated (Kahan) FP32 summation costs four, and the
// sum carries no rounding beyond the terms' own
struct KSum {
  double s;
  __device__ __forceinline__ void zero() { s = 0.0; }
  __device__ __forceinline__ void add(float y) { s += (double)y; }
  __device__ __forceinline__ void add(const KSum& o) { s += o.s; }
  __device__ __forceinline__ double val() const { return s; }
};

struct HConstF {
  float h, inv_h, C;
};
__device__ __forceinline__ HConstF hconst_f(float h) {
  const float hh = h * h;
  return HConstF{h, __builtin_amdgcn_rcpf(h), __builtin_amdgcn_rcpf(hh * hh * (float)kPi)};
}
// FP32 cubic-spline gradient (1/r) dW/dr from r^2 (Functions.cu:76-88);
// UH: every particle has the same h, so h_j = h_i
template <bool UH>
__device__ __forceinline__ float gk_pair_f(float r2, float hj, const HConstF& hc, float& h_) {
  const float rinv = r2 > 0.f ? __builtin_amdgcn_rsqf(r2) : 0.f;
  const float rij = r2 * rinv;
  float inv_h = hc.inv_h, C = hc.C;
  h_ = hc.h;
  if (!UH && hj != hc.h) {
    h_ = 0.5f * (hc.h + hj);
    inv_h = __builtin_amdgcn_rcpf(h_);
    C = __builtin_amdgcn_rcpf(h_ * h_ * h_ * h_ * (float)kPi);
  }
  const float q = rij * inv_h;
  const float t = 2.f - q;
  const float poly = q < 1.f ? (-3.f * q + 2.25f * q * q) : (q < 2.f ? -0.75f * t * t : 0.f);
  return C * rinv * poly;
}

struct Ctx32 {
  int bx, by, bz;
  float cs;
};
__device__ __forceinline__ Ctx32 ctx32(const Fields& f, int t, float cs) {
  int bx, by, bz;
  unpack_cell(__float_as_int(f.p4[t * kTileP].w), bx, by, bz);
  return Ctx32{bx, by, bz, cs};
}
// position of particle j in the tile frame
__device__ __forceinline__ void pos32(const Fields& f, int j, const Ctx32& c, float& x, float& y,
                                      float& z) {
  int cx, cy, cz;
  const float4 p = f.p4[j];
  unpack_cell(__float_as_int(p.w), cx, cy, cz);
  x = p.x + (float)(cx - c.bx) * c.cs;
  y = p.y + (float)(cy - c.by) * c.cs;
  z = p.z + (float)(cz - c.bz) * c.cs;
}

// staged slots per LDS pass: 1024 x 16 B (tile-frame position + index) =
// 16 KB (768 slots send the larger tiles to the chunked path: 17 % slower)
#ifndef WF_STAGE32
#define WF_STAGE32 1024
#endif
constexpr int kStage32 = WF_STAGE32;
#define WF_GATHER32                                                                \
  typedef float PT;                                                                \
  static constexpr int STAGE = kStage32;                                           \
  static constexpr int LANES = WF_LANES32;                                         \
  typedef Ctx32 Ctx;                                                               \
  template <class P>                                                               \
  __device__ static Ctx ctx(const Fields& f, int t, const P& p) {                  \
    return ctx32(f, t, p.cs);                                                      \
  }                                                                                \
  __device__ static void pos(const Fields& f, int j, const Ctx& c, float& x, float& y, \
                             float& z) {                                           \
    pos32(f, j, c, x, y, z);                                                       \
  }

// One staged particle in LDS: its position in the tile frame and its
// particle index, read together by one LDS access per pair (the index
// addresses the neighbour's packed record in global memory).
template <class PT>
struct alignas(16) StagedSlot;
template <>
struct alignas(16) StagedSlot<float> {
  float x, y, z;
  int j;
};
template <>
struct alignas(16) StagedSlot<double> {
  double x, y, z;
  int j, pad;
};

// A pair functor may stage one more 16-byte value per staged particle in LDS
// beside its position (typedef Extra, Extra extra(f, j) reads it when the set
// is staged, put(nb, e) hands it to the neighbour record): a field every pair
// needs then costs an LDS read instead of a global gather per pair.
template <class K, class = void>
struct ExtraOf {
  static constexpr bool on = false;
  typedef int type;
};
template <class K>
struct ExtraOf<K, std::void_t<typename K::Extra>> {
  static constexpr bool on = true;
  typedef typename K::Extra type;
};

// waves per SIMD a pair functor asks the register allocator to leave room
// for (MIN_WAVES; 1 = no request)
template <class K, class = void>
struct WavesOf {
  static constexpr int value = 1;
};
template <class K>
struct WavesOf<K, std::void_t<decltype(K::MIN_WAVES)>> {
  static constexpr int value = K::MIN_WAVES;
};

// the value v holds in lane (this lane ^ m)
template <class T>
__device__ __forceinline__ T lane_value_xor(const T& v, int m) {
  static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
  constexpr int W = sizeof(T) / 4;
  int in[W], out[W];
  __builtin_memcpy(in, &v, sizeof(T));
#pragma unroll
  for (int w = 0; w < W; w++) out[w] = __shfl_xor(in[w], m);
  T r;
  __builtin_memcpy(&r, out, sizeof(T));
  return r;
}

// One workgroup per tile, L = K::LANES consecutive lanes per particle: lanes
// L p .. L p + L - 1 all hold particle p and deal its slot list out by quads
// (lane L p + h takes the quads q with q % L == h); their sums are merged
// (butterfly over the L lanes) before lane L p writes the result.
template <class K>
__global__ __launch_bounds__(4 * kTileP) __attribute__((amdgpu_waves_per_eu(WavesOf<K>::value)))
void k_pair_gather(Fields f, TileLists tl, typename K::Prm prm) {
  constexpr int ST = K::STAGE;
  constexpr int L = K::LANES;
  static_assert(L == 1 || L == 2 || L == 4, "1, 2 or 4 lanes per particle");
  constexpr int NT = L * kTileP;  // threads per tile
  typedef StagedSlot<typename K::PT> Slot;
  __shared__ Slot st[ST];
  constexpr bool XS = ExtraOf<K>::on;
  typedef typename ExtraOf<K>::type XT;
  __shared__ XT sx[XS ? ST : 1];
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int sub = tid & (L - 1);
  const int i = t * kTileP + tid / L;
  const bool act = i < prm.limit;
  const typename K::Ctx cx = K::ctx(f, t, prm);
  typename K::Own own;
  typename K::Acc acc;
  K::init(f, i, act, own, acc, cx);
  const int ub = tl.uoff[t], ucnt = tl.ucnt[t];
  const int* __restrict__ tj = tl.tj + ub;
  const int cnt = act ? tl.count[i] : 0;
#ifndef WF_STAGE_UNROLL
#define WF_STAGE_UNROLL 8
#endif
  // staging: SU slots per thread at a time, their indices, then their
  // positions, then the LDS writes, so that each thread keeps SU index loads
  // and then SU position loads in flight instead of one dependent pair
  auto stage = [&](int c0, int m) {
    constexpr int SU = WF_STAGE_UNROLL;
    for (int s0 = tid; s0 < m; s0 += SU * NT) {
      int jj[SU];
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int s = s0 + u * NT;
        jj[u] = s < m ? tj[c0 + s] : 0;
      }
      typename K::PT x[SU], y[SU], z[SU];
#pragma unroll
      for (int u = 0; u < SU; u++)
        if (s0 + u * NT < m) K::pos(f, jj[u], cx, x[u], y[u], z[u]);
      XT e[XS ? SU : 1];
      if constexpr (XS) {
#pragma unroll
        for (int u = 0; u < SU; u++)
          if (s0 + u * NT < m) e[u] = K::extra(f, jj[u]);
      }
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int s = s0 + u * NT;
        if (s < m) {
          Slot v;
          v.x = x[u];
          v.y = y[u];
          v.z = z[u];
          v.j = jj[u];
          st[s] = v;
          if constexpr (XS) sx[s] = e[u];
        }
      }
    }
  };
  constexpr int NB = K::NBATCH;  // neighbour records held at once (4, or 2)
  if (ucnt <= ST) {
    // the whole staged set fits: one pass, a quad of four slots at a time;
    // the lane's slot word two quads ahead is fetched while the current
    // quad's records are in flight, and the next quad's particle indices
    // come from LDS
    stage(0, ucnt);
    __syncthreads();
    // lanes past the padded particle range read row 0 of the last particle
    // (their lists are empty); every slot word is loaded unconditionally
    // from a row clamped into the list, since a select between a load and a
    // constant turns into a load through a selected pointer with the
    // constant in scratch
    const uint2* __restrict__ q = tl.slotq + min((int64_t)i, tl.npad - 1);
    const int klast = max(cnt - 1, 0);
    auto word = [&](int k) { return q[(int64_t)(min(k, klast) >> 2) * tl.npad]; };
    auto sl_of = [](const uint2 w, int u) {
      const unsigned h = (u & 2) ? w.y : w.x;
      return (int)((u & 1) ? h >> 16 : h & 0xFFFFu);
    };
    constexpr int KS = 4 * L;  // slots from one of the lane's quads to its next
    const int k0 = 4 * sub;
    uint2 w0 = word(k0);
    uint2 w1 = word(k0 + KS);
    int jn[4];
#pragma unroll
    for (int u = 0; u < 4; u++) jn[u] = k0 + u < cnt ? st[sl_of(w0, u)].j : 0;
    for (int k = k0; k < cnt; k += KS) {
      const int nk = min(4, cnt - k);
      int jc[4];
#pragma unroll
      for (int u = 0; u < 4; u++) jc[u] = jn[u];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (k + KS + u < cnt) jn[u] = st[sl_of(w1, u)].j;
      const uint2 w2 = word(k + 2 * KS);
#pragma unroll
      for (int u0 = 0; u0 < 4; u0 += NB) {
        typename K::Nb nb[NB];
#pragma unroll
        for (int u = 0; u < NB; u++)
          if (u0 + u < nk) K::load(f, jc[u0 + u], nb[u]);
        // the batch's staged positions are read before any pair is evaluated
        // (entries past the list read slot 0 and are not evaluated)
        typename K::PT px[NB], py[NB], pz[NB];
        int sl[NB];
#pragma unroll
        for (int u = 0; u < NB; u++) {
          sl[u] = u0 + u < nk ? sl_of(w0, u0 + u) : 0;
          const Slot& sv = st[sl[u]];
          px[u] = sv.x;
          py[u] = sv.y;
          pz[u] = sv.z;
        }
        // a staged extra value is read just before its pair
#pragma unroll
        for (int u = 0; u < NB; u++)
          if (u0 + u < nk) {
            if constexpr (XS) K::put(nb[u], sx[sl[u]]);
            K::pair(px[u], py[u], pz[u], nb[u], own, acc, prm);
          }
      }
      w0 = w1;
      w1 = w2;
    }
  } else {
    SlotReader sr;
    sr.init(tl, i, cnt);
    // this lane's slots: those of the quads q with q % L == sub
    auto mine = [&]() { return ((sr.k >> 2) & (L - 1)) == sub; };
    for (int c0 = 0; c0 < ucnt; c0 += ST) {
      const int m = min(ST, ucnt - c0);
      __syncthreads();
      stage(c0, m);
      __syncthreads();
      const int lim = c0 + m;
      // two neighbours per iteration while both fall in this pass: their
      // gathers are issued before either pair is evaluated
      while (sr.k < sr.cnt) {
        if (!mine()) {
          sr.skip_quad();
          continue;
        }
        const int s0 = sr.slot();
        if (s0 >= lim) break;
        sr.advance();
        const Slot& a0 = st[s0 - c0];
        typename K::Nb a;
        K::load(f, a0.j, a);
        if constexpr (XS) K::put(a, sx[s0 - c0]);
        if (sr.k < sr.cnt && mine()) {
          const int s1 = sr.slot();
          if (s1 < lim) {
            sr.advance();
            const Slot& b0 = st[s1 - c0];
            typename K::Nb b;
            K::load(f, b0.j, b);
            if constexpr (XS) K::put(b, sx[s1 - c0]);
            K::pair(a0.x, a0.y, a0.z, a, own, acc, prm);
            K::pair(b0.x, b0.y, b0.z, b, own, acc, prm);
            continue;
          }
        }
        K::pair(a0.x, a0.y, a0.z, a, own, acc, prm);
      }
    }
  }
  if constexpr (L > 1) {
#pragma unroll
    for (int m = 1; m < L; m <<= 1) {
      const typename K::Acc other = lane_value_xor(acc, m);
      K::merge(acc, other);
    }
  }
  if (sub == 0) K::finish(f, i, act, own, acc, prm);
}

// A: density rate (InteractionAlt.cu:166-246), fused density update writes
// the next {v, m/rho} mirror
template <bool UH>
struct DensG {
  WF_GATHER32
  static constexpr int NBATCH = 4;
  struct Prm {
    int limit, n;
    double dt;
    int fuse;
    float cs;
    int mdj_alt;  // write m/rho into the next conduction record (fused with conduction)
  };
  struct Own {
    float x, y, z, vx, vy, vz;
    HConstF hc;
  };
  struct Nb {
    float4 v;  // v, m/rho
    float h;
  };
  typedef KSum Acc;
  __device__ static void merge(Acc& a, const Acc& b) {
    a.add(b);
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx32& c) {
    a.zero();
    if (!act) return;
    pos32(f, i, c, o.x, o.y, o.z);
    const float4 v = f.v4[i];
    o.vx = v.x; o.vy = v.y; o.vz = v.z;
    o.hc = hconst_f(f.h32[i]);
  }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    nb.v = f.v4[j];
    nb.h = UH ? 0.f : f.h32[j];
  }
  __device__ static void pair(float x, float y, float z, const Nb& nb, const Own& o, Acc& a,
                              const Prm&) {
    const float xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const float r2 = xx * xx + xy * xy + xz * xz;
    float h_;
    const float GK = gk_pair_f<UH>(r2, nb.h, o.hc, h_);
    const float vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    a.add(nb.v.w * (GK * (vx * xx + vy * xy + vz * xz)));
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own&, const Acc& a,
                                const Prm& p) {
    if (act) {
      const double rho = f.rho[i];
      const double drho = rho * a.val();
      f.drho[i] = drho;
      if (p.fuse) {
        const double rn = rho + p.dt * drho;
        f.rho_alt[i] = rn;
        float4 v = f.v4[i];
        v.w = (float)(f.m[i] / rn);
        f.v4_alt[i] = v;
        (p.mdj_alt ? f.tk4_alt : f.tk4)[i].z = v.w;  // m/rho for the conduction loop
      }
    } else if (p.fuse && i < p.n) {
      f.rho_alt[i] = f.rho[i];
      f.v4_alt[i] = f.v4[i];
    }
  }
};

// B: strain and rotation rates (InteractionAlt.cu:262-383)
template <bool UH>
struct RatesG {
  WF_GATHER32
  static constexpr int NBATCH = 4;
  struct Prm {
    int limit;
    double xz_sign;
    float cs;
  };
  typedef typename DensG<UH>::Own Own;
  typedef typename DensG<UH>::Nb Nb;
  struct Acc {
    KSum dxx, dyy, dzz, dxy, dyz, dxz, wxy, wyz, wxz;
  };
  __device__ static void merge(Acc& a, const Acc& b) {
    a.dxx.add(b.dxx); a.dyy.add(b.dyy); a.dzz.add(b.dzz);
    a.dxy.add(b.dxy); a.dyz.add(b.dyz); a.dxz.add(b.dxz);
    a.wxy.add(b.wxy); a.wyz.add(b.wyz); a.wxz.add(b.wxz);
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx32& c) {
    a.dxx.zero(); a.dyy.zero(); a.dzz.zero(); a.dxy.zero(); a.dyz.zero(); a.dxz.zero();
    a.wxy.zero(); a.wyz.zero(); a.wxz.zero();
    if (!act) return;
    pos32(f, i, c, o.x, o.y, o.z);
    const float4 v = f.v4[i];
    o.vx = v.x; o.vy = v.y; o.vz = v.z;
    o.hc = hconst_f(f.h32[i]);
  }
  // the neighbour's {v, m/rho} is staged in LDS with its position
  typedef float4 Extra;
  __device__ static Extra extra(const Fields& f, int j) { return f.v4[j]; }
  __device__ static void put(Nb& nb, const Extra& e) { nb.v = e; }
  __device__ static void load(const Fields& f, int j, Nb& nb) { nb.h = UH ? 0.f : f.h32[j]; }
  __device__ static void pair(float x, float y, float z, const Nb& nb, const Own& o, Acc& a,
                              const Prm&) {
    const float xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const float r2 = xx * xx + xy * xy + xz * xz;
    float h_;
    const float w = gk_pair_f<UH>(r2, nb.h, o.hc, h_) * nb.v.w;
    const float vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    const float ax = w * vx, ay = w * vy, az = w * vz;
    a.dxx.add(ax * xx);
    a.dyy.add(ay * xy);
    a.dzz.add(az * xz);
    a.dxy.add(ax * xy + ay * xx);
    a.dyz.add(ay * xz + az * xy);
    a.dxz.add(ax * xz + az * xx);
    a.wxy.add(ax * xy - ay * xx);
    a.wyz.add(ay * xz - az * xy);
    a.wxz.add(ax * xz - az * xx);
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own&, const Acc& a,
                                const Prm& p) {
    if (!act) return;
    f.strrate32[0][i] = (float)-a.dxx.val(); f.strrate32[1][i] = (float)-a.dyy.val();
    f.strrate32[2][i] = (float)-a.dzz.val();
    f.strrate32[3][i] = (float)(-0.5 * a.dxy.val());
    f.strrate32[4][i] = (float)(-0.5 * a.dyz.val());
    f.strrate32[5][i] = (float)(-0.5 * a.dxz.val());
    f.rotrate32[3][i] = (float)(-0.5 * a.wxy.val());  // the diagonal stays 0
    f.rotrate32[4][i] = (float)(-0.5 * a.wyz.val());
    f.rotrate32[5][i] = (float)(-0.5 * p.xz_sign * a.wxz.val());
  }
};

// D: acceleration with Monaghan viscosity (InteractionAlt.cu:18-150); the
// neighbour's mass is (m/rho) rho
template <bool UH>
#ifndef WF_ACCEL32_NBATCH
#define WF_ACCEL32_NBATCH 4
#endif
#ifndef WF_ACCEL32_WAVES
#define WF_ACCEL32_WAVES 1
#endif
struct AccelG {
  WF_GATHER32
  static constexpr int NBATCH = WF_ACCEL32_NBATCH;
  static constexpr int MIN_WAVES = WF_ACCEL32_WAVES;
  struct Prm {
    int limit;
    float alpha32, beta32;
    float cs;
  };
  struct Own {
    float x, y, z, vx, vy, vz, di, Ci, s0, s1, s2, s3, s4, s5;
    HConstF hc;
  };
  struct Nb {
    float4 v, sa, sb;
    float h;
  };
  struct Acc {
    KSum ax, ay, az;
  };
  __device__ static void merge(Acc& a, const Acc& b) {
    a.ax.add(b.ax); a.ay.add(b.ay); a.az.add(b.az);
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx32& c) {
    a.ax.zero(); a.ay.zero(); a.az.zero();
    if (!act) return;
    pos32(f, i, c, o.x, o.y, o.z);
    const float4 v = f.v4[i], sa = f.arec[rec32_at(i, 0)], sb = f.arec[rec32_at(i, 1)];
    o.vx = v.x; o.vy = v.y; o.vz = v.z;
    o.s0 = sa.x; o.s1 = sa.y; o.s2 = sa.z; o.s3 = sa.w;
    o.s4 = sb.x; o.s5 = sb.y; o.di = sb.z; o.Ci = sb.w;
    o.hc = hconst_f(f.h32[i]);
  }
  // the neighbour's {v, m/rho} is staged in LDS with its position, so a pair
  // gathers only the 32-byte stress record
  typedef float4 Extra;
  __device__ static Extra extra(const Fields& f, int j) { return f.v4[j]; }
  __device__ static void put(Nb& nb, const Extra& e) { nb.v = e; }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    nb.sa = f.arec[rec32_at(j, 0)];
    nb.sb = f.arec[rec32_at(j, 1)];
    nb.h = UH ? 0.f : f.h32[j];
  }
  __device__ static void pair(float x, float y, float z, const Nb& nb, const Own& o, Acc& a,
                              const Prm& p) {
    const float xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const float r2 = xx * xx + xy * xy + xz * xz;
    float h_;
    const float GK = gk_pair_f<UH>(r2, nb.h, o.hc, h_);
    const float vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    const float vdx = vx * xx + vy * xy + vz * xz;
    // Monaghan viscosity on approaching pairs only: (alpha C mu + beta mu^2)
    // / rho_bar with mu = h v/e, v = min(v_ij.x_ij, 0), e = r^2 + 0.01 h^2,
    // as 2 h v (alpha C + beta mu) / (e (rho_i + rho_j)): one reciprocal and
    // no branch (receding pairs get v = 0)
    const float vn = fminf(vdx, 0.f);
    const float rs = o.di + nb.sb.z;
    const float inv = __builtin_amdgcn_rcpf((r2 + 0.01f * h_ * h_) * rs);
    float coef = p.alpha32 * (0.5f * (o.Ci + nb.sb.w));
    if (p.beta32 != 0.f) coef += p.beta32 * (h_ * vn * rs * inv);
    const float PI = 2.f * h_ * vn * coef * inv;
    const float g = GK * (nb.v.w * nb.sb.z);
    const float gx = g * xx, gy = g * xy, gz = g * xz;
    const float mxx = o.s0 + nb.sa.x + PI, myy = o.s1 + nb.sa.y + PI;
    const float mzz = o.s2 + nb.sa.z + PI;
    const float mxy = o.s3 + nb.sa.w, myz = o.s4 + nb.sb.x, mxz = o.s5 + nb.sb.y;
    a.ax.add(mxx * gx + mxy * gy + mxz * gz);
    a.ay.add(mxy * gx + myy * gy + myz * gz);
    a.az.add(mxz * gx + myz * gy + mzz * gz);
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own&, const Acc& a,
                                const Prm&) {
    if (!act) return;
    f.a[0][i] = a.ax.val();
    f.a[1][i] = a.ay.val();
    f.a[2][i] = a.az.val();
  }
};

// conduction (Thermal.cu:101-142), fused convection, plastic-work heat and
// leapfrog; the next {T, k} mirror is written
template <bool UH>
struct CondG {
  WF_GATHER32
  static constexpr int NBATCH = 4;
  struct Prm {
    int limit;
    double T_inf, h_conv, dt;
    int leap, first;
    float cs;
  };
  struct Own {
    float x, y, z, k, T, mdj;
    HConstF hc;
  };
  struct Nb {
    float4 tk;  // T, k, m/rho
    float h;
  };
  typedef KSum Acc;
  __device__ static void merge(Acc& a, const Acc& b) {
    a.add(b);
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx32& c) {
    a.zero();
    if (!act) return;
    pos32(f, i, c, o.x, o.y, o.z);
    const float4 tk = f.tk4[i];
    o.T = tk.x;
    o.k = tk.y;
    o.mdj = tk.z;
    o.hc = hconst_f(f.h32[i]);
  }
  // the neighbour's {T, k, m/rho} is staged in LDS with its position (the
  // pass writes only the next record, tk4_alt)
  typedef float4 Extra;
  __device__ static Extra extra(const Fields& f, int j) { return f.tk4[j]; }
  __device__ static void put(Nb& nb, const Extra& e) { nb.tk = e; }
  __device__ static void load(const Fields& f, int j, Nb& nb) { nb.h = UH ? 0.f : f.h32[j]; }
  __device__ static void pair(float x, float y, float z, const Nb& nb, const Own& o, Acc& a,
                              const Prm&) {
    const float xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const float r2 = xx * xx + xy * xy + xz * xz;
    float h_;
    const float GK = gk_pair_f<UH>(r2, nb.h, o.hc, h_);
    const float kj = nb.tk.y;
    const float kh = (kj == o.k) ? 2.f * o.k : 4.f * o.k * kj * __builtin_amdgcn_rcpf(o.k + kj);
    a.add(nb.tk.z * kh * (o.T - nb.tk.x) * GK);
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own& o, const Acc& a,
                                const Prm& p) {
    if (!act) return;
    const double rhoi = f.rho[i], cpi = f.cp[i], Ti = f.T[i];
    double acc = a.val() * (1. / (rhoi * cpi));
    const int bct = f.BC_T[i];
    if (bct == WF_BCT_CONVECTION) {
      double mi = f.m[i];
      double dS2 = pow(mi / rhoi, 0.666666666);
      double q_conv = p.h_conv * dS2 * (p.T_inf - Ti) / mi;
      acc += 1. / cpi * q_conv;
    }
    double qp = f.qpl[i];
    if (qp != 0.0) acc += qp * (1. / (rhoi * cpi));
    const double qf = f.qfric[i];
    if (qf != 0.0) acc += qf / (f.m[i] * cpi);
    if (bct == WF_BCT_FIXED) acc = 0.0;
    f.dTdt[i] = acc;
    if (p.leap) {
      double Tn;
      if (p.first) {
        f.Ta[i] = Ti - p.dt / 2.0 * acc;
        Tn = Ti;
      } else {
        double ta = f.Ta[i];
        f.Tb[i] = ta;
        double tan = ta + acc * p.dt;
        f.Ta[i] = tan;
        Tn = (tan + ta) / 2.;
      }
      f.T_alt[i] = Tn;
      f.tk4_alt[i] = make_float4((float)Tn, o.k, o.mdj, 0.f);
    }
  }
};

// ---------------------------------------------------------------------------
// S: free-surface normal n_i = sum_j m_j x_ij * N/(M nb_i) and the surface
// flag |n_i| >= 0.25 h_i with nb_i <= 46 (Geometry.cu:26-88).  Rigid
// triangles are not particles here, so every list entry counts.
#ifndef WF_STAGE_SURF
#define WF_STAGE_SURF 1024
#endif
template <bool UM>
struct SurfD {
  // FP64 positions; its own staging size (a pair costs little here, so the
  // chunked path for large staged sets weighs more)
  typedef double PT;
  static constexpr int STAGE = WF_STAGE_SURF;
#ifndef WF_LANES_SURF
#define WF_LANES_SURF 2
#endif
  static constexpr int LANES = WF_LANES_SURF;
  typedef NoCtx Ctx;
  template <class P>
  __device__ static Ctx ctx(const Fields&, int, const P&) {
    return Ctx{};
  }
  __device__ static void pos(const Fields& f, int j, const Ctx&, double& x, double& y,
                             double& z) {
    x = f.x[0][j];
    y = f.x[1][j];
    z = f.x[2][j];
  }
  static constexpr int NBATCH = 4;
  struct Prm {
    int limit;
    int id_free_surf;
    double ntot, totmass;
    double r2cut;  // > 0: lists carry a skin; count only pairs within the radius
    double m_uni;  // every particle's mass (UM)
  };
  struct Own {
    double x, y, z;
  };
  struct Nb {
    double m;
  };
  struct Acc {
    double nx, ny, nz;
    int cnt;
  };
  __device__ static void merge(Acc& a, const Acc& b) {
    a.nx += b.nx; a.ny += b.ny; a.nz += b.nz;
    a.cnt += b.cnt;
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const NoCtx&) {
    a.nx = a.ny = a.nz = 0.0;
    a.cnt = 0;
    if (!act) return;
    o.x = f.x[0][i]; o.y = f.x[1][i]; o.z = f.x[2][i];
  }
  __device__ static void load(const Fields& f, int j, Nb& nb) { nb.m = UM ? 0.0 : f.m[j]; }
  __device__ static void pair(double x, double y, double z, const Nb& nb, const Own& o, Acc& a,
                              const Prm& p) {
    const double dx = o.x - x, dy = o.y - y, dz = o.z - z;
    if (p.r2cut > 0.0 && !(dx * dx + dy * dy + dz * dz < p.r2cut)) return;
    const double mj = UM ? p.m_uni : nb.m;
    a.nx += mj * dx;
    a.ny += mj * dy;
    a.nz += mj * dz;
    a.cnt++;
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own&, const Acc& a,
                                const Prm& p) {
    if (!act) return;
    const double sc = p.ntot / (p.totmass * (double)a.cnt);
    const double nx = a.nx * sc, ny = a.ny * sc, nz = a.nz * sc;
    f.nrm[0][i] = nx;
    f.nrm[1][i] = ny;
    f.nrm[2][i] = nz;
    if (sqrt(nx * nx + ny * ny + nz * nz) >= 0.25 * f.h[i] && a.cnt <= 46 && !f.idlock[i])
      f.ID[i] = p.id_free_surf;
  }
};

static inline int ntiles(int n) { return (n + kTileP - 1) / kTileP; }

// the same in mixed precision: FP32 tile-frame positions, each FP32 term
// m_j x_ij summed in FP64 and the flag decided in FP64 from those sums (a
// flag can differ from the FP64 pass only for |n| within ~1e-7 of 0.25 h)
template <bool UM>
struct SurfG {
  WF_GATHER32
  static constexpr int NBATCH = 4;
  struct Prm {
    int limit;
    int id_free_surf;
    double ntot, totmass;
    float r2cut;  // > 0: lists carry a skin; count only pairs within the radius
    float m_uni;  // every particle's mass (UM)
    float cs;
  };
  struct Own {
    float x, y, z;
  };
  struct Nb {
    float m;
  };
  struct Acc {
    KSum nx, ny, nz;
    int cnt;
  };
  __device__ static void merge(Acc& a, const Acc& b) {
    a.nx.add(b.nx); a.ny.add(b.ny); a.nz.add(b.nz);
    a.cnt += b.cnt;
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx32& c) {
    a.nx.zero(); a.ny.zero(); a.nz.zero();
    a.cnt = 0;
    if (!act) return;
    pos32(f, i, c, o.x, o.y, o.z);
  }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    nb.m = UM ? 0.f : (float)f.m[j];
  }
  __device__ static void pair(float x, float y, float z, const Nb& nb, const Own& o, Acc& a,
                              const Prm& p) {
    const float dx = o.x - x, dy = o.y - y, dz = o.z - z;
    if (p.r2cut > 0.f && !(dx * dx + dy * dy + dz * dz < p.r2cut)) return;
    const float mj = UM ? p.m_uni : nb.m;
    a.nx.add(mj * dx);
    a.ny.add(mj * dy);
    a.nz.add(mj * dz);
    a.cnt++;
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own&, const Acc& a,
                                const Prm& p) {
    if (!act) return;
    const double sc = p.ntot / (p.totmass * (double)a.cnt);
    const double nx = a.nx.val() * sc, ny = a.ny.val() * sc, nz = a.nz.val() * sc;
    f.nrm[0][i] = nx;
    f.nrm[1][i] = ny;
    f.nrm[2][i] = nz;
    if (sqrt(nx * nx + ny * ny + nz * nz) >= 0.25 * f.h[i] && a.cnt <= 46 && !f.idlock[i])
      f.ID[i] = p.id_free_surf;
  }
};

void launch_surface_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                          double ntot, double totmass, int id_free_surf, double r2cut, bool um,
                          double m_value, float cs32) {
  if (solid <= 0) return;
  const float r2 = (float)r2cut;
  if (um) {
    const SurfG<true>::Prm p{solid, id_free_surf, ntot, totmass, r2, (float)m_value, cs32};
    hipLaunchKernelGGL(k_pair_gather<SurfG<true>>, dim3(ntiles(solid)),
                       dim3(kTileP * SurfG<true>::LANES), 0, s, f, tl, p);
  } else {
    const SurfG<false>::Prm p{solid, id_free_surf, ntot, totmass, r2, 0.f, cs32};
    hipLaunchKernelGGL(k_pair_gather<SurfG<false>>, dim3(ntiles(solid)),
                       dim3(kTileP * SurfG<false>::LANES), 0, s, f, tl, p);
  }
}

void launch_surface_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                          double ntot, double totmass, int id_free_surf, double r2cut, bool um,
                          double m_value) {
  if (solid <= 0) return;
  if (um) {
    const SurfD<true>::Prm p{solid, id_free_surf, ntot, totmass, r2cut, m_value};
    hipLaunchKernelGGL(k_pair_gather<SurfD<true>>, dim3(ntiles(solid)), dim3(kTileP * SurfD<true>::LANES), 0, s, f,
                       tl, p);
  } else {
    const SurfD<false>::Prm p{solid, id_free_surf, ntot, totmass, r2cut, 0.0};
    hipLaunchKernelGGL(k_pair_gather<SurfD<false>>, dim3(ntiles(solid)), dim3(kTileP * SurfD<false>::LANES), 0, s, f,
                       tl, p);
  }
}

}  // namespace wf

namespace wf {
// neighbour record of a fused pass, and the first pass's staged extra value
// (if it has one) forwarded to the fused functor
template <class A, class B>
struct CondDensNb {
  typename A::Nb a;
  typename B::Nb b;
};
template <class A, class Nb, bool = ExtraOf<A>::on>
struct FusedExtra {};
template <class A, class Nb>
struct FusedExtra<A, Nb, true> {
  typedef typename A::Extra Extra;
  __device__ static Extra extra(const Fields& f, int j) { return A::extra(f, j); }
  __device__ static void put(Nb& nb, const Extra& e) { A::put(nb.a, e); }
};

// Conduction of one step fused with the density pass of the next: both read
// the positions after integration and the same neighbour lists, and neither
// reads what the other writes (the density pass puts the new m/rho into the
// next conduction record), so one traversal of the lists serves both.  Each
// sum runs over the list in the same order as in its own kernel.
template <class A, class B>
struct CondDens : FusedExtra<A, CondDensNb<A, B>> {
  typedef typename A::PT PT;
  static constexpr int STAGE = A::STAGE;
  static constexpr int LANES = A::LANES;
  static_assert(A::LANES == B::LANES, "fused passes share the lane split");
  static constexpr int NBATCH = A::NBATCH < B::NBATCH ? A::NBATCH : B::NBATCH;
  typedef typename A::Ctx Ctx;
  struct Prm {
    int limit;  // particles of the conduction pass (>= the density pass's)
    typename A::Prm a;
    typename B::Prm b;
  };
  struct Own {
    typename A::Own a;
    typename B::Own b;
  };
  typedef CondDensNb<A, B> Nb;
  struct Acc {
    typename A::Acc a;
    typename B::Acc b;
  };
  template <class P>
  __device__ static Ctx ctx(const Fields& f, int t, const P& p) {
    return A::ctx(f, t, p.a);
  }
  __device__ static void pos(const Fields& f, int j, const Ctx& c, PT& x, PT& y, PT& z) {
    A::pos(f, j, c, x, y, z);
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx& c) {
    A::init(f, i, act, o.a, a.a, c);
    B::init(f, i, act, o.b, a.b, c);
  }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    A::load(f, j, nb.a);
    B::load(f, j, nb.b);
  }
  __device__ static void pair(PT x, PT y, PT z, const Nb& nb, const Own& o, Acc& a,
                              const Prm& p) {
    A::pair(x, y, z, nb.a, o.a, a.a, p.a);
    B::pair(x, y, z, nb.b, o.b, a.b, p.b);
  }
  __device__ static void merge(Acc& a, const Acc& b) {
    A::merge(a.a, b.a);
    B::merge(a.b, b.b);
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own& o, const Acc& a,
                                const Prm& p) {
    A::finish(f, i, act, o.a, a.a, p.a);
    B::finish(f, i, i < p.b.limit, o.b, a.b, p.b);
  }
};
template <bool UH>
using CondDensD = CondDens<CondD<UH>, DensD<UH>>;
template <bool UH>
using CondDensG = CondDens<CondG<UH>, DensG<UH>>;

// launch policy POL with every particle's h equal (UH) or not
template <template <bool> class POL, class... A>
static void launch_gather(hipStream_t s, const Fields& f, const TileLists& tl, int grid, bool uh,
                          A... args) {
  if (uh) {
    const typename POL<true>::Prm p{args...};
    hipLaunchKernelGGL(k_pair_gather<POL<true>>, dim3(grid), dim3(kTileP * POL<true>::LANES), 0, s, f, tl, p);
  } else {
    const typename POL<false>::Prm p{args...};
    hipLaunchKernelGGL(k_pair_gather<POL<false>>, dim3(grid), dim3(kTileP * POL<false>::LANES), 0, s, f, tl, p);
  }
}

void launch_dens_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int n, int solid,
                       double dt, bool fuse, bool uh) {
  if (n <= 0) return;
  launch_gather<DensD>(s, f, tl, ntiles(n), uh, solid, n, dt, fuse ? 1 : 0);
}
void launch_rates_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                        int quirks, bool uh) {
  if (solid <= 0) return;
  const double xz = (quirks & WF_Q_ROT_XZ_SIGN) ? -1.0 : 1.0;
  launch_gather<RatesD>(s, f, tl, ntiles(solid), uh, solid, xz);
}
void launch_accel_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                        double alpha, double beta, bool uh) {
  if (solid <= 0) return;
  launch_gather<AccelD>(s, f, tl, ntiles(solid), uh, solid, alpha, beta);
}
void launch_cond_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int n, double T_inf,
                       double h_conv, double dt, bool leap, bool first, bool uh) {
  if (n <= 0) return;
  launch_gather<CondD>(s, f, tl, ntiles(n), uh, n, T_inf, h_conv, dt, leap ? 1 : 0,
                       first ? 1 : 0);
}
void launch_dens_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int n, int solid,
                       double dt, bool fuse, float cs32, bool uh) {
  if (n <= 0) return;
  launch_gather<DensG>(s, f, tl, ntiles(n), uh, solid, n, dt, fuse ? 1 : 0, cs32);
}
void launch_rates_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                        int quirks, float cs32, bool uh) {
  if (solid <= 0) return;
  const double xz = (quirks & WF_Q_ROT_XZ_SIGN) ? -1.0 : 1.0;
  launch_gather<RatesG>(s, f, tl, ntiles(solid), uh, solid, xz, cs32);
}
void launch_accel_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                        double alpha, double beta, float cs32, bool uh) {
  if (solid <= 0) return;
  launch_gather<AccelG>(s, f, tl, ntiles(solid), uh, solid, (float)alpha, (float)beta, cs32);
}
void launch_cond_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int n, double T_inf,
                       double h_conv, double dt, bool leap, bool first, float cs32, bool uh) {
  if (n <= 0) return;
  launch_gather<CondG>(s, f, tl, ntiles(n), uh, n, T_inf, h_conv, dt, leap ? 1 : 0,
                       first ? 1 : 0, cs32);
}
void launch_cond_dens_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int n,
                            int solid, double T_inf, double h_conv, double dt, bool leap,
                            bool first, bool uh) {
  if (n <= 0) return;
  auto go = [&](auto tag) {
    typedef decltype(tag) K;
    const typename K::Prm p{n, {n, T_inf, h_conv, dt, leap ? 1 : 0, first ? 1 : 0},
                            {solid, n, dt, 1, 1}};
    hipLaunchKernelGGL(k_pair_gather<K>, dim3(ntiles(n)), dim3(kTileP * K::LANES), 0, s, f, tl, p);
  };
  if (uh) go(CondDensD<true>{});
  else go(CondDensD<false>{});
}
void launch_cond_dens_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int n,
                            int solid, double T_inf, double h_conv, double dt, bool leap,
                            bool first, float cs32, bool uh) {
  if (n <= 0) return;
  auto go = [&](auto tag) {
    typedef decltype(tag) K;
    const typename K::Prm p{n, {n, T_inf, h_conv, dt, leap ? 1 : 0, first ? 1 : 0, cs32},
                            {solid, n, dt, 1, cs32, 1}};
    hipLaunchKernelGGL(k_pair_gather<K>, dim3(ntiles(n)), dim3(kTileP * K::LANES), 0, s, f, tl, p);
  };
  if (uh) go(CondDensG<true>{});
  else go(CondDensG<false>{});
}
}  // namespace wf
