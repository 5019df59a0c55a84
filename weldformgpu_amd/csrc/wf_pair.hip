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
#include <hip/hip_ext.h>

#include <climits>
#include <type_traits>

#include "wf_device.hpp"
#include "wf_tiles.hpp"

// straight-line batches (drop(), below) in the FP32 rates and FP64
// acceleration loops: off, measured slower there (759k-particle cylinder:
// rates 0.072 vs 0.059 ms, FP64 acceleration 0.346 vs 0.321 ms)
#ifndef WF_DROP_RATES32
#define WF_DROP_RATES32 0
#endif
#ifndef WF_DROP_ACCEL64
#define WF_DROP_ACCEL64 0
#endif
// cooperative record loads (coload_rec) in the FP64 acceleration, density
// and conduction loops (759k-particle cylinder, yielding: acceleration 0.333
// -> 0.291 ms, density 0.151 -> 0.143, conduction 0.137 -> 0.128); the FP64
// rates loop (0.145 -> 0.149) and the FP32 acceleration loop (0.118 ->
// 0.120) were slower with them and load per lane
#ifndef WF_COLOAD64
#define WF_COLOAD64 1
#endif

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

// Gather of particle j's record from an array of S-byte records (+ off
// bytes) through a 32-bit byte offset: the load then takes the array base
// from scalar registers and the offset costs one shift (a 64-bit index costs
// three instructions).  Domains hold at most kMaxParticles particles, so
// j * S stays below 2^32.
template <class T, unsigned S = sizeof(T)>
__device__ __forceinline__ T gat(const void* base, int j, unsigned off = 0) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (unsigned)j * S + off);
}

// lanes 2p and 2p+1 swap a 32-bit word (DPP quad permutation [1, 0, 3, 2])
__device__ __forceinline__ unsigned swap_lane_pair(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
}

// Cooperative record loads for two lanes per particle: lanes 2p and 2p+1
// each hold one neighbour j, and both load both records, lane s reading the
// 16-byte chunks c with c % 2 == s ^ r of record r (r = 0: lane 2p's
// neighbour, r = 1: lane 2p+1's).  A load instruction then touches one record
// per lane pair instead of one per lane: the address unit's cost per gather
// grows with the distinct cache lines it touches, so this halves that cost
// for the same bytes.  Each lane keeps the half of its own record that it
// read and swaps the other half with its partner.  C: chunks per record of
// rec_bytes = 16 C at base.
template <int C>
__device__ __forceinline__ void coload_rec(const void* base, int j, int sub, uint4 (&out)[C]) {
  static_assert(C % 2 == 0, "whole chunk pairs");
  const int jo = (int)swap_lane_pair((unsigned)j);
  const int j0 = sub ? jo : j, j1 = sub ? j : jo;
  uint4 a[C / 2], b[C / 2];
#pragma unroll
  for (int c = 0; c < C / 2; c++) a[c] = gat<uint4, 16 * C>(base, j0, 16u * (2 * c + sub));
#pragma unroll
  for (int c = 0; c < C / 2; c++) b[c] = gat<uint4, 16 * C>(base, j1, 16u * (2 * c + (sub ^ 1)));
  // lane 2p read the even chunks of its record (a) and the odd ones of its
  // partner's (b); lane 2p+1 the odd chunks of lane 2p's record (a) and the
  // even ones of its own (b)
#pragma unroll
  for (int c = 0; c < C / 2; c++) {
    const uint4 send = sub ? a[c] : b[c];
    out[2 * c] = sub ? b[c] : a[c];
    out[2 * c + 1] = make_uint4(swap_lane_pair(send.x), swap_lane_pair(send.y),
                                swap_lane_pair(send.z), swap_lane_pair(send.w));
  }
}

// The mixed acceleration pass may integrate each particle as soon as its sum
// is done (the step's BC + position/velocity update, integrate_apply):
// other tiles still read the old FP32 mirrors, so the new ones go to the
// spare planes in out, swapped in by the caller after the pass.
struct IntegratePrm {
  double dt;
  BcTable bc;
  Frame32 fr;
  IntegrateOut out;
};

// FP64 policies need no per-tile context
struct NoCtx {};

// ---------------------------------------------------------------------------
// Per-pair cubic-spline gradient GK = (1/r) dW/dr (Functions.cu:76-88) from
// r^2: one refined reciprocal square root instead of sqrt + three IEEE
// divisions.  With hj == hi (the solver sets one h for all particles,
// Domain_d::Set_h) the thread's own 1/h and 3/(pi h^4) are reused, so the
// common case needs no division at all.
//
// The two branches of dW/dq (q < 1: -3q + 9/4 q^2; 1 <= q < 2: -3/4 (2-q)^2;
// else 0) are one branch-free expression: with u = clamp(1 - q, 0, 1) and
// t = clamp(1 - q/2, 0, 1),  dW/dq = 3 (u^2 - t^2) = 3 (u - t)(u + t)
// (for q < 1, 3[(1-q)^2 - (1-q/2)^2] = -3q + 9/4 q^2; for 1 <= q < 2, u = 0
// and -3 t^2 = -3/4 (2-q)^2; for q >= 2 both are 0).  Each clamp folds into
// the instruction that produces it, so the polynomial costs five operations
// and no divergent branch.  Differences from the reference's expression
// order are a few ulp.
struct HConst {
  double h, inv_h, C3;  // h, 1/h, 3/(pi h^4)
};
__device__ __forceinline__ HConst hconst(double h) {
  return HConst{h, 1.0 / h, 3.0 / (h * h * h * h * kPi)};
}
__device__ __forceinline__ double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = y * (1.5 - hx * y * y);
  y = y * (1.5 - hx * y * y);
  return y;
}
// 1/x to within an ulp: the hardware estimate and two Newton steps
__device__ __forceinline__ double rcp_f64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(y, e, y);
}
__device__ __forceinline__ double clamp01(double x) { return fmin(fmax(x, 0.0), 1.0); }
__device__ __forceinline__ float clamp01f(float x) { return fminf(fmaxf(x, 0.f), 1.f); }
__device__ __forceinline__ double gk_pair(double r2, double hj, const HConst& hc, double& rinv,
                                          double& h_) {
  // coincident particles (never in a list built with d2 > 0, but possible
  // after motion) give q = 0, where dW/dq = 0, and contribute nothing
  rinv = rsqrt_f64(fmax(r2, 1e-300));
  const double rij = r2 * rinv;
  double inv_h = hc.inv_h, C3 = hc.C3;
  h_ = hc.h;
  if (hj != hc.h) {
    h_ = (hc.h + hj) / 2.0;
    inv_h = 1.0 / h_;
    C3 = 3.0 / (h_ * h_ * h_ * h_ * kPi);
  }
  const double q = rij * inv_h;
  const double u = clamp01(1.0 - q), t = clamp01(1.0 - 0.5 * q);
  return C3 * rinv * ((u - t) * (u + t));
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
  __device__ static void drop(Nb& nb) { nb.v.w = 0.0; }
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
    nb.v = gat<double4>(f.kv, j);
    nb.h = UH ? 0.0 : gat<double>(f.h, j);
  }
#if WF_COLOAD64
  static constexpr bool CO = true;
  __device__ static void coload(const Fields& f, int j, int sub, Nb& nb) {
    uint4 c[2];
    coload_rec<2>(f.kv, j, sub, c);
    __builtin_memcpy(&nb.v, c, 32);
    nb.h = UH ? 0.0 : gat<double>(f.h, j);
  }
#endif
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
  __device__ static void drop(Nb& nb) { DensD<UH>::drop(nb); }
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
// BV: the beta (quadratic) viscosity term is on
template <bool UH, bool BV = false>
struct AccelD {
  WF_GATHER64
  static constexpr int NBATCH = WF_ACCEL64_NBATCH;
  struct Prm {
    int limit;
    double alpha, beta;
    StatsFold fold;  // the stress pass's statistics, folded by workgroup 0
  };
  struct Own {
    double x, y, z, vx, vy, vz, di, Ci, s0, s1, s2, s3, s4, s5;
    HConst hc;
  };
  struct Nb {
    double4 v, sa, sb;  // {v, m}, sigma/rho^2 xx yy zz xy, {yz, xz, rho, Cs}
    double h;
  };
#if WF_DROP_ACCEL64
  __device__ static void drop(Nb& nb) { nb.v.w = 0.0; }
#endif
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
    nb.v = gat<double4, 96>(f.arec64, j);  // rec_at(f, j, 0..2)
    nb.sa = gat<double4, 96>(f.arec64, j, 32);
    nb.sb = gat<double4, 96>(f.arec64, j, 64);
    nb.h = UH ? 0.0 : gat<double>(f.h, j);
  }
#if WF_COLOAD64
  static constexpr bool CO = true;
  __device__ static void coload(const Fields& f, int j, int sub, Nb& nb) {
    uint4 c[6];
    coload_rec<6>(f.arec64, j, sub, c);
    __builtin_memcpy(&nb.v, &c[0], 32);
    __builtin_memcpy(&nb.sa, &c[2], 32);
    __builtin_memcpy(&nb.sb, &c[4], 32);
    nb.h = UH ? 0.0 : gat<double>(f.h, j);
  }
#endif
  __device__ static void pair(double x, double y, double z, const Nb& nb, const Own& o, Acc& a,
                              const Prm& p) {
    const double xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const double r2 = xx * xx + xy * xy + xz * xz;
    double rinv, h_;
    const double GK = gk_pair(r2, UH ? o.hc.h : nb.h, o.hc, rinv, h_);
    const double vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    const double vdx = vx * xx + vy * xy + vz * xz;
    // Monaghan viscosity on approaching pairs only: (alpha C mu + beta mu^2)
    // / rho_bar with mu = h v/e, v = min(v_ij.x_ij, 0), e = r^2 + 0.01 h^2,
    // as 2 h v (alpha C + beta mu) / (e (rho_i + rho_j)): one refined
    // reciprocal and no branch (receding pairs get v = 0)
    const double vn = fmin(vdx, 0.0);
    const double rs = o.di + nb.sb.z;
    const double inv = rcp_f64((r2 + 0.01 * h_ * h_) * rs);
    double PI;
    if constexpr (BV) {
      const double coef = (0.5 * p.alpha) * (o.Ci + nb.sb.w) + p.beta * (h_ * vn * rs * inv);
      PI = 2.0 * h_ * vn * coef * inv;
    } else {
      PI = vn * ((h_ * p.alpha) * (o.Ci + nb.sb.w)) * inv;
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
  __device__ static void drop(Nb& nb) { nb.tk.z = 0.0; }
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
    nb.tk = gat<double4>(f.tk64, j);
    nb.h = UH ? 0.0 : gat<double>(f.h, j);
  }
#if WF_COLOAD64
  static constexpr bool CO = true;
  __device__ static void coload(const Fields& f, int j, int sub, Nb& nb) {
    uint4 c[2];
    coload_rec<2>(f.tk64, j, sub, c);
    __builtin_memcpy(&nb.tk, c, 32);
    nb.h = UH ? 0.0 : gat<double>(f.h, j);
  }
#endif
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
  float h, inv_h, C3;  // h, 1/h, 3/(pi h^4)
};
__device__ __forceinline__ HConstF hconst_f(float h) {
  const float hh = h * h;
  return HConstF{h, __builtin_amdgcn_rcpf(h), 3.f * __builtin_amdgcn_rcpf(hh * hh * (float)kPi)};
}
// FP32 cubic-spline gradient (1/r) dW/dr from r^2 (Functions.cu:76-88), the
// branch-free form of gk_pair; UH: every particle has the same h, so h_j = h_i
template <bool UH>
__device__ __forceinline__ float gk_pair_f(float r2, float hj, const HConstF& hc, float& h_) {
  const float rinv = __builtin_amdgcn_rsqf(fmaxf(r2, 1e-30f));
  const float rij = r2 * rinv;
  float inv_h = hc.inv_h, C3 = hc.C3;
  h_ = hc.h;
  if (!UH && hj != hc.h) {
    h_ = 0.5f * (hc.h + hj);
    inv_h = __builtin_amdgcn_rcpf(h_);
    C3 = 3.f * __builtin_amdgcn_rcpf(h_ * h_ * h_ * h_ * (float)kPi);
  }
  const float q = rij * inv_h;
  const float u = clamp01f(1.f - q), t = clamp01f(1.f - 0.5f * q);
  return C3 * rinv * ((u - t) * (u + t));
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
  const float4 p = gat<float4>(f.p4, j);
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

// A pair functor may sum a batch's pair terms in a narrower partial sum
// (typedef Part, pzero(part), flush(acc, part)): pair() then adds into the
// partial, which is added to the functor's sums once per batch of up to
// NBATCH pairs.  Without a Part, pair() adds into the sums directly.
template <class K, class = void>
struct PartOf {
  static constexpr bool on = false;
};
template <class K>
struct PartOf<K, std::void_t<typename K::Part>> {
  static constexpr bool on = true;
};

// A pair functor whose every term scales with one neighbour weight (m_j or
// m_j/rho_j) may say how to zero it (drop(nb)): the batch then evaluates
// all of its pairs in straight-line code, entries past the list with a zero
// weight (adding exact zeros), so that the compiler keeps every record load
// ahead of the first pair and each pair waits only for its own record.
template <class K, class = void>
struct DropOf {
  static constexpr bool on = false;
};
template <class K>
struct DropOf<K, std::void_t<decltype(K::drop(std::declval<typename K::Nb&>()))>> {
  static constexpr bool on = true;
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

// A pair functor may load its neighbour records cooperatively (static
// constexpr bool CO = true, coload(f, j, sub, nb)); with two lanes per
// particle both lanes then run the same iterations of the list walk.
template <class K, class = void>
struct CoOf {
  static constexpr bool on = false;
};
template <class K>
struct CoOf<K, std::void_t<decltype(K::CO)>> {
  static constexpr bool on = K::CO;
};

// A pass whose Prm has a StatsFold (fold) folds the preceding stress pass's
// block statistics in its first workgroup, after that workgroup's tile, so
// that the fold overlaps the other workgroups instead of taking a launch.
template <class K, class = void>
struct FoldOf {
  static constexpr bool on = false;
};
template <class K>
struct FoldOf<K, std::void_t<decltype(std::declval<typename K::Prm&>().fold)>> {
  static constexpr bool on = true;
};

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
    // 32-bit word index: rows x npad stays below 2^32 (kMaxParticles)
    const unsigned np = (unsigned)tl.npad;
    auto word = [&](int k) { return q[(unsigned)(min(k, klast) >> 2) * np]; };
    auto sl_of = [](const uint2 w, int u) {
      const unsigned h = (u & 2) ? w.y : w.x;
      return (int)((u & 1) ? h >> 16 : h & 0xFFFFu);
    };
    constexpr int KS = 4 * L;  // slots from one of the lane's quads to its next
    const int k0 = 4 * sub;
    // with cooperative loads both lanes of a particle take the same number
    // of steps (a lane past its list loads for its partner and adds nothing)
    constexpr bool CO = CoOf<K>::on && L == 2;
    const int kend = CO ? (cnt + KS - 1) / KS * KS + k0 : cnt;
    uint2 w0 = word(k0);
    uint2 w1 = word(k0 + KS);
    int jn[4];
#pragma unroll
    for (int u = 0; u < 4; u++) jn[u] = k0 + u < cnt ? st[sl_of(w0, u)].j : 0;
    for (int k = k0; k < kend; k += KS) {
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
        // with straight-line batches (drop()) every record of the batch is
        // loaded, also for entries past the list: their index is the
        // particle index an earlier entry held (or particle 0), so the load
        // stays in bounds, and each pair waits only for its own record;
        // otherwise only the list's entries are loaded (the FP64
        // acceleration loop is bound by its gathers)
        typename K::Nb nb[NB];
#pragma unroll
        for (int u = 0; u < NB; u++) {
          if constexpr (CO) K::coload(f, jc[u0 + u], sub, nb[u]);
          else if (DropOf<K>::on || u0 + u == 0 || u0 + u < nk) K::load(f, jc[u0 + u], nb[u]);
        }
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
        auto batch = [&](auto& sum) {
#pragma unroll
          for (int u = 0; u < NB; u++) {
            if constexpr (DropOf<K>::on) {
              if constexpr (XS) K::put(nb[u], sx[sl[u]]);
              if (u0 + u >= nk) K::drop(nb[u]);
              K::pair(px[u], py[u], pz[u], nb[u], own, sum, prm);
            } else if ((!CO && u0 + u == 0) || u0 + u < nk) {
              // (the quad's first entry is in the list unless the lane only
              // loads for its partner: k < cnt)
              if constexpr (XS) K::put(nb[u], sx[sl[u]]);
              K::pair(px[u], py[u], pz[u], nb[u], own, sum, prm);
            }
          }
        };
        if constexpr (PartOf<K>::on) {
          typename K::Part part;
          K::pzero(part);
          batch(part);
          K::flush(acc, part);
        } else {
          batch(acc);
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
        // one or two pairs, summed like a batch of the one-pass loop
        auto eval = [&](auto&& body) {
          if constexpr (PartOf<K>::on) {
            typename K::Part part;
            K::pzero(part);
            body(part);
            K::flush(acc, part);
          } else {
            body(acc);
          }
        };
        if (sr.k < sr.cnt && mine()) {
          const int s1 = sr.slot();
          if (s1 < lim) {
            sr.advance();
            const Slot& b0 = st[s1 - c0];
            typename K::Nb b;
            K::load(f, b0.j, b);
            if constexpr (XS) K::put(b, sx[s1 - c0]);
            eval([&](auto& sum) {
              K::pair(a0.x, a0.y, a0.z, a, own, sum, prm);
              K::pair(b0.x, b0.y, b0.z, b, own, sum, prm);
            });
            continue;
          }
        }
        eval([&](auto& sum) { K::pair(a0.x, a0.y, a0.z, a, own, sum, prm); });
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
  if constexpr (FoldOf<K>::on) {
    if (blockIdx.x == 0 && prm.fold.nb > 0) {
      __syncthreads();  // the staged set is read no more: its LDS is scratch
      block_fold_stats(prm.fold, reinterpret_cast<unsigned long long*>(st));
    }
  }
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
  __device__ static void drop(Nb& nb) { nb.v.w = 0.f; }
  typedef KSum Acc;
  typedef float Part;  // a batch's terms in FP32, then added to the FP64 sum
  __device__ static void pzero(Part& q) { q = 0.f; }
  __device__ static void flush(Acc& a, const Part& q) { a.add(q); }
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
    nb.v = gat<float4>(f.v4, j);
    nb.h = UH ? 0.f : gat<float>(f.h32, j);
  }
  __device__ static void pair(float x, float y, float z, const Nb& nb, const Own& o, Part& a,
                              const Prm&) {
    const float xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const float r2 = xx * xx + xy * xy + xz * xz;
    float h_;
    const float GK = gk_pair_f<UH>(r2, nb.h, o.hc, h_);
    const float vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    a = __builtin_fmaf(nb.v.w, GK * (vx * xx + vy * xy + vz * xz), a);
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

// B: strain and rotation rates (InteractionAlt.cu:262-383).  With
// a = w v_ij (w = (m_j/rho_j) GK) the rates are sums of the nine products
// a_k x_l: the diagonal takes a_x x_x, a_y x_y, a_z x_z, each off-diagonal
// pair (k, l) the sum (strain) and the difference (rotation) of a_k x_l and
// a_l x_k.  The loop sums the nine products (one FMA each) and forms the
// sums and differences once per particle.  Summing four pairs' products in
// FP32 before adding them to FP64 was as fast but doubled the plastic-strain
// and temperature error of a 3,000-step Johnson-Cook run against FP64
// (5.6e-5 vs 2.8e-5).
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
#if WF_DROP_RATES32
  __device__ static void drop(Nb& nb) { DensG<UH>::drop(nb); }
#endif
  // the neighbour's {v, m/rho} is staged in LDS with its position
  typedef float4 Extra;
  __device__ static Extra extra(const Fields& f, int j) { return gat<float4>(f.v4, j); }
  __device__ static void put(Nb& nb, const Extra& e) { nb.v = e; }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    nb.h = UH ? 0.f : gat<float>(f.h32, j);
  }
  // products a_k x_l: xx, yy, zz, then (x y, y x), (y z, z y), (x z, z x)
  // each product of two FP32 factors is exact in FP64 and added by one FP64
  // FMA: 6 conversions and 9 FMAs a pair.  Widening each FP32 product and
  // adding it costs 27 operations, and this loop is bound by its vector ALU:
  // on a 759k-particle cylinder 0.082 -> 0.066 ms, with the same error
  // against the FP64 rates (4.7e-7 relative L2; the FP32 factors set it)
  struct Acc {
    double xx, yy, zz, xy, yx, yz, zy, xz, zx;
  };
  __device__ static void merge(Acc& a, const Acc& b) {
    a.xx += b.xx; a.yy += b.yy; a.zz += b.zz;
    a.xy += b.xy; a.yx += b.yx; a.yz += b.yz;
    a.zy += b.zy; a.xz += b.xz; a.zx += b.zx;
  }
  __device__ static void init(const Fields& f, int i, bool act, Own& o, Acc& a, const Ctx32& c) {
    a = Acc{0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!act) return;
    pos32(f, i, c, o.x, o.y, o.z);
    const float4 v = f.v4[i];
    o.vx = v.x; o.vy = v.y; o.vz = v.z;
    o.hc = hconst_f(f.h32[i]);
  }
  __device__ static void pair(float x, float y, float z, const Nb& nb, const Own& o, Acc& a,
                              const Prm&) {
    const float xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const float r2 = xx * xx + xy * xy + xz * xz;
    float h_;
    const float w = gk_pair_f<UH>(r2, nb.h, o.hc, h_) * nb.v.w;
    const float vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    const double ax = w * vx, ay = w * vy, az = w * vz;
    const double dx = xx, dy = xy, dz = xz;
    a.xx = __builtin_fma(ax, dx, a.xx);
    a.yy = __builtin_fma(ay, dy, a.yy);
    a.zz = __builtin_fma(az, dz, a.zz);
    a.xy = __builtin_fma(ax, dy, a.xy);
    a.yx = __builtin_fma(ay, dx, a.yx);
    a.yz = __builtin_fma(ay, dz, a.yz);
    a.zy = __builtin_fma(az, dy, a.zy);
    a.xz = __builtin_fma(ax, dz, a.xz);
    a.zx = __builtin_fma(az, dx, a.zx);
  }
  __device__ static double val(double v) { return v; }
  __device__ static void finish(const Fields& f, int i, bool act, const Own&, const Acc& a,
                                const Prm& p) {
    if (!act) return;
    const double xy = val(a.xy), yx = val(a.yx), yz = val(a.yz), zy = val(a.zy);
    const double xz = val(a.xz), zx = val(a.zx);
    f.strrate32[0][i] = (float)-val(a.xx); f.strrate32[1][i] = (float)-val(a.yy);
    f.strrate32[2][i] = (float)-val(a.zz);
    f.strrate32[3][i] = (float)(-0.5 * (xy + yx));
    f.strrate32[4][i] = (float)(-0.5 * (yz + zy));
    f.strrate32[5][i] = (float)(-0.5 * (xz + zx));
    f.rotrate32[3][i] = (float)(-0.5 * (xy - yx));  // the diagonal stays 0
    f.rotrate32[4][i] = (float)(-0.5 * (yz - zy));
    f.rotrate32[5][i] = (float)(-0.5 * p.xz_sign * (xz - zx));
  }
};

// D: acceleration with Monaghan viscosity (InteractionAlt.cu:18-150)
#ifndef WF_ACCEL32_NBATCH
#define WF_ACCEL32_NBATCH 4
#endif
#ifndef WF_ACCEL32_WAVES
#define WF_ACCEL32_WAVES 1
#endif
// BV: the beta (quadratic) viscosity term is on; IT: each particle is
// integrated as soon as its sum is done (IntegratePrm), its integration
// inputs loaded before the pair loop
template <bool UH, bool BV = false, bool IT = false>
struct AccelG {
  WF_GATHER32
  static constexpr int NBATCH = WF_ACCEL32_NBATCH;
  static constexpr int MIN_WAVES = WF_ACCEL32_WAVES;
  struct Prm {
    int limit;
    float alpha32, beta32;
    float cs;
    IntegratePrm ip;  // integrate each particle after its sum (IT)
    StatsFold fold;   // the stress pass's statistics, folded by workgroup 0
  };
  struct NoIn {};
  struct Own {
    float x, y, z, vx, vy, vz, di, Ci, s0, s1, s2, s3, s4, s5;
    HConstF hc;
    std::conditional_t<IT, IntegrateIn, NoIn> in;
  };
  struct Nb {
    float4 v, sa, sb;  // {v, m}, sigma/rho^2 {xx, yy, zz, xy}, {yz, xz, rho, Cs}
    float h;
  };
  __device__ static void drop(Nb& nb) { nb.v.w = 0.f; }
  struct Acc {
    KSum ax, ay, az;
  };
  // a batch's pair terms summed in FP32, then added to the FP64 sums
  struct Part {
    float ax, ay, az;
  };
  __device__ static void pzero(Part& q) { q = Part{0.f, 0.f, 0.f}; }
  __device__ static void flush(Acc& a, const Part& q) {
    a.ax.add(q.ax); a.ay.add(q.ay); a.az.add(q.az);
  }
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
    if constexpr (IT) integrate_load<kRec32>(f, i, o.in);
  }
  // the neighbour's {v, m} (m = (m/rho) rho, formed once per staged
  // particle) is staged in LDS with its position, so a pair gathers only the
  // 32-byte stress record
  typedef float4 Extra;
  __device__ static Extra extra(const Fields& f, int j) {
    float4 v = gat<float4>(f.v4, j);
    v.w *= gat<float4, 32>(f.arec, j, 16).z;  // rec32_at(j, 1)
    return v;
  }
  __device__ static void put(Nb& nb, const Extra& e) { nb.v = e; }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    nb.sa = gat<float4, 32>(f.arec, j);  // rec32_at(j, 0..1)
    nb.sb = gat<float4, 32>(f.arec, j, 16);
    nb.h = UH ? 0.f : gat<float>(f.h32, j);
  }

  __device__ static void pair(float x, float y, float z, const Nb& nb, const Own& o, Part& a,
                              const Prm& p) {
    const float xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const float r2 = xx * xx + xy * xy + xz * xz;
    float h_;
    const float GK = gk_pair_f<UH>(r2, nb.h, o.hc, h_);
    const float vx = o.vx - nb.v.x, vy = o.vy - nb.v.y, vz = o.vz - nb.v.z;
    const float vdx = vx * xx + vy * xy + vz * xz;
    // Monaghan viscosity on approaching pairs only: (alpha C mu + beta mu^2)
    // / rho_bar with mu = h v/e, v = min(v_ij.x_ij, 0), e = r^2 + 0.01 h^2,
    // as 2 h v (alpha C + beta mu) / (e (rho_i + rho_j)) = v h alpha (C_i +
    // C_j) / (e (rho_i + rho_j)) without beta: one reciprocal and no branch
    // (receding pairs get v = 0)
    const float vn = fminf(vdx, 0.f);
    const float rs = o.di + nb.sb.z;
    const float inv = __builtin_amdgcn_rcpf((r2 + 0.01f * h_ * h_) * rs);
    float PI;
    if constexpr (BV) {
      const float coef =
          (0.5f * p.alpha32) * (o.Ci + nb.sb.w) + p.beta32 * (h_ * vn * rs * inv);
      PI = 2.f * h_ * vn * coef * inv;
    } else {
      PI = vn * ((h_ * p.alpha32) * (o.Ci + nb.sb.w)) * inv;
    }
    const float g = GK * nb.v.w;
    const float gx = g * xx, gy = g * xy, gz = g * xz;
    const float mxx = o.s0 + nb.sa.x + PI, myy = o.s1 + nb.sa.y + PI;
    const float mzz = o.s2 + nb.sa.z + PI;
    const float mxy = o.s3 + nb.sa.w, myz = o.s4 + nb.sb.x, mxz = o.s5 + nb.sb.y;
    a.ax = __builtin_fmaf(mxz, gz, __builtin_fmaf(mxy, gy, __builtin_fmaf(mxx, gx, a.ax)));
    a.ay = __builtin_fmaf(myz, gz, __builtin_fmaf(myy, gy, __builtin_fmaf(mxy, gx, a.ay)));
    a.az = __builtin_fmaf(mzz, gz, __builtin_fmaf(myz, gy, __builtin_fmaf(mxz, gx, a.az)));
  }
  __device__ static void finish(const Fields& f, int i, bool act, const Own& o, const Acc& a,
                                const Prm& p) {
    if (!act) return;
    if constexpr (IT) {
      integrate_apply<kRec32>(f, i, a.ax.val(), a.ay.val(), a.az.val(), p.ip.dt, p.ip.bc,
                              p.ip.fr, p.ip.out, o.in);
      return;
    }
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
  __device__ static void drop(Nb& nb) { nb.tk.z = 0.f; }
  typedef KSum Acc;
  typedef float Part;  // a batch's terms in FP32, then added to the FP64 sum
  __device__ static void pzero(Part& q) { q = 0.f; }
  __device__ static void flush(Acc& a, const Part& q) { a.add(q); }
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
  __device__ static Extra extra(const Fields& f, int j) { return gat<float4>(f.tk4, j); }
  __device__ static void put(Nb& nb, const Extra& e) { nb.tk = e; }
  __device__ static void load(const Fields& f, int j, Nb& nb) {
    nb.h = UH ? 0.f : gat<float>(f.h32, j);
  }
  __device__ static void pair(float x, float y, float z, const Nb& nb, const Own& o, Part& a,
                              const Prm&) {
    const float xx = o.x - x, xy = o.y - y, xz = o.z - z;
    const float r2 = xx * xx + xy * xy + xz * xz;
    float h_;
    const float GK = gk_pair_f<UH>(r2, nb.h, o.hc, h_);
    const float kj = nb.tk.y;
    const float kh = (kj == o.k) ? 2.f * o.k : 4.f * o.k * kj * __builtin_amdgcn_rcpf(o.k + kj);
    a = __builtin_fmaf(nb.tk.z * kh, (o.T - nb.tk.x) * GK, a);
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
// the fused pass sums a batch in partial sums when both passes do
template <class A, class B, bool = PartOf<A>::on && PartOf<B>::on>
struct FusedPart {};
template <class A, class B>
struct FusedPart<A, B, true> {
  struct Part {
    typename A::Part a;
    typename B::Part b;
  };
  __device__ static void pzero(Part& q) {
    A::pzero(q.a);
    B::pzero(q.b);
  }
  template <class Acc>
  __device__ static void flush(Acc& s, const Part& q) {
    A::flush(s.a, q.a);
    B::flush(s.b, q.b);
  }
};

// and evaluates a batch in straight-line code when both passes can
template <class A, class B, bool = DropOf<A>::on && DropOf<B>::on>
struct FusedDrop {};
template <class A, class B>
struct FusedDrop<A, B, true> {
  __device__ static void drop(CondDensNb<A, B>& nb) {
    A::drop(nb.a);
    B::drop(nb.b);
  }
};

template <class A, class B>
struct CondDens : FusedExtra<A, CondDensNb<A, B>>, FusedPart<A, B>, FusedDrop<A, B> {
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
    // (may be null) the plastic-zone trigger's largest squared hot
    // displacement, as FP64 bits, and skin^2: when 4 d2 > skin^2 a rebuild
    // starts the next step, and the density half writes nothing
    const unsigned long long* pz_d2;
    double pz_skin2;
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
  static constexpr bool CO = CoOf<A>::on && CoOf<B>::on;
  __device__ static void coload(const Fields& f, int j, int sub, Nb& nb) {
    if constexpr (CO) {
      A::coload(f, j, sub, nb.a);
      B::coload(f, j, sub, nb.b);
    }
  }
  // S: the sums (Acc) or the batch's partial sums (Part)
  template <class S>
  __device__ static void pair(PT x, PT y, PT z, const Nb& nb, const Own& o, S& a,
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
    if (p.pz_d2 && 4.0 * __longlong_as_double((long long)*p.pz_d2) > p.pz_skin2) return;
    B::finish(f, i, i < p.b.limit, o.b, a.b, p.b);
  }
};
template <bool UH>
using CondDensD = CondDens<CondD<UH>, DensD<UH>>;
template <bool UH>
using CondDensG = CondDens<CondG<UH>, DensG<UH>>;

// the acceleration loops with the beta viscosity term
template <bool UH>
using AccelDB = AccelD<UH, true>;
template <bool UH>
using AccelGB = AccelG<UH, true>;

LaunchEvents g_next_launch = {nullptr, nullptr};

// one pair pass: grid tiles of kTileP particles, K::LANES threads each (with
// its dispatch timestamps in g_next_launch's events when those are set)
template <class K>
static void launch_pair(hipStream_t s, int grid, const Fields& f, const TileLists& tl,
                        const typename K::Prm& p) {
  if (g_next_launch.start) {
    hipExtLaunchKernelGGL(k_pair_gather<K>, dim3(grid), dim3(kTileP * K::LANES), 0, s,
                          g_next_launch.start, g_next_launch.stop, 0, f, tl, p);
    g_next_launch = LaunchEvents{nullptr, nullptr};
  } else {
    hipLaunchKernelGGL(k_pair_gather<K>, dim3(grid), dim3(kTileP * K::LANES), 0, s, f, tl, p);
  }
}

// launch policy POL with every particle's h equal (UH) or not
template <template <bool> class POL, class... A>
static void launch_gather(hipStream_t s, const Fields& f, const TileLists& tl, int grid, bool uh,
                          A... args) {
  if (uh) launch_pair<POL<true>>(s, grid, f, tl, typename POL<true>::Prm{args...});
  else launch_pair<POL<false>>(s, grid, f, tl, typename POL<false>::Prm{args...});
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
                        double alpha, double beta, bool uh, const StatsFold& fold) {
  if (solid <= 0) return;
  if (beta != 0.0) {
    launch_gather<AccelDB>(s, f, tl, ntiles(solid), uh, solid, alpha, beta, fold);
  } else {
    launch_gather<AccelD>(s, f, tl, ntiles(solid), uh, solid, alpha, beta, fold);
  }
}
void launch_accel_integrate_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                                  double alpha, double beta, float cs32, bool uh, double dt,
                                  const BcTable& bc, const Frame32& fr, const IntegrateOut& out,
                                  const StatsFold& fold) {
  if (solid <= 0) return;
  const IntegratePrm ip{dt, bc, fr, out};
  auto go = [&](auto tag) {
    typedef decltype(tag) K;
    launch_pair<K>(s, ntiles(solid), f, tl,
                   typename K::Prm{solid, (float)alpha, (float)beta, cs32, ip, fold});
  };
  if (beta != 0.0) uh ? go(AccelG<true, true, true>{}) : go(AccelG<false, true, true>{});
  else uh ? go(AccelG<true, false, true>{}) : go(AccelG<false, false, true>{});
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
                        double alpha, double beta, float cs32, bool uh, const StatsFold& fold) {
  if (solid <= 0) return;
  if (beta != 0.0)
    launch_gather<AccelGB>(s, f, tl, ntiles(solid), uh, solid, (float)alpha, (float)beta, cs32,
                           IntegratePrm{}, fold);
  else
    launch_gather<AccelG>(s, f, tl, ntiles(solid), uh, solid, (float)alpha, (float)beta, cs32,
                          IntegratePrm{}, fold);
}
void launch_cond_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int n, double T_inf,
                       double h_conv, double dt, bool leap, bool first, float cs32, bool uh) {
  if (n <= 0) return;
  launch_gather<CondG>(s, f, tl, ntiles(n), uh, n, T_inf, h_conv, dt, leap ? 1 : 0,
                       first ? 1 : 0, cs32);
}
void launch_cond_dens_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int n,
                            int solid, double T_inf, double h_conv, double dt, bool leap,
                            bool first, bool uh, const unsigned long long* pz_d2,
                            double pz_skin2) {
  if (n <= 0) return;
  auto go = [&](auto tag) {
    typedef decltype(tag) K;
    const typename K::Prm p{n, {n, T_inf, h_conv, dt, leap ? 1 : 0, first ? 1 : 0},
                            {solid, n, dt, 1, 1}, pz_d2, pz_skin2};
    launch_pair<K>(s, ntiles(n), f, tl, p);
  };
  if (uh) go(CondDensD<true>{});
  else go(CondDensD<false>{});
}
void launch_cond_dens_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int n,
                            int solid, double T_inf, double h_conv, double dt, bool leap,
                            bool first, float cs32, bool uh, const unsigned long long* pz_d2,
                            double pz_skin2) {
  if (n <= 0) return;
  auto go = [&](auto tag) {
    typedef decltype(tag) K;
    const typename K::Prm p{n, {n, T_inf, h_conv, dt, leap ? 1 : 0, first ? 1 : 0, cs32},
                            {solid, n, dt, 1, cs32, 1}, pz_d2, pz_skin2};
    launch_pair<K>(s, ntiles(n), f, tl, p);
  };
  if (uh) go(CondDensG<true>{});
  else go(CondDensG<false>{});
}
}  // namespace wf
