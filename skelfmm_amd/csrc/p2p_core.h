// p2p_core.h -- kernel closed forms and the one-sided near-field kernel, shared by the
// library (skelfmm_hip.hip, compiled ahead of time for gfx950) and by near-field kernels of
// user-defined operators compiled at run time with hiprtc (skelfmm_amd/jit.py), which
// include this header after specializing OpT<> and kernel_value<> for their operator id.
#ifndef SKELFMM_P2P_CORE_H
#define SKELFMM_P2P_CORE_H

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include "skelfmm_hip.h"
#else
// run-time compilation: the ids and flags this header needs (values as in skelfmm_hip.h)
enum skf_kernel {
  SKF_LAPLACE_2D = 0,
  SKF_LAPLACE_3D = 1,
  SKF_HELMHOLTZ_2D = 2,
  SKF_HELMHOLTZ_3D = 3,
  SKF_LAPLACE_DL_3D = 4,
  SKF_LAPLACE_ADL_3D = 5,
  SKF_HELMHOLTZ_DL_3D = 6,
  SKF_HELMHOLTZ_ADL_3D = 7,
  SKF_USER_2D = 16,
  SKF_USER_3D = 17
};
#define SKF_SRC_NEG 1
#define SKF_SRC_SELF 2
#endif
// ---------------------------------------------------------- kernel algebra
// Closed forms of skelfmm/util_batched.py (torch path): distances are clamped
// from below (r^2 >= 1e-15 for the cdist-based kernels, >= 1e-30 for the
// Laplace double layers in FP64), self pairs (i == j inside a self segment)
// contribute zero. The FP32 form uses r^2 + 1e-20 for every kernel (the bias is
// below FP32 resolution for any r^2 > 1e-12 and keeps coincident points finite).

template <typename R>
struct Cplx {
  R re, im;
};

template <typename R>
struct Mth;

template <>
struct Mth<double> {
  static constexpr bool fp32_bias = false;
  static constexpr double clamp_cd = 1e-15;
  static constexpr double clamp_dl = 1e-30;
  // 1/sqrt(x) for normal x > 0: hardware estimate (relative error ~5e-8) refined by one Newton
  // step with its second-order term, y (1 + e/2 + 3e^2/8), e = 1 - x y^2 (measured ~1 ulp).
  __device__ static inline double rsq(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
  }
  __device__ static inline double lg(double x) { return log(x); }
  __device__ static inline double half_lg(double x) { return 0.5 * log(x); }
  __device__ static inline void sc(double x, double* s, double* c) { sincos(x, s, c); }
  __device__ static inline double bj0(double x) { return j0(x); }
  __device__ static inline double by0(double x) { return y0(x); }
};

// FP32 sine and cosine of a bounded phase, without a branch or a reduction. The host checks that
// every kernel argument kappa r is at most kTrigDirectMax and then selects these instances; their
// kernel constant is kappa / (2 pi) (kernel_kappa), so the phase comes out in revolutions,
// t = r kappa / (2 pi): one rounding of the constant and one of the product, as forming kappa r
// has. The hardware v_sin_f32 / v_cos_f32 take revolutions and reduce them exactly themselves
// within their input range |t| <= 256 (kTrigDirectMax keeps |t| below 255): measured max absolute
// error 1.25e-7 against FP64 on an MI355X for every |t| up to 300, the same as on [-1/2, 1/2].
// Other arguments use sincosf. The FP64 instances (sincos_reduced) take kappa r up to
// kTrigFastMax.
constexpr float kTrigFastMax = 65536.0f;
constexpr float kTrigDirectMax = 1600.0f;
constexpr double kInvTwoPi = 0.15915494309189535;  // 1 / (2 pi)
constexpr float kTwoPiF = 6.28318530717958648f;

__device__ inline void sincos_rev(float t, float* sp, float* cp) {
  *sp = __builtin_amdgcn_sinf(t);
  *cp = __builtin_amdgcn_cosf(t);
}

// The same for FP64 (the plain form), from the mathematics alone: Cody-Waite reduction
// x = n (pi/2) + y with pi/2 split in two FP64 parts (each step one FMA, so one rounding;
// the omitted third part is below 1e-28 for |n| < 2^16), Taylor polynomials of sin (to y^15)
// and cos (to y^16) on |y| <= pi/4 (truncation below 5e-17), and the quadrant of n: about
// 2e-16 absolute, no branches.
__device__ inline void sincos_reduced(double x, double* sp, double* cp) {
  const double n = rint(x * 0.6366197723675814);
  double y = fma(-n, 1.5707963267948966, x);
  y = fma(-n, 6.123233995736766e-17, y);
  const double z = y * y;
  double ps = fma(z, -7.647163731819816e-13, 1.6059043836821613e-10);
  ps = fma(z, ps, -2.505210838544172e-08);
  ps = fma(z, ps, 2.7557319223985893e-06);
  ps = fma(z, ps, -0.0001984126984126984);
  ps = fma(z, ps, 0.008333333333333333);
  ps = fma(z, ps, -0.16666666666666666);
  const double sn = fma(y * z, ps, y);
  double pc = fma(z, 4.779477332387385e-14, -1.1470745597729725e-11);
  pc = fma(z, pc, 2.08767569878681e-09);
  pc = fma(z, pc, -2.755731922398589e-07);
  pc = fma(z, pc, 2.48015873015873e-05);
  pc = fma(z, pc, -0.001388888888888889);
  pc = fma(z, pc, 0.041666666666666664);
  pc = fma(z, pc, -0.5);
  const double cs = fma(z, pc, 1.0);
  const int q = int(n);
  const double sv = (q & 1) ? cs : sn;
  const double cv = (q & 1) ? sn : cs;
  *sp = (q & 2) ? -sv : sv;
  *cp = ((q + 1) & 2) ? -cv : cv;
}

template <>
struct Mth<float> {
  static constexpr bool fp32_bias = true;
  static constexpr float clamp_cd = 1e-20f;
  static constexpr float clamp_dl = 1e-20f;
  __device__ static inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
  // natural log from the hardware log2 (x >= 1e-20 here: normal, no denormal scaling needed)
  __device__ static inline float lg(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }
  __device__ static inline float half_lg(float x) { return __builtin_amdgcn_logf(x) * 0.34657359027997265f; }
  __device__ static inline void sc(float x, float* s, float* c) { sincosf(x, s, c); }
  __device__ static inline float bj0(float x) { return j0f(x); }
  __device__ static inline float by0(float x) { return y0f(x); }
};

// Traits per operator: spatial dimension, complex value, where normals live.
template <int OP>
struct OpT;
template <> struct OpT<SKF_LAPLACE_2D> { static constexpr int dim = 2, cplx = 0, geom = 0; };
template <> struct OpT<SKF_LAPLACE_3D> { static constexpr int dim = 3, cplx = 0, geom = 0; };
template <> struct OpT<SKF_HELMHOLTZ_2D> { static constexpr int dim = 2, cplx = 1, geom = 0; };
template <> struct OpT<SKF_HELMHOLTZ_3D> { static constexpr int dim = 3, cplx = 1, geom = 0; };
template <> struct OpT<SKF_LAPLACE_DL_3D> { static constexpr int dim = 3, cplx = 0, geom = 1; };
template <> struct OpT<SKF_LAPLACE_ADL_3D> { static constexpr int dim = 3, cplx = 0, geom = 2; };
template <> struct OpT<SKF_HELMHOLTZ_DL_3D> { static constexpr int dim = 3, cplx = 1, geom = 1; };
template <> struct OpT<SKF_HELMHOLTZ_ADL_3D> { static constexpr int dim = 3, cplx = 1, geom = 2; };

// K(x, y) for dx = x - y and the normal dot product rn = dx . n (source or target normal).
// FT: the Helmholtz arguments are bounded by kTrigDirectMax (FP32) or kTrigFastMax (FP64) (the
// host checks it and selects these
// instances)
template <typename R, bool FT>
__device__ inline void sincos_k(R x, R* s, R* c) {
  if constexpr (FT)
    sincos_reduced(x, s, c);
  else
    Mth<R>::sc(x, s, c);
}

// FP32 instances with bounded arguments take the phase in revolutions (sincos_rev)
template <typename R, bool FT>
struct PhaseRev {
  static constexpr bool on = FT && sizeof(R) == 4;
};

// The kernel constant a near-field kernel hands to kernel_value<R, OP, FT>: kappa, or kappa / (2 pi)
// for the instances that take the phase in revolutions.
template <typename R, bool FT>
__device__ inline R kernel_kappa(double kappa) {
  return PhaseRev<R, FT>::on ? R(kappa * kInvTwoPi) : R(kappa);
}

// instances with the reduced sincos exist for the 3D Helmholtz kernels (both precisions)
template <typename R, int OP>
struct HasFastTrig {
  static constexpr bool value = OP == SKF_HELMHOLTZ_3D || OP == SKF_HELMHOLTZ_DL_3D || OP == SKF_HELMHOLTZ_ADL_3D;
};

template <typename R, int OP, bool FT = false>
__device__ inline Cplx<R> kernel_value(R dx, R dy, R dz, R rn, R kappa) {
  using M = Mth<R>;
  constexpr R inv4pi = R(0.079577471545947667884441881686257181);
  Cplx<R> k;
  R r2;
  if (M::fp32_bias) {
    // FP32 form: r^2 is accumulated on top of the clamp (3 FMAs, no separate max)
    r2 = fma(dx, dx, M::clamp_cd);
    r2 = fma(dy, dy, r2);
    if (OpT<OP>::dim == 3) r2 = fma(dz, dz, r2);
  } else {
    r2 = dx * dx + dy * dy;
    if (OpT<OP>::dim == 3) r2 += dz * dz;
  }
  if (OP == SKF_LAPLACE_2D) {
    k.re = M::half_lg(M::fp32_bias ? r2 : fmax(r2, M::clamp_cd));
    k.im = 0;
  } else if (OP == SKF_LAPLACE_3D) {
    k.re = M::rsq(M::fp32_bias ? r2 : fmax(r2, M::clamp_cd));
    k.im = 0;
  } else if (OP == SKF_HELMHOLTZ_2D) {
    R r = sqrt(M::fp32_bias ? r2 : fmax(r2, M::clamp_cd));
    k.re = M::bj0(kappa * r);
    k.im = M::by0(kappa * r);
  } else if (OP == SKF_HELMHOLTZ_3D) {
    const R rc = M::fp32_bias ? r2 : fmax(r2, M::clamp_cd);
    R ir = M::rsq(rc);
    R r = rc * ir;
    R s, c;
    if constexpr (PhaseRev<R, FT>::on)
      sincos_rev(r * kappa, &s, &c);  // kappa / (2 pi) here
    else
      sincos_k<R, FT>(kappa * r, &s, &c);
    k.re = c * ir;
    k.im = s * ir;
  } else if (OP == SKF_LAPLACE_DL_3D || OP == SKF_LAPLACE_ADL_3D) {
    R ir = M::rsq(M::fp32_bias ? r2 : fmax(r2, M::clamp_dl));
    R v = inv4pi * rn * ir * ir * ir;
    k.re = (OP == SKF_LAPLACE_DL_3D) ? -v : v;
    k.im = 0;
  } else {  // Helmholtz double layers: (1/4pi) rn e^{ikr} (ikr - 1) / r^3, sign flipped for the adjoint
    R rc = M::fp32_bias ? r2 : fmax(r2, M::clamp_cd);
    R ir = M::rsq(rc);
    R r = rc * ir;
    R s, c, kr;
    if constexpr (PhaseRev<R, FT>::on) {
      const R t = r * kappa;  // kappa / (2 pi) here: revolutions
      kr = t * kTwoPiF;
      sincos_rev(t, &s, &c);
    } else {
      kr = kappa * r;
      sincos_k<R, FT>(kr, &s, &c);
    }
    R f = inv4pi * rn * ir * ir * ir;
    if (OP == SKF_HELMHOLTZ_ADL_3D) f = -f;
    // e^{ikr}(ikr - 1) = (c + i s)(-1 + i kr) = (-c - kr s) + i (kr c - s)
    k.re = f * (-c - kr * s);
    k.im = f * (kr * c - s);
  }
  return k;
}

// --------------------------------------------------------------- P2P kernel

constexpr int kWG = 256;   // 4 wave64 per workgroup
constexpr int kTCH = 64;   // targets per work item (one per lane)

struct P2PDev {
  const void* px;
  const void* py;
  const void* pz;
  const void* nx;
  const void* ny;
  const void* nz;
  const double* centers;
  const void* q;
  void* u;
  const int* g_pt;
  const int* g_n;
  const int* g_out;
  const int* g_box;
  const int* e_pt;
  const int* e_n;
  const int* e_q;
  const int* e_box;
  const int* e_flags;
  const int4* items;
  int width;
  int nch;  // channels (complex channels for complex kernels)
  double kappa;
  double* urev;  // reverse outputs of the symmetric kernel (FP64; u itself in the FP64 form)
  int fast_trig;  // every kernel argument kappa r is within kTrigDirectMax (FP32) / kTrigFastMax (FP64)
  // precision form of the double layers: FP64 rows [7][geo64_n] (x, y, z, nx, ny, nz, n.x)
  const double* geo64;
  long long geo64_n;
  // precision form: per entry the shift of its points into its group's box frame (3 floats), or
  // nullptr (then formed from the box centres in the kernel)
  const float* e_shift;
  // every output row belongs to one work item and holds zero: the matrix-core kernel stores
  int store_rows;
};

// Shift of entry s's points into its group's box frame (precision form): the host's table when
// there is one (one load level less at the start of a work item), else the FP64 difference of
// the two box centres rounded once -- the same value the host stores.
template <typename R>
__device__ inline void entry_shift(const P2PDev& d, int s, double ctx, double cty, double ctz, R& ox, R& oy, R& oz) {
  if (d.e_shift) {
    ox = R(d.e_shift[3 * s + 0]);
    oy = R(d.e_shift[3 * s + 1]);
    oz = R(d.e_shift[3 * s + 2]);
  } else {
    const int eb = d.e_box[s];
    ox = R(ctx - d.centers[3 * eb + 0]);
    oy = R(cty - d.centers[3 * eb + 1]);
    oz = R(ctz - d.centers[3 * eb + 2]);
  }
}

// Normal projection rn = (x - y) . n of the double layers in the precision form. On a surface
// rn ~ r^2 / 2 for close pairs while the FP32 box-centred coordinates carry ~6e-8 of a box side
// of rounding each, so rn is formed in FP64 from FP64 data and rounded once:
//   source normal (geom 1): rn = x . n_y - (n_y . y)   target values x, source values (n_y, n_y . y)
//   target normal (geom 2): rn = (n_x . x) - n_x . y   target values (n_x, n_x . x), source values y
template <int OP>
struct Geo64 {
  static constexpr bool on = OpT<OP>::geom != 0;
  // the four values a staged source carries
  __device__ static inline void source(const double* g, long long n, int sl, double* v) {
    if (OpT<OP>::geom == 1) {
      v[0] = g[3 * n + sl];
      v[1] = g[4 * n + sl];
      v[2] = g[5 * n + sl];
      v[3] = g[6 * n + sl];
    } else {
      v[0] = g[sl];
      v[1] = g[n + sl];
      v[2] = g[2 * n + sl];
      v[3] = 0.0;
    }
  }
  // the four values a target holds
  __device__ static inline void target(const double* g, long long n, int sl, double* v) {
    if (OpT<OP>::geom == 1) {
      v[0] = g[sl];
      v[1] = g[n + sl];
      v[2] = g[2 * n + sl];
      v[3] = 0.0;
    } else {
      v[0] = g[3 * n + sl];
      v[1] = g[4 * n + sl];
      v[2] = g[5 * n + sl];
      v[3] = g[6 * n + sl];
    }
  }
  __device__ static inline float rn(const double* t, const double* sv) {
    double v;
    if (OpT<OP>::geom == 1)
      v = fma(t[0], sv[0], fma(t[1], sv[1], fma(t[2], sv[2], -sv[3])));
    else
      v = fma(-t[0], sv[0], fma(-t[1], sv[1], fma(-t[2], sv[2], t[3])));
    return float(v);
  }
};

__device__ inline void atomic_acc(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ inline void atomic_acc(float* p, float v) { unsafeAtomicAdd(p, v); }

// Values per staged source: coordinates, optional source normal, NCH charges (x2 if complex).
// G64: the precision form of a double layer, whose normal projection comes from FP64 rows staged
// beside the tile (Geo64), so the tile holds no FP32 normals.
template <int OP, int NCH, bool G64 = false>
struct Stage {
  static constexpr int dim = OpT<OP>::dim;
  static constexpr int cw = OpT<OP>::cplx ? 2 : 1;
  static constexpr int ng = (OpT<OP>::geom == 1 && !G64) ? 3 : 0;
  static constexpr int nraw = dim + ng + NCH * cw;
  static constexpr int nv = (nraw + 1) & ~1;  // even -> 16-byte aligned rows
};

// Sum over `cnt` staged sources of one wave's tile for this lane's target. SELF skips the
// source whose local index equals the target's (the self pair's zero diagonal).
template <typename R, int OP, int NCH, bool SELF, bool FT, bool G64 = false>
__device__ inline void tile_sum(const R* __restrict__ tw, int cnt, int self_k, R xs0, R ys0, R zs0, R gxt, R gyt,
                                R gzt, R kappa, R* part, const double* __restrict__ gw = nullptr,
                                const double* t64 = nullptr) {
  using S = Stage<OP, NCH, G64>;
  constexpr int NV = S::nv;
  constexpr int CW = S::cw;
#pragma unroll 4
  for (int k = 0; k < cnt; ++k) {
    const R* row = tw + k * NV;
    R dx = xs0 - row[0];
    R dy = ys0 - row[1];
    R dz = (S::dim == 3) ? zs0 - row[2] : R(0);
    R rn = 0;
    if constexpr (G64) {
      rn = R(Geo64<OP>::rn(t64, gw + 4 * k));
    } else {
      if (OpT<OP>::geom == 1) rn = dx * row[S::dim] + dy * row[S::dim + 1] + dz * row[S::dim + 2];
      if (OpT<OP>::geom == 2) rn = dx * gxt + dy * gyt + dz * gzt;
    }
    Cplx<R> kv = kernel_value<R, OP, FT>(dx, dy, dz, rn, kappa);
    if (SELF && k == self_k) {
      kv.re = 0;
      kv.im = 0;
    }
    const R* qq = row + S::dim + S::ng;
    if (CW == 1) {
#pragma unroll
      for (int c = 0; c < NCH; ++c) part[c] += kv.re * qq[c];
    } else {
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        part[2 * c] = fma(kv.re, qq[2 * c], fma(-kv.im, qq[2 * c + 1], part[2 * c]));
        part[2 * c + 1] = fma(kv.re, qq[2 * c + 1], fma(kv.im, qq[2 * c], part[2 * c + 1]));
      }
    }
  }
}

// One work item: up to 64 targets of one group against a range of its source entries.
// The 4 waves of the workgroup split each 256-source tile (64 sources per wave);
// partial sums are combined through LDS and added to the output with atomics.
template <typename R, int OP, int NCH, bool PREC, bool FT = false>
__global__ __launch_bounds__(kWG) void p2p_kernel(P2PDev d) {
  constexpr bool G64 = PREC && Geo64<OP>::on;
  using S = Stage<OP, NCH, G64>;
  constexpr int NV = S::nv;
  constexpr int CW = S::cw;
  constexpr int NA = NCH * CW;
  __shared__ R tile[kWG * NV];
  __shared__ double red[4][kTCH][NA];
  __shared__ double gtile[G64 ? kWG * 4 : 1];  // FP64 normal-projection data of the staged sources

  const int4 it = d.items[blockIdx.x];
  const int g = it.x, t0 = it.y, sb = it.z, se = it.w;
  if (se <= sb) return;  // padding item of an XCD-ordered table (uniform: the whole workgroup)
  const int ch0 = blockIdx.y * NCH;
  const int nchv = min(NCH, d.nch - ch0);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ng = d.g_n[g];
  const int nt = min(kTCH, ng - t0);
  const int tloc = t0 + lane;
  const bool tvalid = lane < nt;
  const int tslot = d.g_pt[g] + (tvalid ? tloc : t0);

  const R* px = static_cast<const R*>(d.px);
  const R* py = static_cast<const R*>(d.py);
  const R* pz = static_cast<const R*>(d.pz);
  const R* qv = static_cast<const R*>(d.q);
  R xt = px[tslot], yt = py[tslot], zt = (S::dim == 3) ? pz[tslot] : R(0);
  R gxt = 0, gyt = 0, gzt = 0;
  if (OpT<OP>::geom == 2 && !G64) {
    gxt = static_cast<const R*>(d.nx)[tslot];
    gyt = static_cast<const R*>(d.ny)[tslot];
    gzt = static_cast<const R*>(d.nz)[tslot];
  }
  double t64[4] = {0.0, 0.0, 0.0, 0.0};
  if constexpr (G64) Geo64<OP>::target(d.geo64, d.geo64_n, tslot, t64);
  double ctx = 0, cty = 0, ctz = 0;
  if (PREC && !d.e_shift) {
    const int gb = d.g_box[g];
    ctx = d.centers[3 * gb + 0];
    cty = d.centers[3 * gb + 1];
    ctz = d.centers[3 * gb + 2];
  }
  const R kappa = kernel_kappa<R, FT>(d.kappa);

  double acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = 0.0;

  for (int s = sb; s < se; ++s) {
    const int ps = d.e_pt[s], ns = d.e_n[s], qo = d.e_q[s], fl = d.e_flags[s];
    const bool self = (fl & SKF_SRC_SELF) != 0;
    const double sgn = (fl & SKF_SRC_NEG) ? -1.0 : 1.0;
    // target coordinates relative to this source box (precision form)
    R xs0 = xt, ys0 = yt, zs0 = zt;
    if (PREC) {
      R ox, oy, oz;
      entry_shift<R>(d, s, ctx, cty, ctz, ox, oy, oz);
      xs0 = xt + ox;
      ys0 = yt + oy;
      zs0 = zt + oz;
    }
    for (int base = 0; base < ns; base += kWG) {
      // stage 256 sources (coordinates, normals, charges of this channel block)
      {
        const int j = base + tid;
        R* row = &tile[tid * NV];
        if (j < ns) {
          const int sl = ps + j;
          row[0] = px[sl];
          row[1] = py[sl];
          if (S::dim == 3) row[2] = pz[sl];
          if (S::ng) {
            row[S::dim + 0] = static_cast<const R*>(d.nx)[sl];
            row[S::dim + 1] = static_cast<const R*>(d.ny)[sl];
            row[S::dim + 2] = static_cast<const R*>(d.nz)[sl];
          }
          const R* qs = qv + (long long)(qo + j) * d.width + ch0 * CW;
#pragma unroll
          for (int c = 0; c < NA; ++c) row[S::dim + S::ng + c] = (c < nchv * CW) ? qs[c] : R(0);
          if constexpr (G64) Geo64<OP>::source(d.geo64, d.geo64_n, sl, &gtile[tid * 4]);
        } else {
#pragma unroll
          for (int c = 0; c < NV; ++c) row[c] = R(0);
          if constexpr (G64) {
#pragma unroll
            for (int c = 0; c < 4; ++c) gtile[tid * 4 + c] = 0.0;
          }
        }
      }
      __syncthreads();
      const int wb = base + wave * 64;  // wave-uniform
      const int cnt = __builtin_amdgcn_readfirstlane(max(0, min(64, ns - wb)));
      R part[NA];
#pragma unroll
      for (int a = 0; a < NA; ++a) part[a] = 0;
      const R* tw = &tile[wave * 64 * NV];
      const double* gw = &gtile[G64 ? wave * 64 * 4 : 0];
      // the diagonal of a self segment falls in this tile only for targets tloc in [wb, wb + cnt)
      if (self && tloc - wb >= 0 && tloc - wb < 64)
        tile_sum<R, OP, NCH, true, FT, G64>(tw, cnt, tloc - wb, xs0, ys0, zs0, gxt, gyt, gzt, kappa, part, gw, t64);
      else
        tile_sum<R, OP, NCH, false, FT, G64>(tw, cnt, -1, xs0, ys0, zs0, gxt, gyt, gzt, kappa, part, gw, t64);
#pragma unroll
      for (int a = 0; a < NA; ++a) acc[a] += sgn * double(part[a]);
      __syncthreads();
    }
  }
#pragma unroll
  for (int a = 0; a < NA; ++a) red[wave][lane][a] = acc[a];
  __syncthreads();
  if (wave == 0 && tvalid) {
    R* uo = static_cast<R*>(d.u) + (long long)(d.g_out[g] + tloc) * d.width + ch0 * CW;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      if (a < nchv * CW) {
        double v = red[0][lane][a] + red[1][lane][a] + red[2][lane][a] + red[3][lane][a];
        atomic_acc(&uo[a], R(v));
      }
    }
  }
}

#endif  // SKELFMM_P2P_CORE_H
