// skelfmm_hip.hip -- gfx950 kernels and C ABI for the skeletonized FMM apply.
//
// Apply = charge gather -> upward interpolation (per level) -> near-field
// translations (one launch over every interaction list) -> downward
// interpolation (per level) -> result scatter. The algebra follows
// skelfmm/batched_fmm.py:539-564 (and the serial statement in
// skelfmm/recursive_skel.py:784-1084); the data layout is ragged (no
// (nboxes, max_bs) padding) and every stage runs on the GPU.
//
// Build: hipcc --offload-arch=gfx950 -O3 -fPIC -shared -std=c++17
//        -Xclang -target-feature -Xclang -packed-fp32-ops -munsafe-fp-atomics
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "skelfmm_hip.h"

// ------------------------------------------------------------------ errors
static thread_local char g_err[512] = "";

static int set_err(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define SKF_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) return set_err(SKF_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define SKF_LAUNCHED()                                                                  \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess) return set_err(SKF_ERR_HIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

extern "C" const char* skf_last_error(void) { return g_err; }
extern "C" int skf_version(void) { return 1; }

// hash of the sources and flags this library was compiled from (set by the build script), also
// kept as a tagged string in the file so a loader can check it before loading
#ifndef SKF_SOURCE_HASH
#define SKF_SOURCE_HASH "0000000000000000"
#endif
extern "C" __attribute__((used, visibility("default"))) const char skf_build_tag[] =
    "skelfmm-build-id:" SKF_SOURCE_HASH;
extern "C" const char* skf_build_id(void) { return SKF_SOURCE_HASH; }

#include "p2p_core.h"

namespace {

// ------------------------------------------- symmetric near field (pairs of lists)
// For two point sets U, V that interact in both directions with the same sign, one evaluation per
// point pair serves u_U[t] += K(t, s) q_V[s] and u_V[s] += K(s, t) q_U[t]. Single layers are
// symmetric, K(s, t) = K(t, s). Double layers are K(x, y) = g(|x - y|) rn with the normal
// projection rn linear in the normal (source normal: (x - y) . n_y; target normal: (x - y) . n_x),
// so g is evaluated once (kernel_value at rn = 1) and scaled by the two directions' projections.
// A work item holds up to kSymTargets targets of one group, taken in passes
// of 64 * TPL (TPL per lane, in registers, with their own charges q[g_out + t]; a last pass of
// 1-32 or 65-96 targets in rows of 32 that both halves of a wave hold, each half taking its own
// sources), and a run of its source entries; the
// 4 waves split every 256-source tile. Forward sums stay in the lane (FP64 across tiles, combined
// over the waves through LDS at the end). Reverse sums: every lane writes its partial for each
// source of a block of SUB sources to a per-wave LDS table; the wave then sums the table's
// columns (64 / SUB lanes per source, a butterfly over those lanes) and adds each source's total
// with an FP64 atomic at urev[e_q + s].
constexpr int kSymTargets = 512;  // targets per work item (the host's SYM_TARGETS_PER_ITEM)

// targets per lane and pass: 256-target passes fit the typical leaf (a few hundred points)
// without idle lanes and keep the registers at four waves per SIMD or more; double layers hold
// each target's geometry too and take two per lane
template <typename R, int OP>
struct SymTPL {
  static constexpr int n = OpT<OP>::geom ? 2 : 4;
};

// Staged source row of the symmetric kernel: coordinates, charges and, for the plain form of a
// double layer, the normal (both directions' projections need both points' normals). The
// precision form of a double layer keeps each point's FP64 row (x, y, z, nx, ny, nz, n . x) in
// registers (targets) or in an LDS tile of its own (sources) instead.
template <int OP, bool PREC>
struct SymRow {
  static constexpr int dim = OpT<OP>::dim;
  static constexpr int cw = OpT<OP>::cplx ? 2 : 1;
  static constexpr bool g64 = PREC && OpT<OP>::geom != 0;
  static constexpr int nrm = (OpT<OP>::geom != 0 && !PREC) ? 3 : 0;  // normal at dim + cw
  static constexpr int nv = (dim + cw + nrm + 1) & ~1;
  // offsets of a point's "target values" and "source values" in its FP64 row (Geo64<OP>::rn)
  static constexpr int toff = OpT<OP>::geom == 1 ? 0 : 3;
  static constexpr int soff = OpT<OP>::geom == 1 ? 3 : 0;
};
constexpr int kGeoRow = 7;  // FP64 geometry row: x, y, z, nx, ny, nz, n . x

// sources per reverse block: the per-wave table (n x 64 partials) stays at 4 KiB
template <typename R, int CW>
struct SymSub {
  static constexpr int n = 16 * 4 / int(sizeof(R) * CW);
  static constexpr int lanes = 64 / n;  // lanes summing one source's column
};

// Sum over aligned groups of N = 2, 4, 8 or 16 lanes (every lane of a group gets it), by
// exchanges inside a row of 16 lanes (DPP): quad permutations [1,0,3,2] and [2,3,0,1], then the
// half-row mirror (lane i with 7 - i) and the row mirror (i with 15 - i), each pairing lanes of
// two subgroups that already hold equal partial sums.
// (rows outside ROWS receive 0)
template <int CTRL, int ROWS = 0xF>
__device__ inline float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS = 0xF>
__device__ inline double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int N, typename R>
__device__ inline R group_sum(R v) {
  static_assert(N == 1 || N == 2 || N == 4 || N == 8 || N == 16, "groups inside one DPP row");
  if constexpr (N >= 2) v += dpp_mov<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  if constexpr (N >= 4) v += dpp_mov<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  if constexpr (N >= 8) v += dpp_mov<0x141>(v);  // row_half_mirror
  if constexpr (N >= 16) v += dpp_mov<0x140>(v); // row_mirror
  return v;
}

// Sum over the whole wave, returned to every lane: each row of 16 as group_sum, then lane 15 of
// rows 0 and 2 added into rows 1 and 3 (row_bcast:15) and lane 31 into rows 2 and 3
// (row_bcast:31), so lane 63 holds the total, read out as a wave-uniform value.
__device__ inline float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
__device__ inline double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(b), 63);
  const int hi = __builtin_amdgcn_readlane(int(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <typename R>
__device__ inline R wave_sum(R v) {
  v = group_sum<16>(v);
  v += dpp_mov<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_mov<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return lane63(v);
}

// Orders one wave's LDS accesses around its reverse table: LDS instructions of a wave execute
// in order, so only the compiler must not move accesses across this point (no wait on the
// global atomics in flight).
__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// Entry table of a work item (LDS), for items with at most kSymEnt source entries: stream
// offsets (prefix sums of the entry sizes), first point, charge offset, flags and, in the
// precision form, the shift into the target box's frame. Items with more entries stage entry
// by entry instead.
constexpr int kSymEnt = 128;
template <typename R>
struct SymEnt {
  int start[kSymEnt + 1];
  int pt[kSymEnt], q[kSymEnt], fl[kSymEnt];
  R sh[kSymEnt][3];
};

// One pass of the symmetric kernel over targets tp .. tp + 64 TPL - 1 of the item (TPL per lane).
// HALF: rows of 32 targets, tp .. tp + 32 TPL - 1, held by both halves of the wave; each half takes
// its own SUB sources of every block of 2 SUB (a last pass that would leave more than half of a
// 64-lane row idle).
template <typename R, int OP, bool PREC, int UNR, int TPL, bool FT, bool HALF = false>
__device__ __forceinline__ void sym_pass(const P2PDev& d, R* __restrict__ tile, double* __restrict__ gtile,
                                         R* __restrict__ rvl, double* __restrict__ red, int* __restrict__ sidx,
                                         const SymEnt<R>* et, bool etab, int nent, int tot, int tp, int tend, int t0,
                                         int sb, int se, double ctx, double cty, double ctz) {
  using S = SymRow<OP, PREC>;
  constexpr int NV = S::nv;
  constexpr int CW = S::cw;
  constexpr int GEOM = OpT<OP>::geom;
  constexpr bool G64 = S::g64;
  constexpr int NRM = S::nrm;
  constexpr int SUB = SymSub<R, CW>::n;
  constexpr int LPS = SymSub<R, CW>::lanes;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int TROW = HALF ? 32 : 64;                     // targets per row
  const int tl = HALF ? (lane & 31) : lane;                // this lane's target in the row
  const int hoff = HALF ? (lane >> 5) * SUB : 0;           // this lane's sources in a block
  const int g = reinterpret_cast<const int4*>(d.items)[blockIdx.x].x;
  const int gpt = d.g_pt[g], gout = d.g_out[g];
  const R* px = static_cast<const R*>(d.px);
  const R* py = static_cast<const R*>(d.py);
  const R* pz = static_cast<const R*>(d.pz);
  const R* qv = static_cast<const R*>(d.q);
  R* uo = static_cast<R*>(d.u);
  const int w = d.width;
  const R kappa = kernel_kappa<R, FT>(d.kappa);
  // targets tp + tl + TROW k; padding targets repeat the first one with zero charge
  R qt[TPL * CW], xt[TPL], yt[TPL], zt[TPL];
  R nt[NRM ? TPL : 1][3];            // plain form of a double layer: target normals
  double gt[G64 ? TPL : 1][kGeoRow];  // precision form of a double layer: target FP64 rows
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int t = tp + tl + TROW * k;
    const int sl = gpt + (t < tend ? t : t0);
    xt[k] = px[sl];
    yt[k] = py[sl];
    zt[k] = (S::dim == 3) ? pz[sl] : R(0);
#pragma unroll
    for (int c = 0; c < CW; ++c) qt[k * CW + c] = (t < tend) ? qv[(long long)(gout + t) * w + c] : R(0);
    if constexpr (NRM != 0) {
      nt[k][0] = static_cast<const R*>(d.nx)[sl];
      nt[k][1] = static_cast<const R*>(d.ny)[sl];
      nt[k][2] = static_cast<const R*>(d.nz)[sl];
    }
    if constexpr (G64) {
#pragma unroll
      for (int c = 0; c < kGeoRow; ++c) gt[k][c] = d.geo64[c * d.geo64_n + sl];
    }
  }
  double acc[TPL * CW];
#pragma unroll
  for (int a = 0; a < TPL * CW; ++a) acc[a] = 0.0;

  // evaluation of one staged tile of nsc sources (4 waves x 64); dst(j, c, v) adds the reverse
  // sum v of channel c of the tile's source j; returns with the forward partials in part[]
  R part[TPL * CW];
  auto tile_eval = [&](int nsc, auto&& dst) {
    const int wb = wave * 64;  // wave-uniform, within the tile
    const int cnt = __builtin_amdgcn_readfirstlane(max(0, min(64, nsc - wb)));
    const R* tw = &tile[wave * 64 * NV];
    const double* gw = G64 ? &gtile[wave * 64 * kGeoRow] : nullptr;
#pragma unroll
    for (int a = 0; a < TPL * CW; ++a) part[a] = 0;
    for (int sub = 0; sub < cnt; sub += (HALF ? 2 : 1) * SUB) {
      // only the block's valid rows are evaluated; table rows past them are never summed (half
      // rows: the second half evaluates rows sub + SUB + jj, zero-charge padding past the tile's
      // sources where those end first)
      const int nj = min(SUB, cnt - sub);
#pragma unroll UNR
      for (int jj = 0; jj < nj; ++jj) {
        const R* row = tw + (sub + hoff + jj) * NV;
        const R xs = row[0];
        const R ys = row[1];
        const R zs = (S::dim == 3) ? row[2] : R(0);
        R r0 = 0, r1 = 0;
#pragma unroll
        for (int k = 0; k < TPL; ++k) {
          const R dx = xt[k] - xs;
          const R dy = yt[k] - ys;
          const R dz = (S::dim == 3) ? zt[k] - zs : R(0);
          // single layer: one value both ways; double layer: g(r) (the value at rn = 1) scaled by
          // the forward projection (point t against source s) and the reverse one (s against t)
          const Cplx<R> kv = kernel_value<R, OP, FT>(dx, dy, dz, GEOM ? R(1) : R(0), kappa);
          Cplx<R> kf = kv, kr = kv;
          if constexpr (GEOM != 0) {
            R rf, rr;
            if constexpr (G64) {
              const double* gs = gw + (sub + hoff + jj) * kGeoRow;
              rf = Geo64<OP>::rn(gt[k] + S::toff, gs + S::soff);
              rr = Geo64<OP>::rn(gs + S::toff, gt[k] + S::soff);
            } else {
              const R* ns = row + S::dim + CW;
              const R pns = dx * ns[0] + dy * ns[1] + dz * ns[2];                // (t - s) . n_s
              const R pnt = dx * nt[k][0] + dy * nt[k][1] + dz * nt[k][2];       // (t - s) . n_t
              rf = GEOM == 1 ? pns : pnt;
              rr = GEOM == 1 ? -pnt : -pns;
            }
            kf.re = kv.re * rf;
            kf.im = kv.im * rf;
            kr.re = kv.re * rr;
            kr.im = kv.im * rr;
          }
          if (CW == 1) {
            part[k] += kf.re * row[S::dim];
            r0 += kr.re * qt[k];
          } else {
            // complex multiply-accumulates as two FMAs each (no separate product and add)
            const R sr = row[S::dim], si = row[S::dim + 1];
            part[2 * k] = fma(kf.re, sr, fma(-kf.im, si, part[2 * k]));
            part[2 * k + 1] = fma(kf.re, si, fma(kf.im, sr, part[2 * k + 1]));
            const R tr = qt[2 * k], ti = qt[2 * k + 1];
            r0 = fma(kr.re, tr, fma(-kr.im, ti, r0));
            r1 = fma(kr.re, ti, fma(kr.im, tr, r1));
          }
        }
        rvl[((wave * SUB + jj) * CW + 0) * 64 + lane] = r0;
        if (CW == 2) rvl[((wave * SUB + jj) * CW + CW - 1) * 64 + lane] = r1;
      }
      wave_sync_lds();
      // column sums: lanes j * LH .. j * LH + LH - 1 sum source j's partials, PER each (half
      // rows: source j < SUB is row j, lanes 0-31; source SUB + j row j, lanes 32-63)
      constexpr int LH = HALF ? LPS / 2 : LPS;
      constexpr int PER = 64 / LPS;
      const int j = lane / LH, pp = lane % LH;
      const int trow = HALF ? j % SUB : j, tcol = HALF ? (j / SUB) * 32 : 0;
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        R v = 0;
#pragma unroll
        for (int i = 0; i < PER; ++i) v += rvl[((wave * SUB + trow) * CW + c) * 64 + tcol + pp * PER + i];
        v = group_sum<LH>(v);
        if (pp == 0 && sub + j < cnt) dst(wb + sub + j, c, double(v));
      }
      wave_sync_lds();
    }
  };
  // a staged source's geometry (double layers): its normal in the row (plain form) or its FP64 row
  auto stage_geom = [&](R* row, int i, int sl) {
    if constexpr (NRM != 0) {
      row[S::dim + CW + 0] = static_cast<const R*>(d.nx)[sl];
      row[S::dim + CW + 1] = static_cast<const R*>(d.ny)[sl];
      row[S::dim + CW + 2] = static_cast<const R*>(d.nz)[sl];
    }
    if constexpr (G64) {
#pragma unroll
      for (int c = 0; c < kGeoRow; ++c) gtile[i * kGeoRow + c] = d.geo64[c * d.geo64_n + sl];
    }
  };
  // padding rows: zero charges; a finite geometry (its terms are multiplied by zero)
  auto stage_geom_pad = [&](int i) {
    if constexpr (G64) {
#pragma unroll
      for (int c = 0; c < kGeoRow; ++c) gtile[i * kGeoRow + c] = 0.0;
    }
  };
  if constexpr (sizeof(R) == 8 && !OpT<OP>::cplx) etab = false;  // entry by entry (see the kernel)
  if (etab) {
    // sources as one stream across the item's entries, cut into tiles of 256 (small skeleton
    // entries share a tile); each staged source carries its entry's sign in its charges and
    // its reverse-sum slot (complemented for subtracted entries) in sidx
    int ce = 0;  // entry of the tile start (uniform)
    for (int base = 0; base < tot; base += kWG) {
      const int nsc = min(kWG, tot - base);
      {
        const int pos = base + tid;
        R* row = &tile[tid * NV];
        if (pos < tot) {
          int e = ce;
          while (et->start[e + 1] <= pos) ++e;
          const int j = pos - et->start[e];
          const int sl = et->pt[e] + j, qs0 = et->q[e] + j;
          const bool neg = (et->fl[e] & SKF_SRC_NEG) != 0;
          row[0] = PREC ? px[sl] - et->sh[e][0] : px[sl];
          row[1] = PREC ? py[sl] - et->sh[e][1] : py[sl];
          if (S::dim == 3) row[2] = PREC ? pz[sl] - et->sh[e][2] : pz[sl];
          const R* qs = qv + (long long)qs0 * w;
#pragma unroll
          for (int c = 0; c < CW; ++c) row[S::dim + c] = neg ? -qs[c] : qs[c];
          stage_geom(row, tid, sl);
          sidx[tid] = neg ? ~qs0 : qs0;
        } else {
#pragma unroll
          for (int c = 0; c < NV; ++c) row[c] = R(0);
          stage_geom_pad(tid);
        }
      }
      __syncthreads();
      while (ce < nent && et->start[ce + 1] <= base + kWG) ++ce;
      tile_eval(nsc, [&](int jt, int c, double v) {
        const int ix = sidx[jt];
        unsafeAtomicAdd(&d.urev[(long long)(ix < 0 ? ~ix : ix) * w + c], ix < 0 ? -v : v);
      });
#pragma unroll
      for (int a = 0; a < TPL * CW; ++a) acc[a] += double(part[a]);
      __syncthreads();
    }
  } else {
    for (int s = sb; s < se; ++s) {
      const int ps = d.e_pt[s], ns = d.e_n[s], qo = d.e_q[s], fl = d.e_flags[s];
      const double sgn = (fl & SKF_SRC_NEG) ? -1.0 : 1.0;
      // precision form: sources are shifted into the target box's frame (minus the difference of
      // the centres) as they are staged, once per source for all targets of the item
      R ox = 0, oy = 0, oz = 0;
      if (PREC) {
        entry_shift<R>(d, s, ctx, cty, ctz, ox, oy, oz);
      }
      for (int base = 0; base < ns; base += kWG) {
        const int nsc = min(kWG, ns - base);
        {
          const int j = base + tid;
          R* row = &tile[tid * NV];
          if (j < ns) {
            const int sl = ps + j;
            row[0] = PREC ? px[sl] - ox : px[sl];
            row[1] = PREC ? py[sl] - oy : py[sl];
            if (S::dim == 3) row[2] = PREC ? pz[sl] - oz : pz[sl];
            const R* qs = qv + (long long)(qo + j) * w;
#pragma unroll
            for (int c = 0; c < CW; ++c) row[S::dim + c] = qs[c];
            stage_geom(row, tid, sl);
          } else {
#pragma unroll
            for (int c = 0; c < NV; ++c) row[c] = R(0);
            stage_geom_pad(tid);
          }
        }
        __syncthreads();
        tile_eval(nsc, [&](int jt, int c, double v) {
          unsafeAtomicAdd(&d.urev[(long long)(qo + base + jt) * w + c], sgn * v);
        });
#pragma unroll
        for (int a = 0; a < TPL * CW; ++a) acc[a] += sgn * double(part[a]);
        __syncthreads();
      }
    }
  }
  // forward sums: combine the 4 waves (and the two halves of half rows), one target slot at a time
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
#pragma unroll
    for (int c = 0; c < CW; ++c) red[(wave * kTCH + lane) * CW + c] = acc[k * CW + c];
    __syncthreads();
    const int t = tp + tl + TROW * k;
    if (wave == 0 && lane < TROW && t < tend) {
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        double v = red[(0 * kTCH + lane) * CW + c] + red[(1 * kTCH + lane) * CW + c] + red[(2 * kTCH + lane) * CW + c] +
                   red[(3 * kTCH + lane) * CW + c];
        if (HALF)
          v += red[(0 * kTCH + lane + 32) * CW + c] + red[(1 * kTCH + lane + 32) * CW + c] +
               red[(2 * kTCH + lane + 32) * CW + c] + red[(3 * kTCH + lane + 32) * CW + c];
        atomic_acc(&uo[(long long)(gout + t) * w + c], R(v));
      }
    }
    __syncthreads();
  }
}

template <typename R, int OP, bool PREC, int UNR, bool FT = false>
__global__ __launch_bounds__(kWG) void p2p_sym_kernel(P2PDev d) {
  using S = SymRow<OP, PREC>;
  constexpr int NV = S::nv;
  constexpr int CW = S::cw;
  constexpr int TPL = SymTPL<R, OP>::n;
  constexpr int SUB = SymSub<R, CW>::n;
  __shared__ R tile[kWG * NV];
  __shared__ double gtile[S::g64 ? kWG * kGeoRow : 1];  // precision form of a double layer
  __shared__ R rvl[4 * SUB * CW * 64];  // reverse partials of one block of sources, per wave
  __shared__ double red[4 * kTCH * CW];
  __shared__ SymEnt<R> et;
  __shared__ int sidx[kWG];  // reverse-sum slot of each staged source (complemented if subtracted)

  const int4 it = d.items[blockIdx.x];
  const int g = it.x, t0 = it.y, sb = it.z, se = it.w;
  if (se <= sb) return;  // padding item of an XCD-ordered table (uniform: the whole workgroup)
  const int ng = d.g_n[g];
  double ctx = 0, cty = 0, ctz = 0;
  if (PREC && !d.e_shift) {
    const int gb = d.g_box[g];
    ctx = d.centers[3 * gb + 0];
    cty = d.centers[3 * gb + 1];
    ctz = d.centers[3 * gb + 2];
  }
  const int nent = se - sb;
  // the source stream across entries pays where entries are small against the 4 x 64 tile; the
  // FP64 real kernels stage entry by entry (measured faster there: the stream path's registers
  // halve their occupancy)
  constexpr bool STREAM = sizeof(R) == 4 || OpT<OP>::cplx;
  const bool etab = STREAM && nent <= kSymEnt;  // uniform
  int tot = 0;
  if (etab) {
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid < 64) {
      // wave 0: prefix sums of the entry sizes, two per lane
      int n2[2], loc = 0;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = lane * 2 + u;
        n2[u] = e < nent ? d.e_n[sb + e] : 0;
        loc += n2[u];
      }
      int inc = loc;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
      }
      int run = inc - loc;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        et.start[lane * 2 + u] = run;
        run += n2[u];
      }
      if (lane == 63) et.start[kSymEnt] = run;
    }
    for (int e = tid; e < nent; e += kWG) {
      const int s = sb + e;
      et.pt[e] = d.e_pt[s];
      et.q[e] = d.e_q[s];
      et.fl[e] = d.e_flags[s];
      R ox = 0, oy = 0, oz = 0;
      if (PREC) {
        entry_shift<R>(d, s, ctx, cty, ctz, ox, oy, oz);
      }
      et.sh[e][0] = ox;
      et.sh[e][1] = oy;
      et.sh[e][2] = oz;
    }
    __syncthreads();
    tot = et.start[nent];
  }
  const int tend = min(ng, t0 + kSymTargets);
  // wave-uniform passes; the last one takes as many targets per lane as it needs (no idle slots
  // beyond the next multiple of 64), in rows of 32 where those pad less (1 or 3 per lane)
  for (int tp = t0; tp < tend;) {
    const int tpl = min(TPL, (tend - tp + 63) / 64);
    const int tph = (tend - tp + 31) / 32;
    if (tph == 1 || (TPL >= 3 && tph == 3)) {  // then 32 tph < 64 tpl: the last pass
      if (tph == 1)
        sym_pass<R, OP, PREC, UNR, 1, FT, true>(d, tile, gtile, rvl, red, sidx, &et, etab, nent, tot, tp, tend, t0, sb,
                                                se, ctx, cty, ctz);
      else
        sym_pass<R, OP, PREC, UNR, (TPL >= 3 ? 3 : 1), FT, true>(d, tile, gtile, rvl, red, sidx, &et, etab, nent, tot,
                                                                 tp, tend, t0, sb, se, ctx, cty, ctz);
      break;
    }
    if (TPL >= 4 && tpl >= 4)
      sym_pass<R, OP, PREC, UNR, (TPL >= 4 ? 4 : 1), FT>(d, tile, gtile, rvl, red, sidx, &et, etab, nent, tot, tp, tend,
                                                          t0, sb, se, ctx, cty, ctz);
    else if (TPL >= 3 && tpl == 3)
      sym_pass<R, OP, PREC, UNR, (TPL >= 3 ? 3 : 1), FT>(d, tile, gtile, rvl, red, sidx, &et, etab, nent, tot, tp, tend,
                                                          t0, sb, se, ctx, cty, ctz);
    else if (tpl >= 2)
      sym_pass<R, OP, PREC, UNR, 2, FT>(d, tile, gtile, rvl, red, sidx, &et, etab, nent, tot, tp, tend, t0, sb, se, ctx,
                                        cty, ctz);
    else
      sym_pass<R, OP, PREC, UNR, 1, FT>(d, tile, gtile, rvl, red, sidx, &et, etab, nent, tot, tp, tend, t0, sb, se, ctx,
                                        cty, ctz);
    tp += 64 * tpl;
  }
}

template <typename R>
__global__ void fold_rev_kernel(R* __restrict__ u, const double* __restrict__ urev, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] += R(urev[i]);
}

// Multi-RHS near field on the matrix cores (real kernels, FP32 or FP64): for every 4 sources
// a lane evaluates one kernel value K(t, s) with t = its wave's target (lane % 16) and
// s = source (lane / 16) -- the A-operand layout of v_mfma_{f32,f64}_16x16x4 (A[i][k] in
// lane i + 16k) -- and four MFMAs multiply it into 16 x 64 right-hand sides read from LDS
// (B[k][j] in lane j + 16k). Wave w owns targets t0 + 16w .. t0 + 16w + 15. Accumulator r
// of a lane holds D[4 (lane/16) + r][lane % 16] (f32) or D[lane/16 + 4 r][lane % 16] (f64).
// The MFMA sums of one 64-source tile are added into FP64 accumulators after every tile.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kMB = 64;         // channels per launch block (4 MFMA column blocks of 16)
constexpr int kQS = kMB + 16;   // padded LDS row: rows 4 apart land on distinct banks

template <typename R>
struct Mfma;
template <>
struct Mfma<float> {
  typedef v4f V;
  __device__ static inline V mac(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct Mfma<double> {
  typedef v4d V;
  __device__ static inline V mac(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <typename R, int OP, bool PREC>
__global__ __launch_bounds__(kWG) void p2p_mfma_kernel(P2PDev d) {
  static_assert(OpT<OP>::cplx == 0, "real kernels only");
  constexpr bool G64 = PREC && Geo64<OP>::on;  // double layer, precision form: rn from FP64 rows
  using S = Stage<OP, 1, G64>;
  using MF = Mfma<R>;
  typedef typename MF::V V;
  constexpr int DIM = S::dim;
  constexpr int NG = S::ng;
  __shared__ R qt[64 * kQS];
  __shared__ R pt[64 * 8];   // x, y, z, nx, ny, nz per source (precision form: in the target box frame)
  __shared__ double pg[G64 ? 64 * 4 : 1];  // FP64 normal-projection data per source (Geo64)
  __shared__ int pself[64];  // local index of a self-segment source (its diagonal target), else -1
  // the item's entries (up to kEnt): stream offsets and per-entry data, so that finding the entry
  // of a staged source is a walk through LDS rather than a chain of dependent global loads
  constexpr int kEnt = sizeof(R) == 4 ? 256 : 128;
  __shared__ int e_start[kEnt + 1];
  __shared__ int e_meta[kEnt][3];  // charge offset, first point, flags
  __shared__ R e_shift[kEnt][3];   // precision form: centre difference into the target box frame
  const int4 it = d.items[blockIdx.x];
  const int g = it.x, t0 = it.y, sb = it.z, se = it.w;
  if (se <= sb) return;  // padding item of an XCD-ordered table (uniform: the whole workgroup)
  const int c0 = blockIdx.y * kMB;
  const int ncv = min(kMB, d.nch - c0);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ng = d.g_n[g];
  // target blocks of 16 per item: with at most 32 targets the waves split the sources as well
  // (TB blocks x SPL source parts; wave = part * TB + block), so small groups keep all four
  // waves on matrix work; with more, one block per wave and waves past the last block only stage
  const int nti = min(64, ng - t0);
  const int TB = nti <= 16 ? 1 : nti <= 32 ? 2 : 4;
  const int SPL = 4 / TB;
  const int tb = wave % TB, spart = wave / TB;
  const bool wave_busy = t0 + tb * 16 < ng;
  const int tloc = t0 + tb * 16 + (lane & 15);  // target of this lane's A rows
  const bool tvalid = tloc < ng;
  const int tslot = d.g_pt[g] + (tvalid ? tloc : t0);
  const R* px = static_cast<const R*>(d.px);
  const R* py = static_cast<const R*>(d.py);
  const R* pz = static_cast<const R*>(d.pz);
  const R* qv = static_cast<const R*>(d.q);
  const R xt = px[tslot], yt = py[tslot], zt = (DIM == 3) ? pz[tslot] : R(0);
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
  double acc[4][4];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[nb][r] = 0.0;

  // The item's entries form one stream of sources; tiles of 64 are cut from it across entry
  // boundaries (small skeleton entries fill a tile together). Each staged source carries its
  // entry's sign (folded into its charges), self flag and, in the precision form, the shift
  // into the target box's frame.
  const int nent = se - sb;
  const bool etab = nent <= kEnt;  // uniform
  int tot = 0;
  if (etab) {
    // wave 0 scans the entry sizes (4 per lane), every thread fills the per-entry table
    if (wave == 0) {
      int n4[4], loc = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = lane * 4 + u;
        n4[u] = e < nent ? d.e_n[sb + e] : 0;
        loc += n4[u];
      }
      int inc = loc;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
      }
      int run = inc - loc;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = lane * 4 + u;
        if (e <= kEnt) e_start[e] = run;
        run += n4[u];
      }
      if (lane == 63 && 4 * 64 <= kEnt) e_start[kEnt] = run;
    }
    for (int e = tid; e < nent; e += kWG) {
      const int s = sb + e;
      e_meta[e][0] = d.e_q[s];
      e_meta[e][1] = d.e_pt[s];
      e_meta[e][2] = d.e_flags[s];
      R ox = 0, oy = 0, oz = 0;
      if (PREC) {
        entry_shift<R>(d, s, ctx, cty, ctz, ox, oy, oz);
      }
      e_shift[e][0] = ox;
      e_shift[e][1] = oy;
      e_shift[e][2] = oz;
    }
    __syncthreads();
    tot = e_start[nent];
  } else {
    for (int s = sb; s < se; ++s) tot += d.e_n[s];
  }
  int cs = sb, cj = 0;  // stream cursor of the tile start (uniform; global path)
  int ce = 0;           // entry of the tile start (LDS path)
  // staging of one tile: global loads into registers, then LDS stores (loading the next tile
  // during the matrix work was measured slower: the extra registers cost a wave per SIMD)
  const int sr = tid >> 2, cq = (tid & 3) * 16;
  R rq[16], rp[6];
  double rg[4] = {0.0, 0.0, 0.0, 0.0};
  int rself = -1;
  auto stage_load = [&](int tbase, int nsrc) {
    if (sr < nsrc) {
      int s, j, eq, ept, fl;
      R ox = 0, oy = 0, oz = 0;
      if (etab) {
        const int pos = tbase + sr;
        int e = ce;
        while (e_start[e + 1] <= pos) ++e;
        j = pos - e_start[e];
        s = sb + e;
        eq = e_meta[e][0];
        ept = e_meta[e][1];
        fl = e_meta[e][2];
        ox = e_shift[e][0];
        oy = e_shift[e][1];
        oz = e_shift[e][2];
      } else {
        s = cs;
        j = cj + sr;
        while (j >= d.e_n[s]) {
          j -= d.e_n[s];
          ++s;
        }
        eq = d.e_q[s];
        ept = d.e_pt[s];
        fl = d.e_flags[s];
        if (PREC) {
          entry_shift<R>(d, s, ctx, cty, ctz, ox, oy, oz);
        }
      }
      const R sgn = (fl & SKF_SRC_NEG) ? R(-1) : R(1);
      const R* src = qv + (long long)(eq + j) * d.width + c0 + cq;
#pragma unroll
      for (int c = 0; c < 16; ++c) rq[c] = (cq + c < ncv) ? sgn * src[c] : R(0);
      if ((tid & 3) == 0) {
        const int sl = ept + j;
        rp[0] = px[sl] - ox;
        rp[1] = py[sl] - oy;
        rp[2] = (DIM == 3) ? pz[sl] - oz : R(0);
        rp[3] = NG ? static_cast<const R*>(d.nx)[sl] : R(0);
        rp[4] = NG ? static_cast<const R*>(d.ny)[sl] : R(0);
        rp[5] = NG ? static_cast<const R*>(d.nz)[sl] : R(0);
        if constexpr (G64) Geo64<OP>::source(d.geo64, d.geo64_n, sl, rg);
        rself = (fl & SKF_SRC_SELF) ? j : -1;
      }
    } else {
      // far away, zero charge: contributes exactly zero
#pragma unroll
      for (int c = 0; c < 16; ++c) rq[c] = R(0);
      rp[0] = rp[1] = rp[2] = R(1e18);
      rp[3] = rp[4] = rp[5] = R(0);
      rg[0] = rg[1] = rg[2] = rg[3] = 0.0;
      rself = -1;
    }
    // advance the cursor by the tile (uniform)
    if (etab) {
      const int next = tbase + nsrc;
      while (ce < nent && e_start[ce + 1] <= next) ++ce;
    } else {
      cj += nsrc;
      while (cs < se && cj >= d.e_n[cs]) {
        cj -= d.e_n[cs];
        ++cs;
      }
    }
  };
  auto stage_store = [&]() {
    R* dst = &qt[sr * kQS + cq];
#pragma unroll
    for (int c = 0; c < 16; ++c) dst[c] = rq[c];
    if ((tid & 3) == 0) {
      R* row = &pt[sr * 8];
#pragma unroll
      for (int c = 0; c < 6; ++c) row[c] = rp[c];
      if constexpr (G64) {
#pragma unroll
        for (int c = 0; c < 4; ++c) pg[sr * 4 + c] = rg[c];
      }
      pself[sr] = rself;
    }
  };
  if (tot > 0) {
    stage_load(0, min(64, tot));
    stage_store();
  }
  for (int base = 0; base < tot; base += 64) {
    const int nsrc = min(64, tot - base);
    __syncthreads();
    const bool more = base + 64 < tot;
    if (wave_busy) {
      V D[4];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) D[nb] = V{0, 0, 0, 0};
      const int ksteps = (nsrc + 3) >> 2;
      for (int kk = spart; kk < ksteps; kk += SPL) {
        const int sj = kk * 4 + (lane >> 4);
        const R* row = &pt[sj * 8];
        const R dx = xt - row[0], dy = yt - row[1], dz = (DIM == 3) ? zt - row[2] : R(0);
        R rn = 0;
        if constexpr (G64) {
          rn = R(Geo64<OP>::rn(t64, &pg[sj * 4]));
        } else {
          if (OpT<OP>::geom == 1) rn = dx * row[3] + dy * row[4] + dz * row[5];
          if (OpT<OP>::geom == 2) rn = dx * gxt + dy * gyt + dz * gzt;
        }
        R a = kernel_value<R, OP>(dx, dy, dz, rn, R(0)).re;
        if (pself[sj] == tloc) a = R(0);
        const R* qrow = &qt[sj * kQS + (lane & 15)];
        // all four column blocks: channels past the last are zero in the staged tile, so the
        // matrix ops stay branch-free and the accumulators stay in AGPRs
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) D[nb] = MF::mac(a, qrow[nb * 16], D[nb]);
      }
#pragma unroll
      for (int nb = 0; nb < 4; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[nb][r] += double(D[nb][r]);
    }
    __syncthreads();
    if (more) {
      stage_load(base + 64, min(64, tot - base - 64));
      stage_store();
    }
  }
  R* uo = static_cast<R*>(d.u);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const int ch = nb * 16 + (lane & 15);
    if (ch >= ncv) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int tl = t0 + tb * 16 + MF::row(lane, r);
      if (wave_busy && tl < ng) atomic_acc(&uo[(long long)(d.g_out[g] + tl) * d.width + c0 + ch], R(acc[nb][r]));
    }
  }
}

// The matrix-core near field with the charge tiles staged by LDS-direct loads
// (global_load_lds_dwordx4: lane l's 16 bytes land at the instruction's LDS base + 16 l) into
// two LDS buffers: the loads of tile i + 1 are issued before the matrix work of tile i and
// waited for after it, so the staging latency overlaps the matrix work without holding the
// charges in registers. A tile is TR sources x 64 channels (16 KiB: TR = 64 in FP32, 32 in
// FP64), row-major. Column j of matrix block nb is channel 16 E (nb / E) + E j + nb mod E, E the
// scalars per 16 bytes (4 in FP32, 2 in FP64), so the B operands a lane needs in a matrix step
// come in 16-byte LDS reads (one in FP32, two in FP64) at offsets that are constants of the
// unrolled loop, and the 16 lanes of a row read 256 contiguous bytes. Staged charges are raw:
// the entry sign (times
// the kernel's constant factor, 0.5 ln 2 for the FP32 2D Laplace kernel) multiplies the kernel
// value, and rows past the tile's sources load row 0 with factor 0. The two buffers are distinct
// arrays and the tile loop is unrolled by two, so the compiler can tell the LDS reads of the
// current tile from the loads in flight. Needs a width that is a multiple of 16 bytes' worth of
// scalars and 16-byte aligned charges.
typedef __attribute__((address_space(3))) void* lds_void_ptr;
typedef __attribute__((address_space(1))) void* global_void_ptr;

// vmcnt(0), other counters unconstrained (gfx9 s_waitcnt encoding)
__device__ inline void wait_vm_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// lambdas of the kernels below are always inlined (a call would put their state on the stack)
#define SKF_INLINE __attribute__((always_inline))

template <int N>
struct IntC {
  static constexpr int value = N;
};

// FP32 matrix sums run over at most this many steps (4 sources each) before they are added into
// the FP64 accumulators
constexpr int kFoldSteps = 32;

// FP32: four waves per SIMD (the LDS of four workgroups fits a CU)
template <typename R, int OP, bool PREC>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 1)))
void p2p_mfma_dma_kernel(P2PDev d) {
  static_assert(OpT<OP>::cplx == 0, "real kernels only");
  constexpr bool G64 = PREC && Geo64<OP>::on;  // double layer, precision form: rn from FP64 rows
  using S = Stage<OP, 1, G64>;
  using MF = Mfma<R>;
  typedef typename MF::V V;
  constexpr int DIM = S::dim;
  constexpr int NG = S::ng;
  // per-source row: x, y, z, factor (+ normals); the self index rides in a free slot of the row
  // (z in 2D, the eighth value with normals), else in its own array
  constexpr int PS = NG ? 8 : 4;
  constexpr int SG = NG ? 6 : 3;                           // position of the factor in the row
  constexpr int SELFP = DIM == 2 ? 2 : (NG ? 7 : -1);      // position of the self index, or -1
  // FP32 2D Laplace: the kernel value is log2(r^2) times the staged factor (0.5 ln 2 folded in)
  constexpr bool L2F = OP == SKF_LAPLACE_2D && sizeof(R) == 4;
  constexpr R KSCALE = L2F ? R(0.34657359027997265) : R(1);
  constexpr int TR = sizeof(R) == 4 ? 64 : 32;             // sources per tile
  constexpr int EPC = 16 / int(sizeof(R));                 // scalars per 16-byte chunk
  constexpr int CPR = kMB / EPC;                           // chunks per tile row
  constexpr int RPI = 64 / CPR;                            // rows filled by one load instruction
  constexpr int RW = TR / 4;                               // rows staged per wave
  static_assert(4 * RPI == RW, "four load instructions cover a wave's rows");
  typedef R QV __attribute__((ext_vector_type(EPC)));      // one 16-byte LDS read of B operands
  __shared__ __attribute__((aligned(32))) R qa[TR * kMB];
  __shared__ __attribute__((aligned(32))) R qb[TR * kMB];
  __shared__ __attribute__((aligned(16))) R pa[TR * PS];
  __shared__ __attribute__((aligned(16))) R pb[TR * PS];
  __shared__ double g64a[G64 ? TR * 4 : 1];  // FP64 normal-projection data per source (Geo64)
  __shared__ double g64b[G64 ? TR * 4 : 1];
  __shared__ int sa[SELFP < 0 ? TR : 1];
  __shared__ int sbf[SELFP < 0 ? TR : 1];
  constexpr int kEnt = 128;
  __shared__ int e_start[kEnt + 1];
  __shared__ int e_meta[kEnt][3];  // charge offset, first point, flags
  __shared__ R e_shift[kEnt][3];
  const int4 it = d.items[blockIdx.x];
  const int g = it.x, t0 = it.y, sb = it.z, se = it.w;
  if (se <= sb) return;  // padding item of an XCD-ordered table (uniform: the whole workgroup)
  const int c0 = blockIdx.y * kMB;
  const int ncv = min(kMB, d.nch - c0);
  const int cmax = d.width - c0 - EPC;  // last 16-byte chunk start inside the row (width % EPC == 0)
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ng = d.g_n[g];
  const int nti = min(64, ng - t0);
  const int TB = nti <= 16 ? 1 : nti <= 32 ? 2 : 4;
  const int SPL = 4 / TB;
  const int tb = wave % TB, spart = wave / TB;
  const bool wave_busy = t0 + tb * 16 < ng;
  const int tloc = t0 + tb * 16 + (lane & 15);
  const R tlocr = R(tloc);
  const bool tvalid = tloc < ng;
  const int tslot = d.g_pt[g] + (tvalid ? tloc : t0);
  const R* px = static_cast<const R*>(d.px);
  const R* py = static_cast<const R*>(d.py);
  const R* pz = static_cast<const R*>(d.pz);
  const R* qv = static_cast<const R*>(d.q);
  const R xt = px[tslot], yt = py[tslot], zt = (DIM == 3) ? pz[tslot] : R(0);
  R gxt = 0, gyt = 0, gzt = 0;
  if (OpT<OP>::geom == 2 && !G64) {
    gxt = static_cast<const R*>(d.nx)[tslot];
    gyt = static_cast<const R*>(d.ny)[tslot];
    gzt = static_cast<const R*>(d.nz)[tslot];
  }
  double t64[4] = {0.0, 0.0, 0.0, 0.0};
  if constexpr (G64) Geo64<OP>::target(d.geo64, d.geo64_n, tslot, t64);
  // precision form: sources are shifted into the target box's frame (the FP64 form has no centres)
  double ctx = 0, cty = 0, ctz = 0;
  if (PREC && !d.e_shift) {
    const int gb = d.g_box[g];
    ctx = d.centers[3 * gb + 0];
    cty = d.centers[3 * gb + 1];
    ctz = d.centers[3 * gb + 2];
  }
  // FP32: matrix sums over at most 4 kFoldSteps sources per wave, then added into FP64 accumulators; FP64:
  // the matrix accumulators themselves run across all tiles
  constexpr bool FOLD = sizeof(R) == 4;
  double acc[FOLD ? 4 : 1][4];
  V D[4];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) D[nb] = V{0, 0, 0, 0};
  if constexpr (FOLD) {
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[nb][r] = 0.0;
  }
  // this lane's chunk of a staged row (clamped inside the row for narrow widths: channels past
  // the last are never written out)
  const int col = min((lane % CPR) * EPC, cmax);

  const int nent = se - sb;
  const bool etab = nent <= kEnt;  // uniform
  int tot = 0;
  if (etab) {
    if (wave == 0) {
      int n2[2], loc = 0;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = lane * 2 + u;
        n2[u] = e < nent ? d.e_n[sb + e] : 0;
        loc += n2[u];
      }
      int inc = loc;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
      }
      int run = inc - loc;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = lane * 2 + u;
        e_start[e] = run;
        run += n2[u];
      }
      if (lane == 63) e_start[kEnt] = run;
    }
    for (int e = tid; e < nent; e += kWG) {
      const int s = sb + e;
      e_meta[e][0] = d.e_q[s];
      e_meta[e][1] = d.e_pt[s];
      e_meta[e][2] = d.e_flags[s];
      R ox = 0, oy = 0, oz = 0;
      if (PREC) {
        entry_shift<R>(d, s, ctx, cty, ctz, ox, oy, oz);
      }
      e_shift[e][0] = ox;
      e_shift[e][1] = oy;
      e_shift[e][2] = oz;
    }
    __syncthreads();
    tot = e_start[nent];
  } else {
    for (int s = sb; s < se; ++s) tot += d.e_n[s];
  }

  // row `off` of the tile at stream position tbase -> (entry, local index): through the LDS table
  // from the tile's first entry, or through the global lists from the tile's cursor
  int ce = 0, cs = sb, cj = 0;  // tile start: entry (LDS path); list position and offset (global path)
  auto locate = [&](int tbase, int off, int& s, int& j) SKF_INLINE {
    if (etab) {
      const int pos = tbase + off;
      int e = ce;
      while (e_start[e + 1] <= pos) ++e;
      s = e;
      j = pos - e_start[e];
    } else {
      s = cs;
      j = cj + off;
      while (j >= d.e_n[s]) {
        j -= d.e_n[s];
        ++s;
      }
    }
  };
  // move the cursors to the next tile (TR positions on)
  auto advance = [&](int next) SKF_INLINE {
    if (etab) {
      while (ce < nent && e_start[ce + 1] <= next) ++ce;
    } else {
      cj += TR;
      while (cs < se && cj >= d.e_n[cs]) {
        cj -= d.e_n[cs];
        ++cs;
      }
    }
  };
  // Staging of the tile at stream position tbase (nsrc sources). Every LDS read that locates
  // rows (entry table walks) happens before the tile's LDS-direct loads are issued: a later LDS
  // read the compiler cannot tell apart from the loads in flight would wait for them.
  // Wave w stages rows RW w .. RW w + RW - 1: lane l < RW locates row RW w + l once (one walk of
  // the entry table per wave) and stages its per-source data; load instruction ii of the wave
  // fills rows RW w + RPI ii .. + RPI - 1, lane l loading chunk l mod CPR of row
  // RW w + RPI ii + l / CPR at the charge offset that lane RPI ii + l / CPR located (a shuffle).
  const int sr = wave * RW + (lane % RW);
  const bool meta_thread = lane < RW;
  R rp[PS], rsh[3];
  double rg[4] = {0.0, 0.0, 0.0, 0.0};
  int rself = -1;
  auto stage_issue = [&](R* qdst, int tbase, int nsrc) SKF_INLINE {
    int s, j;
    locate(tbase, sr < nsrc ? sr : 0, s, j);
    const int qrow = (etab ? e_meta[s][0] : d.e_q[s]) + j;
    const R* src[4];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
      const int qr = __shfl(qrow, ii * RPI + lane / CPR, 64);
      src[ii] = qv + (long long)qr * d.width + c0 + col;
    }
    int sl = 0, fl = 0;
    R ox = 0, oy = 0, oz = 0;
    const bool have = meta_thread && sr < nsrc;
    if (have) {  // row sr: the entry walk above
      if (etab) {
        sl = e_meta[s][1] + j;
        fl = e_meta[s][2];
        ox = e_shift[s][0];
        oy = e_shift[s][1];
        oz = e_shift[s][2];
      } else {
        sl = d.e_pt[s] + j;
        fl = d.e_flags[s];
        if (PREC) {
          entry_shift<R>(d, s, ctx, cty, ctz, ox, oy, oz);
        }
      }
      rself = (fl & SKF_SRC_SELF) ? j : -1;
    } else {
      rself = -1;
    }
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
      __builtin_amdgcn_global_load_lds((global_void_ptr)src[ii], (lds_void_ptr)(qdst + (wave * RW + ii * RPI) * kMB),
                                       16, 0, 0);
    // raw loads only: their values are used after the matrix work (meta_store), so no wait on
    // them (and on the loads issued before them) is needed here
    if (meta_thread) {
#pragma unroll
      for (int c = 0; c < PS; ++c) rp[c] = R(0);
      if (have) {
        rp[0] = px[sl];
        rp[1] = py[sl];
        if (DIM == 3) rp[2] = pz[sl];
        if (NG) {
          rp[3] = static_cast<const R*>(d.nx)[sl];
          rp[4] = static_cast<const R*>(d.ny)[sl];
          rp[5] = static_cast<const R*>(d.nz)[sl];
        }
        rp[SG] = (fl & SKF_SRC_NEG) ? -KSCALE : KSCALE;
        rsh[0] = ox;
        rsh[1] = oy;
        rsh[2] = oz;
        if constexpr (G64) Geo64<OP>::source(d.geo64, d.geo64_n, sl, rg);
      } else {
        rp[0] = rp[1] = rp[2] = R(1e18);  // far away with factor 0: contributes exactly zero
        rsh[0] = rsh[1] = rsh[2] = R(0);
        rg[0] = rg[1] = rg[2] = rg[3] = 0.0;
      }
    }
  };
  auto meta_store = [&](R* __restrict__ pdst, int* __restrict__ sdst, double* __restrict__ gdst) SKF_INLINE {
    if (!meta_thread) return;
    R row[PS];
    row[0] = rp[0] - rsh[0];
    row[1] = rp[1] - rsh[1];
    row[2] = rp[2] - rsh[2];
#pragma unroll
    for (int c = 3; c < PS; ++c) row[c] = rp[c];
    // self index: exact in R (local indices stay far below 2^24)
    if constexpr (SELFP >= 0) row[SELFP] = R(rself);
#pragma unroll
    for (int c = 0; c < PS; ++c) pdst[sr * PS + c] = row[c];
    if constexpr (G64) {
#pragma unroll
      for (int c = 0; c < 4; ++c) gdst[sr * 4 + c] = rg[c];
    }
    if constexpr (SELFP < 0) sdst[sr] = rself;
  };
  // FP32: add the matrix sums into the FP64 accumulators and restart them
  auto fold = [&]() SKF_INLINE {
    if constexpr (FOLD) {
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[nb][r] += double(D[nb][r]);
        D[nb] = V{0, 0, 0, 0};
      }
    }
  };
  int kacc = 0;  // matrix steps summed in D since the last fold (FP32)
  // matrix work of one staged tile
  // (restrict parameters: the reads carry alias scopes, so they need not wait for the loads in flight)
  // (qcur2: channels 32 .. 63 of the tile for the second FP64 read, a pointer of its own so that
  // this read too carries the scope that tells it apart from the loads in flight)
  // (SP: source parts of the item, a template constant of the whole tile loop so that the matrix
  // accumulators keep their registers from tile to tile)
  auto compute = [&](auto spc, const R* __restrict__ qcur, const R* __restrict__ qcur2,
                     const R* __restrict__ pcur, const int* __restrict__ scur, const double* __restrict__ gcur,
                     int nsrc) SKF_INLINE {
    constexpr int SP = decltype(spc)::value;
    if (!wave_busy) return;
    const int ksteps = (nsrc + 3) >> 2;
    // this wave's steps kk = spart, spart + SP, ... (uniform)
    const int nk = ksteps > spart ? (ksteps - spart + SP - 1) / SP : 0;
    if constexpr (FOLD) {
      if (kacc + nk > kFoldSteps) {  // at most kFoldSteps steps per FP32 sum
        fold();
        kacc = 0;
      }
      kacc += nk;
    }
    const int k = lane >> 4;
    const R* prow = pcur + (spart * 4 + k) * PS;
    const R* qrow = qcur + (spart * 4 + k) * kMB + EPC * (lane & 15);
    const R* qrow2 = qcur2 + (spart * 4 + k) * kMB + EPC * (lane & 15);
    const int* srow = scur + (spart * 4 + k);
    const double* grow = gcur + (spart * 4 + k) * 4;
    // steps in groups of G, unrolled, at offsets that are constants: a group's extra steps read
    // rows past the tile's sources (factor 0) and never rows past the tile
    {
      constexpr int G = (TR / 4) / SP < 4 ? (TR / 4) / SP : 4;
      for (int i = 0; i < nk; i += G) {
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int o = (i + u) * SP * 4;  // row offset of this step
          const R* row = prow + o * PS;
          const R dx = xt - row[0], dy = yt - row[1], dz = (DIM == 3) ? zt - row[2] : R(0);
          R kv;
          if constexpr (L2F) {
            kv = __builtin_amdgcn_logf(fma(dy, dy, fma(dx, dx, Mth<R>::clamp_cd)));
          } else {
            R rn = 0;
            if constexpr (G64) {
              rn = R(Geo64<OP>::rn(t64, grow + o * 4));
            } else {
              if (OpT<OP>::geom == 1) rn = dx * row[3] + dy * row[4] + dz * row[5];
              if (OpT<OP>::geom == 2) rn = dx * gxt + dy * gyt + dz * gzt;
            }
            kv = kernel_value<R, OP>(dx, dy, dz, rn, R(0)).re;
          }
          R a = kv * row[SG];
          if constexpr (SELFP >= 0) {
            if (row[SELFP] == tlocr) a = R(0);
          } else {
            if (srow[o] == tloc) a = R(0);
          }
#pragma unroll
          for (int rd = 0; rd < 4 / EPC; ++rd) {
            const QV v = *reinterpret_cast<const QV*>((rd == 0 ? qrow : qrow2) + o * kMB);
#pragma unroll
            for (int e = 0; e < EPC; ++e) D[rd * EPC + e] = MF::mac(a, v[e], D[rd * EPC + e]);
          }
        }
      }
    }
  };
  // one tile: prefetch the next into one buffer while computing the other. The two buffer
  // orders stay separate code (distinct markers): merged, the LDS reads would go through a
  // pointer that may be either buffer and would wait for the loads in flight.
  auto step = [&](auto spc, auto bufc, int base) SKF_INLINE {
    constexpr bool A = decltype(bufc)::value == 0;
    R* qc = A ? qa : qb;
    R* pc = A ? pa : pb;
    int* sc = A ? sa : sbf;
    double* gc = A ? g64a : g64b;
    R* qn = A ? qb : qa;
    R* pn = A ? pb : pa;
    int* sn = A ? sbf : sa;
    double* gn = A ? g64b : g64a;
    const int nsrc = min(TR, tot - base);
    const bool more = base + TR < tot;
    int nnext = 0;
    if (more) {
      advance(base + TR);
      nnext = min(TR, tot - base - TR);
      stage_issue(qn, base + TR, nnext);
    }
    // keep the staging ahead of the matrix work (the compiler may otherwise sink it below)
    if constexpr (A)
      asm volatile("; tile in buffer a" ::: "memory");
    else
      asm volatile("; tile in buffer b" ::: "memory");
    compute(spc, qc, qc + 16 * EPC, pc, sc, gc, nsrc);
    if (more) meta_store(pn, sn, gn);
    wait_vm_loads();
    __syncthreads();
  };
  if (tot > 0) {
    stage_issue(qa, 0, min(TR, tot));
    meta_store(pa, sa, g64a);
    __syncthreads();
  }
  // on every path: the target data loaded above is complete, so the tile loop's only
  // outstanding loads are the staging of the next tile
  wait_vm_loads();
  __syncthreads();
  auto tiles = [&](auto spc) SKF_INLINE {
    for (int base = 0; base < tot; base += 2 * TR) {
      step(spc, IntC<0>{}, base);
      if (base + TR < tot) step(spc, IntC<1>{}, base + TR);
    }
  };
  if (SPL == 1)
    tiles(IntC<1>{});
  else if (SPL == 2)
    tiles(IntC<2>{});
  else
    tiles(IntC<4>{});
  // The wave's 16 x 64 results go through LDS (every wave has passed the last barrier, so the
  // tiles are no longer read; each wave takes a region of its own), so that each store or atomic
  // instruction covers one target's 64 consecutive channels.
  auto region = [&](int w) SKF_INLINE {
    return sizeof(R) == 4 ? qa + w * 16 * kMB : (w < 2 ? qa : qb) + (w & 1) * 16 * kMB;
  };
  auto to_lds = [&]() SKF_INLINE {
    fold();
    R* tr = region(wave);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      R* trow = tr + MF::row(lane, r) * kMB + EPC * (lane & 15);
#pragma unroll
      for (int rd = 0; rd < 4 / EPC; ++rd) {
        QV v;
#pragma unroll
        for (int e = 0; e < EPC; ++e) v[e] = FOLD ? R(acc[FOLD ? rd * EPC + e : 0][r]) : R(D[rd * EPC + e][r]);
        *reinterpret_cast<QV*>(trow + rd * 16 * EPC) = v;
      }
    }
  };
  R* uo = static_cast<R*>(d.u);
  const int tb0 = t0 + tb * 16;
  if (d.store_rows) {
    // each output row belongs to this item alone and holds zero: the source parts of a target
    // block are summed in LDS and the first part's wave stores the rows (no atomics)
    if (wave_busy) to_lds();
    __syncthreads();
    if (!wave_busy || spart != 0) return;
#pragma unroll 4
    for (int t = 0; t < 16; ++t) {
      R v = R(0);
      for (int p = 0; p < SPL; ++p) v += region(p * TB + tb)[t * kMB + lane];
      if (tb0 + t < ng && lane < ncv) uo[(long long)(d.g_out[g] + tb0 + t) * d.width + c0 + lane] = v;
    }
    return;
  }
  if (!wave_busy) return;
  to_lds();
  const R* tr = region(wave);
#pragma unroll 4
  for (int t = 0; t < 16; ++t) {
    const R v = tr[t * kMB + lane];
    if (tb0 + t < ng && lane < ncv) atomic_acc(&uo[(long long)(d.g_out[g] + tb0 + t) * d.width + c0 + lane], v);
  }
}

template <typename R, int OP, bool PREC>
int launch_p2p_mfma(const P2PDev& d, int nitems, hipStream_t st) {
  dim3 grid(nitems, (d.nch + kMB - 1) / kMB);
  // 16-byte LDS-direct loads of charge rows: width a multiple of 16 bytes' worth, aligned base
  if (d.width % (16 / int(sizeof(R))) == 0 && (reinterpret_cast<unsigned long long>(d.q) & 15ull) == 0) {
    hipLaunchKernelGGL((p2p_mfma_dma_kernel<R, OP, PREC>), grid, dim3(kWG), 0, st, d);
    SKF_LAUNCHED();
    return SKF_OK;
  }
  hipLaunchKernelGGL((p2p_mfma_kernel<R, OP, PREC>), grid, dim3(kWG), 0, st, d);
  SKF_LAUNCHED();
  return SKF_OK;
}

// static LDS of one p2p_kernel instance (tile + cross-wave reduction + FP64 double-layer rows of
// the precision form); kept within 64 KiB
template <typename R, int OP, int NCH, bool PREC>
constexpr int p2p_lds_bytes() {
  constexpr bool G64 = PREC && Geo64<OP>::on;
  using S = Stage<OP, NCH, G64>;
  return kWG * S::nv * int(sizeof(R)) + 4 * kTCH * NCH * S::cw * 8 + (G64 ? kWG * 4 * 8 : 8);
}

template <typename R, int OP, int NCH, bool PREC>
int launch_p2p_t(const P2PDev& d, int nitems, hipStream_t st) {
  if constexpr (p2p_lds_bytes<R, OP, NCH, PREC>() <= 65536) {
    dim3 grid(nitems, (d.nch + NCH - 1) / NCH);
    if constexpr (HasFastTrig<R, OP>::value) {
      if (d.fast_trig) {
        hipLaunchKernelGGL((p2p_kernel<R, OP, NCH, PREC, true>), grid, dim3(kWG), 0, st, d);
        return SKF_OK;
      }
    }
    hipLaunchKernelGGL((p2p_kernel<R, OP, NCH, PREC>), grid, dim3(kWG), 0, st, d);
    return SKF_OK;
  } else {
    return set_err(SKF_ERR_UNSUPPORTED, "channel block too large for LDS");
  }
}

template <typename R, bool PREC, int NCH>
int launch_p2p_op(int op, const P2PDev& d, int nitems, hipStream_t st) {
  int rc;
  switch (op) {
    case SKF_LAPLACE_2D: rc = launch_p2p_t<R, SKF_LAPLACE_2D, NCH, PREC>(d, nitems, st); break;
    case SKF_LAPLACE_3D: rc = launch_p2p_t<R, SKF_LAPLACE_3D, NCH, PREC>(d, nitems, st); break;
    case SKF_HELMHOLTZ_2D: rc = launch_p2p_t<R, SKF_HELMHOLTZ_2D, NCH, PREC>(d, nitems, st); break;
    case SKF_HELMHOLTZ_3D: rc = launch_p2p_t<R, SKF_HELMHOLTZ_3D, NCH, PREC>(d, nitems, st); break;
    case SKF_LAPLACE_DL_3D: rc = launch_p2p_t<R, SKF_LAPLACE_DL_3D, NCH, PREC>(d, nitems, st); break;
    case SKF_LAPLACE_ADL_3D: rc = launch_p2p_t<R, SKF_LAPLACE_ADL_3D, NCH, PREC>(d, nitems, st); break;
    case SKF_HELMHOLTZ_DL_3D: rc = launch_p2p_t<R, SKF_HELMHOLTZ_DL_3D, NCH, PREC>(d, nitems, st); break;
    case SKF_HELMHOLTZ_ADL_3D: rc = launch_p2p_t<R, SKF_HELMHOLTZ_ADL_3D, NCH, PREC>(d, nitems, st); break;
    default: return set_err(SKF_ERR_ARG, "unknown kernel id %d", op);
  }
  if (rc) return rc;
  SKF_LAUNCHED();
  return SKF_OK;
}

// channels evaluated per kernel value: the widest block whose LDS fits (FP64: 8 real / 4 complex,
// FP32: 16 real / 8 complex); multi-RHS real kernels (>= mfma_min channels) go to the matrix-core kernel
template <typename R, bool PREC>
int launch_p2p_prec(int op, const P2PDev& d, int nitems, hipStream_t st, int mfma_min) {
  const bool cx = (op == SKF_HELMHOLTZ_2D || op == SKF_HELMHOLTZ_3D || op == SKF_HELMHOLTZ_DL_3D ||
                   op == SKF_HELMHOLTZ_ADL_3D);
  if (!cx && mfma_min > 0 && d.nch >= mfma_min) {
    switch (op) {
      case SKF_LAPLACE_2D: return launch_p2p_mfma<R, SKF_LAPLACE_2D, PREC>(d, nitems, st);
      case SKF_LAPLACE_3D: return launch_p2p_mfma<R, SKF_LAPLACE_3D, PREC>(d, nitems, st);
      case SKF_LAPLACE_DL_3D: return launch_p2p_mfma<R, SKF_LAPLACE_DL_3D, PREC>(d, nitems, st);
      case SKF_LAPLACE_ADL_3D: return launch_p2p_mfma<R, SKF_LAPLACE_ADL_3D, PREC>(d, nitems, st);
      default: break;
    }
  }
  if (d.nch == 1) return launch_p2p_op<R, PREC, 1>(op, d, nitems, st);
  if (d.nch <= 2) return launch_p2p_op<R, PREC, 2>(op, d, nitems, st);
  if (d.nch <= 4 || (sizeof(R) == 8 && cx)) return launch_p2p_op<R, PREC, 4>(op, d, nitems, st);
  if (d.nch <= 8 || sizeof(R) == 8 || cx) return launch_p2p_op<R, PREC, 8>(op, d, nitems, st);
  return launch_p2p_op<R, PREC, 16>(op, d, nitems, st);
}

template <typename R, bool PREC, int UNR>
int launch_p2p_sym_u(int op, const P2PDev& d, int nitems, hipStream_t st) {
  const dim3 grid(nitems), block(kWG);
  switch (op) {
    case SKF_LAPLACE_2D:
      hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_LAPLACE_2D, PREC, UNR>), grid, block, 0, st, d);
      break;
    case SKF_LAPLACE_3D:
      hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_LAPLACE_3D, PREC, UNR>), grid, block, 0, st, d);
      break;
    case SKF_HELMHOLTZ_2D:
      hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_HELMHOLTZ_2D, PREC, UNR>), grid, block, 0, st, d);
      break;
    case SKF_HELMHOLTZ_3D:
      if (HasFastTrig<R, SKF_HELMHOLTZ_3D>::value && d.fast_trig)
        hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_HELMHOLTZ_3D, PREC, UNR, HasFastTrig<R, SKF_HELMHOLTZ_3D>::value>),
                           grid, block, 0, st, d);
      else
        hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_HELMHOLTZ_3D, PREC, UNR>), grid, block, 0, st, d);
      break;
    case SKF_LAPLACE_DL_3D:
      hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_LAPLACE_DL_3D, PREC, UNR>), grid, block, 0, st, d);
      break;
    case SKF_LAPLACE_ADL_3D:
      hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_LAPLACE_ADL_3D, PREC, UNR>), grid, block, 0, st, d);
      break;
    case SKF_HELMHOLTZ_DL_3D:
      if (d.fast_trig)
        hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_HELMHOLTZ_DL_3D, PREC, UNR, true>), grid, block, 0, st, d);
      else
        hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_HELMHOLTZ_DL_3D, PREC, UNR>), grid, block, 0, st, d);
      break;
    case SKF_HELMHOLTZ_ADL_3D:
      if (d.fast_trig)
        hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_HELMHOLTZ_ADL_3D, PREC, UNR, true>), grid, block, 0, st, d);
      else
        hipLaunchKernelGGL((p2p_sym_kernel<R, SKF_HELMHOLTZ_ADL_3D, PREC, UNR>), grid, block, 0, st, d);
      break;
    default: return set_err(SKF_ERR_ARG, "kernel %d has no symmetric form", op);
  }
  SKF_LAUNCHED();
  return SKF_OK;
}

// unroll: sources evaluated per step of the symmetric kernel's inner loop (1, 2 or 4; 0 = 2)
template <typename R, bool PREC>
int launch_p2p_sym(int op, const P2PDev& d, int nitems, hipStream_t st, int unroll) {
  if (nitems <= 0) return SKF_OK;
  if (unroll == 1) return launch_p2p_sym_u<R, PREC, 1>(op, d, nitems, st);
  if (unroll == 4) return launch_p2p_sym_u<R, PREC, 4>(op, d, nitems, st);
  if (unroll == 0 || unroll == 2) return launch_p2p_sym_u<R, PREC, 2>(op, d, nitems, st);
  return set_err(SKF_ERR_ARG, "symmetric unroll must be 0, 1, 2 or 4");
}

// kernels whose pairs of lists share one evaluation per point pair (p2p_sym_kernel)
bool op_symmetric(int op) {
  return op == SKF_LAPLACE_2D || op == SKF_LAPLACE_3D || op == SKF_HELMHOLTZ_2D || op == SKF_HELMHOLTZ_3D ||
         op == SKF_LAPLACE_DL_3D || op == SKF_LAPLACE_ADL_3D || op == SKF_HELMHOLTZ_DL_3D || op == SKF_HELMHOLTZ_ADL_3D;
}

int check_lists(const skf_p2p_lists* L) {
  if (L->ngroups < 0 || L->nentries < 0 || L->nitems < 0) return set_err(SKF_ERR_ARG, "bad split list sizes");
  if (L->nitems > 0 && (!L->items || !L->g_pt || !L->g_n || !L->g_out || !L->g_box || !L->e_pt || !L->e_n ||
                        !L->e_q || !L->e_box || !L->e_flags))
    return set_err(SKF_ERR_ARG, "missing split list arrays");
  return SKF_OK;
}

void use_lists(P2PDev& d, const skf_p2p_lists* L) {
  d.g_pt = L->g_pt; d.g_n = L->g_n; d.g_out = L->g_out; d.g_box = L->g_box;
  d.e_pt = L->e_pt; d.e_n = L->e_n; d.e_q = L->e_q; d.e_box = L->e_box; d.e_flags = L->e_flags;
  d.items = reinterpret_cast<const int4*>(L->items);
  d.e_shift = L->e_shift;  // precision form: the lists' own table (or none)
  d.store_rows = 0;         // the split lists' rows have several writers
}

bool op_is_complex(int op) {
  return op == SKF_HELMHOLTZ_2D || op == SKF_HELMHOLTZ_3D || op == SKF_HELMHOLTZ_DL_3D ||
         op == SKF_HELMHOLTZ_ADL_3D;
}
bool op_dim3(int op) { return !(op == SKF_LAPLACE_2D || op == SKF_HELMHOLTZ_2D || op == SKF_USER_2D); }
bool op_user(int op) { return op == SKF_USER_2D || op == SKF_USER_3D; }

// near field of a user-defined operator: p2p_kernel instances compiled at run time (hiprtc,
// skf_jit_compile) for channel blocks of 1, 2, 4 and 8, launched through the module API with
// the same argument block as the library's own instances
int launch_user(const skf_p2p_args* a, const P2PDev& d, int nitems, hipStream_t st) {
  static const int nchb[4] = {1, 2, 4, 8};
  const int k = d.nch == 1 ? 0 : d.nch <= 2 ? 1 : d.nch <= 4 ? 2 : 3;
  hipFunction_t fn = static_cast<hipFunction_t>(a->user_fn[k]);
  if (!fn) return set_err(SKF_ERR_ARG, "no run-time compiled near-field kernel for blocks of %d channels", nchb[k]);
  const unsigned gy = unsigned((d.nch + nchb[k] - 1) / nchb[k]);
  if (gy > 65535u) return set_err(SKF_ERR_ARG, "too many channels");
  P2PDev dd = d;
  void* args[] = {&dd};
  SKF_HIP(hipModuleLaunchKernel(fn, unsigned(nitems), gy, 1, kWG, 1, 1, 0, st, args, nullptr));
  return SKF_OK;
}
bool op_geom(int op) { return op >= SKF_LAPLACE_DL_3D && op <= SKF_HELMHOLTZ_ADL_3D; }

int run_p2p(const skf_p2p_args* a, const void* q, void* u, hipStream_t st) {
  if (a->nitems == 0) return SKF_OK;
  if (a->width <= 0) return set_err(SKF_ERR_ARG, "width must be positive");
  const bool cx = op_is_complex(a->kernel);
  if (cx && (a->width % 2)) return set_err(SKF_ERR_ARG, "complex kernel needs an even width");
  if (!a->px || !a->py || (op_dim3(a->kernel) && !a->pz)) return set_err(SKF_ERR_ARG, "missing coordinates");
  if (op_geom(a->kernel) && (!a->nx || !a->ny || !a->nz)) return set_err(SKF_ERR_ARG, "missing normals");
  if (a->precision == 1 && !a->centers) return set_err(SKF_ERR_ARG, "precision form needs box centres");
  if (a->precision == 1 && op_geom(a->kernel) && (!a->geo64 || a->geo64_n <= 0))
    return set_err(SKF_ERR_ARG, "precision form of a double layer needs the FP64 geometry rows (geo64)");
  P2PDev d;
  d.px = a->px; d.py = a->py; d.pz = a->pz;
  d.nx = a->nx; d.ny = a->ny; d.nz = a->nz;
  d.centers = a->centers;
  d.q = q; d.u = u;
  d.g_pt = a->g_pt; d.g_n = a->g_n; d.g_out = a->g_out; d.g_box = a->g_box;
  d.e_pt = a->e_pt; d.e_n = a->e_n; d.e_q = a->e_q; d.e_box = a->e_box; d.e_flags = a->e_flags;
  d.items = reinterpret_cast<const int4*>(a->items);
  d.width = a->width;
  d.nch = cx ? a->width / 2 : a->width;
  d.kappa = a->kappa;
  d.urev = nullptr;
  d.geo64 = a->geo64;
  d.geo64_n = a->geo64_n;
  d.e_shift = a->e_shift;
  d.store_rows = a->store_rows;
  // the 3D Helmholtz kernels take the branch-free reduced sincos (FP32: hardware sine/cosine after
  // the reduction; FP64: polynomials) when the host bounds every argument kappa r (kappa times
  // the largest pair distance) within its range
  const double trig_max = a->precision == 1 ? double(kTrigDirectMax) : double(kTrigFastMax);
  d.fast_trig = (a->max_distance > 0.0 && std::fabs(a->kappa) * a->max_distance <= trig_max) ? 1 : 0;
  const int mfma_min = a->mfma_min_channels == 0 ? 16 : a->mfma_min_channels;
  if (a->precision != 0 && a->precision != 1) return set_err(SKF_ERR_ARG, "precision must be 0 or 1");
  if (op_user(a->kernel)) return launch_user(a, d, a->nitems, st);
  const bool mfma = !cx && mfma_min > 0 && d.nch >= mfma_min;
  const bool sym = a->one && a->sym && d.nch == 1 && op_symmetric(a->kernel) && !mfma &&
                   (a->precision == 0 || (a->urev && a->urev_len > 0));
  if (!sym) {
    if (a->precision == 0) return launch_p2p_prec<double, false>(a->kernel, d, a->nitems, st, mfma_min);
    return launch_p2p_prec<float, true>(a->kernel, d, a->nitems, st, mfma_min);
  }
  int rc;
  if ((rc = check_lists(a->one)) || (rc = check_lists(a->sym))) return rc;
  P2PDev d1 = d, d2 = d;
  use_lists(d1, a->one);
  use_lists(d2, a->sym);
  if (a->precision == 0) {
    d2.urev = static_cast<double*>(u);
    if (a->one->nitems > 0 && (rc = launch_p2p_prec<double, false>(a->kernel, d1, a->one->nitems, st, mfma_min)))
      return rc;
    return launch_p2p_sym<double, false>(a->kernel, d2, a->sym->nitems, st, a->sym_unroll);
  }
  d2.urev = a->urev;
  if ((rc = skf_fill_zero(a->urev, a->urev_len, 0, st))) return rc;
  if (a->one->nitems > 0 && (rc = launch_p2p_prec<float, true>(a->kernel, d1, a->one->nitems, st, mfma_min)))
    return rc;
  if ((rc = launch_p2p_sym<float, true>(a->kernel, d2, a->sym->nitems, st, a->sym_unroll))) return rc;
  const long long n = a->urev_len;
  hipLaunchKernelGGL(fold_rev_kernel<float>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, static_cast<float*>(u),
                     a->urev, n);
  SKF_LAUNCHED();
  return SKF_OK;
}

// ---------------------------------------------------- interpolation kernels
template <typename R>
__global__ void fill_zero_kernel(R* p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = R(0);
}

// zero fill with 16-byte stores (16-byte aligned base), scalars past the last whole vector
template <typename R>
__global__ void fill_zero16_kernel(R* p, long long n) {
  typedef unsigned int U4 __attribute__((ext_vector_type(4)));
  constexpr long long E = 16 / sizeof(R);
  const long long nv = n / E;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  U4* pv = reinterpret_cast<U4*>(p);
  for (long long i = id; i < nv; i += stride) pv[i] = U4{0u, 0u, 0u, 0u};
  for (long long i = nv * E + id; i < n; i += stride) p[i] = R(0);
}

// dst[ddst[i]*w + c] = src[dsrc[i]*w + c]  (or dst[ddst] with user/slot maps)
// Row gathers dst[idst[i]][:] = src[isrc[i]][:] of w scalars, V scalars per thread (V divides
// w, 16-byte accesses where the alignment allows) with 32-bit index arithmetic when the
// element count allows (one division per V elements instead of a 64-bit division per element).
template <typename R, int V>
struct VecOf {
  typedef R T __attribute__((ext_vector_type(V)));
};
template <typename R>
struct VecOf<R, 1> {
  typedef R T;
};

template <typename R, int V>
__global__ void map_copy_kernel(const R* __restrict__ src, R* __restrict__ dst, const int* __restrict__ isrc,
                                const int* __restrict__ idst, int n, int w) {
  typedef typename VecOf<R, V>::T VT;
  const unsigned wv = unsigned(w / V);
  const unsigned tot = unsigned(n) * wv;
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += stride) {
    const unsigned i = t / wv, c = (t - i * wv) * V;
    *reinterpret_cast<VT*>(dst + (long long)idst[i] * w + c) =
        *reinterpret_cast<const VT*>(src + (long long)isrc[i] * w + c);
  }
}

// element counts beyond 32-bit index arithmetic
template <typename R>
__global__ void map_copy_kernel64(const R* __restrict__ src, R* __restrict__ dst, const int* __restrict__ isrc,
                                  const int* __restrict__ idst, int n, int w) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long tot = (long long)n * w;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; t < tot; t += stride) {
    int i = int(t / w), c = int(t - (long long)i * w);
    dst[(long long)idst[i] * w + c] = src[(long long)isrc[i] * w + c];
  }
}

// downward parent step: uskel[dst] += u[src]; u[dst] += uskel[dst]
template <typename R, int V>
__global__ void down_parent_kernel(R* __restrict__ uvec, R* __restrict__ uskel, const int* __restrict__ idst,
                                   const int* __restrict__ isrc, int n, int w) {
  typedef typename VecOf<R, V>::T VT;
  const unsigned wv = unsigned(w / V);
  const unsigned tot = unsigned(n) * wv;
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += stride) {
    const unsigned i = t / wv, c = (t - i * wv) * V;
    const long long dd = (long long)idst[i] * w + c;
    VT* pus = reinterpret_cast<VT*>(uskel + dd);
    VT* puv = reinterpret_cast<VT*>(uvec + dd);
    const VT us = *pus + *reinterpret_cast<const VT*>(uvec + (long long)isrc[i] * w + c);
    *pus = us;
    *puv = *puv + us;
  }
}

template <typename R>
__global__ void down_parent_kernel64(R* __restrict__ uvec, R* __restrict__ uskel, const int* __restrict__ idst,
                                     const int* __restrict__ isrc, int n, int w) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long tot = (long long)n * w;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; t < tot; t += stride) {
    int i = int(t / w), c = int(t - (long long)i * w);
    long long dd = (long long)idst[i] * w + c;
    R us = uskel[dd] + uvec[(long long)isrc[i] * w + c];
    uskel[dd] = us;
    uvec[dd] += us;
  }
}

// workgroups of a grid-stride launch of 256 threads (one pass over up to 2^28 elements)
constexpr long long kMaxGrid = 1 << 20;
int grid_for(long long n);

// scalars per thread for rows of w scalars at these addresses: 16-byte vectors when w and the
// base addresses allow, else 8-byte (FP32 pairs), else 1; 0 when 32-bit indexing would overflow
template <typename R>
int copy_width(int n, int w, const void* a, const void* b, const void* c = nullptr) {
  const unsigned long long al = (unsigned long long)a | (unsigned long long)b | (unsigned long long)c;
  int v = 1;
  if (w % int(16 / sizeof(R)) == 0 && al % 16 == 0) v = int(16 / sizeof(R));
  else if (sizeof(R) == 4 && w % 2 == 0 && al % 8 == 0) v = 2;
  const unsigned long long tot = (unsigned long long)n * (unsigned long long)(w / v);
  if (tot + (unsigned long long)kMaxGrid * 256ull >= (1ull << 32)) return 0;
  return v;
}

template <typename R>
int launch_map_copy(const R* src, R* dst, const int* isrc, const int* idst, int n, int w, hipStream_t st) {
  const int v = copy_width<R>(n, w, src, dst);
  const dim3 g(grid_for((long long)n * w / (v > 0 ? v : 1))), blk(256);
  if (v == 4) hipLaunchKernelGGL((map_copy_kernel<R, 4>), g, blk, 0, st, src, dst, isrc, idst, n, w);
  else if (v == 2) hipLaunchKernelGGL((map_copy_kernel<R, 2>), g, blk, 0, st, src, dst, isrc, idst, n, w);
  else if (v == 1) hipLaunchKernelGGL((map_copy_kernel<R, 1>), g, blk, 0, st, src, dst, isrc, idst, n, w);
  else hipLaunchKernelGGL((map_copy_kernel64<R>), g, blk, 0, st, src, dst, isrc, idst, n, w);
  SKF_LAUNCHED();
  return SKF_OK;
}

template <typename R>
int launch_down_parent(R* uv, R* us, const int* idst, const int* isrc, int n, int w, hipStream_t st) {
  const int v = copy_width<R>(n, w, uv, us);
  const dim3 g(grid_for((long long)n * w / (v > 0 ? v : 1))), blk(256);
  if (v == 4) hipLaunchKernelGGL((down_parent_kernel<R, 4>), g, blk, 0, st, uv, us, idst, isrc, n, w);
  else if (v == 2) hipLaunchKernelGGL((down_parent_kernel<R, 2>), g, blk, 0, st, uv, us, idst, isrc, n, w);
  else if (v == 1) hipLaunchKernelGGL((down_parent_kernel<R, 1>), g, blk, 0, st, uv, us, idst, isrc, n, w);
  else hipLaunchKernelGGL((down_parent_kernel64<R>), g, blk, 0, st, uv, us, idst, isrc, n, w);
  SKF_LAUNCHED();
  return SKF_OK;
}

constexpr int kRowsPerWave = 4;
constexpr int kRowsPerItem = 4 * kRowsPerWave;
constexpr int kColsPerItem = 256;
constexpr int kIB = 4;  // channels per pass in the interpolation kernels

template <typename R, bool CX>
__device__ inline void cmac(R& ar, R& ai, R tr, R ti, R xr, R xi) {
  if (CX) {
    ar += tr * xr - ti * xi;
    ai += tr * xi + ti * xr;
  } else {
    ar += tr * xr;
  }
}

// Upward: qskel[off+j] = q[off+j] + sum_i T[j,i] q[off+r+i] for the item's rows j.
// One wave per row: lanes stride the row of T (coalesced), a wave reduction finishes it.
template <typename R, bool CX>
__global__ __launch_bounds__(256) void up_gemv_kernel(const R* __restrict__ T, const int* __restrict__ box_off,
                                                      const int* __restrict__ box_bs, const int* __restrict__ box_rank,
                                                      const long long* __restrict__ box_toff,
                                                      const int2* __restrict__ items, const R* __restrict__ qvec,
                                                      R* __restrict__ qskel, int w) {
  constexpr int CWd = CX ? 2 : 1;
  const int2 it = items[blockIdx.x];
  const int b = it.x;
  const int off = box_off[b], bs = box_bs[b], r = box_rank[b];
  const long long toff = box_toff[b];
  const int nred = bs - r;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nchan = w / CWd;
  for (int rr = 0; rr < kRowsPerWave; ++rr) {
    const int j = it.y + wave * kRowsPerWave + rr;
    if (j >= r) break;
    const R* Trow = T + (toff + (long long)j * nred) * CWd;
    for (int c0 = 0; c0 < nchan; c0 += kIB) {
      R ar[kIB], ai[kIB];
#pragma unroll
      for (int c = 0; c < kIB; ++c) ar[c] = ai[c] = 0;
      for (int i = lane; i < nred; i += 64) {
        const R tr = Trow[(long long)i * CWd];
        const R ti = CX ? Trow[(long long)i * CWd + 1] : R(0);
        const R* x = qvec + (long long)(off + r + i) * w;
#pragma unroll
        for (int c = 0; c < kIB; ++c) {
          if (c0 + c < nchan) {
            const R xr = x[(c0 + c) * CWd];
            const R xi = CX ? x[(c0 + c) * CWd + 1] : R(0);
            cmac<R, CX>(ar[c], ai[c], tr, ti, xr, xi);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < kIB; ++c) {
        ar[c] = wave_sum(ar[c]);
        if (CX) ai[c] = wave_sum(ai[c]);
      }
      if (lane == 0) {
        const R* x = qvec + (long long)(off + j) * w;
        R* y = qskel + (long long)(off + j) * w;
#pragma unroll
        for (int c = 0; c < kIB; ++c) {
          if (c0 + c < nchan) {
            y[(c0 + c) * CWd] = x[(c0 + c) * CWd] + ar[c];
            if (CX) y[(c0 + c) * CWd + 1] = x[(c0 + c) * CWd + 1] + ai[c];
          }
        }
      }
    }
  }
}

// One channel (one real or complex right-hand side): each wave takes its 4 rows of the item
// together, so every load of the shared vector slice serves 4 rows and 5 loads are in flight per
// step; one wave reduction per row at the end.
template <typename R, bool CX, int UNR = 2>
__global__ __launch_bounds__(256) void up_gemv1_kernel(const R* __restrict__ T, const int* __restrict__ box_off,
                                                       const int* __restrict__ box_bs, const int* __restrict__ box_rank,
                                                       const long long* __restrict__ box_toff,
                                                       const int2* __restrict__ items, const R* __restrict__ qvec,
                                                       R* __restrict__ qskel, int w) {
  constexpr int CWd = CX ? 2 : 1;
  const int2 it = items[blockIdx.x];
  const int b = it.x;
  const int off = box_off[b], bs = box_bs[b], r = box_rank[b];
  const long long toff = box_toff[b];
  const int nred = bs - r;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j0 = it.y + wave * kRowsPerWave;  // wave-uniform
  if (j0 >= r) return;
  R ar[kRowsPerWave], ai[kRowsPerWave];
  const R* Trow[kRowsPerWave];
#pragma unroll
  for (int rr = 0; rr < kRowsPerWave; ++rr) {
    ar[rr] = ai[rr] = 0;
    Trow[rr] = T + (toff + (long long)min(j0 + rr, r - 1) * nred) * CWd;
  }
  const R* x = qvec + (long long)(off + r) * w;
#pragma unroll UNR
  for (int i = lane; i < nred; i += 64) {
    const R xr = x[(long long)i * w];
    const R xi = CX ? x[(long long)i * w + 1] : R(0);
#pragma unroll
    for (int rr = 0; rr < kRowsPerWave; ++rr) {
      const R tr = Trow[rr][(long long)i * CWd];
      const R ti = CX ? Trow[rr][(long long)i * CWd + 1] : R(0);
      cmac<R, CX>(ar[rr], ai[rr], tr, ti, xr, xi);
    }
  }
#pragma unroll
  for (int rr = 0; rr < kRowsPerWave; ++rr) {
    ar[rr] = wave_sum(ar[rr]);
    if (CX) ai[rr] = wave_sum(ai[rr]);
  }
  if (lane < kRowsPerWave) {
    // lane rr writes row j0 + rr (static register selection below)
    R vr = ar[0], vi = ai[0];
#pragma unroll
    for (int rr = 1; rr < kRowsPerWave; ++rr)
      if (lane == rr) {
        vr = ar[rr];
        vi = ai[rr];
      }
    const int j = j0 + lane;
    if (j < r) {
      const R* xs = qvec + (long long)(off + j) * w;
      R* y = qskel + (long long)(off + j) * w;
      y[0] = xs[0] + vr;
      if (CX) y[1] = xs[1] + vi;
    }
  }
}

// One channel, downward: thread per column, rows unrolled by RB (RB independent loads in flight);
// the skeleton potentials of a chunk of rows are staged in LDS and read as broadcasts.
template <typename R, bool CX, int RB = 8>
__global__ __launch_bounds__(256) void down_gemv1_kernel(const R* __restrict__ T, const int* __restrict__ box_off,
                                                         const int* __restrict__ box_bs,
                                                         const int* __restrict__ box_rank,
                                                         const long long* __restrict__ box_toff,
                                                         const int2* __restrict__ items, R* __restrict__ uvec,
                                                         const R* __restrict__ uskel, int w) {
  constexpr int CWd = CX ? 2 : 1;
  constexpr int JB = 256;
  __shared__ R us[JB * CWd];
  const int2 it = items[blockIdx.x];
  const int b = it.x;
  const int off = box_off[b], bs = box_bs[b], r = box_rank[b];
  const long long toff = box_toff[b];
  const int nred = bs - r;
  const int i = it.y + threadIdx.x;
  const bool act = i < nred;
  R ar0 = 0, ai0 = 0, ar1 = 0, ai1 = 0;
  const R* tcol = T + (toff + (act ? i : 0)) * CWd;
  for (int j0 = 0; j0 < r; j0 += JB) {
    const int nj = min(JB, r - j0);
    __syncthreads();
    for (int t = threadIdx.x; t < nj; t += blockDim.x) {
      const R* xs = uskel + (long long)(off + j0 + t) * w;
      us[t * CWd] = xs[0];
      if (CX) us[t * CWd + 1] = xs[1];
    }
    __syncthreads();
    if (act) {
      int jj = 0;
      for (; jj + RB <= nj; jj += RB) {
        R tr[RB], ti[RB];
#pragma unroll
        for (int k = 0; k < RB; ++k) {
          const R* tp = tcol + (long long)(j0 + jj + k) * nred * CWd;
          tr[k] = tp[0];
          ti[k] = CX ? tp[1] : R(0);
        }
#pragma unroll
        for (int k = 0; k < RB; k += 2) {
          cmac<R, CX>(ar0, ai0, tr[k], ti[k], us[(jj + k) * CWd], CX ? us[(jj + k) * CWd + 1] : R(0));
          cmac<R, CX>(ar1, ai1, tr[k + 1], ti[k + 1], us[(jj + k + 1) * CWd], CX ? us[(jj + k + 1) * CWd + 1] : R(0));
        }
      }
      for (; jj < nj; ++jj) {
        const R* tp = tcol + (long long)(j0 + jj) * nred * CWd;
        cmac<R, CX>(ar0, ai0, tp[0], CX ? tp[1] : R(0), us[jj * CWd], CX ? us[jj * CWd + 1] : R(0));
      }
    }
  }
  if (act) {
    R* y = uvec + (long long)(off + r + i) * w;
    y[0] += ar0 + ar1;
    if (CX) y[1] += ai0 + ai1;
  }
}

// Downward: u[off+r+i] += sum_j T[j,i] uskel[off+j] for the item's columns i (thread per column).
template <typename R, bool CX>
__global__ __launch_bounds__(256) void down_gemv_kernel(const R* __restrict__ T, const int* __restrict__ box_off,
                                                        const int* __restrict__ box_bs,
                                                        const int* __restrict__ box_rank,
                                                        const long long* __restrict__ box_toff,
                                                        const int2* __restrict__ items, R* __restrict__ uvec,
                                                        const R* __restrict__ uskel, int w) {
  constexpr int CWd = CX ? 2 : 1;
  constexpr int JB = 128;
  __shared__ R us[JB * kIB * CWd];
  const int2 it = items[blockIdx.x];
  const int b = it.x;
  const int off = box_off[b], bs = box_bs[b], r = box_rank[b];
  const long long toff = box_toff[b];
  const int nred = bs - r;
  const int i = it.y + threadIdx.x;
  const int nchan = w / CWd;
  for (int c0 = 0; c0 < nchan; c0 += kIB) {
    R ar[kIB], ai[kIB];
#pragma unroll
    for (int c = 0; c < kIB; ++c) ar[c] = ai[c] = 0;
    for (int j0 = 0; j0 < r; j0 += JB) {
      const int nj = min(JB, r - j0);
      __syncthreads();
      for (int t = threadIdx.x; t < nj * kIB; t += blockDim.x) {
        const int jj = t / kIB, c = t % kIB;
        const R* x = uskel + (long long)(off + j0 + jj) * w;
        us[(jj * kIB + c) * CWd] = (c0 + c < nchan) ? x[(c0 + c) * CWd] : R(0);
        if (CX) us[(jj * kIB + c) * CWd + 1] = (c0 + c < nchan) ? x[(c0 + c) * CWd + 1] : R(0);
      }
      __syncthreads();
      if (i < nred) {
        for (int jj = 0; jj < nj; ++jj) {
          const R* tp = T + (toff + (long long)(j0 + jj) * nred + i) * CWd;
          const R tr = tp[0];
          const R ti = CX ? tp[1] : R(0);
#pragma unroll
          for (int c = 0; c < kIB; ++c)
            cmac<R, CX>(ar[c], ai[c], tr, ti, us[(jj * kIB + c) * CWd], CX ? us[(jj * kIB + c) * CWd + 1] : R(0));
        }
      }
    }
    if (i < nred) {
      R* y = uvec + (long long)(off + r + i) * w;
#pragma unroll
      for (int c = 0; c < kIB; ++c) {
        if (c0 + c < nchan) {
          y[(c0 + c) * CWd] += ar[c];
          if (CX) y[(c0 + c) * CWd + 1] += ai[c];
        }
      }
    }
  }
}

int grid_for(long long n) {
  long long g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > kMaxGrid) g = kMaxGrid;
  return int(g);
}

// Multi-RHS interpolation (>= 16 channels): lanes run over right-hand-side channels, so every
// vector row is read coalesced and each T entry is one wave-uniform load used by 64 lanes.
// Upward: a wave owns 4 rows j of the item and accumulates them together (each x row read once
// per 4 rows).
template <typename R, bool CX>
__global__ __launch_bounds__(256) void up_gemm_kernel(const R* __restrict__ T, const int* __restrict__ box_off,
                                                      const int* __restrict__ box_bs, const int* __restrict__ box_rank,
                                                      const long long* __restrict__ box_toff,
                                                      const int2* __restrict__ items, const R* __restrict__ qvec,
                                                      R* __restrict__ qskel, int w) {
  constexpr int CWd = CX ? 2 : 1;
  const int2 it = items[blockIdx.x];
  const int b = it.x;
  const int off = box_off[b], bs = box_bs[b], r = box_rank[b];
  const long long toff = box_toff[b];
  const int nred = bs - r;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j0 = it.y + wave * kRowsPerWave;
  const int nrow = max(0, min(kRowsPerWave, r - j0));
  if (nrow == 0) return;
  const int nchan = w / CWd;
  for (int c0 = 0; c0 < nchan; c0 += 64) {
    const int c = c0 + lane;
    const bool cv = c < nchan;
    R ar[kRowsPerWave], ai[kRowsPerWave];
#pragma unroll
    for (int k = 0; k < kRowsPerWave; ++k) ar[k] = ai[k] = 0;
    for (int i = 0; i < nred; ++i) {
      const R* x = qvec + (long long)(off + r + i) * w + (long long)c * CWd;
      const R xr = cv ? x[0] : R(0);
      const R xi = (CX && cv) ? x[1] : R(0);
#pragma unroll
      for (int k = 0; k < kRowsPerWave; ++k) {
        if (k < nrow) {
          const R* tp = T + (toff + (long long)(j0 + k) * nred + i) * CWd;
          cmac<R, CX>(ar[k], ai[k], tp[0], CX ? tp[1] : R(0), xr, xi);
        }
      }
    }
    if (cv) {
#pragma unroll
      for (int k = 0; k < kRowsPerWave; ++k) {
        if (k < nrow) {
          const long long o = (long long)(off + j0 + k) * w + (long long)c * CWd;
          qskel[o] = qvec[o] + ar[k];
          if (CX) qskel[o + 1] = qvec[o + 1] + ai[k];
        }
      }
    }
  }
}

// Downward: a wave owns a quarter of the item's columns and handles CG of them at a time.
template <typename R, bool CX>
__global__ __launch_bounds__(256) void down_gemm_kernel(const R* __restrict__ T, const int* __restrict__ box_off,
                                                        const int* __restrict__ box_bs,
                                                        const int* __restrict__ box_rank,
                                                        const long long* __restrict__ box_toff,
                                                        const int2* __restrict__ items, R* __restrict__ uvec,
                                                        const R* __restrict__ uskel, int w) {
  constexpr int CWd = CX ? 2 : 1;
  // columns per group: each skeleton row read feeds CG columns (real FP32: 16, else 8)
  constexpr int CG = (sizeof(R) == 4 && !CX) ? 16 : 8;
  const int2 it = items[blockIdx.x];
  const int b = it.x;
  const int off = box_off[b], bs = box_bs[b], r = box_rank[b];
  const long long toff = box_toff[b];
  const int nred = bs - r;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int cols_per_wave = kColsPerItem / 4;
  const int ibeg = it.y + wave * cols_per_wave;
  const int iend = min(nred, ibeg + cols_per_wave);
  const int nchan = w / CWd;
  for (int c0 = 0; c0 < nchan; c0 += 64) {
    const int c = c0 + lane;
    const bool cv = c < nchan;
    for (int i0 = ibeg; i0 < iend; i0 += CG) {
      const int ncol = min(CG, iend - i0);
      R ar[CG], ai[CG];
#pragma unroll
      for (int k = 0; k < CG; ++k) ar[k] = ai[k] = 0;
      for (int j = 0; j < r; ++j) {
        const R* x = uskel + (long long)(off + j) * w + (long long)c * CWd;
        const R xr = cv ? x[0] : R(0);
        const R xi = (CX && cv) ? x[1] : R(0);
        const R* tp = T + (toff + (long long)j * nred + i0) * CWd;
#pragma unroll
        for (int k = 0; k < CG; ++k)
          if (k < ncol) cmac<R, CX>(ar[k], ai[k], tp[k * CWd], CX ? tp[k * CWd + 1] : R(0), xr, xi);
      }
      if (cv) {
#pragma unroll
        for (int k = 0; k < CG; ++k) {
          if (k < ncol) {
            R* y = uvec + (long long)(off + r + i0 + k) * w + (long long)c * CWd;
            y[0] += ar[k];
            if (CX) y[1] += ai[k];
          }
        }
      }
    }
  }
}

// FP32 interpolation passes with many right-hand sides (real vectors, >= 16 channels) on the
// matrix cores: v_mfma_f32_16x16x4f32 over 16 rows of T (upward: T[j][i], downward: T^T) and 16
// channels; the four waves of a workgroup take the four 16-channel blocks of a 64-channel block
// (grid.y). A and B operands are loaded straight from T and the vectors (4 bytes a lane per
// operand and step; the loads of kInterpSteps steps are issued before their matrix ops, so they
// are in flight together), accumulated in FP32 like
// the VALU kernels. Accumulator r of a lane holds D[4 (lane / 16) + r][lane % 16].
constexpr int kInterpSteps = 8;  // matrix steps whose operand loads are issued together
template <bool UP>
__global__ __launch_bounds__(256) void interp_mfma_kernel(const float* __restrict__ T, const int* __restrict__ box_off,
                                                          const int* __restrict__ box_bs,
                                                          const int* __restrict__ box_rank,
                                                          const long long* __restrict__ box_toff,
                                                          const int2* __restrict__ items, const float* __restrict__ src,
                                                          float* __restrict__ dst, int w) {
  const int2 it = items[blockIdx.x];
  const int b = it.x;
  const int off = box_off[b], bs = box_bs[b], r = box_rank[b];
  const long long toff = box_toff[b];
  const int nred = bs - r;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int m = lane & 15, k = lane >> 4;
  const float* Tb = T + toff;
  if (UP) {
    const int cb = blockIdx.y * 64 + wave * 16;  // first channel of the wave's block
    if (cb >= w) return;
    const int ch = cb + m;
    const bool cv = ch < w;
    // q_skel[j] = q[j] + sum_i T[j][i] q[r + i] for the item's rows j0 .. j0 + 15
    const int j0 = it.y;
    const bool rv = j0 + m < r;
    const float* ta = Tb + (long long)(rv ? j0 + m : j0) * nred;
    const float* xb = src + (long long)(off + r) * w + (cv ? ch : cb);
    v4f D = v4f{0, 0, 0, 0};
    for (int i0 = 0; i0 < nred; i0 += 4 * kInterpSteps) {
      float a[kInterpSteps], x[kInterpSteps];
#pragma unroll
      for (int u = 0; u < kInterpSteps; ++u) {
        const int i = i0 + 4 * u + k;
        const bool iv = i < nred;
        a[u] = (rv && iv) ? ta[i] : 0.0f;
        x[u] = (iv && cv) ? xb[(long long)i * w] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < kInterpSteps; ++u) D = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], x[u], D, 0, 0, 0);
    }
    if (cv) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = j0 + 4 * k + q;
        if (j < r) {
          const long long o = (long long)(off + j) * w + ch;
          dst[o] = src[o] + D[q];
        }
      }
    }
  } else {
    // u[r + i] += sum_j T[j][i] u_skel[j]: wave w takes the item's 16-column blocks w, w + 4, ...
    // with all 64 channels of the launch block as four matrix blocks that share the T operand.
    // Column n of matrix block c is channel c0 + 4 n + c, so a lane's four B operands of a step
    // are four consecutive channels (one 16-byte load; widths are multiples of 4 here) and its
    // results of one row are four consecutive channels too.
    const int ibeg = it.y, iend = min(nred, it.y + kColsPerItem);
    const int c4 = blockIdx.y * 64 + 4 * m;  // this lane's first channel
    const bool cv = c4 < w;
    const float* xb = src + (long long)off * w + (cv ? c4 : 0);
    for (int ib = ibeg + 16 * wave; ib < iend; ib += 64) {
      const int ic = ib + m;
      const bool iv = ic < iend;
      v4f D[4] = {v4f{0, 0, 0, 0}, v4f{0, 0, 0, 0}, v4f{0, 0, 0, 0}, v4f{0, 0, 0, 0}};
      for (int j0 = 0; j0 < r; j0 += 4 * kInterpSteps) {
        float a[kInterpSteps];
        v4f x[kInterpSteps];
#pragma unroll
        for (int u = 0; u < kInterpSteps; ++u) {
          const int j = j0 + 4 * u + k;
          const bool jv = j < r;
          a[u] = (iv && jv) ? Tb[(long long)j * nred + ic] : 0.0f;
          x[u] = (jv && cv) ? *reinterpret_cast<const v4f*>(xb + (long long)j * w) : v4f{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < kInterpSteps; ++u)
#pragma unroll
          for (int c = 0; c < 4; ++c) D[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], x[u][c], D[c], 0, 0, 0);
      }
      if (cv) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = ib + 4 * k + q;
          if (i < iend) {
            v4f* o = reinterpret_cast<v4f*>(dst + (long long)(off + r + i) * w + c4);
            v4f v = *o;
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] += D[c][q];
            *o = v;
          }
        }
      }
    }
  }
}

template <typename R>
int run_up_level(const skf_level_args* L, const void* T, int cx, void* qvec, void* qskel, int w, hipStream_t st) {
  R* qv = static_cast<R*>(qvec);
  R* qs = static_cast<R*>(qskel);
  if (L->ngather > 0) {
    int rc = launch_map_copy<R>(qs, qv, L->gather_src, L->gather_dst, L->ngather, w, st);
    if (rc) return rc;
  }
  if (L->nitems > 0) {
    const int2* items = reinterpret_cast<const int2*>(L->items);
    const int nchan = cx ? w / 2 : w;
    if (nchan >= 16) {
      const dim3 g2(L->nitems, (nchan + 63) / 64);
      if constexpr (sizeof(R) == 4) {
        if (!cx && L->variant != 1) {
          hipLaunchKernelGGL((interp_mfma_kernel<true>), g2, dim3(256), 0, st, static_cast<const float*>(T), L->box_off,
                             L->box_bs, L->box_rank, L->box_toff, items, qv, qs, w);
          SKF_LAUNCHED();
          return SKF_OK;
        }
      }
      if (cx)
        hipLaunchKernelGGL((up_gemm_kernel<R, true>), dim3(L->nitems), dim3(256), 0, st, static_cast<const R*>(T),
                           L->box_off, L->box_bs, L->box_rank, L->box_toff, items, qv, qs, w);
      else
        hipLaunchKernelGGL((up_gemm_kernel<R, false>), dim3(L->nitems), dim3(256), 0, st, static_cast<const R*>(T),
                           L->box_off, L->box_bs, L->box_rank, L->box_toff, items, qv, qs, w);
    } else if (nchan == 1) {
      // variant: vector loads in flight per lane and step (unroll 2 by default, 1, 4)
      const dim3 g(L->nitems), blk(256);
      const R* Tr = static_cast<const R*>(T);
#define SKF_UP1(CXV, U) \
  hipLaunchKernelGGL((up_gemv1_kernel<R, CXV, U>), g, blk, 0, st, Tr, L->box_off, L->box_bs, L->box_rank, L->box_toff, \
                     items, qv, qs, w)
      if (cx) {
        if (L->variant == 1) SKF_UP1(true, 1); else if (L->variant == 2) SKF_UP1(true, 4); else SKF_UP1(true, 2);
      } else {
        if (L->variant == 1) SKF_UP1(false, 1); else if (L->variant == 2) SKF_UP1(false, 4); else SKF_UP1(false, 2);
      }
#undef SKF_UP1
    } else if (cx)
      hipLaunchKernelGGL((up_gemv_kernel<R, true>), dim3(L->nitems), dim3(256), 0, st, static_cast<const R*>(T),
                         L->box_off, L->box_bs, L->box_rank, L->box_toff, items, qv, qs, w);
    else
      hipLaunchKernelGGL((up_gemv_kernel<R, false>), dim3(L->nitems), dim3(256), 0, st, static_cast<const R*>(T),
                         L->box_off, L->box_bs, L->box_rank, L->box_toff, items, qv, qs, w);
    SKF_LAUNCHED();
  }
  return SKF_OK;
}

template <typename R>
int run_down_level(const skf_level_args* L, const void* T, int cx, void* uvec, void* uskel, int w, hipStream_t st) {
  R* uv = static_cast<R*>(uvec);
  R* us = static_cast<R*>(uskel);
  if (L->ngather > 0) {
    int rc = launch_down_parent<R>(uv, us, L->gather_dst, L->gather_src, L->ngather, w, st);
    if (rc) return rc;
  }
  if (L->nitems > 0) {
    const int2* items = reinterpret_cast<const int2*>(L->items);
    const int nchan = cx ? w / 2 : w;
    if (nchan >= 16) {
      if constexpr (sizeof(R) == 4) {
        // 16-byte vector accesses: a width that is a multiple of 4 and aligned vectors
        const bool v4ok = w % 4 == 0 && (reinterpret_cast<unsigned long long>(uv) & 15ull) == 0 &&
                          (reinterpret_cast<unsigned long long>(us) & 15ull) == 0;
        if (!cx && L->variant != 1 && v4ok) {
          hipLaunchKernelGGL((interp_mfma_kernel<false>), dim3(L->nitems, (nchan + 63) / 64), dim3(256), 0, st,
                             static_cast<const float*>(T), L->box_off, L->box_bs, L->box_rank, L->box_toff, items, us,
                             uv, w);
          SKF_LAUNCHED();
          return SKF_OK;
        }
      }
      if (cx)
        hipLaunchKernelGGL((down_gemm_kernel<R, true>), dim3(L->nitems), dim3(256), 0, st, static_cast<const R*>(T),
                           L->box_off, L->box_bs, L->box_rank, L->box_toff, items, uv, us, w);
      else
        hipLaunchKernelGGL((down_gemm_kernel<R, false>), dim3(L->nitems), dim3(256), 0, st, static_cast<const R*>(T),
                           L->box_off, L->box_bs, L->box_rank, L->box_toff, items, uv, us, w);
    } else if (nchan == 1) {
      // variant: column loads in flight per thread and step (8 rows by default, 4, 16)
      const dim3 g(L->nitems), blk(256);
      const R* Tr = static_cast<const R*>(T);
#define SKF_DN1(CXV, B) \
  hipLaunchKernelGGL((down_gemv1_kernel<R, CXV, B>), g, blk, 0, st, Tr, L->box_off, L->box_bs, L->box_rank, \
                     L->box_toff, items, uv, us, w)
      if (cx) {
        if (L->variant == 1) SKF_DN1(true, 4); else if (L->variant == 2) SKF_DN1(true, 16); else SKF_DN1(true, 8);
      } else {
        if (L->variant == 1) SKF_DN1(false, 4); else if (L->variant == 2) SKF_DN1(false, 16); else SKF_DN1(false, 8);
      }
#undef SKF_DN1
    } else if (cx)
      hipLaunchKernelGGL((down_gemv_kernel<R, true>), dim3(L->nitems), dim3(256), 0, st, static_cast<const R*>(T),
                         L->box_off, L->box_bs, L->box_rank, L->box_toff, items, uv, us, w);
    else
      hipLaunchKernelGGL((down_gemv_kernel<R, false>), dim3(L->nitems), dim3(256), 0, st, static_cast<const R*>(T),
                         L->box_off, L->box_bs, L->box_rank, L->box_toff, items, uv, us, w);
    SKF_LAUNCHED();
  }
  return SKF_OK;
}

int check_level(const skf_level_args* L) {
  if (!L) return set_err(SKF_ERR_ARG, "null level");
  if (L->ngather < 0 || L->nitems < 0 || L->nboxes < 0) return set_err(SKF_ERR_ARG, "negative level sizes");
  if (L->variant < 0 || L->variant > 2) return set_err(SKF_ERR_ARG, "level variant must be 0, 1 or 2");
  if (L->ngather > 0 && (!L->gather_dst || !L->gather_src)) return set_err(SKF_ERR_ARG, "missing gather maps");
  if (L->nitems > 0 && (!L->items || !L->box_off || !L->box_bs || !L->box_rank || !L->box_toff))
    return set_err(SKF_ERR_ARG, "missing box tables");
  return SKF_OK;
}

}  // namespace

// ---------------------------------------------------------------- C entries
extern "C" int skf_fill_zero(void* ptr, long long n, int precision, void* stream) {
  if (n <= 0) return SKF_OK;
  if (!ptr) return set_err(SKF_ERR_ARG, "null pointer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t rs = precision == 0 ? sizeof(double) : sizeof(float);
  if ((reinterpret_cast<unsigned long long>(ptr) & 15ull) == 0) {
    const dim3 g(grid_for((n * (long long)rs + 15) / 16));
    if (precision == 0)
      hipLaunchKernelGGL((fill_zero16_kernel<double>), g, dim3(256), 0, st, static_cast<double*>(ptr), n);
    else
      hipLaunchKernelGGL((fill_zero16_kernel<float>), g, dim3(256), 0, st, static_cast<float*>(ptr), n);
  } else if (precision == 0) {
    hipLaunchKernelGGL((fill_zero_kernel<double>), dim3(grid_for(n)), dim3(256), 0, st, static_cast<double*>(ptr), n);
  } else {
    hipLaunchKernelGGL((fill_zero_kernel<float>), dim3(grid_for(n)), dim3(256), 0, st, static_cast<float*>(ptr), n);
  }
  SKF_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_gather_charges(const void* x, void* qvec, const int* leaf_slot, const int* leaf_user, int n,
                                  int width, int precision, void* stream) {
  if (n <= 0) return SKF_OK;
  if (!x || !qvec || !leaf_slot || !leaf_user || width <= 0) return set_err(SKF_ERR_ARG, "bad gather arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (precision == 0)
    return launch_map_copy<double>(static_cast<const double*>(x), static_cast<double*>(qvec), leaf_user, leaf_slot, n,
                                   width, st);
  return launch_map_copy<float>(static_cast<const float*>(x), static_cast<float*>(qvec), leaf_user, leaf_slot, n,
                                width, st);
}

extern "C" int skf_scatter_result(const void* uvec, void* y, const int* leaf_slot, const int* leaf_user, int n,
                                  int width, int precision, void* stream) {
  if (n <= 0) return SKF_OK;
  if (!y || !uvec || !leaf_slot || !leaf_user || width <= 0) return set_err(SKF_ERR_ARG, "bad scatter arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (precision == 0)
    return launch_map_copy<double>(static_cast<const double*>(uvec), static_cast<double*>(y), leaf_slot, leaf_user, n,
                                   width, st);
  return launch_map_copy<float>(static_cast<const float*>(uvec), static_cast<float*>(y), leaf_slot, leaf_user, n,
                                width, st);
}

extern "C" int skf_p2p(const skf_p2p_args* a, void* stream) {
  if (!a) return set_err(SKF_ERR_ARG, "null p2p args");
  if (!a->q || !a->u) return set_err(SKF_ERR_ARG, "null vectors");
  return run_p2p(a, a->q, a->u, static_cast<hipStream_t>(stream));
}

// ---------------------------------------------- near field without allocation
// The FP32 symmetric near field needs an FP64 buffer for its reverse sums (urev, one scalar per
// potential index). skf_p2p_noalloc takes it from this file-scope array instead of from the
// caller: every GPU the library runs on holds its own copy for the life of the process. In host
// code `scratch` names a host-side stand-in, so its device address is looked up once per device
// (hipGetSymbolAddress with that device current) and only that address is used.
static __device__ __align__(256) double scratch[SKF_NOALLOC_MAX_UREV];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(double** out) {
  int dev = -1;
  SKF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_err(SKF_ERR_ARG, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(SKF_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SKF_OK;
}
}  // namespace

extern "C" int skf_p2p_noalloc(const skf_p2p_args* a, void* stream) {
  if (!a) return set_err(SKF_ERR_ARG, "null p2p args");
  if (!a->q || !a->u) return set_err(SKF_ERR_ARG, "null vectors");
  if (a->urev) return set_err(SKF_ERR_ARG, "skf_p2p_noalloc takes no urev buffer (it uses the library's own)");
  if (a->urev_len < 0 || a->urev_len > (long long)SKF_NOALLOC_MAX_UREV)
    return set_err(SKF_ERR_ARG, "urev_len %lld outside [0, %lld]: the library's reverse-sum buffer holds %lld doubles",
                   a->urev_len, (long long)SKF_NOALLOC_MAX_UREV, (long long)SKF_NOALLOC_MAX_UREV);
  skf_p2p_args b = *a;
  if (a->urev_len > 0) {
    int rc = scratch_address(&b.urev);
    if (rc) return rc;
  }
  // run_p2p zeroes the first urev_len scalars with a fill kernel before the symmetric kernel
  // accumulates into them, so nothing left by an earlier call is read
  return run_p2p(&b, a->q, a->u, static_cast<hipStream_t>(stream));
}

extern "C" int skf_up_level(const skf_level_args* L, const void* T, int complex_t, void* qvec, void* qskel,
                            int width, int precision, void* stream) {
  int rc = check_level(L);
  if (rc) return rc;
  if (complex_t && (width % 2)) return set_err(SKF_ERR_ARG, "complex T needs an even width");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return precision == 0 ? run_up_level<double>(L, T, complex_t, qvec, qskel, width, st)
                        : run_up_level<float>(L, T, complex_t, qvec, qskel, width, st);
}

extern "C" int skf_down_level(const skf_level_args* L, const void* T, int complex_t, void* uvec, void* uskel,
                              int width, int precision, void* stream) {
  int rc = check_level(L);
  if (rc) return rc;
  if (complex_t && (width % 2)) return set_err(SKF_ERR_ARG, "complex T needs an even width");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return precision == 0 ? run_down_level<double>(L, T, complex_t, uvec, uskel, width, st)
                        : run_down_level<float>(L, T, complex_t, uvec, uskel, width, st);
}

// ------------------------------------------------------------- whole apply
struct skf_plan {
  skf_plan_desc d;
  std::vector<skf_level_args> up;
  std::vector<skf_level_args> down;
};

extern "C" int skf_plan_create(const skf_plan_desc* d, skf_plan** out) {
  if (!d || !out) return set_err(SKF_ERR_ARG, "null plan arguments");
  if (d->precision != 0 && d->precision != 1) return set_err(SKF_ERR_ARG, "precision must be 0 or 1");
  if (d->width <= 0 || d->nslots < 0 || d->npoints < 0) return set_err(SKF_ERR_ARG, "bad sizes");
  if (!d->qbuf || !d->ubuf) return set_err(SKF_ERR_ARG, "missing work vectors");
  if (d->nup < 0 || d->ndown < 0) return set_err(SKF_ERR_ARG, "bad stage counts");
  for (int i = 0; i < d->nup; ++i) {
    int rc = check_level(&d->up[i]);
    if (rc) return rc;
  }
  for (int i = 0; i < d->ndown; ++i) {
    int rc = check_level(&d->down[i]);
    if (rc) return rc;
  }
  skf_plan* p = new (std::nothrow) skf_plan;
  if (!p) return set_err(SKF_ERR_ARG, "out of host memory");
  p->d = *d;
  p->up.assign(d->up, d->up + d->nup);
  p->down.assign(d->down, d->down + d->ndown);
  p->d.up = p->up.data();
  p->d.down = p->down.data();
  *out = p;
  return SKF_OK;
}

extern "C" int skf_plan_destroy(skf_plan* p) {
  delete p;
  return SKF_OK;
}

extern "C" int skf_apply(skf_plan* p, const void* x, void* y, void* stream) {
  if (!p || !x || !y) return set_err(SKF_ERR_ARG, "null apply arguments");
  const skf_plan_desc& d = p->d;
  const int w = d.width, prec = d.precision;
  const long long half = (long long)d.nslots * w;
  const size_t rs = prec == 0 ? sizeof(double) : sizeof(float);
  char* qb = static_cast<char*>(d.qbuf);
  char* ub = static_cast<char*>(d.ubuf);
  void* qvec = qb;
  void* qskel = qb + half * rs;
  void* uvec = ub;
  void* uskel = ub + half * rs;
  int rc;
  // qbuf needs no clearing: the gathers and upward stages write every charge slot before any
  // kernel reads it (a layout invariant the host tests check); the potentials accumulate
  if ((rc = skf_fill_zero(d.ubuf, 2 * half, prec, stream))) return rc;
  if ((rc = skf_gather_charges(x, qvec, d.leaf_slot, d.leaf_user, d.npoints, w, prec, stream))) return rc;
  for (int i = 0; i < d.nup; ++i)
    if ((rc = skf_up_level(&d.up[i], d.T, d.complex_t, qvec, qskel, w, prec, stream))) return rc;
  skf_p2p_args a = d.p2p;
  a.precision = prec;
  a.width = w;
  if ((rc = run_p2p(&a, d.qbuf, d.ubuf, static_cast<hipStream_t>(stream)))) return rc;
  for (int i = 0; i < d.ndown; ++i)
    if ((rc = skf_down_level(&d.down[i], d.T, d.complex_t, uvec, uskel, w, prec, stream))) return rc;
  return skf_scatter_result(uvec, y, d.leaf_slot, d.leaf_user, d.npoints, w, prec, stream);
}

// ------------------------------------------------------------ dense blocks
namespace {

// u[b, i, :] (+)= sum_j K(xt[b,i], xs[b,j]) q[b,j,:]; one workgroup per (block, 64 targets),
// AoS inputs, no LDS staging (library seam for KernelOperator.torch_apply, not the apply hot path).
template <typename R, int OP>
__global__ __launch_bounds__(64) void blocks_kernel(int m, int n, int ndim, const R* __restrict__ xt,
                                                    const R* __restrict__ xs, const R* __restrict__ gt,
                                                    const R* __restrict__ gs, const R* __restrict__ q,
                                                    R* __restrict__ u, int w, int self_int, R kappa) {
  constexpr int CWd = OpT<OP>::cplx ? 2 : 1;
  const int b = blockIdx.y;
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= m) return;
  const R* xi = xt + ((long long)b * m + i) * ndim;
  const R x0 = xi[0], x1 = xi[1], x2 = ndim == 3 ? xi[2] : R(0);
  R g0 = 0, g1 = 0, g2 = 0;
  if (OpT<OP>::geom == 2) {
    const R* gi = gt + ((long long)b * m + i) * 3;
    g0 = gi[0]; g1 = gi[1]; g2 = gi[2];
  }
  const int nchan = w / CWd;
  for (int c = 0; c < nchan; ++c) {
    double ar = 0, ai = 0;
    for (int j = 0; j < n; ++j) {
      const R* yj = xs + ((long long)b * n + j) * ndim;
      R dx = x0 - yj[0], dy = x1 - yj[1], dz = ndim == 3 ? x2 - yj[2] : R(0);
      R rn = 0;
      if (OpT<OP>::geom == 1) {
        const R* gj = gs + ((long long)b * n + j) * 3;
        rn = dx * gj[0] + dy * gj[1] + dz * gj[2];
      }
      if (OpT<OP>::geom == 2) rn = dx * g0 + dy * g1 + dz * g2;
      Cplx<R> kv = kernel_value<R, OP>(dx, dy, dz, rn, kappa);
      if (self_int && i == j) kv.re = kv.im = 0;
      const R* qj = q + ((long long)b * n + j) * w + c * CWd;
      if (CWd == 1) {
        ar += double(kv.re * qj[0]);
      } else {
        ar += double(kv.re * qj[0] - kv.im * qj[1]);
        ai += double(kv.re * qj[1] + kv.im * qj[0]);
      }
    }
    R* ui = u + ((long long)b * m + i) * w + c * CWd;
    ui[0] = R(ar);
    if (CWd == 2) ui[1] = R(ai);
  }
}

template <typename R>
int launch_blocks(int op, int nblocks, int m, int n, int ndim, const void* xt, const void* xs, const void* gt,
                  const void* gs, const void* q, void* u, int w, int self_int, double kappa, hipStream_t st) {
  dim3 grid((m + 63) / 64, nblocks);
  const R* a0 = static_cast<const R*>(xt);
  const R* a1 = static_cast<const R*>(xs);
  const R* a2 = static_cast<const R*>(gt);
  const R* a3 = static_cast<const R*>(gs);
  const R* a4 = static_cast<const R*>(q);
  R* a5 = static_cast<R*>(u);
  R k = R(kappa);
#define SKF_BLK(OPC) \
  hipLaunchKernelGGL((blocks_kernel<R, OPC>), grid, dim3(64), 0, st, m, n, ndim, a0, a1, a2, a3, a4, a5, w, self_int, k)
  switch (op) {
    case SKF_LAPLACE_2D: SKF_BLK(SKF_LAPLACE_2D); break;
    case SKF_LAPLACE_3D: SKF_BLK(SKF_LAPLACE_3D); break;
    case SKF_HELMHOLTZ_2D: SKF_BLK(SKF_HELMHOLTZ_2D); break;
    case SKF_HELMHOLTZ_3D: SKF_BLK(SKF_HELMHOLTZ_3D); break;
    case SKF_LAPLACE_DL_3D: SKF_BLK(SKF_LAPLACE_DL_3D); break;
    case SKF_LAPLACE_ADL_3D: SKF_BLK(SKF_LAPLACE_ADL_3D); break;
    case SKF_HELMHOLTZ_DL_3D: SKF_BLK(SKF_HELMHOLTZ_DL_3D); break;
    case SKF_HELMHOLTZ_ADL_3D: SKF_BLK(SKF_HELMHOLTZ_ADL_3D); break;
    default: return set_err(SKF_ERR_ARG, "unknown kernel id %d", op);
  }
#undef SKF_BLK
  SKF_LAUNCHED();
  return SKF_OK;
}

}  // namespace

extern "C" int skf_p2p_blocks(int kernel, int precision, int width, double kappa, int nblocks, int m, int n,
                              int ndim, const void* xt, const void* xs, const void* nt, const void* ns,
                              const void* q, void* u, int self_interaction, void* stream) {
  if (nblocks <= 0 || m <= 0) return SKF_OK;
  if (!xt || !xs || !q || !u || width <= 0 || n < 0) return set_err(SKF_ERR_ARG, "bad block arguments");
  if ((op_dim3(kernel) ? 3 : 2) != ndim) return set_err(SKF_ERR_ARG, "dimension does not match the kernel");
  if (op_is_complex(kernel) && (width % 2)) return set_err(SKF_ERR_ARG, "complex kernel needs an even width");
  if (kernel == SKF_LAPLACE_DL_3D || kernel == SKF_HELMHOLTZ_DL_3D) {
    if (!ns) return set_err(SKF_ERR_ARG, "source normals required");
  }
  if (kernel == SKF_LAPLACE_ADL_3D || kernel == SKF_HELMHOLTZ_ADL_3D) {
    if (!nt) return set_err(SKF_ERR_ARG, "target normals required");
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (precision == 0)
    return launch_blocks<double>(kernel, nblocks, m, n, ndim, xt, xs, nt, ns, q, u, width, self_interaction, kappa, st);
  if (precision == 1)
    return launch_blocks<float>(kernel, nblocks, m, n, ndim, xt, xs, nt, ns, q, u, width, self_interaction, kappa, st);
  return set_err(SKF_ERR_ARG, "precision must be 0 or 1");
}

// ------------------------------------------------- run-time compiled kernels
// Near-field kernels of user-defined operators: hiprtc compiles a source that includes
// p2p_core.h (passed in as a named header) with the operator's OpT<> / kernel_value<>
// specializations, for gfx950 with the library's own code-generation flags; the requested
// template instances are resolved by name expression and loaded as one module.
#include <hip/hiprtc.h>

static std::string g_jit_log;

extern "C" const char* skf_jit_log(void) { return g_jit_log.c_str(); }

extern "C" int skf_jit_compile(const char* source, const char* header_name, const char* header_source,
                               const char* const* names, int nnames, void** module, void** functions,
                               long long* code_bytes) {
  if (!source || nnames < 0 || (nnames > 0 && !names)) return set_err(SKF_ERR_ARG, "bad jit arguments");
  if (module && !functions && nnames > 0) return set_err(SKF_ERR_ARG, "functions array needed");
  g_jit_log.clear();
  hiprtcProgram prog;
  const char* hsrc[1] = {header_source};
  const char* hname[1] = {header_name};
  const int nh = (header_name && header_source) ? 1 : 0;
  if (hiprtcCreateProgram(&prog, source, "skf_user_p2p.hip", nh, hsrc, hname) != HIPRTC_SUCCESS)
    return set_err(SKF_ERR_HIP, "hiprtcCreateProgram failed");
  for (int i = 0; i < nnames; ++i) {
    if (hiprtcAddNameExpression(prog, names[i]) != HIPRTC_SUCCESS) {
      hiprtcDestroyProgram(&prog);
      return set_err(SKF_ERR_ARG, "bad kernel name expression %s", names[i]);
    }
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-munsafe-fp-atomics",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(prog, int(sizeof(opts) / sizeof(opts[0])), opts);
  size_t ls = 0;
  if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
    g_jit_log.assign(ls, '\0');
    hiprtcGetProgramLog(prog, &g_jit_log[0]);
    g_jit_log.resize(strlen(g_jit_log.c_str()));
  }
  if (cr != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return set_err(SKF_ERR_ARG, "run-time compilation failed (skf_jit_log holds the compiler log)");
  }
  size_t cs = 0;
  if (hiprtcGetCodeSize(prog, &cs) != HIPRTC_SUCCESS || cs == 0) {
    hiprtcDestroyProgram(&prog);
    return set_err(SKF_ERR_HIP, "no code from run-time compilation");
  }
  if (code_bytes) *code_bytes = (long long)cs;
  if (!module) {  // compile-only check (no device needed)
    hiprtcDestroyProgram(&prog);
    return SKF_OK;
  }
  std::vector<char> code(cs);
  std::vector<std::string> lowered(nnames);
  bool ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
  for (int i = 0; ok && i < nnames; ++i) {
    const char* ln = nullptr;
    ok = hiprtcGetLoweredName(prog, names[i], &ln) == HIPRTC_SUCCESS && ln;
    if (ok) lowered[i] = ln;
  }
  hiprtcDestroyProgram(&prog);
  if (!ok) return set_err(SKF_ERR_HIP, "run-time compilation: code or kernel names unavailable");
  hipModule_t mod;
  SKF_HIP(hipModuleLoadData(&mod, code.data()));
  for (int i = 0; i < nnames; ++i) {
    hipFunction_t fn;
    const hipError_t e = hipModuleGetFunction(&fn, mod, lowered[i].c_str());
    if (e != hipSuccess) {
      (void)hipModuleUnload(mod);
      return set_err(SKF_ERR_HIP, "hipModuleGetFunction(%s): %s", names[i], hipGetErrorString(e));
    }
    functions[i] = reinterpret_cast<void*>(fn);
  }
  *module = reinterpret_cast<void*>(mod);
  return SKF_OK;
}

extern "C" int skf_jit_unload(void* module) {
  if (!module) return SKF_OK;
  SKF_HIP(hipModuleUnload(static_cast<hipModule_t>(module)));
  return SKF_OK;
}
