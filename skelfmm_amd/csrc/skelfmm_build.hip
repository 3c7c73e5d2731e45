// skelfmm_build.hip -- gfx950 kernels of the GPU skeletonization build and of the learned rank
// predictor (part of libskelfmm_hip.so).
//
// Build, per tree level and batched over boxes (reference algorithm: skelfmm/util.py:13-51,87-124):
//   skf_proxy_matrices  A_b = proxy matrix of box b (row-major, m_b x n_b): K_BP^T for the symmetric
//                       point kernels, [K_PB; K_BP^T] with proxy normals for the double layers
//   skf_pqr_batched     column-pivoted Householder QR of every A_b in place, stopped at the first
//                       |R_ii| <= tol |R_00| (adaptive rank, util.py:25-31) or after kmax_b steps
//   skf_trsm_batched    T_b = R11^{-1} R12 (util.py:43-49)
//   skf_id_residual     proxy-residual check of an ID on an independent sketch (learned-rank build)
// Learned rank predictor (8 -> 64 -> 64 -> 1 MLP, GELU, FP32):
//   skf_mlp_forward / skf_mlp_backward / skf_adam_step
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "skelfmm_hip.h"

namespace {

thread_local char b_err[512] = "";
int berr(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b_err, sizeof(b_err), fmt, ap);
  va_end(ap);
  return code;
}

#define BLD_LAUNCHED()                                                                 \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess) return berr(SKF_ERR_HIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------ scalar types
struct zd {
  double re, im;
};
__device__ inline zd operator+(zd a, zd b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline zd operator-(zd a, zd b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline zd operator*(zd a, zd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline zd operator*(double s, zd a) { return {s * a.re, s * a.im}; }

template <typename S>
struct Sc;
template <>
struct Sc<double> {
  __device__ static double zero() { return 0.0; }
  __device__ static double conj(double a) { return a; }
  __device__ static double abs2(double a) { return a * a; }
  __device__ static double real(double a) { return a; }
  __device__ static double div(double a, double b) { return a / b; }
  __device__ static double from(double r, double) { return r; }
};
template <>
struct Sc<zd> {
  __device__ static zd zero() { return {0.0, 0.0}; }
  __device__ static zd conj(zd a) { return {a.re, -a.im}; }
  __device__ static double abs2(zd a) { return a.re * a.re + a.im * a.im; }
  __device__ static double real(zd a) { return a.re; }
  __device__ static zd div(zd a, zd b) {
    const double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
  }
  __device__ static zd from(double r, double i) { return {r, i}; }
};

constexpr int kT = 256;

__device__ inline double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// ----------------------------------------------------------- proxy matrices
// kernel values in FP64, exact distances (the proxies never coincide with box points)
__device__ inline zd kval(int op, double kappa, const double* x, const double* y, const double* nt, const double* ns,
                          int dim) {
  const double inv4pi = 0.079577471545947667884441881686257181;
  double dx = x[0] - y[0], dy = x[1] - y[1], dz = dim == 3 ? x[2] - y[2] : 0.0;
  double r2 = dx * dx + dy * dy + dz * dz;
  double r = sqrt(r2);
  zd k = {0.0, 0.0};
  switch (op) {
    case SKF_LAPLACE_2D: k.re = log(r); break;
    case SKF_LAPLACE_3D: k.re = 1.0 / r; break;
    case SKF_HELMHOLTZ_2D: k.re = j0(kappa * r); k.im = y0(kappa * r); break;
    case SKF_HELMHOLTZ_3D: {
      double s, c;
      sincos(kappa * r, &s, &c);
      k.re = c / r;
      k.im = s / r;
      break;
    }
    case SKF_LAPLACE_DL_3D: k.re = -inv4pi * (dx * ns[0] + dy * ns[1] + dz * ns[2]) / (r2 * r); break;
    case SKF_LAPLACE_ADL_3D: k.re = inv4pi * (dx * nt[0] + dy * nt[1] + dz * nt[2]) / (r2 * r); break;
    default: {  // Helmholtz double layers
      const bool adj = op == SKF_HELMHOLTZ_ADL_3D;
      const double rn = adj ? (dx * nt[0] + dy * nt[1] + dz * nt[2]) : (dx * ns[0] + dy * ns[1] + dz * ns[2]);
      double s, c;
      const double kr = kappa * r;
      sincos(kr, &s, &c);
      double f = inv4pi * rn / (r2 * r);
      if (adj) f = -f;
      k.re = f * (-c - kr * s);
      k.im = f * (kr * c - s);
    }
  }
  return k;
}

template <typename S>
__device__ inline void store(S* p, zd v);
template <>
__device__ inline void store<double>(double* p, zd v) { *p = v.re; }
template <>
__device__ inline void store<zd>(zd* p, zd v) { *p = v; }

// A_b[i][j], i over proxy rows, j over the box's candidate points (row-major, m_b x n_b)
template <typename S>
__global__ __launch_bounds__(kT) void proxy_kernel(int op, int dim, double kappa, const double* __restrict__ pts,
                                                   const double* __restrict__ nrm, const int* __restrict__ idx,
                                                   const long long* __restrict__ idx_off, const int* __restrict__ nn,
                                                   const double* __restrict__ ctr, const double* __restrict__ side,
                                                   const double* __restrict__ unit, int nproxy, int stacked,
                                                   S* __restrict__ A, const long long* __restrict__ a_off, int lda0) {
  const int b = blockIdx.y;
  const int n = nn[b];
  const int lda = lda0 > 0 ? lda0 : n;
  const int m = stacked ? 2 * nproxy : nproxy;
  const long long tot = (long long)m * n;
  const double c[3] = {ctr[3 * b], ctr[3 * b + 1], ctr[3 * b + 2]};
  const double sd = side[b];
  const int* ib = idx + idx_off[b];
  S* a = A + a_off[b];
  for (long long t = (long long)blockIdx.x * kT + threadIdx.x; t < tot; t += (long long)gridDim.x * kT) {
    const int i = int(t / n), j = int(t - (long long)i * n);
    const int ip = i % nproxy;
    const double* u = unit + 3 * ip;
    double p[3] = {c[0] + sd * u[0], c[1] + sd * u[1], dim == 3 ? c[2] + sd * u[2] : 0.0};
    const double un = sqrt(u[0] * u[0] + u[1] * u[1] + (dim == 3 ? u[2] * u[2] : 0.0));
    double pn[3] = {u[0] / un, u[1] / un, dim == 3 ? u[2] / un : 0.0};  // outward proxy normal (sphere)
    const int pj = ib[j];
    const double* x = pts + (long long)3 * pj;
    const double* xn = nrm ? nrm + (long long)3 * pj : pn;
    zd v;
    if (!stacked) {
      v = kval(op, kappa, x, p, xn, pn, dim);  // K(x_j, p_i): row i of K_BP^T
    } else if (i < nproxy) {
      v = kval(op, kappa, p, x, pn, xn, dim);  // K(p_i, x_j): proxy as target, box as source
    } else {
      v = kval(op, kappa, x, p, xn, pn, dim);  // K(x_j, p_i): box as target, proxy as source
    }
    store<S>(&a[(long long)i * lda + j], v);
  }
}

// ------------------------------------------------------- pivoted QR (batched)
// One workgroup per box; A_b row-major (m x n). Threads own columns (coalesced row reads).
// Shared: col norms^2 (n), reference norms^2 (n), Householder vector (m).
template <typename S>
__global__ __launch_bounds__(kT) void pqr_kernel(S* __restrict__ A, const long long* __restrict__ a_off,
                                                 const int* __restrict__ mm, const int* __restrict__ nn,
                                                 const int* __restrict__ kmax, double tol, int* __restrict__ rank_out,
                                                 int* __restrict__ conv_out, int* __restrict__ perm,
                                                 const long long* __restrict__ p_off, double* __restrict__ rdiag,
                                                 const int* __restrict__ lda_b) {
  using C = Sc<S>;
  extern __shared__ double sh[];
  __shared__ double red[4];
  __shared__ int redi[4];
  __shared__ double redv[4];
  const int b = blockIdx.x;
  const int m = mm[b], n = nn[b];
  const int kcap = min(kmax[b], min(m, n));
  const long long n_ = lda_b ? lda_b[b] : n;  // row stride
  S* a = A + a_off[b];
  int* pv = perm + p_off[b];
  double* rd = rdiag + p_off[b];
  double* nrm2 = sh;                     // n
  double* nref = sh + n;                 // n
  S* v = reinterpret_cast<S*>(sh + 2 * n);  // m
  const int tid = threadIdx.x;
  for (int j = tid; j < n; j += kT) {
    double s = 0.0;
    for (int r = 0; r < m; ++r) s += C::abs2(a[r * n_ + j]);
    nrm2[j] = s;
    nref[j] = s;
    pv[j] = j;
  }
  __syncthreads();
  double r00 = 0.0;
  int rank = kcap;
  bool converged = (kcap == min(m, n));
  for (int i = 0; i < kcap; ++i) {
    // pivot: argmax of the remaining column norms
    double best = -1.0;
    int bj = i;
    for (int j = i + tid; j < n; j += kT) {
      if (nrm2[j] > best) {
        best = nrm2[j];
        bj = j;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oj = __shfl_xor(bj, o, 64);
      if (ob > best || (ob == best && oj < bj)) {
        best = ob;
        bj = oj;
      }
    }
    if ((tid & 63) == 0) {
      redv[tid >> 6] = best;
      redi[tid >> 6] = bj;
    }
    __syncthreads();
    int p = redi[0];
    double pb = redv[0];
    for (int w = 1; w < 4; ++w)
      if (redv[w] > pb || (redv[w] == pb && redi[w] < p)) {
        pb = redv[w];
        p = redi[w];
      }
    __syncthreads();
    if (p != i) {
      for (int r = tid; r < m; r += kT) {
        const S t = a[r * n_ + i];
        a[r * n_ + i] = a[r * n_ + p];
        a[r * n_ + p] = t;
      }
      if (tid == 0) {
        double t = nrm2[i];
        nrm2[i] = nrm2[p];
        nrm2[p] = t;
        t = nref[i];
        nref[i] = nref[p];
        nref[p] = t;
        const int ti = pv[i];
        pv[i] = pv[p];
        pv[p] = ti;
      }
    }
    __syncthreads();
    // Householder reflector for column i, rows i..m-1 (zlarfg: beta real, tau = (beta - alpha)/beta);
    // alpha is read before the reduction's barriers, so no thread reads it after it is overwritten
    const S alpha = a[i * n_ + i];
    double sig = 0.0;
    for (int r = i + 1 + tid; r < m; r += kT) sig += C::abs2(a[r * n_ + i]);
    sig = block_sum(sig, red);
    const double normx = sqrt(C::abs2(alpha) + sig);
    if (i == 0) r00 = normx;
    if (normx <= tol * r00 || normx == 0.0) {
      rank = i;
      converged = true;
      break;
    }
    const double beta = C::real(alpha) >= 0.0 ? -normx : normx;
    const S denom = alpha - C::from(beta, 0.0);
    const S tau = C::div(C::from(beta, 0.0) - alpha, C::from(beta, 0.0));
    for (int r = i + tid; r < m; r += kT) v[r] = (r == i) ? C::from(1.0, 0.0) : C::div(a[r * n_ + i], denom);
    if (tid == 0) {
      a[i * n_ + i] = C::from(beta, 0.0);
      rd[i] = normx;
    }
    __syncthreads();
    // apply H^H = I - conj(tau) v v^H to the trailing columns
    const S ctau = C::conj(tau);
    for (int j = i + 1 + tid; j < n; j += kT) {
      S w = C::zero();
      for (int r = i; r < m; ++r) w = w + C::conj(v[r]) * a[r * n_ + j];
      w = ctau * w;
      for (int r = i; r < m; ++r) a[r * n_ + j] = a[r * n_ + j] - v[r] * w;
      // downdate the column norm; recompute when cancellation eats half the digits
      const double nr = nrm2[j] - C::abs2(a[i * n_ + j]);
      if (nr <= 1.5e-8 * nref[j]) {
        double s = 0.0;
        for (int r = i + 1; r < m; ++r) s += C::abs2(a[r * n_ + j]);
        nrm2[j] = s;
        nref[j] = s;
      } else {
        nrm2[j] = nr;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    rank_out[b] = rank;
    conv_out[b] = converged ? 1 : 0;
  }
}

// T_b = R11^{-1} R12 by back substitution; one thread per column of R12, T row-major (r x (n - r)).
template <typename S>
__global__ __launch_bounds__(kT) void trsm_kernel(const S* __restrict__ A, const long long* __restrict__ a_off,
                                                  const int* __restrict__ nn, const int* __restrict__ rank,
                                                  S* __restrict__ T, const long long* __restrict__ t_off,
                                                  const int* __restrict__ lda_b) {
  using C = Sc<S>;
  const int b = blockIdx.y;
  const int n = nn[b], r = rank[b];
  const int nc = n - r;
  const int c = blockIdx.x * kT + threadIdx.x;
  if (r <= 0 || c >= nc) return;
  const long long lda = lda_b ? lda_b[b] : n;
  const S* a = A + a_off[b];
  S* t = T + t_off[b];
  for (int i = r - 1; i >= 0; --i) {
    S s = a[i * lda + r + c];
    for (int l = i + 1; l < r; ++l) s = s - a[i * lda + l] * t[(long long)l * nc + c];
    t[(long long)i * nc + c] = C::div(s, a[i * lda + i]);
  }
}

// Proxy-residual check of an ID from an independent sketch Z_b (c rows, row stride ldz):
// res[b] = max_j |Z[:, p[r + j]] - sum_i Z[:, p[i]] T[i][j]|_2 and znorm[b] = max_j |Z[:, j]|_2.
// One workgroup per box, threads over the redundant columns; block max through LDS.
template <typename S>
__global__ __launch_bounds__(kT) void id_residual_kernel(const S* __restrict__ Z, const long long* __restrict__ z_off,
                                                         const int* __restrict__ ldz_b, int c, const int* __restrict__ nn,
                                                         const int* __restrict__ rank, const int* __restrict__ perm,
                                                         const long long* __restrict__ p_off, const S* __restrict__ T,
                                                         const long long* __restrict__ t_off,
                                                         double* __restrict__ res, double* __restrict__ znorm) {
  using C = Sc<S>;
  __shared__ double red[2][4];
  const int b = blockIdx.x;
  const int n = nn[b], r = rank[b];
  const int nc = n - r;
  const S* z = Z + z_off[b];
  const long long ldz = ldz_b ? ldz_b[b] : n;
  const int* pv = perm + p_off[b];
  const S* t = T + t_off[b];
  double emax = 0.0, zmax = 0.0;
  for (int j = threadIdx.x; j < n; j += kT) {
    double s = 0.0;
    for (int row = 0; row < c; ++row) s += C::abs2(z[(long long)row * ldz + j]);
    zmax = fmax(zmax, s);
  }
  for (int j = threadIdx.x; j < nc; j += kT) {
    const int cj = pv[r + j];
    double s = 0.0;
    for (int row = 0; row < c; ++row) {
      const S* zr = z + (long long)row * ldz;
      S e = zr[cj];
      for (int i = 0; i < r; ++i) e = e - zr[pv[i]] * t[(long long)i * nc + j];
      s += C::abs2(e);
    }
    emax = fmax(emax, s);
  }
  for (int o = 32; o > 0; o >>= 1) {
    emax = fmax(emax, __shfl_xor(emax, o, 64));
    zmax = fmax(zmax, __shfl_xor(zmax, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = emax;
    red[1][threadIdx.x >> 6] = zmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    res[b] = sqrt(fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3])));
    znorm[b] = sqrt(fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3])));
  }
}

// --------------------------------------------------------------------- MLP
// params: W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3 (4801 floats)
constexpr int kIn = 8, kH = 64, kBatch = 32;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH;
constexpr int kNP = oB3 + 1;

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Forward for a batch of kBatch samples held by one workgroup: Z1, Z2 in LDS ([unit][sample]).
__device__ void mlp_fwd_block(const float* __restrict__ P, const float* xs, float* z1, float* z2, float* y, int nb) {
  const int tid = threadIdx.x;
  for (int t = tid; t < kH * kBatch; t += kT) {
    const int u = t / kBatch, s = t % kBatch;
    float acc = P[oB1 + u];
    for (int i = 0; i < kIn; ++i) acc += P[oW1 + u * kIn + i] * xs[s * kIn + i];
    z1[u * kBatch + s] = acc;
  }
  __syncthreads();
  for (int t = tid; t < kH * kBatch; t += kT) {
    const int u = t / kBatch, s = t % kBatch;
    float acc = P[oB2 + u];
    for (int v = 0; v < kH; ++v) acc += P[oW2 + u * kH + v] * gelu(z1[v * kBatch + s]);
    z2[u * kBatch + s] = acc;
  }
  __syncthreads();
  if (tid < kBatch) {
    float acc = P[oB3];
    for (int u = 0; u < kH; ++u) acc += P[oW3 + u] * gelu(z2[u * kBatch + tid]);
    y[tid] = acc;
  }
  __syncthreads();
  (void)nb;
}

__global__ __launch_bounds__(kT) void mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                         int n, float* __restrict__ Y) {
  __shared__ float xs[kBatch * kIn], z1[kH * kBatch], z2[kH * kBatch], y[kBatch];
  const int s0 = blockIdx.x * kBatch;
  const int nb = min(kBatch, n - s0);
  for (int t = threadIdx.x; t < kBatch * kIn; t += kT) xs[t] = (t / kIn < nb) ? X[(long long)s0 * kIn + t] : 0.f;
  __syncthreads();
  mlp_fwd_block(P, xs, z1, z2, y, nb);
  if (threadIdx.x < nb) Y[s0 + threadIdx.x] = y[threadIdx.x];
}

// Mean-squared-error gradient of one batch, added into grad (atomics across workgroups).
__global__ __launch_bounds__(kT) void mlp_backward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                          const float* __restrict__ Tg, int n, float inv_n,
                                                          float* __restrict__ grad, float* __restrict__ loss) {
  __shared__ float xs[kBatch * kIn], z1[kH * kBatch], z2[kH * kBatch], y[kBatch], dy[kBatch];
  __shared__ float d2[kH * kBatch], d1[kH * kBatch];
  const int tid = threadIdx.x;
  const int s0 = blockIdx.x * kBatch;
  const int nb = min(kBatch, n - s0);
  for (int t = tid; t < kBatch * kIn; t += kT) xs[t] = (t / kIn < nb) ? X[(long long)s0 * kIn + t] : 0.f;
  __syncthreads();
  mlp_fwd_block(P, xs, z1, z2, y, nb);
  if (tid < kBatch) {
    const float e = tid < nb ? y[tid] - Tg[s0 + tid] : 0.f;
    dy[tid] = 2.f * e * inv_n;
    if (tid < nb) atomicAdd(loss, e * e * inv_n);
  }
  __syncthreads();
  // output layer
  if (tid < kH) {
    float g = 0.f;
    for (int s = 0; s < kBatch; ++s) g += dy[s] * gelu(z2[tid * kBatch + s]);
    atomicAdd(&grad[oW3 + tid], g);
  }
  if (tid == kH) {
    float g = 0.f;
    for (int s = 0; s < kBatch; ++s) g += dy[s];
    atomicAdd(&grad[oB3], g);
  }
  for (int t = tid; t < kH * kBatch; t += kT) {
    const int u = t / kBatch, s = t % kBatch;
    d2[t] = dy[s] * P[oW3 + u] * gelu_d(z2[t]);
  }
  __syncthreads();
  // hidden layer 2: dW2[u][v] = sum_s d2[u][s] gelu(z1[v][s]); db2[u] = sum_s d2[u][s]
  for (int t = tid; t < kH * kH; t += kT) {
    const int u = t / kH, v = t % kH;
    float g = 0.f;
    for (int s = 0; s < kBatch; ++s) g += d2[u * kBatch + s] * gelu(z1[v * kBatch + s]);
    atomicAdd(&grad[oW2 + t], g);
  }
  if (tid < kH) {
    float g = 0.f;
    for (int s = 0; s < kBatch; ++s) g += d2[tid * kBatch + s];
    atomicAdd(&grad[oB2 + tid], g);
  }
  // d1[v][s] = gelu'(z1[v][s]) sum_u W2[u][v] d2[u][s]
  for (int t = tid; t < kH * kBatch; t += kT) {
    const int v = t / kBatch, s = t % kBatch;
    float acc = 0.f;
    for (int u = 0; u < kH; ++u) acc += P[oW2 + u * kH + v] * d2[u * kBatch + s];
    d1[t] = acc * gelu_d(z1[t]);
  }
  __syncthreads();
  for (int t = tid; t < kH * kIn; t += kT) {
    const int v = t / kIn, i = t % kIn;
    float g = 0.f;
    for (int s = 0; s < kBatch; ++s) g += d1[v * kBatch + s] * xs[s * kIn + i];
    atomicAdd(&grad[oW1 + t], g);
  }
  if (tid < kH) {
    float g = 0.f;
    for (int s = 0; s < kBatch; ++s) g += d1[tid * kBatch + s];
    atomicAdd(&grad[oB1 + tid], g);
  }
}

__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, int np, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  P[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" const char* skf_build_last_error(void) { return b_err; }

extern "C" int skf_mlp_param_count(void) { return kNP; }

extern "C" int skf_proxy_matrices(int kernel, int ndim, double kappa, const double* pts, const double* normals,
                                  const int* idx, const long long* idx_off, const int* ncols, const double* centers,
                                  const double* side, const double* unit_proxy, int nproxy, int stacked,
                                  int complex_out, void* A, const long long* a_off, int lda, int nbox, int max_cols,
                                  void* stream) {
  if (nbox <= 0) return SKF_OK;
  if (!pts || !idx || !idx_off || !ncols || !centers || !side || !unit_proxy || !A || !a_off || nproxy <= 0)
    return berr(SKF_ERR_ARG, "bad proxy-matrix arguments");
  if (ndim != 2 && ndim != 3) return berr(SKF_ERR_ARG, "ndim must be 2 or 3");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long per = (long long)(stacked ? 2 : 1) * nproxy * (max_cols > 0 ? max_cols : 1);
  int gx = int((per + kT - 1) / kT);
  if (gx > 1024) gx = 1024;
  dim3 grid(gx, nbox);
  if (complex_out)
    hipLaunchKernelGGL((proxy_kernel<zd>), grid, dim3(kT), 0, st, kernel, ndim, kappa, pts, normals, idx, idx_off,
                       ncols, centers, side, unit_proxy, nproxy, stacked, static_cast<zd*>(A), a_off, lda);
  else
    hipLaunchKernelGGL((proxy_kernel<double>), grid, dim3(kT), 0, st, kernel, ndim, kappa, pts, normals, idx, idx_off,
                       ncols, centers, side, unit_proxy, nproxy, stacked, static_cast<double*>(A), a_off, lda);
  BLD_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_pqr_batched(void* A, const long long* a_off, const int* lda, const int* m, const int* n,
                               const int* kmax, double tol, int* rank, int* converged, int* perm,
                               const long long* p_off, double* rdiag, int nbox, int max_m, int max_n, int complex_a,
                               void* stream) {
  if (nbox <= 0) return SKF_OK;
  if (!A || !a_off || !m || !n || !kmax || !rank || !converged || !perm || !p_off || !rdiag)
    return berr(SKF_ERR_ARG, "bad QR arguments");
  const size_t shm = sizeof(double) * (2 * (size_t)max_n) + (complex_a ? sizeof(zd) : sizeof(double)) * (size_t)max_m;
  if (shm > 64 * 1024) return berr(SKF_ERR_UNSUPPORTED, "QR workspace of %zu bytes exceeds 64 KiB of LDS", shm);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (complex_a) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(&pqr_kernel<zd>), hipFuncAttributeMaxDynamicSharedMemorySize,
                        int(shm));
    hipLaunchKernelGGL((pqr_kernel<zd>), dim3(nbox), dim3(kT), shm, st, static_cast<zd*>(A), a_off, m, n, kmax, tol,
                       rank, converged, perm, p_off, rdiag, lda);
  } else {
    hipFuncSetAttribute(reinterpret_cast<const void*>(&pqr_kernel<double>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, int(shm));
    hipLaunchKernelGGL((pqr_kernel<double>), dim3(nbox), dim3(kT), shm, st, static_cast<double*>(A), a_off, m, n,
                       kmax, tol, rank, converged, perm, p_off, rdiag, lda);
  }
  BLD_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_trsm_batched(const void* A, const long long* a_off, const int* lda, const int* n, const int* rank,
                                void* T, const long long* t_off, int nbox, int max_cols, int complex_a,
                                void* stream) {
  if (nbox <= 0) return SKF_OK;
  if (!A || !a_off || !n || !rank || !T || !t_off) return berr(SKF_ERR_ARG, "bad TRSM arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((max_cols + kT - 1) / kT > 0 ? (max_cols + kT - 1) / kT : 1, nbox);
  if (complex_a)
    hipLaunchKernelGGL((trsm_kernel<zd>), grid, dim3(kT), 0, st, static_cast<const zd*>(A), a_off, n, rank,
                       static_cast<zd*>(T), t_off, lda);
  else
    hipLaunchKernelGGL((trsm_kernel<double>), grid, dim3(kT), 0, st, static_cast<const double*>(A), a_off, n, rank,
                       static_cast<double*>(T), t_off, lda);
  BLD_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_id_residual(const void* Z, const long long* z_off, const int* ldz, int c, const int* n, const int* rank,
                               const int* perm, const long long* p_off, const void* T, const long long* t_off,
                               int nbox, int complex_a, double* res, double* znorm, void* stream) {
  if (nbox <= 0) return SKF_OK;
  if (!Z || !z_off || !n || !rank || !perm || !p_off || !T || !t_off || !res || !znorm || c <= 0)
    return berr(SKF_ERR_ARG, "bad ID-residual arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (complex_a)
    hipLaunchKernelGGL((id_residual_kernel<zd>), dim3(nbox), dim3(kT), 0, st, static_cast<const zd*>(Z), z_off, ldz, c,
                       n, rank, perm, p_off, static_cast<const zd*>(T), t_off, res, znorm);
  else
    hipLaunchKernelGGL((id_residual_kernel<double>), dim3(nbox), dim3(kT), 0, st, static_cast<const double*>(Z), z_off,
                       ldz, c, n, rank, perm, p_off, static_cast<const double*>(T), t_off, res, znorm);
  BLD_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_mlp_forward(const float* params, const float* X, int n, float* y, void* stream) {
  if (n <= 0) return SKF_OK;
  if (!params || !X || !y) return berr(SKF_ERR_ARG, "bad MLP arguments");
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + kBatch - 1) / kBatch), dim3(kT), 0,
                     static_cast<hipStream_t>(stream), params, X, n, y);
  BLD_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_mlp_backward(const float* params, const float* X, const float* target, int n, float* grad,
                                float* loss, void* stream) {
  if (n <= 0) return SKF_OK;
  if (!params || !X || !target || !grad || !loss) return berr(SKF_ERR_ARG, "bad MLP arguments");
  hipLaunchKernelGGL(mlp_backward_kernel, dim3((n + kBatch - 1) / kBatch), dim3(kT), 0,
                     static_cast<hipStream_t>(stream), params, X, target, n, 1.0f / float(n), grad, loss);
  BLD_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_adam_step(float* params, const float* grad, float* m1, float* m2, int np, float lr, float beta1,
                             float beta2, float eps, int step, void* stream) {
  if (np <= 0) return SKF_OK;
  if (!params || !grad || !m1 || !m2 || step < 1) return berr(SKF_ERR_ARG, "bad Adam arguments");
  const float c1 = 1.f - powf(beta1, float(step));
  const float c2 = 1.f - powf(beta2, float(step));
  hipLaunchKernelGGL(adam_kernel, dim3((np + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), params, grad,
                     m1, m2, np, lr, beta1, beta2, eps, c1, c2);
  BLD_LAUNCHED();
  return SKF_OK;
}
