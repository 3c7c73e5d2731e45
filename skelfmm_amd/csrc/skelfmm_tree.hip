// skelfmm_tree.hip -- gfx950 kernels of the tree, neighbour-list and pair-list construction (part of
// libskelfmm_hip.so).
//
// Reference: the adaptive 2^d-tree the reference imports from simpletree (skelfmm/skel_fmm.py:150;
// its contract as recursive_skel.py:54-68,142-166,622-780 use it) and the five pair lists of
// recursive_skel.py:622-712 (get_parallel_lists). The host statement of the same tree is
// skelfmm_amd/tree.py; every integer these kernels produce equals it bit for bit:
//   skf_tree_keys           Morton key of every point at depth D (bit b of axis k -> bit b d + k)
//   skf_tree_sort_pairs     stable radix sort of (key, point index) pairs (rocPRIM)
//   skf_tree_child_ranges   the 2^d child cells of split boxes and their point ranges in key order
//   skf_tree_balance_marks  leaves that a 2:1 level restriction forces to split
//   skf_tree_box_tables     children, parent, level and integer coordinates of every box
//   skf_tree_colleagues     same-level touching boxes (self included), ascending
//   skf_tree_coarse         touching leaves on coarser levels, ascending
//   skf_tree_pairs          leaf, tree and coarse pair lists of levels above the root (count / fill)
// Boxes are numbered level by level and in Morton order inside a level; a level's codes are
// sorted, so every lookup is a binary search inside that level's range.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "skelfmm_hip.h"

namespace {

thread_local char t_err[512] = "";
int terr(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err, sizeof(t_err), fmt, ap);
  va_end(ap);
  return code;
}

#define TREE_LAUNCHED()                                                                    \
  do {                                                                                     \
    hipError_t e_ = hipGetLastError();                                                     \
    if (e_ != hipSuccess) return terr(SKF_ERR_HIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

typedef unsigned long long u64;

constexpr int kBlock = 256;

inline unsigned blocks_for(long long n) { return unsigned((n + kBlock - 1) / kBlock); }

__device__ inline long long interleave(const long long* q, int d, int nbits) {
  long long code = 0;
  for (int b = 0; b < nbits; ++b)
    for (int k = 0; k < d; ++k) code |= ((q[k] >> b) & 1LL) << (b * d + k);
  return code;
}

__device__ inline void deinterleave(long long code, int d, int nbits, long long* q) {
  for (int k = 0; k < d; ++k) q[k] = 0;
  for (int b = 0; b < nbits; ++b)
    for (int k = 0; k < d; ++k) q[k] |= ((code >> (b * d + k)) & 1LL) << b;
}

// index of `c` in the sorted codes[lo, hi), or -1
__device__ inline long long find_code(const long long* __restrict__ codes, long long lo, long long hi, long long c) {
  const long long end = hi;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (codes[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < end && codes[lo] == c) ? lo : -1;
}

// first index in keys[lo, hi) whose key is >= k
__device__ inline int lower_bound_u64(const u64* __restrict__ keys, int lo, int hi, u64 k) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------- point keys
// q_k = floor(((x_k - lo_k) / L0) * 2^D), clipped to [0, 2^D - 1]: the same FP64 operations in the
// same order as the host statement, so the keys agree bit for bit
__global__ void keys_kernel(const double* __restrict__ pts, int n, int d, double lo0, double lo1, double lo2, double L0,
                            double scale, long long qmax, int nbits, u64* __restrict__ keys, int* __restrict__ idx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double lo[3] = {lo0, lo1, lo2};
  long long q[3] = {0, 0, 0};
  for (int k = 0; k < d; ++k) {
    const double v = floor((pts[i * d + k] - lo[k]) / L0 * scale);
    long long qi = (long long)v;
    qi = qi < 0 ? 0 : (qi > qmax ? qmax : qi);
    q[k] = qi;
  }
  keys[i] = (u64)interleave(q, d, nbits);
  idx[i] = int(i);
}

// ------------------------------------------------------------- child ranges
// thread t = (parent p, child k): child code p_code * 2^d + k at level child_level and its point
// range [lower_bound(code << shift), lower_bound((code + 1) << shift)) inside the parent's range
__global__ void child_ranges_kernel(const u64* __restrict__ keys, int d, int D, int child_level, int np,
                                    const long long* __restrict__ p_code, const int* __restrict__ p_start,
                                    const int* __restrict__ p_end, long long* __restrict__ c_code,
                                    int* __restrict__ c_start, int* __restrict__ c_end) {
  const int nc = 1 << d;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)np * nc) return;
  const int p = int(t / nc), k = int(t % nc);
  const long long code = p_code[p] * nc + k;
  const int shift = d * (D - child_level);
  const int s = p_start[p], e = p_end[p];
  c_code[t] = code;
  c_start[t] = lower_bound_u64(keys, s, e, (u64)code << shift);
  c_end[t] = lower_bound_u64(keys, s, e, (u64)(code + 1) << shift);
}

// ---------------------------------------------------------------- 2:1 balance
// Boxes of level l: every touching cell of the same size inside the domain; the deepest existing
// box containing it, if it sits more than one level above l and has no children, must split.
__global__ void balance_marks_kernel(int d, int l, const long long* __restrict__ codes,
                                     const long long* __restrict__ sep, const int* __restrict__ split,
                                     int* __restrict__ mark) {
  const long long b = sep[l] + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= sep[l + 1]) return;
  long long ic[3];
  deinterleave(codes[b], d, l, ic);
  const long long side = 1LL << l;
  const int noff = d == 3 ? 27 : 9;
  for (int o = 0; o < noff; ++o) {
    long long nb[3] = {0, 0, 0};
    int t = o;
    bool zero = true, inside = true;
    for (int k = d - 1; k >= 0; --k) {
      const int off = t % 3 - 1;
      t /= 3;
      zero = zero && off == 0;
      nb[k] = ic[k] + off;
      inside = inside && nb[k] >= 0 && nb[k] < side;
    }
    if (zero || !inside) continue;
    for (int lp = l; lp >= 0; --lp) {
      long long sub[3];
      for (int k = 0; k < d; ++k) sub[k] = nb[k] >> (l - lp);
      const long long c = lp > 0 ? interleave(sub, d, lp) : 0;
      const long long id = find_code(codes, sep[lp], sep[lp + 1], c);
      if (id >= 0) {
        if (lp < l - 1 && !split[id]) mark[id] = 1;
        break;
      }
    }
  }
}

// ----------------------------------------------------------------- box tables
__global__ void box_level_kernel(const long long* __restrict__ sep, int nlevels, long long nb, int d,
                                 const long long* __restrict__ codes, int* __restrict__ level,
                                 int* __restrict__ icoord) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int l = 0;
  while (l + 1 < nlevels && sep[l + 1] <= b) ++l;
  level[b] = l;
  long long ic[3];
  deinterleave(codes[b], d, l, ic);
  for (int k = 0; k < d; ++k) icoord[b * d + k] = int(ic[k]);
}

// thread (box, k): children[b][k] = id of code * 2^d + k on the next level or -1; parent of it = b
__global__ void children_kernel(const long long* __restrict__ sep, int nlevels, long long nb, int d,
                                const long long* __restrict__ codes, const int* __restrict__ level,
                                long long* __restrict__ children, long long* __restrict__ parent) {
  const int nc = 1 << d;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * nc) return;
  const long long b = t / nc;
  const int k = int(t % nc);
  const int l = level[b];
  long long id = -1;
  if (l + 1 < nlevels) id = find_code(codes, sep[l + 1], sep[l + 2], codes[b] * nc + k);
  children[t] = id;
  if (id >= 0) parent[id] = b;
}

// ------------------------------------------------------------ neighbour rows
__device__ inline void sort_small(long long* v, int n) {
  for (int i = 1; i < n; ++i) {
    const long long x = v[i];
    int j = i - 1;
    while (j >= 0 && v[j] > x) {
      v[j + 1] = v[j];
      --j;
    }
    v[j + 1] = x;
  }
}

// rows [b0, b1): same-level touching boxes including the box itself, ascending, padded with -1
__global__ void colleagues_kernel(int d, const long long* __restrict__ codes, const long long* __restrict__ sep,
                                  const int* __restrict__ level, long long b0, long long b1,
                                  long long* __restrict__ out) {
  const long long b = b0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= b1) return;
  const int noff = d == 3 ? 27 : 9;
  const int l = level[b];
  long long ic[3];
  deinterleave(codes[b], d, l, ic);
  const long long side = 1LL << l;
  long long v[27];
  int nv = 0;
  for (int o = 0; o < noff; ++o) {
    long long nb[3] = {0, 0, 0};
    int t = o;
    bool inside = true;
    for (int k = d - 1; k >= 0; --k) {
      nb[k] = ic[k] + (t % 3 - 1);
      t /= 3;
      inside = inside && nb[k] >= 0 && nb[k] < side;
    }
    if (!inside) continue;
    const long long c = l > 0 ? interleave(nb, d, l) : 0;
    const long long id = find_code(codes, sep[l], sep[l + 1], c);
    if (id >= 0) v[nv++] = id;
  }
  sort_small(v, nv);
  long long* row = out + (b - b0) * noff;
  for (int i = 0; i < noff; ++i) row[i] = i < nv ? v[i] : -1;
}

// rows [b0, b1): leaves on coarser levels touching the box (the deepest existing box containing a
// touching cell of the box's size, kept when it is a leaf above the box's level), ascending, unique
__global__ void coarse_kernel(int d, const long long* __restrict__ codes, const long long* __restrict__ sep,
                              const int* __restrict__ level, const unsigned char* __restrict__ is_leaf, long long b0,
                              long long b1, long long* __restrict__ out) {
  const long long b = b0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= b1) return;
  const int noff = d == 3 ? 27 : 9;
  const int l = level[b];
  long long* row = out + (b - b0) * noff;
  long long v[27];
  int nv = 0;
  if (l > 0) {
    long long ic[3];
    deinterleave(codes[b], d, l, ic);
    const long long side = 1LL << l;
    for (int o = 0; o < noff; ++o) {
      long long nb[3] = {0, 0, 0};
      int t = o;
      bool zero = true, inside = true;
      for (int k = d - 1; k >= 0; --k) {
        const int off = t % 3 - 1;
        t /= 3;
        zero = zero && off == 0;
        nb[k] = ic[k] + off;
        inside = inside && nb[k] >= 0 && nb[k] < side;
      }
      if (zero || !inside) continue;
      for (int lp = l; lp >= 0; --lp) {
        long long sub[3];
        for (int k = 0; k < d; ++k) sub[k] = nb[k] >> (l - lp);
        const long long c = lp > 0 ? interleave(sub, d, lp) : 0;
        const long long id = find_code(codes, sep[lp], sep[lp + 1], c);
        if (id >= 0) {
          if (is_leaf[id] && level[id] < l) v[nv++] = id;
          break;
        }
      }
    }
    sort_small(v, nv);
    int u = 0;
    for (int i = 0; i < nv; ++i)
      if (u == 0 || v[i] != v[u - 1]) v[u++] = v[i];
    nv = u;
  }
  for (int i = 0; i < noff; ++i) row[i] = i < nv ? v[i] : -1;
}

// ------------------------------------------------------------------ pair lists
// rows of boxes [b0, b1): leaf pairs (both boxes leaves), tree pairs (other colleagues), coarse
// pairs, in box order and row order. fill = 0: counts[3 (b - b0) + list]; fill = 1: write the
// pairs at the exclusive offsets offs[3 (b - b0) + list]
__global__ void pairs_kernel(int width, const long long* __restrict__ coll, const long long* __restrict__ coarse,
                             const unsigned char* __restrict__ is_leaf, long long b0, long long b1, int fill,
                             long long* __restrict__ counts, const long long* __restrict__ offs,
                             long long* __restrict__ leaf_pairs, long long* __restrict__ tree_pairs,
                             long long* __restrict__ coarse_pairs) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= b1 - b0) return;
  const long long b = b0 + r;
  const long long* crow = coll + r * width;
  const long long* xrow = coarse + r * width;
  long long nl = 0, nt = 0, nx = 0;
  long long ol = 0, ot = 0, ox = 0;
  if (fill) {
    ol = offs[3 * r];
    ot = offs[3 * r + 1];
    ox = offs[3 * r + 2];
  }
  for (int i = 0; i < width; ++i) {
    const long long c = crow[i];
    if (c < 0) continue;
    if (is_leaf[b] && is_leaf[c]) {
      if (fill) {
        leaf_pairs[2 * (ol + nl)] = b;
        leaf_pairs[2 * (ol + nl) + 1] = c;
      }
      ++nl;
    } else {
      if (fill) {
        tree_pairs[2 * (ot + nt)] = b;
        tree_pairs[2 * (ot + nt) + 1] = c;
      }
      ++nt;
    }
  }
  for (int i = 0; i < width; ++i) {
    const long long c = xrow[i];
    if (c < 0) continue;
    if (fill) {
      coarse_pairs[2 * (ox + nx)] = b;
      coarse_pairs[2 * (ox + nx) + 1] = c;
    }
    ++nx;
  }
  if (!fill) {
    counts[3 * r] = nl;
    counts[3 * r + 1] = nt;
    counts[3 * r + 2] = nx;
  }
}

int check_dim(int d) { return (d == 2 || d == 3) ? SKF_OK : terr(SKF_ERR_ARG, "dimension must be 2 or 3"); }

}  // namespace

extern "C" const char* skf_tree_last_error(void) { return t_err; }

extern "C" int skf_tree_keys(const double* pts, int n, int ndim, const double* lo, double L0, int depth, void* keys,
                             int* idx, void* stream) {
  int rc = check_dim(ndim);
  if (rc) return rc;
  if (n < 0 || !lo || !(L0 > 0.0)) return terr(SKF_ERR_ARG, "bad key arguments");
  if (depth < 1 || depth * ndim > 63) return terr(SKF_ERR_ARG, "key depth %d out of range", depth);
  if (n == 0) return SKF_OK;
  if (!pts || !keys || !idx) return terr(SKF_ERR_ARG, "null key buffers");
  const double scale = double(1LL << depth);
  hipLaunchKernelGGL(keys_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), pts, n, ndim,
                     lo[0], lo[1], ndim == 3 ? lo[2] : 0.0, L0, scale, (1LL << depth) - 1, depth,
                     static_cast<u64*>(keys), idx);
  TREE_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_tree_sort_pairs(void* temp, unsigned long long* temp_bytes, const void* keys_in, void* keys_out,
                                   const int* vals_in, int* vals_out, int n, int end_bit, void* stream) {
  if (!temp_bytes) return terr(SKF_ERR_ARG, "null temp_bytes");
  if (n < 0 || end_bit < 1 || end_bit > 64) return terr(SKF_ERR_ARG, "bad sort arguments");
  size_t bytes = size_t(*temp_bytes);
  hipError_t e = rocprim::radix_sort_pairs(temp, bytes, static_cast<const u64*>(keys_in), static_cast<u64*>(keys_out),
                                           vals_in, vals_out, size_t(n), 0u, unsigned(end_bit),
                                           static_cast<hipStream_t>(stream));
  *temp_bytes = (unsigned long long)bytes;
  if (e != hipSuccess) return terr(SKF_ERR_HIP, "radix sort: %s", hipGetErrorString(e));
  return SKF_OK;
}

extern "C" int skf_tree_child_ranges(const void* keys, int ndim, int depth, int child_level, int nparent,
                                     const long long* p_code, const int* p_start, const int* p_end, long long* c_code,
                                     int* c_start, int* c_end, void* stream) {
  int rc = check_dim(ndim);
  if (rc) return rc;
  if (child_level < 1 || child_level > depth || nparent < 0) return terr(SKF_ERR_ARG, "bad level arguments");
  if (nparent == 0) return SKF_OK;
  const long long nt = (long long)nparent << ndim;
  hipLaunchKernelGGL(child_ranges_kernel, dim3(blocks_for(nt)), dim3(kBlock), 0, static_cast<hipStream_t>(stream),
                     static_cast<const u64*>(keys), ndim, depth, child_level, nparent, p_code, p_start, p_end, c_code,
                     c_start, c_end);
  TREE_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_tree_balance_marks(int ndim, int level, long long nbox_level, const long long* codes,
                                      const long long* sep, const int* split, int* mark, void* stream) {
  int rc = check_dim(ndim);
  if (rc) return rc;
  if (level < 0 || nbox_level < 0) return terr(SKF_ERR_ARG, "bad balance arguments");
  if (nbox_level == 0) return SKF_OK;
  hipLaunchKernelGGL(balance_marks_kernel, dim3(blocks_for(nbox_level)), dim3(kBlock), 0,
                     static_cast<hipStream_t>(stream), ndim, level, codes, sep, split, mark);
  TREE_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_tree_box_tables(int ndim, int nlevels, long long nboxes, const long long* codes,
                                   const long long* sep, int* level, int* icoord, long long* children,
                                   long long* parent, void* stream) {
  int rc = check_dim(ndim);
  if (rc) return rc;
  if (nlevels < 1 || nboxes < 1) return terr(SKF_ERR_ARG, "bad box-table arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(box_level_kernel, dim3(blocks_for(nboxes)), dim3(kBlock), 0, st, sep, nlevels, nboxes, ndim, codes,
                     level, icoord);
  TREE_LAUNCHED();
  hipLaunchKernelGGL(children_kernel, dim3(blocks_for(nboxes << ndim)), dim3(kBlock), 0, st, sep, nlevels, nboxes,
                     ndim, codes, level, children, parent);
  TREE_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_tree_colleagues(int ndim, const long long* codes, const long long* sep, const int* level,
                                   long long b0, long long b1, long long* out, void* stream) {
  int rc = check_dim(ndim);
  if (rc) return rc;
  if (b0 < 0 || b1 < b0) return terr(SKF_ERR_ARG, "bad box range");
  if (b1 == b0) return SKF_OK;
  hipLaunchKernelGGL(colleagues_kernel, dim3(blocks_for(b1 - b0)), dim3(kBlock), 0, static_cast<hipStream_t>(stream),
                     ndim, codes, sep, level, b0, b1, out);
  TREE_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_tree_coarse(int ndim, const long long* codes, const long long* sep, const int* level,
                               const unsigned char* is_leaf, long long b0, long long b1, long long* out, void* stream) {
  int rc = check_dim(ndim);
  if (rc) return rc;
  if (b0 < 0 || b1 < b0) return terr(SKF_ERR_ARG, "bad box range");
  if (b1 == b0) return SKF_OK;
  hipLaunchKernelGGL(coarse_kernel, dim3(blocks_for(b1 - b0)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), ndim,
                     codes, sep, level, is_leaf, b0, b1, out);
  TREE_LAUNCHED();
  return SKF_OK;
}

extern "C" int skf_tree_pairs(int width, const long long* coll, const long long* coarse, const unsigned char* is_leaf,
                              long long b0, long long b1, int fill, long long* counts, const long long* offsets,
                              long long* leaf_pairs, long long* tree_pairs, long long* coarse_pairs, void* stream) {
  if (width != 9 && width != 27) return terr(SKF_ERR_ARG, "row width must be 9 or 27");
  if (b0 < 0 || b1 < b0) return terr(SKF_ERR_ARG, "bad box range");
  if (b1 == b0) return SKF_OK;
  if (fill ? (!offsets || !leaf_pairs || !tree_pairs || !coarse_pairs) : !counts)
    return terr(SKF_ERR_ARG, "null pair-list buffers");
  hipLaunchKernelGGL(pairs_kernel, dim3(blocks_for(b1 - b0)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), width,
                     coll, coarse, is_leaf, b0, b1, fill, counts, offsets, leaf_pairs, tree_pairs, coarse_pairs);
  TREE_LAUNCHED();
  return SKF_OK;
}
