/*
 * skelfmm_hip.h -- C ABI of the MI355X (gfx950) apply library libskelfmm_hip.so.
 *
 * All device buffers are owned by the caller (allocated through PyTorch's
 * caching allocator or hipMalloc) and passed as plain pointers. The library
 * allocates nothing on the apply path. Every entry point returns 0 on success
 * and a nonzero skf_status otherwise; skf_last_error() then holds a message.
 * Launches are asynchronous on the stream passed in (a hipStream_t cast to
 * void*, NULL = default stream); the library never synchronizes the host.
 *
 * Reference interfaces replaced (files of annayesy/skelFMM):
 *   skf_apply        SkelFMMBatched.matvec            skelfmm/batched_fmm.py:539-564
 *                    (get_qorig 441-447, compute_qskel 450-467, addto_uvec 496-511,
 *                     subtract_from_uskel 514-534, resultvec_from_uskel 469-483,
 *                     get_result 486-490)
 *   skf_p2p          translation()/translation_helper()/kernel_apply()
 *                                                     skelfmm/batched_fmm.py:24-44,181-372
 *                    and the operator kernels         skelfmm/util_batched.py:226-656
 *   skf_up_level     apply_proj_leaves/levtree(up) + addto_skel_lev(up)
 *                                                     skelfmm/batched_fmm.py:381-438
 *   skf_down_level   addto_skel_lev(down) + apply_proj_levtree/leaves(down)
 *                                                     skelfmm/batched_fmm.py:381-438,469-483
 *   skf_gather_charges / skf_scatter_result
 *                    get_qorig / get_result           skelfmm/batched_fmm.py:441-447,486-490
 *   skf_p2p_blocks   KernelOperator.torch_apply on (B, m, d) blocks
 *                                                     skelfmm/operators.py:69-89,
 *                                                     skelfmm/util_batched.py:240-656
 */
#ifndef SKELFMM_HIP_H
#define SKELFMM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* operator codes (skelfmm_amd.operators.KID_*) */
enum skf_kernel {
  SKF_LAPLACE_2D = 0,
  SKF_LAPLACE_3D = 1,
  SKF_HELMHOLTZ_2D = 2,
  SKF_HELMHOLTZ_3D = 3,
  SKF_LAPLACE_DL_3D = 4,
  SKF_LAPLACE_ADL_3D = 5,
  SKF_HELMHOLTZ_DL_3D = 6,
  SKF_HELMHOLTZ_ADL_3D = 7,
  /* user-defined real point kernels, near field compiled at run time (skf_jit_compile) */
  SKF_USER_2D = 16,
  SKF_USER_3D = 17
};

enum skf_status {
  SKF_OK = 0,
  SKF_ERR_ARG = 1,      /* bad argument / shape */
  SKF_ERR_HIP = 2,      /* a HIP runtime call failed */
  SKF_ERR_UNSUPPORTED = 3
};

/* source-entry flag bits */
#define SKF_SRC_NEG 1  /* subtract this source segment's contribution */
#define SKF_SRC_SELF 2 /* the segment is the target segment itself: skip i == j */

/*
 * Near-field work: target segments ("groups") of consecutive point slots, each
 * with a list of source segments. Every work item covers up to 64 targets of
 * one group and a contiguous range of that group's source entries.
 *
 *   group g : points pt[g] .. pt[g]+n[g]-1, results accumulated into
 *             u[out[g] + i] (i < n[g]); box id box[g] (precision form centre)
 *   entry s : points pt[s] .. pt[s]+n[s]-1 with charges q[qoff[s] + j];
 *             flags SKF_SRC_*; box id box[s]
 *   item    : int4 {group, first target, first entry, end entry}; an item
 *             with no entries (end entry <= first entry) is skipped, so a
 *             launch order may pad the table with them
 *
 * Vectors hold `width` real scalars per slot ([slot][width]); for complex
 * kernels `width` = 2*nrhs (interleaved re/im). Real kernels treat complex
 * vectors as 2*nrhs real channels.
 */
/*
 * Split near-field work for the kernels evaluated once per point pair (the single layers,
 * Laplace / Helmholtz in 2D and 3D, and the 3D double layers): the same group / entry / item
 * tables as below, once for the pairs of lists that interact in both directions ("sym": one
 * representative entry per pair, items of up to 512 targets) and once for the rest ("one": self
 * segments and unpaired entries, items of up to 64 targets). A sym entry s of group g adds
 * K(t, s) q[e_q + j] into u[g_out + i] and K(s, t) q[g_out + i] into u[e_q + j] (K(s, t) =
 * K(t, s) for the single layers; for the double layers the same g(|t - s|) with the other
 * point's normal projection): the reverse output of an entry is its charge offset (a set's
 * charges and potentials share slots and kind).
 */
typedef struct skf_p2p_lists {
  int ngroups;
  const int* g_pt;
  const int* g_n;
  const int* g_out;
  const int* g_box;
  int nentries;
  const int* e_pt;
  const int* e_n;
  const int* e_q;
  const int* e_box;
  const int* e_flags;
  int nitems;
  const int* items;
  const float* e_shift; /* precision form, optional: as skf_p2p_args.e_shift, for these entries */
} skf_p2p_lists;

typedef struct skf_p2p_args {
  int kernel;         /* enum skf_kernel */
  int precision;      /* 0: FP64 points/vectors, 1: FP32 (box-centred points, FP64 centres) */
  int width;          /* real scalars per vector slot */
  double kappa;       /* wavenumber (Helmholtz), 0 otherwise */
  const void* px;     /* point coordinates, SoA, Real[nslots] each */
  const void* py;
  const void* pz;     /* NULL in 2D */
  const void* nx;     /* normals (double-layer kernels) or NULL */
  const void* ny;
  const void* nz;
  const double* centers; /* (nbox, 3) box centres, precision form only */
  const void* q;      /* charges, Real[...][width] */
  void* u;            /* potentials, Real[...][width], accumulated */
  int ngroups;
  const int* g_pt;
  const int* g_n;
  const int* g_out;
  const int* g_box;
  int nentries;
  const int* e_pt;
  const int* e_n;
  const int* e_q;
  const int* e_box;
  const int* e_flags;
  int nitems;
  const int* items;   /* int4 per item */
  int mfma_min_channels; /* real kernels (FP32 or FP64) with at least this many channels use the
                            matrix-core kernel; 0 = default (16), negative = never */
  /* optional symmetric split (one channel, symmetric kernel): used instead of the lists above
     when both are set; the FP32 form needs `urev`, an FP64 buffer of urev_len scalars covering
     every u index (2 * nslots * width in the apply), zeroed and folded into u by the call */
  const skf_p2p_lists* one;
  const skf_p2p_lists* sym;
  double* urev;
  long long urev_len;
  int sym_unroll;              /* sources per inner step of the symmetric kernel: 1, 2, 4 (0 = 2) */
  /* kernel = SKF_USER_2D / SKF_USER_3D: run-time compiled near-field kernels (hipFunction_t from
     skf_jit_compile) for channel blocks of 1, 2, 4 and 8, in this argument's precision */
  void* user_fn[4];
  /* an upper bound of every pair distance (e.g. the point cloud's diameter), 0 = unknown; with
     it the 3D Helmholtz kernels evaluate e^{i kappa r} with a branch-free sine/cosine (FP32:
     the hardware sine/cosine of the phase in revolutions when kappa * max_distance <= 1600;
     FP64: reduction and polynomials when kappa * max_distance <= 65536) */
  double max_distance;
  /* precision form of the double-layer kernels (required there, ignored otherwise): FP64 rows
     [7][geo64_n] per slot: absolute x, y, z, normal nx, ny, nz, and n . x. The normal projection
     rn = (x - y) . n of a pair is formed from these in FP64: on a surface rn ~ r^2 / 2 for close
     pairs, far below the FP32 rounding of box-centred coordinates (~6e-8 of a box side) */
  const double* geo64;
  long long geo64_n;
  /* precision form, optional: per entry the shift of its points into its group's box frame,
     3 floats (centre of the group's box minus centre of the entry's box, formed in FP64 and
     rounded once); NULL: the kernels form it from `centers` (one more dependent load when a work
     item starts) */
  const float* e_shift;
  /* 1: every output row of `items` belongs to exactly one work item (no group's sources are
     split over several items) and u holds zero there before the call; the matrix-core kernel
     (>= mfma_min_channels real channels, LDS-direct staging) then stores its sums instead of
     adding them with atomics. 0: accumulate (the general contract) */
  int store_rows;
} skf_p2p_args;

/* one level of the upward or downward interpolation pass */
typedef struct skf_level_args {
  int nboxes;                 /* boxes of the level in the tables below */
  const int* box_off;         /* first slot of the box */
  const int* box_bs;          /* candidate count */
  const int* box_rank;        /* skeleton count */
  const long long* box_toff;  /* element offset of T (rank x (bs-rank), row-major) */
  int ngather;                /* upward: q[dst] = qskel[src]; downward: uskel[dst] += u[src], u[dst] += uskel[dst] */
  const int* gather_dst;
  const int* gather_src;
  int nitems;                 /* int2 {box index, first row (up) / column (down)} */
  const int* items;
  int variant;                /* launch shape (autotuned per level and width; 0 = default):
                                 one channel: up 0/1/2 = 2/1/4 vector loads per lane and step,
                                 down 0/1/2 = 8/4/16 rows per step; >= 16 real FP32 channels:
                                 0 = matrix cores (the downward stage needs a width that is
                                 a multiple of 4 and 16-byte aligned vectors, else VALU GEMM),
                                 1 = VALU GEMM; ignored otherwise */
} skf_level_args;

typedef struct skf_plan_desc {
  int precision;              /* 0 FP64, 1 FP32 */
  int complex_t;              /* T entries complex (Helmholtz kernels) */
  int width;                  /* real scalars per slot */
  int nslots;                 /* total candidate slots */
  int npoints;                /* N */
  const int* leaf_slot;       /* N entries: slot of the user's point i ... */
  const int* leaf_user;       /* ... and its user index */
  const void* T;              /* interpolation matrices, Real (or complex Real pairs) */
  void* qbuf;                 /* Real[2*nslots*width]: q_vec then q_skel (every slot the apply
                                 reads is written by it first: no clearing needed) */
  void* ubuf;                 /* Real[2*nslots*width]: u_vec then u_skel */
  int nup;                    /* upward stages, deepest level first (root gather last) */
  const skf_level_args* up;
  int ndown;                  /* downward stages, shallowest skeletonized level first */
  const skf_level_args* down;
  skf_p2p_args p2p;           /* q/u inside are ignored: qbuf/ubuf are used */
} skf_plan_desc;

typedef struct skf_plan skf_plan;

const char* skf_last_error(void);
int skf_version(void);
/* 16 hex digits: hash of the sources and compile flags the library was built from (no reference
   counterpart; lets a loader prove that the binary matches the source tree) */
const char* skf_build_id(void);

/* zero `n` real scalars of the given precision */
int skf_fill_zero(void* ptr, long long n, int precision, void* stream);

/* q_vec[leaf_slot[i]][:] = x[leaf_user[i]][:] ; y[leaf_user[i]][:] = u_vec[leaf_slot[i]][:] */
int skf_gather_charges(const void* x, void* qvec, const int* leaf_slot, const int* leaf_user, int n,
                       int width, int precision, void* stream);
int skf_scatter_result(const void* uvec, void* y, const int* leaf_slot, const int* leaf_user, int n,
                       int width, int precision, void* stream);

/* near-field translations; q and u as in skf_p2p_args */
int skf_p2p(const skf_p2p_args* a, void* stream);

/*
 * skf_p2p without any allocation by or for the library: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The FP32 symmetric near field's FP64 reverse-sum
 * buffer (urev) comes from a file-scope device array of SKF_NOALLOC_MAX_UREV doubles (32 MiB)
 * that each GPU holds for the life of the process; a->urev must be NULL and a->urev_len (the
 * scalars needed, 2 * nslots * width in the apply) at most SKF_NOALLOC_MAX_UREV, else
 * SKF_ERR_ARG. Calls issue no hipMalloc, hipFree or allocating library routine. The array is
 * shared by every call on a GPU and zeroed by each call before use: calls that overlap on one
 * GPU (other streams, other threads) race on it.
 */
#define SKF_NOALLOC_MAX_UREV (1LL << 22)
int skf_p2p_noalloc(const skf_p2p_args* a, void* stream);

/* one upward level: gather children's skeleton charges, then qskel = q[:r] + T q[r:bs] */
int skf_up_level(const skf_level_args* L, const void* T, int complex_t, void* qvec, void* qskel, int width,
                 int precision, void* stream);
/* one downward level: uskel += parent's u, u[:r] += uskel, u[r:bs] += T^T uskel */
int skf_down_level(const skf_level_args* L, const void* T, int complex_t, void* uvec, void* uskel, int width,
                   int precision, void* stream);

/* whole apply: y = K~ x (x, y in the caller's point order, [N][width]) */
int skf_plan_create(const skf_plan_desc* d, skf_plan** out);
int skf_plan_destroy(skf_plan* p);
int skf_apply(skf_plan* p, const void* x, void* y, void* stream);

/*
 * Dense blocks: u[b, i, :] (+)= sum_j K(xt[b,i], xs[b,j]) q[b, j, :] for B blocks of
 * m targets and n sources, AoS coordinates (B, m, 3|2), normals (B, m|n, 3) or NULL.
 * self_interaction zeroes i == j. accumulate = 0 overwrites u.
 */
int skf_p2p_blocks(int kernel, int precision, int width, double kappa, int nblocks, int m, int n, int ndim,
                   const void* xt, const void* xs, const void* nt, const void* ns, const void* q, void* u,
                   int self_interaction, void* stream);

/*
 * Run-time compiled near-field kernels for user-defined operators (SKF_USER_2D / _3D).
 * Reference: the KernelOperator plugin (skelfmm/operators.py:17-108) whose lazy kernels the
 * reference evaluates through PyKeOps formulas (skelfmm/util_batched.py:189-224); here the
 * operator's formula becomes a kernel_value<> specialization compiled with hiprtc.
 * source: the translation unit (it includes header_name); names: template instances to load
 * (e.g. "p2p_kernel<double, 17, 1, false>"). module = NULL compiles only (no device needed)
 * and reports the code size; otherwise the module is loaded and functions[i] receives the
 * hipFunction_t of names[i]. skf_jit_log() holds the compiler log of the last call.
 */
int skf_jit_compile(const char* source, const char* header_name, const char* header_source,
                    const char* const* names, int nnames, void** module, void** functions,
                    long long* code_bytes);
const char* skf_jit_log(void);
int skf_jit_unload(void* module);

/* ---------------------------------------------------------------- GPU build
 * Reference: the per-box interpolative decomposition of the build,
 * skelfmm/util.py:87-124 (build_stacked_proxy_matrix) and util.py:13-51
 * (_skeleton_id_pivoted_qr), batched over the boxes of one tree level
 * (the per-box loop of skelfmm/recursive_skel.py:412-491).
 * Complex (Helmholtz) matrices are interleaved (re, im) doubles.
 */
const char* skf_build_last_error(void);

/* A_b (row-major m_b x n_b with row stride lda, or n_b when lda = 0, at a_off[b] elements):
   rows = proxies, columns = the box's candidate
   points idx[idx_off[b] .. + ncols[b]] of pts (N x 3, z = 0 in 2D); proxies = centers[b] +
   side[b] * unit_proxy (nproxy x 3); stacked = 1 for [K_PB; K_BP^T] (m_b = 2 nproxy) with outward
   proxy normals, else K_BP^T (m_b = nproxy). */
int skf_proxy_matrices(int kernel, int ndim, double kappa, const double* pts, const double* normals, const int* idx,
                       const long long* idx_off, const int* ncols, const double* centers, const double* side,
                       const double* unit_proxy, int nproxy, int stacked, int complex_out, void* A,
                       const long long* a_off, int lda, int nbox, int max_cols, void* stream);

/* In-place column-pivoted Householder QR of every A_b (row stride lda[b], or n[b] when lda is NULL)
   for at most kmax[b] steps, stopping at the
   first |R_ii| <= tol |R_00|: rank[b] = that i (converged[b] = 1), or kmax[b] with converged[b] = 0
   when the cutoff was not reached within kmax[b] < min(m, n) steps. perm (at p_off[b]) holds the
   column pivots, rdiag the |R_ii|. */
int skf_pqr_batched(void* A, const long long* a_off, const int* lda, const int* m, const int* n, const int* kmax,
                    double tol, int* rank, int* converged, int* perm, const long long* p_off, double* rdiag, int nbox,
                    int max_m, int max_n, int complex_a, void* stream);

/* T_b = R11^{-1} R12 (rank x (n - rank), row-major at t_off[b]) from the factored A_b. */
int skf_trsm_batched(const void* A, const long long* a_off, const int* lda, const int* n, const int* rank, void* T,
                     const long long* t_off, int nbox, int max_cols, int complex_a, void* stream);

/* Proxy-residual check of the IDs (perm, rank, T as produced above) on sketches Z_b (c rows,
   row stride ldz[b] or n[b] when ldz is NULL, at z_off[b]) of the proxy matrices: res[b] = max_j |Z_red,j - Z_skel T_j|_2,
   znorm[b] = max_j |Z_j|_2. Used by the learned-rank build to accept or reject a sketched ID;
   the reference always factors the whole proxy matrix (util.py:13-51). */
int skf_id_residual(const void* Z, const long long* z_off, const int* ldz, int c, const int* n, const int* rank,
                    const int* perm, const long long* p_off, const void* T, const long long* t_off, int nbox,
                    int complex_a, double* res, double* znorm, void* stream);

/* ------------------------------------------------- learned rank predictor
 * No counterpart in the reference (which always runs the full adaptive ID):
 * an 8 -> 64 -> 64 -> 1 GELU MLP (FP32) mapping box features to skeleton rank.
 * params: W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3 (skf_mlp_param_count() floats).
 */
int skf_mlp_param_count(void);
int skf_mlp_forward(const float* params, const float* X, int n, float* y, void* stream);
/* adds d(mean squared error)/d(params) into grad and the mean squared error into *loss */
int skf_mlp_backward(const float* params, const float* X, const float* target, int n, float* grad, float* loss,
                     void* stream);
int skf_adam_step(float* params, const float* grad, float* m1, float* m2, int np, float lr, float beta1, float beta2,
                  float eps, int step, void* stream);

/* ------------------------------------------------------------- GPU tree and lists
 * Reference: the adaptive level-restricted 2^d-tree the reference imports from simpletree
 * (skelfmm/skel_fmm.py:150, BalancedTree; its contract as skelfmm/recursive_skel.py:54-68,142-166
 * uses it), the colleague / coarse neighbour queries get_boxes_colleague_neigh_fullvec and
 * get_boxes_coarse_neigh_fullvec, and the pair lists of get_parallel_lists
 * (skelfmm/recursive_skel.py:622-712). Boxes are numbered level by level, Morton order inside a
 * level; `codes` holds every box's Morton code (level-major, sorted inside a level) and `sep` the
 * first box of each level (nlevels + 1 entries). Keys are 64-bit unsigned.
 */
const char* skf_tree_last_error(void);
/* key of point i: Morton interleave of q_k = clip(floor(((x_k - lo_k) / L0) 2^depth)); idx[i] = i */
int skf_tree_keys(const double* pts, int n, int ndim, const double* lo, double L0, int depth, void* keys, int* idx,
                  void* stream);
/* stable radix sort of (key, value) pairs over key bits [0, end_bit); temp = NULL stores the
   workspace size in *temp_bytes */
int skf_tree_sort_pairs(void* temp, unsigned long long* temp_bytes, const void* keys_in, void* keys_out,
                        const int* vals_in, int* vals_out, int n, int end_bit, void* stream);
/* children of nparent boxes (codes, point ranges in sorted-key order) at child_level: entry
   p 2^ndim + k = code p_code[p] 2^ndim + k and its point range (empty when start == end) */
int skf_tree_child_ranges(const void* keys, int ndim, int depth, int child_level, int nparent,
                          const long long* p_code, const int* p_start, const int* p_end, long long* c_code,
                          int* c_start, int* c_end, void* stream);
/* mark[b] = 1 for every box b without children (split[b] == 0) that a box of `level` touches
   through a same-size cell while sitting more than one level above it (2:1 restriction) */
int skf_tree_balance_marks(int ndim, int level, long long nbox_level, const long long* codes, const long long* sep,
                           const int* split, int* mark, void* stream);
/* level[b], icoord[b][ndim], children[b][2^ndim] (-1 absent) and parent[b] (caller fills -1) */
int skf_tree_box_tables(int ndim, int nlevels, long long nboxes, const long long* codes, const long long* sep,
                        int* level, int* icoord, long long* children, long long* parent, void* stream);
/* rows of boxes [b0, b1), 3^ndim wide, padded with -1: same-level touching boxes incl. self */
int skf_tree_colleagues(int ndim, const long long* codes, const long long* sep, const int* level, long long b0,
                        long long b1, long long* out, void* stream);
/* rows of boxes [b0, b1): touching leaves on coarser levels, ascending */
int skf_tree_coarse(int ndim, const long long* codes, const long long* sep, const int* level,
                    const unsigned char* is_leaf, long long b0, long long b1, long long* out, void* stream);
/* pair lists of boxes [b0, b1) from their colleague and coarse rows: fill = 0 writes counts
   [3 per box: leaf pairs, tree pairs, coarse pairs]; fill = 1 writes (box, other) pairs at the
   exclusive offsets (3 per box) */
int skf_tree_pairs(int width, const long long* coll, const long long* coarse, const unsigned char* is_leaf,
                   long long b0, long long b1, int fill, long long* counts, const long long* offsets,
                   long long* leaf_pairs, long long* tree_pairs, long long* coarse_pairs, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SKELFMM_HIP_H */
