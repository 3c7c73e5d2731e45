/*
 * p2p_ref.c -- C restatement of the dense kernel sums (oracle; test
 * infrastructure only, never linked into the product).
 *
 * out[i][c] = sum_j K(xt_i, xs_j) q[j][c] in FP64 with the closed forms of
 * skelfmm/util.py:128-377 (K = 0 where the two points coincide, no clamp).
 * Complex kernels (Helmholtz) take complex charges as interleaved (re, im)
 * pairs: width = 2 * nrhs. Real kernels treat every real channel separately.
 * Used for the FP64 direct-sum error at sampled targets of full-size problems.
 *
 * Build: gcc -O2 -fopenmp -fPIC -shared p2p_ref.c -o liboracle_p2p.so -lm
 */
#include <math.h>
#include <stddef.h>

#define INV4PI 0.079577471545947667884441881686257181

/* kernel ids as in include/skelfmm_hip.h */
static int is_complex(int k) { return k == 2 || k == 3 || k == 6 || k == 7; }

static void kval(int k, double dx, double dy, double dz, double rn, double kappa, double* re, double* im) {
  double r2 = dx * dx + dy * dy + dz * dz;
  double r = sqrt(r2);
  *re = 0.0;
  *im = 0.0;
  if (r == 0.0) return;
  switch (k) {
    case 0: *re = log(r); break;
    case 1: *re = 1.0 / r; break;
    case 2: *re = j0(kappa * r); *im = y0(kappa * r); break;
    case 3: *re = cos(kappa * r) / r; *im = sin(kappa * r) / r; break;
    case 4: *re = -INV4PI * rn / (r2 * r); break;
    case 5: *re = INV4PI * rn / (r2 * r); break;
    case 6:
    case 7: {
      double kr = kappa * r, c = cos(kr), s = sin(kr);
      double f = INV4PI * rn / (r2 * r);
      if (k == 7) f = -f;
      *re = f * (-c - kr * s);
      *im = f * (kr * c - s);
      break;
    }
    default: break;
  }
}

/* xt: (nt, ndim), xs: (ns, ndim), gt: (nt, 3) or NULL, gs: (ns, 3) or NULL,
   q: (ns, width), out: (nt, width). Returns 0, or -1 on bad arguments. */
int oracle_direct_sum(int kernel, int ndim, long nt, const double* xt, long ns, const double* xs,
                      const double* gt, const double* gs, const double* q, int width, double kappa,
                      double* out) {
  if (ndim != 2 && ndim != 3) return -1;
  if (width <= 0 || width > 256) return -1;
  const int cx = is_complex(kernel);
  if (cx && (width % 2)) return -1;
  if ((kernel == 4 || kernel == 6) && !gs) return -1;
  if ((kernel == 5 || kernel == 7) && !gt) return -1;
#pragma omp parallel for schedule(dynamic, 4)
  for (long i = 0; i < nt; ++i) {
    double acc[256];
    for (int c = 0; c < width; ++c) acc[c] = 0.0;
    const double* x = xt + i * ndim;
    for (long j = 0; j < ns; ++j) {
      const double* y = xs + j * ndim;
      double dx = x[0] - y[0], dy = x[1] - y[1], dz = ndim == 3 ? x[2] - y[2] : 0.0;
      double rn = 0.0;
      if (kernel == 4 || kernel == 6) rn = dx * gs[3 * j] + dy * gs[3 * j + 1] + dz * gs[3 * j + 2];
      if (kernel == 5 || kernel == 7) rn = dx * gt[3 * i] + dy * gt[3 * i + 1] + dz * gt[3 * i + 2];
      double re, im;
      kval(kernel, dx, dy, dz, rn, kappa, &re, &im);
      const double* qj = q + j * width;
      if (cx) {
        for (int c = 0; c < width; c += 2) {
          acc[c] += re * qj[c] - im * qj[c + 1];
          acc[c + 1] += re * qj[c + 1] + im * qj[c];
        }
      } else {
        for (int c = 0; c < width; ++c) acc[c] += re * qj[c];
      }
    }
    for (int c = 0; c < width; ++c) out[i * width + c] = acc[c];
  }
  return 0;
}
