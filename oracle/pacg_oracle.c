/*
 * pacg oracle -- CPU restatement of the permutation-avoiding convolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
 * else.  The product library (libpacg.so) never links, loads or calls it.
 *
 * It restates, in plain double-precision C, what the reference computes:
 *
 *   twiddles      w_k^{mj} = exp(-2 pi i m j / k)          twiddle_factors.c:17-131
 *   A             DIT stages q = 0..t-1, leg m * w^{mj}, F_r  butterflies.c:12-39,99-156,277-389
 *   conj(A)       same with conj(w), conj(F_r)              butterflies.c:41-68,158-215,391-503
 *   A^T           DIF stages q = t-1..0, F_r then * w^{mj}   butterflies.c:70-97,217-275,505-632
 *   2D/3D         per-axis application, contiguous axis first butterflies_2d.c:12-49, butterflies_3d.c:12-78
 *   digit reversal base-r, composed per axis                 index_reversals.c:13-72, _2d.c:14-62, _3d.c:14-72
 *   fft/fft2/ufft/ifft/uifft                                 fft.c:12-88, fft_2d.c, fft_3d.c
 *   PA convolution  A^T x, A^T h, (x*h)/n, conj(A)           convolution.c:31-72, _2d.c:35-70, _3d.c:36-72
 *   FFT-based convolution                                    convolution.c:12-29, _2d.c:12-33, _3d.c:12-34
 *   max error     max_i hypot(dre, dim)                      misc.c:20-30
 *
 * Generalisation: a stage schedule sched[0..t-1] (sched[q] = radix of stage q,
 * stage 0 has unit stride) replaces the single radix r.  With every entry
 * equal to r it reduces to the reference's radix-r kernels; mixed schedules
 * (e.g. 2048 = 4^5 * 2) are what the reference cannot run (its stage loop
 * needs n = r^t exactly, twiddle_factors.c:40-45).
 *
 * Mixed-radix digit reversal: with l_q = prod_{s<q} sched[s] and
 * k_q = l_q * sched[q], position p has digits d_q = (p / l_q) mod sched[q] and
 * rho(p) = sum_q d_q * (n / k_q).  For a pure radix this is
 * reverse_index_r{2,4,8} (index_reversals.c:13-20,30-37,52-59).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAXT 64

enum { OR_A = 0, OR_CONJ = 1, OR_TRANS = 2 };

/* exp(-2 pi i p / r) for r in {2,4,8}; the values are exact except the
 * +-1/sqrt(2) entries, which are the correctly rounded double (the reference
 * forms them as 1./sqrt(2.), butterflies.c:286-290). */
static void unit_root(int r, int p, double* c, double* s) {
  static const double h = 0.70710678118654752440;
  static const double C[8] = {1.0, h, 0.0, -h, -1.0, -h, 0.0, h};
  static const double S[8] = {0.0, -h, -1.0, -h, 0.0, h, 1.0, h};
  p %= r;
  if (p < 0) p += r;
  int e = p * (8 / r);
  *c = C[e];
  *s = S[e];
}

/* In-place r-point DFT, natural order in and out; conj selects exp(+...). */
static void small_dft(int r, double* xr, double* xi, int conj) {
  double yr[8], yi[8];
  for (int a = 0; a < r; ++a) {
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < r; ++b) {
      double c, s;
      unit_root(r, a * b, &c, &s);
      if (conj) s = -s;
      sr += c * xr[b] - s * xi[b];
      si += c * xi[b] + s * xr[b];
    }
    yr[a] = sr;
    yi[a] = si;
  }
  memcpy(xr, yr, (size_t)r * sizeof(double));
  memcpy(xi, yi, (size_t)r * sizeof(double));
}

/* One stage q on a contiguous length-n vector (twiddle_factors.c:30-35 angle,
 * butterflies.c loop nest q / b / j). */
static void one_stage(double* re, double* im, long n, int r, long ell, int kind) {
  long k = (long)r * ell;
  double* wr = (double*)malloc((size_t)(r * ell) * sizeof(double));
  double* wi = (double*)malloc((size_t)(r * ell) * sizeof(double));
  for (int m = 0; m < r; ++m)
    for (long j = 0; j < ell; ++j) {
      double angle = -2.0 * M_PI * (double)(m * j) / (double)k;
      wr[m * ell + j] = cos(angle);
      wi[m * ell + j] = sin(angle);
    }
  double xr[8], xi[8];
  for (long b = 0; b < n; b += k) {
    for (long j = 0; j < ell; ++j) {
      for (int m = 0; m < r; ++m) {
        xr[m] = re[b + m * ell + j];
        xi[m] = im[b + m * ell + j];
      }
      if (kind == OR_TRANS) {
        small_dft(r, xr, xi, 0);
        for (int m = 1; m < r; ++m) {
          double c = wr[m * ell + j], s = wi[m * ell + j];
          double a = xr[m], d = xi[m];
          xr[m] = c * a - s * d;
          xi[m] = c * d + s * a;
        }
      } else {
        double sg = (kind == OR_CONJ) ? -1.0 : 1.0;
        for (int m = 1; m < r; ++m) {
          double c = wr[m * ell + j], s = sg * wi[m * ell + j];
          double a = xr[m], d = xi[m];
          xr[m] = c * a - s * d;
          xi[m] = c * d + s * a;
        }
        small_dft(r, xr, xi, kind == OR_CONJ);
      }
      for (int m = 0; m < r; ++m) {
        re[b + m * ell + j] = xr[m];
        im[b + m * ell + j] = xi[m];
      }
    }
  }
  free(wr);
  free(wi);
}

/* A, conj(A) or A^T on a contiguous vector of length n = prod sched. */
void or_butterflies(double* re, double* im, long n, const int* sched, int t, int kind) {
  long ell[OR_MAXT];
  long l = 1;
  for (int q = 0; q < t; ++q) {
    ell[q] = l;
    l *= sched[q];
  }
  if (kind == OR_TRANS) {
    for (int q = t - 1; q >= 0; --q) one_stage(re, im, n, sched[q], ell[q], kind);
  } else {
    for (int q = 0; q < t; ++q) one_stage(re, im, n, sched[q], ell[q], kind);
  }
}

/* Row-major nd array (nd <= 3), dims[0] outermost; scheds is nd x OR_MAXT.
 * Per-axis application, innermost (contiguous) axis first, the other axes by
 * gather / 1D / scatter of strided lines (butterflies_3d.c:22-74). */
void or_butterflies_nd(double* re, double* im, int nd, const long* dims, const int* scheds,
                       const int* ts, int kind) {
  long total = 1;
  for (int a = 0; a < nd; ++a) total *= dims[a];
  long maxd = 1;
  for (int a = 0; a < nd; ++a)
    if (dims[a] > maxd) maxd = dims[a];
  double* tr = (double*)malloc((size_t)maxd * sizeof(double));
  double* ti = (double*)malloc((size_t)maxd * sizeof(double));
  for (int a = nd - 1; a >= 0; --a) {
    long len = dims[a];
    long inner = 1;
    for (int b = a + 1; b < nd; ++b) inner *= dims[b];
    long outer = total / (len * inner);
    const int* sched = scheds + a * OR_MAXT;
    for (long o = 0; o < outer; ++o) {
      for (long i = 0; i < inner; ++i) {
        long base = o * len * inner + i;
        if (inner == 1) {
          or_butterflies(re + base, im + base, len, sched, ts[a], kind);
          continue;
        }
        for (long p = 0; p < len; ++p) {
          tr[p] = re[base + p * inner];
          ti[p] = im[base + p * inner];
        }
        or_butterflies(tr, ti, len, sched, ts[a], kind);
        for (long p = 0; p < len; ++p) {
          re[base + p * inner] = tr[p];
          im[base + p * inner] = ti[p];
        }
      }
    }
  }
  free(tr);
  free(ti);
}

/* reverse_index_r{2,4,8} (index_reversals.c:13-20,30-37,52-59): t base-r digits. */
int or_reverse_index_r(int x, int t, int r) {
  int bits = (r == 2) ? 1 : (r == 4) ? 2 : 3;
  int y = 0;
  for (int i = 0; i < t; ++i) {
    y = (y << bits) | (x & (r - 1));
    x >>= bits;
  }
  return y;
}

/* Mixed-radix digit reversal of position p under a stage schedule. */
long or_rev_index_mixed(long p, long n, const int* sched, int t) {
  long out = 0, l = 1;
  for (int q = 0; q < t; ++q) {
    long d = (p / l) % sched[q];
    long k = l * sched[q];
    out += d * (n / k);
    l = k;
  }
  return out;
}

/* rho for an nd array: rho[(i1*n2+i2)*n3+i3] = (r1(i1)*n2 + r2(i2))*n3 + r3(i3)
 * (index_reversals_2d.c:14-35, index_reversals_3d.c:14-39). */
void or_perm_nd(int nd, const long* dims, const int* scheds, const int* ts, int64_t* rho) {
  long total = 1;
  for (int a = 0; a < nd; ++a) total *= dims[a];
  for (long idx = 0; idx < total; ++idx) {
    long rem = idx, acc = 0, mul = 1;
    for (int a = nd - 1; a >= 0; --a) {
      long i = rem % dims[a];
      rem /= dims[a];
      acc += or_rev_index_mixed(i, dims[a], scheds + a * OR_MAXT, ts[a]) * mul;
      mul *= dims[a];
    }
    rho[idx] = acc;
  }
}

/* Gather x'[p] = x[rho[p]] (the permutation P with A P = F, valid for mixed radix). */
void or_gather(double* re, double* im, long n, const int64_t* rho) {
  double* tr = (double*)malloc((size_t)n * sizeof(double));
  double* ti = (double*)malloc((size_t)n * sizeof(double));
  for (long p = 0; p < n; ++p) {
    tr[p] = re[rho[p]];
    ti[p] = im[rho[p]];
  }
  memcpy(re, tr, (size_t)n * sizeof(double));
  memcpy(im, ti, (size_t)n * sizeof(double));
  free(tr);
  free(ti);
}

/* Scatter x'[rho[p]] = x[p] (inverse of or_gather). */
void or_scatter(double* re, double* im, long n, const int64_t* rho) {
  double* tr = (double*)malloc((size_t)n * sizeof(double));
  double* ti = (double*)malloc((size_t)n * sizeof(double));
  for (long p = 0; p < n; ++p) {
    tr[rho[p]] = re[p];
    ti[rho[p]] = im[p];
  }
  memcpy(re, tr, (size_t)n * sizeof(double));
  memcpy(im, ti, (size_t)n * sizeof(double));
  free(tr);
  free(ti);
}

/* Reference twiddle tables (twiddle_factors.c:17-131), jagged levels
 * flattened: level q starts at sum_{q'<q} (r-1) r^{q'}; entry m*ell+j holds
 * w_k^{(m+1) j}, k = r^{q+1}, ell = r^q.  Returns the entry count (n-1). */
long or_twiddles_ref_layout(long n, int r, double* wre, double* wim) {
  long off = 0;
  for (long ell = 1; ell * r <= n; ell *= r) {
    long k = ell * r;
    for (int m = 0; m < r - 1; ++m)
      for (long j = 0; j < ell; ++j) {
        double angle = -2.0 * M_PI * (double)((m + 1) * j) / (double)k;
        wre[off + m * ell + j] = cos(angle);
        wim[off + m * ell + j] = sin(angle);
      }
    off += (r - 1) * ell;
  }
  return off;
}

static long total_of(int nd, const long* dims) {
  long n = 1;
  for (int a = 0; a < nd; ++a) n *= dims[a];
  return n;
}

/* which: 0 fft (P then A), 1 fft2 (A^T then P^-1), 2 ufft (A), 3 ifft (P, conj A, /n),
 * 4 uifft (conj A, /n).  fft.c:12-88, fft_2d.c:12-114, fft_3d.c:12-129. */
void or_transform_nd(double* re, double* im, int nd, const long* dims, const int* scheds,
                     const int* ts, int which) {
  long n = total_of(nd, dims);
  int64_t* rho = NULL;
  if (which == 0 || which == 1 || which == 3) {
    rho = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    or_perm_nd(nd, dims, scheds, ts, rho);
  }
  switch (which) {
    case 0:
      or_gather(re, im, n, rho);
      or_butterflies_nd(re, im, nd, dims, scheds, ts, OR_A);
      break;
    case 1:
      or_butterflies_nd(re, im, nd, dims, scheds, ts, OR_TRANS);
      or_scatter(re, im, n, rho);
      break;
    case 2:
      or_butterflies_nd(re, im, nd, dims, scheds, ts, OR_A);
      break;
    case 3:
    case 4:
      if (which == 3) or_gather(re, im, n, rho);
      or_butterflies_nd(re, im, nd, dims, scheds, ts, OR_CONJ);
      for (long i = 0; i < n; ++i) {
        re[i] /= (double)n;
        im[i] /= (double)n;
      }
      break;
    default:
      break;
  }
  free(rho);
}

/* permutation_avoiding_convolution{,_2d,_3d} (convolution.c:31-72):
 * x <- A^T x, h <- A^T h (left unscaled in h), x <- (x*h)/n, x <- conj(A) x. */
void or_pa_convolution_nd(double* re, double* im, double* hre, double* him, int nd,
                          const long* dims, const int* scheds, const int* ts) {
  long n = total_of(nd, dims);
  or_butterflies_nd(re, im, nd, dims, scheds, ts, OR_TRANS);
  or_butterflies_nd(hre, him, nd, dims, scheds, ts, OR_TRANS);
  for (long i = 0; i < n; ++i) {
    double a = re[i], b = im[i], c = hre[i], d = him[i];
    re[i] = (a * c - b * d) / (double)n;
    im[i] = (a * d + b * c) / (double)n;
  }
  or_butterflies_nd(re, im, nd, dims, scheds, ts, OR_CONJ);
}

/* fft_based_convolution{,_2d,_3d} (convolution.c:12-29): h <- F h, x <- ifft(Fx * Fh). */
void or_fft_based_convolution_nd(double* re, double* im, double* hre, double* him, int nd,
                                 const long* dims, const int* scheds, const int* ts) {
  long n = total_of(nd, dims);
  or_transform_nd(re, im, nd, dims, scheds, ts, 0);
  or_transform_nd(hre, him, nd, dims, scheds, ts, 0);
  for (long i = 0; i < n; ++i) {
    double a = re[i], b = im[i], c = hre[i], d = him[i];
    re[i] = a * c - b * d;
    im[i] = a * d + b * c;
  }
  or_transform_nd(re, im, nd, dims, scheds, ts, 3);
}

/* get_max_error (misc.c:20-30): reference stored interleaved (re, im, re, im, ...). */
double or_max_error(const double* re, const double* im, const double* ref_flat, long n) {
  double mx = 0.0;
  for (long i = 0; i < n; ++i) {
    double e = hypot(re[i] - ref_flat[2 * i], im[i] - ref_flat[2 * i + 1]);
    if (e > mx) mx = e;
  }
  return mx;
}
