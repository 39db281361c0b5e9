/*
 * ibfs_oracle.c -- CPU restatement of the reference IBFS multi-domain time
 * step, FP64 throughout.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library
 * and the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Every routine follows a reference routine (paths relative to the reference
 * tree) and keeps its loop structure and operation order where that is
 * cheap, so that the oracle differs from the reference only by the rounding
 * of the sine transform:
 *   advance        src/operators.f90:62-188
 *   vort2flux      src/operators.f90:191-240
 *   curl           src/operators.f90:243-288
 *   rot            src/operators.f90:291-304
 *   coarsify       src/operators.f90:307-333
 *   get_bc         src/operators.f90:336-367
 *   apply_bc       src/operators.f90:370-390
 *   a_times        src/operators.f90:444-462
 *   nonlinear      src/operators.f90:465-510
 *   reg / regT     src/operators.f90:513-546
 *   write_slip     src/operators.f90:549-580   (the slip value only)
 *   write_force    src/operators.f90:583-659   (the force values only)
 *   write_circ     src/operators.f90:662-699   (the circulation values only)
 *   choldc/cholsl  src/operators.f90:766-819
 *   cjgr           src/operators.f90:393-441   (stationary = F force solve)
 *   move + setup_reg src/operators.f90:142-144 (new point positions)
 *   preprocess     src/operators.f90:29-59     (no file cache here)
 *   setup_fft      src/myfft.f90:21-70
 *   dst/ctci/ainv  src/myfft.f90:83-124
 *   setup indices  src/grid.f90:64-134
 *   setup_reg      src/grid.f90:142-230
 *   deltafnc       src/grid.f90:233-249
 *   IC             src/variables.f90:54-67
 *
 * The reference computes its 2-D DST-I with FFTW's RODFT00 (an un-vendored,
 * unpinned dependency).  Here the DST-I is the closed form
 *     Y_k = 2 sum_{j=1}^{N-1} x_j sin(pi j k / N)
 * evaluated in O(N log N) by the odd-extension identity: two real rows a, b
 * are packed as z = a + i b, extended oddly to length 2N, and one complex FFT
 * of length 2N gives Y(b) = Re Z_k and Y(a) = -Im Z_k.  The FFT is the
 * Stockham autosort formulation of Govindaraju et al., "High Performance
 * Discrete Fourier Transforms on Graphics Processors" (SC08), radices 4/2/3/5
 * plus a direct DFT for other prime factors.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI_REF 3.1415926535897932384 /* parameters.f90:35 */

/* ------------------------------------------------------------------------ */
/* complex FFT (Stockham, SC08 formulation)                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  int nst;
  int rad[48];
  double *tw; /* tw[2k], tw[2k+1] = cos, sin of (-2 pi k / n) */
} fft_plan;

static void fft_plan_init(fft_plan *p, int n) {
  p->n = n;
  p->nst = 0;
  int r = n;
  while (r % 4 == 0) { p->rad[p->nst++] = 4; r /= 4; }
  while (r % 2 == 0) { p->rad[p->nst++] = 2; r /= 2; }
  while (r % 3 == 0) { p->rad[p->nst++] = 3; r /= 3; }
  while (r % 5 == 0) { p->rad[p->nst++] = 5; r /= 5; }
  for (int f = 7; r > 1; f += 2)
    while (r % f == 0) { p->rad[p->nst++] = f; r /= f; }
  p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  for (int k = 0; k < n; k++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    p->tw[2 * k] = (double)cosl(a);
    p->tw[2 * k + 1] = (double)sinl(a);
  }
}

static void fft_plan_free(fft_plan *p) { free(p->tw); p->tw = NULL; }

/* forward DFT (exp(-2 pi i jk/n)) of x (interleaved complex, length n); y is
 * scratch of the same size.  Returns the buffer that holds the result. */
static double *fft_exec(const fft_plan *p, double *x, double *y) {
  const int n = p->n;
  int ns = 1;
  double *src = x, *dst = y;
  double vr[64], vi[64], wr[64], wi[64];
  for (int s = 0; s < p->nst; s++) {
    const int R = p->rad[s];
    const int nr = n / R;
    const int twstep = n / (ns * R);
    for (int j = 0; j < nr; j++) {
      const int k = j % ns;
      for (int r = 0; r < R; r++) {
        double ar = src[2 * (j + r * nr)], ai = src[2 * (j + r * nr) + 1];
        int ti = k * r * twstep; /* < n */
        double cr = p->tw[2 * ti], ci = p->tw[2 * ti + 1];
        vr[r] = ar * cr - ai * ci;
        vi[r] = ar * ci + ai * cr;
      }
      if (R == 4) {
        double t0r = vr[0] + vr[2], t0i = vi[0] + vi[2];
        double t1r = vr[0] - vr[2], t1i = vi[0] - vi[2];
        double t2r = vr[1] + vr[3], t2i = vi[1] + vi[3];
        double t3r = vr[1] - vr[3], t3i = vi[1] - vi[3];
        wr[0] = t0r + t2r; wi[0] = t0i + t2i;
        wr[2] = t0r - t2r; wi[2] = t0i - t2i;
        /* -i * t3 */
        wr[1] = t1r + t3i; wi[1] = t1i - t3r;
        wr[3] = t1r - t3i; wi[3] = t1i + t3r;
      } else if (R == 2) {
        wr[0] = vr[0] + vr[1]; wi[0] = vi[0] + vi[1];
        wr[1] = vr[0] - vr[1]; wi[1] = vi[0] - vi[1];
      } else {
        /* direct DFT of prime length R: w_q = sum_r v_r exp(-2 pi i q r / R) */
        const int st = n / R;
        for (int q = 0; q < R; q++) {
          double sr = 0.0, si = 0.0;
          for (int r = 0; r < R; r++) {
            int ti = ((q * r) % R) * st;
            double cr = p->tw[2 * ti], ci = p->tw[2 * ti + 1];
            sr += vr[r] * cr - vi[r] * ci;
            si += vr[r] * ci + vi[r] * cr;
          }
          wr[q] = sr; wi[q] = si;
        }
      }
      const int idxd = (j / ns) * ns * R + k;
      for (int r = 0; r < R; r++) {
        dst[2 * (idxd + r * ns)] = wr[r];
        dst[2 * (idxd + r * ns) + 1] = wi[r];
      }
    }
    double *t = src; src = dst; dst = t;
    ns *= R;
  }
  return src;
}

/* DST-I (unnormalised, RODFT00) of each of nrow rows of length ncol stored
 * contiguously with row stride ncol; in place.  p is the plan of length
 * 2*(ncol+1).  work holds 4*(ncol+1) complex. */
static void dst_rows(double *a, int nrow, int ncol, const fft_plan *p, double *work) {
  const int N = ncol + 1, n2 = 2 * N;
  double *z = work, *y = work + 2 * n2;
  for (int r0 = 0; r0 < nrow; r0 += 2) {
    const double *ra = a + (size_t)r0 * ncol;
    const double *rb = (r0 + 1 < nrow) ? a + (size_t)(r0 + 1) * ncol : NULL;
    z[0] = z[1] = 0.0;
    z[2 * N] = z[2 * N + 1] = 0.0;
    for (int j = 1; j < N; j++) {
      double xa = ra[j - 1], xb = rb ? rb[j - 1] : 0.0;
      z[2 * j] = xa; z[2 * j + 1] = xb;
      z[2 * (n2 - j)] = -xa; z[2 * (n2 - j) + 1] = -xb;
    }
    double *Z = fft_exec(p, z, y);
    double *wa = a + (size_t)r0 * ncol;
    double *wb = rb ? a + (size_t)(r0 + 1) * ncol : NULL;
    for (int k = 1; k < N; k++) {
      wa[k - 1] = -Z[2 * k + 1];
      if (wb) wb[k - 1] = Z[2 * k];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* state                                                                      */
/* ------------------------------------------------------------------------ */
typedef struct orc {
  int m, n, L, nb, nf, nq, nw, nbc, nsmear;
  int left, right, bottom, top;
  int bmin, bmax;
  double dt, re, len, offx, offy, delta, normalize;
  double *xb, *yb;
  int *bdyno;
  int *ismear;   /* (nq, nsmear) column-major, 1-based force index */
  double *smear; /* (nq, nsmear) column-major */
  double *lam, *laminv, *lam1, *lam1i; /* (m-1)(n-1) [x L for lam1/lam1i] */
  double *con1, *con2, *vfac;
  double *omega, *rhs_old, *q, *q0, *fb;
  double *cholmat, *cholvec;
  fft_plan pm, pn;
  int dst_order; /* 0: along i then j; 1: along j then i (same transform, other rounding) */
  double *work, *tbuf; /* dst scratch, transpose buffer */
  /* force solve (parameters.f90:19-20,31): stationary = Cholesky factor,
   * else cjgr warm-started from the previous fb; cg_iters = count of the
   * last cjgr call */
  int stationary, cg_max_iter, cg_iters;
  double cgtol;
} orc_t;

#define OM(a, i, j) (a)[((i)-2) + (size_t)((j)-2) * (o->m - 1)]
#define UIX(i, j) ((size_t)((i)-1) + (size_t)((j)-1) * (o->m + 1))
#define VIX(i, j) ((size_t)(o->m + 1) * o->n + (size_t)((i)-1) + (size_t)((j)-1) * o->m)
#define BC(a, pos) (a)[(pos)-1]

/* deltafnc: grid.f90:233-249 (Roma, Peskin & Berger 1999) */
static double deltafnc(double r, double r0, double dr) {
  double d = r - r0;
  if (fabs(d) <= 0.5 * dr) {
    return (1.0 + sqrt(-3.0 * (d / dr) * (d / dr) + 1.0)) / (3.0 * dr);
  } else if (fabs(d) <= 1.5 * dr) {
    double a = 1.0 - fabs(d / dr);
    return (5.0 - 3.0 * fabs(d / dr) - sqrt(-3.0 * a * a + 1.0)) / (6.0 * dr);
  }
  return 0.0;
}

/* setup_reg: grid.f90:142-230 (ELL tables, padding index 1 weight 0) */
static void setup_reg(orc_t *o) {
  const int m = o->m, n = o->n, nb = o->nb;
  const double delta = o->delta, d2 = 0.5 * delta, sup = 3.0 * d2;
  int ns = 0;
  for (int i = 1; i <= m + 1; i++)
    for (int j = 1; j <= n; j++) {
      double x = delta * (double)(i - 1) - o->offx;
      double y = delta * (double)(j - 1) - o->offy + d2;
      int c = 0;
      for (int k = 0; k < nb; k++)
        if (fabs(x - o->xb[k]) < sup && fabs(y - o->yb[k]) < sup) c++;
      if (c > ns) ns = c;
    }
  for (int i = 1; i <= m; i++)
    for (int j = 1; j <= n + 1; j++) {
      double x = delta * (double)(i - 1) - o->offx + d2;
      double y = delta * (double)(j - 1) - o->offy;
      int c = 0;
      for (int k = 0; k < nb; k++)
        if (fabs(x - o->xb[k]) < sup && fabs(y - o->yb[k]) < sup) c++;
      if (c > ns) ns = c;
    }
  o->nsmear = ns;
  size_t tot = (size_t)o->nq * (ns > 0 ? ns : 1);
  free(o->ismear); free(o->smear);
  o->ismear = (int *)malloc(sizeof(int) * tot);
  o->smear = (double *)calloc(tot, sizeof(double));
  for (size_t t = 0; t < tot; t++) o->ismear[t] = 1;
  for (int i = 1; i <= m + 1; i++)
    for (int j = 1; j <= n; j++) {
      double x = delta * (double)(i - 1) - o->offx;
      double y = delta * (double)(j - 1) - o->offy + d2;
      int c = 0;
      for (int k = 0; k < nb; k++)
        if (fabs(x - o->xb[k]) < sup && fabs(y - o->yb[k]) < sup) {
          size_t e = UIX(i, j) + (size_t)c * o->nq;
          o->ismear[e] = k + 1;
          o->smear[e] = (delta * delta) * deltafnc(x, o->xb[k], delta) * deltafnc(y, o->yb[k], delta);
          c++;
        }
    }
  for (int i = 1; i <= m; i++)
    for (int j = 1; j <= n + 1; j++) {
      double x = delta * (double)(i - 1) - o->offx + d2;
      double y = delta * (double)(j - 1) - o->offy;
      int c = 0;
      for (int k = 0; k < nb; k++)
        if (fabs(x - o->xb[k]) < sup && fabs(y - o->yb[k]) < sup) {
          size_t e = VIX(i, j) + (size_t)c * o->nq;
          o->ismear[e] = nb + k + 1;
          o->smear[e] = (delta * delta) * deltafnc(x, o->xb[k], delta) * deltafnc(y, o->yb[k], delta);
          c++;
        }
    }
}

/* setup_fft tables: myfft.f90:21-70 */
static void setup_fft(orc_t *o) {
  const int m = o->m, n = o->n, L = o->L;
  const size_t nw = (size_t)(m - 1) * (n - 1);
  o->normalize = 4.0 * (double)(float)((long)m * (long)n);
  o->con1 = (double *)malloc(sizeof(double) * L);
  o->con2 = (double *)malloc(sizeof(double) * L);
  o->vfac = (double *)malloc(sizeof(double) * L);
  o->lam = (double *)malloc(sizeof(double) * nw);
  o->laminv = (double *)malloc(sizeof(double) * nw);
  o->lam1 = (double *)malloc(sizeof(double) * nw * L);
  o->lam1i = (double *)malloc(sizeof(double) * nw * L);
  double del2 = o->delta * o->delta;
  for (int k = 1; k <= L; k++) {
    double del22 = del2 * pow(4.0, k - 1);
    o->vfac[k - 1] = 0.5 * o->dt / o->re / del22 / o->normalize;
    o->con1[k - 1] = 1.5 * o->dt / del22 / o->normalize;
    o->con2[k - 1] = -0.5 * o->dt / del22 / o->normalize;
  }
  for (int j = 1; j <= n - 1; j++)
    for (int i = 1; i <= m - 1; i++) {
      size_t e = (size_t)(i - 1) + (size_t)(j - 1) * (m - 1);
      double lam = -2.0 * (cos(PI_REF * (double)(float)i / (double)(float)m) +
                           cos(PI_REF * (double)(float)j / (double)(float)n) - 2.0);
      o->lam[e] = lam;
      o->laminv[e] = 1.0 / lam / o->normalize;
      for (int k = 1; k <= L; k++) {
        double del22 = del2 * pow(4.0, k - 1);
        o->lam1[e + nw * (k - 1)] = (1.0 - 0.5 * o->dt * lam / del22 / o->re) / o->normalize;
        o->lam1i[e + nw * (k - 1)] = 1.0 / (1.0 + 0.5 * o->dt * lam / del22 / o->re);
      }
    }
  fft_plan_init(&o->pm, 2 * m);
  fft_plan_init(&o->pn, 2 * n);
  int mx = m > n ? m : n;
  o->work = (double *)malloc(sizeof(double) * 8 * (size_t)(mx + 1) * 2);
  o->tbuf = (double *)malloc(sizeof(double) * nw);
}

/* ------------------------------------------------------------------------ */
/* public API                                                                 */
/* ------------------------------------------------------------------------ */
int orc_nlevmax(int m, int n) {
  /* grid.f90:79 -- single-precision LOG of REAL(MIN(m,n)/2) */
  int mn = (m < n ? m : n) / 2;
  double v = (double)logf((float)mn) / log(2.0);
  return (int)floor(v) - 1;
}

orc_t *orc_create(int m, int n, int L, double dt, double re, double len, double offx, double offy,
                  int nb, const double *xb, const double *yb, const int *bdyno, int unif_flow) {
  if (m % 4 != 0 || n % 4 != 0 || m < 8 || n < 8 || L < 1) return NULL;
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->m = m; o->n = n; o->L = L; o->nb = nb; o->nf = 2 * nb;
  o->nq = (m + 1) * n + (n + 1) * m;
  o->nw = (m - 1) * (n - 1);
  o->nbc = 2 * (m + 1) + 2 * (n + 1);
  o->left = 0; o->right = n + 1; o->bottom = 2 * (n + 1); o->top = 2 * (n + 1) + m + 1;
  o->dt = dt; o->re = re; o->len = len; o->offx = offx; o->offy = offy;
  o->delta = len / (double)(float)m; /* grid.f90:64 */
  o->xb = (double *)malloc(sizeof(double) * (nb > 0 ? nb : 1));
  o->yb = (double *)malloc(sizeof(double) * (nb > 0 ? nb : 1));
  o->bdyno = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
  o->bmin = 0; o->bmax = 0;
  for (int k = 0; k < nb; k++) {
    o->xb[k] = xb[k]; o->yb[k] = yb[k]; o->bdyno[k] = bdyno ? bdyno[k] : 0;
    if (k == 0 || o->bdyno[k] < o->bmin) o->bmin = o->bdyno[k];
    if (k == 0 || o->bdyno[k] > o->bmax) o->bmax = o->bdyno[k];
  }
  setup_reg(o);
  setup_fft(o);
  o->stationary = 1;
  o->cgtol = 1.0e-7;
  o->cg_max_iter = 3000;
  size_t nw = (size_t)o->nw, nq = (size_t)o->nq, nf = (size_t)o->nf;
  o->omega = (double *)calloc(nw * L, sizeof(double));
  o->rhs_old = (double *)calloc(nw * L, sizeof(double));
  o->q = (double *)calloc(nq * L, sizeof(double));
  o->q0 = (double *)calloc(nq * L, sizeof(double));
  o->fb = (double *)calloc(nf > 0 ? nf : 1, sizeof(double));
  o->cholmat = (double *)calloc(nf * nf > 0 ? nf * nf : 1, sizeof(double));
  o->cholvec = (double *)calloc(nf > 0 ? nf : 1, sizeof(double));
  if (unif_flow) { /* variables.f90:60-64 */
    for (int k = 1; k <= L; k++) {
      double v = o->delta * pow(2.0, k - 1);
      for (size_t e = 0; e < (size_t)(m + 1) * n; e++) o->q0[e + nq * (k - 1)] = v;
    }
  }
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  free(o->xb); free(o->yb); free(o->bdyno); free(o->ismear); free(o->smear);
  free(o->lam); free(o->laminv); free(o->lam1); free(o->lam1i);
  free(o->con1); free(o->con2); free(o->vfac);
  free(o->omega); free(o->rhs_old); free(o->q); free(o->q0); free(o->fb);
  free(o->cholmat); free(o->cholvec);
  fft_plan_free(&o->pm); fft_plan_free(&o->pn);
  free(o->work); free(o->tbuf);
  free(o);
}

void orc_dims(const orc_t *o, int *out) {
  out[0] = o->m; out[1] = o->n; out[2] = o->L; out[3] = o->nb; out[4] = o->nf;
  out[5] = o->nq; out[6] = o->nw; out[7] = o->nbc; out[8] = o->nsmear;
  out[9] = o->bmin; out[10] = o->bmax;
}

double *orc_ptr(orc_t *o, int which) {
  switch (which) {
    case 0: return o->omega;
    case 1: return o->rhs_old;
    case 2: return o->q;
    case 3: return o->q0;
    case 4: return o->fb;
    case 5: return o->cholmat;
    case 6: return o->cholvec;
    case 7: return o->smear;
    case 8: return o->lam;
    case 9: return o->lam1;
    case 10: return o->lam1i;
    case 11: return o->laminv;
    default: return NULL;
  }
}
int *orc_ismear(orc_t *o) { return o->ismear; }
double orc_scalar(const orc_t *o, int which) {
  switch (which) {
    case 0: return o->delta;
    case 1: return o->normalize;
    default: return 0.0;
  }
}
void orc_factors(const orc_t *o, double *con1, double *con2, double *vfac) {
  for (int k = 0; k < o->L; k++) { con1[k] = o->con1[k]; con2[k] = o->con2[k]; vfac[k] = o->vfac[k]; }
}

/* dst: myfft.f90:83-94 (2-D RODFT00 of the (m-1)x(n-1) field, i fastest) */
void orc_dst(orc_t *o, const double *in, double *out) {
  const int m1 = o->m - 1, n1 = o->n - 1;
  if (out != in) memcpy(out, in, sizeof(double) * (size_t)m1 * n1);
  if (o->dst_order == 0) dst_rows(out, n1, m1, &o->pm, o->work); /* along i */
  for (int j = 0; j < n1; j++)
    for (int i = 0; i < m1; i++) o->tbuf[j + (size_t)i * n1] = out[i + (size_t)j * m1];
  dst_rows(o->tbuf, m1, n1, &o->pn, o->work);                    /* along j */
  for (int i = 0; i < m1; i++)
    for (int j = 0; j < n1; j++) out[i + (size_t)j * m1] = o->tbuf[j + (size_t)i * n1];
  if (o->dst_order == 1) dst_rows(out, n1, m1, &o->pm, o->work); /* along i */
}

/* Selects the axis order of the 2-D transform.  Both orders compute the same
 * DST-I exactly; their rounding differs, which gives the noise floor that
 * any two valid implementations (FFTW's included) differ by. */
void orc_set_dst_order(orc_t *o, int order) { o->dst_order = order ? 1 : 0; }

/* ctci: myfft.f90:97-109 */
void orc_ctci(orc_t *o, const double *in, double *out) {
  orc_dst(o, in, out);
  for (int e = 0; e < o->nw; e++) out[e] = o->laminv[e] * out[e];
  orc_dst(o, out, out);
}

/* ainv: myfft.f90:112-124 */
void orc_ainv(orc_t *o, const double *in, double *out) {
  orc_dst(o, in, out);
  for (int e = 0; e < o->nw; e++) out[e] = o->lam1i[e] * out[e] / o->normalize;
  orc_dst(o, out, out);
}

/* curl: operators.f90:243-288 */
void orc_curl(orc_t *o, const double *x, const double *xbc, double *c) {
  const int m = o->m, n = o->n;
  for (int i = 2; i <= m; i++) {
    for (int j = 2; j <= n - 1; j++) c[UIX(i, j)] = OM(x, i, j + 1) - OM(x, i, j);
    c[UIX(i, 1)] = OM(x, i, 2) - BC(xbc, o->bottom + i);
    c[UIX(i, n)] = BC(xbc, o->top + i) - OM(x, i, n);
  }
  for (int j = 1; j <= n; j++) c[UIX(1, j)] = BC(xbc, o->left + j + 1) - BC(xbc, o->left + j);
  for (int j = 1; j <= n; j++) c[UIX(m + 1, j)] = BC(xbc, o->right + j + 1) - BC(xbc, o->right + j);
  for (int j = 2; j <= n; j++) {
    for (int i = 2; i <= m - 1; i++) c[VIX(i, j)] = OM(x, i, j) - OM(x, i + 1, j);
    c[VIX(1, j)] = -OM(x, 2, j) + BC(xbc, o->left + j);
    c[VIX(m, j)] = -BC(xbc, o->right + j) + OM(x, m, j);
  }
  for (int i = 1; i <= m; i++) c[VIX(i, 1)] = BC(xbc, o->bottom + i) - BC(xbc, o->bottom + i + 1);
  for (int i = 1; i <= m; i++) c[VIX(i, n + 1)] = BC(xbc, o->top + i) - BC(xbc, o->top + i + 1);
}

/* rot: operators.f90:291-304 */
void orc_rot(orc_t *o, const double *x, double *r) {
  const int m = o->m, n = o->n;
  for (int i = 2; i <= m; i++)
    for (int j = 2; j <= n; j++)
      OM(r, i, j) = x[VIX(i, j)] - x[VIX(i - 1, j)] - x[UIX(i, j)] + x[UIX(i, j - 1)];
}

/* coarsify: operators.f90:307-333 (crhs and arhs may alias) */
void orc_coarsify(orc_t *o, const double *crhs, const double *rhs, double *arhs) {
  const int m = o->m, n = o->n;
  if (arhs != crhs) memcpy(arhs, crhs, sizeof(double) * (size_t)o->nw);
  for (int i = m / 4 + 2; i <= 3 * m / 4; i++) {
    int ii = 2 * i - m / 2 - 1;
    for (int j = n / 4 + 2; j <= 3 * n / 4; j++) {
      int jj = 2 * j - n / 2 - 1;
      OM(arhs, i, j) = OM(rhs, ii, jj) +
                       0.50 * (OM(rhs, ii - 1, jj) + OM(rhs, ii + 1, jj) + OM(rhs, ii, jj - 1) + OM(rhs, ii, jj + 1)) +
                       0.25 * (OM(rhs, ii - 1, jj - 1) + OM(rhs, ii - 1, jj + 1) + OM(rhs, ii + 1, jj - 1) +
                               OM(rhs, ii + 1, jj + 1));
    }
  }
}

/* get_bc: operators.f90:336-367.  The dummy r is assumed-shape, so r(a,b)
 * is omega(a+1,b+1). */
#define R1(a, b) OM(r, (a) + 1, (b) + 1)
void orc_get_bc(orc_t *o, const double *r, double *rbc, double fac) {
  const int m = o->m, n = o->n;
  for (int i = 0; i <= m; i += 2) {
    BC(rbc, o->bottom + i + 1) = R1(m / 4 + i / 2, n / 4);
    BC(rbc, o->top + i + 1) = R1(m / 4 + i / 2, 3 * n / 4);
  }
  for (int i = 1; i <= m - 1; i += 2) {
    BC(rbc, o->bottom + i + 1) = 0.5 * (R1(m / 4 + (i + 1) / 2, n / 4) + R1(m / 4 - 1 + (i + 1) / 2, n / 4));
    BC(rbc, o->top + i + 1) = 0.5 * (R1(m / 4 + (i + 1) / 2, 3 * n / 4) + R1(m / 4 - 1 + (i + 1) / 2, 3 * n / 4));
  }
  for (int j = 0; j <= n; j += 2) {
    BC(rbc, o->left + j + 1) = R1(m / 4, n / 4 + j / 2);
    BC(rbc, o->right + j + 1) = R1(3 * m / 4, n / 4 + j / 2);
  }
  for (int j = 1; j <= n - 1; j += 2) {
    BC(rbc, o->left + j + 1) = 0.5 * (R1(m / 4, n / 4 + (j + 1) / 2) + R1(m / 4, n / 4 - 1 + (j + 1) / 2));
    BC(rbc, o->right + j + 1) = 0.5 * (R1(3 * m / 4, n / 4 + (j + 1) / 2) + R1(3 * m / 4, n / 4 - 1 + (j + 1) / 2));
  }
  for (int p = 0; p < o->nbc; p++) rbc[p] = rbc[p] * fac;
}
#undef R1

/* apply_bc: operators.f90:370-390 (assumed-shape r(i,j) = omega(i+1,j+1)) */
void orc_apply_bc(orc_t *o, double *r, const double *rbc, double fac) {
  const int m = o->m, n = o->n;
  for (int i = 1; i <= m - 1; i++) {
    OM(r, i + 1, 2) = OM(r, i + 1, 2) + fac * BC(rbc, o->bottom + i + 1);
    OM(r, i + 1, n) = OM(r, i + 1, n) + fac * BC(rbc, o->top + i + 1);
  }
  for (int j = 1; j <= n - 1; j++) {
    OM(r, 2, j + 1) = OM(r, 2, j + 1) + fac * BC(rbc, o->left + j + 1);
    OM(r, m, j + 1) = OM(r, m, j + 1) + fac * BC(rbc, o->right + j + 1);
  }
}

/* nonlinear: operators.f90:465-510.  Boundary faces that rot never reads
 * are set to zero here (the reference leaves them undefined). */
void orc_nonlinear(orc_t *o, const double *omega, const double *q, const double *q0, const double *bc, double *fq) {
  const int m = o->m, n = o->n;
  double *uavg = (double *)malloc(sizeof(double) * (size_t)(m + 1) * (n - 1)); /* (1:m+1, 2:n) */
  double *vavg = (double *)malloc(sizeof(double) * (size_t)(m - 1) * (n + 1)); /* (2:m, 1:n+1) */
#define UA(i, j) uavg[((i)-1) + (size_t)((j)-2) * (m + 1)]
#define VA(i, j) vavg[((i)-2) + (size_t)((j)-1) * (m - 1)]
  for (int j = 2; j <= n; j++)
    for (int i = 1; i <= m + 1; i++)
      UA(i, j) = 0.5 * (q[UIX(i, j)] + q[UIX(i, j - 1)] + q0[UIX(i, j)] + q0[UIX(i, j - 1)]);
  for (int j = 1; j <= n + 1; j++)
    for (int i = 2; i <= m; i++)
      VA(i, j) = 0.5 * (q[VIX(i, j)] + q[VIX(i - 1, j)] + q0[VIX(i, j)] + q0[VIX(i - 1, j)]);
  memset(fq, 0, sizeof(double) * (size_t)o->nq);
  for (int i = 2; i <= m; i++) {
    for (int j = 2; j <= n - 1; j++)
      fq[UIX(i, j)] = 0.5 * (VA(i, j + 1) * OM(omega, i, j + 1) + VA(i, j) * OM(omega, i, j));
    fq[UIX(i, 1)] = 0.5 * (VA(i, 2) * OM(omega, i, 2) + VA(i, 1) * BC(bc, o->bottom + i));
    fq[UIX(i, n)] = 0.5 * (VA(i, n + 1) * BC(bc, o->top + i) + VA(i, n) * OM(omega, i, n));
  }
  for (int j = 2; j <= n; j++) {
    for (int i = 2; i <= m - 1; i++)
      fq[VIX(i, j)] = -0.5 * (UA(i + 1, j) * OM(omega, i + 1, j) + UA(i, j) * OM(omega, i, j));
    fq[VIX(1, j)] = -0.5 * (UA(2, j) * OM(omega, 2, j) + UA(1, j) * BC(bc, o->left + j));
    fq[VIX(m, j)] = -0.5 * (UA(m + 1, j) * BC(bc, o->right + j) + UA(m, j) * OM(omega, m, j));
  }
#undef UA
#undef VA
  free(uavg); free(vavg);
}

/* reg: operators.f90:513-528 */
void orc_reg(orc_t *o, const double *h0, double *h) {
  const size_t nq = (size_t)o->nq;
  for (size_t i = 0; i < nq; i++) {
    double s = 0.0;
    for (int k = 0; k < o->nsmear; k++) s = s + o->smear[i + k * nq] * h0[o->ismear[i + k * nq] - 1];
    h[i] = s;
  }
}

/* regT: operators.f90:531-546 */
void orc_regT(orc_t *o, const double *h, double *h0) {
  const size_t nq = (size_t)o->nq;
  for (int f = 0; f < o->nf; f++) h0[f] = 0.0;
  for (size_t i = 0; i < nq; i++)
    for (int k = 0; k < o->nsmear; k++) {
      int f = o->ismear[i + k * nq] - 1;
      h0[f] = h0[f] + o->smear[i + k * nq] * h[i];
    }
}

/* vort2flux: operators.f90:191-240.  omega (W x L) is modified in place
 * (coarse centres overwritten); vel is (Nq x nlev). */
void orc_vort2flux(orc_t *o, double *omega, double *vel, int nlev) {
  const int L = o->L;
  const size_t nw = (size_t)o->nw, nq = (size_t)o->nq;
  double *s = (double *)malloc(sizeof(double) * nw);
  double *vort = (double *)malloc(sizeof(double) * nw);
  double *sbc = (double *)calloc((size_t)o->nbc, sizeof(double));
  for (int k = 2; k <= L; k++) orc_coarsify(o, omega + nw * (k - 1), omega + nw * (k - 2), omega + nw * (k - 1));
  memcpy(vort, omega + nw * (L - 1), sizeof(double) * nw);
  orc_ctci(o, vort, s);
  if (nlev >= L) orc_curl(o, s, sbc, vel + nq * (L - 1));
  for (int k = L - 1; k >= 1; k--) {
    orc_get_bc(o, s, sbc, 1.0);
    memcpy(vort, omega + nw * (k - 1), sizeof(double) * nw);
    orc_apply_bc(o, vort, sbc, 1.0);
    orc_ctci(o, vort, s);
    if (nlev >= k) orc_curl(o, s, sbc, vel + nq * (k - 1));
  }
  free(s); free(vort); free(sbc);
}

/* a_times: operators.f90:444-462 */
void orc_a_times(orc_t *o, const double *x, double *out) {
  const size_t nw = (size_t)o->nw, nq = (size_t)o->nq;
  double *vort = (double *)calloc(nw * o->L, sizeof(double));
  double *h = (double *)malloc(sizeof(double) * nq);
  double *r = (double *)malloc(sizeof(double) * nw);
  orc_reg(o, x, h);
  orc_rot(o, h, r);
  orc_ainv(o, r, vort);
  orc_vort2flux(o, vort, h, 1);
  orc_regT(o, h, out);
  free(vort); free(h); free(r);
}

/* choldc: operators.f90:766-791 (upper triangle keeps A, strict lower = L,
 * diagonal of L in cholvec).  Returns 0, or i (1-based) where it failed. */
int orc_choldc(orc_t *o) {
  const int nf = o->nf;
  double *a = o->cholmat;
#define CM(i, j) a[((i)-1) + (size_t)((j)-1) * nf]
  for (int i = 1; i <= nf; i++)
    for (int j = i; j <= nf; j++) {
      double sum = CM(i, j);
      for (int k = i - 1; k >= 1; k--) sum = sum - CM(i, k) * CM(j, k);
      if (i == j) {
        if (sum <= 0.0) return i;
        o->cholvec[i - 1] = sqrt(sum);
      } else {
        CM(j, i) = sum / o->cholvec[i - 1];
      }
    }
  return 0;
}

/* cholsl: operators.f90:794-819 */
void orc_cholsl(orc_t *o, const double *b, double *x) {
  const int nf = o->nf;
  const double *a = o->cholmat;
  for (int i = 1; i <= nf; i++) {
    double sum = b[i - 1];
    for (int k = i - 1; k >= 1; k--) sum = sum - CM(i, k) * x[k - 1];
    x[i - 1] = sum / o->cholvec[i - 1];
  }
  for (int i = nf; i >= 1; i--) {
    double sum = x[i - 1];
    for (int k = i + 1; k <= nf; k++) sum = sum - CM(k, i) * x[k - 1];
    x[i - 1] = sum / o->cholvec[i - 1];
  }
#undef CM
}

/* DOT_PRODUCT of two Nf vectors, summed in index order */
static double dotf(int nf, const double *a, const double *b) {
  double s = 0.0;
  for (int f = 0; f < nf; f++) s = s + a[f] * b[f];
  return s;
}

/* cjgr: operators.f90:393-441.  Conjugate gradients on a_times, no
 * preconditioner; x holds the initial guess on entry (the previous fb: the
 * dummy is INTENT(OUT) but read before it is written, operators.f90:402,
 * 409) and the solution on exit.  The residual is recomputed from b - A x
 * every 50 iterations.  Returns the iteration count. */
int orc_cjgr(orc_t *o, double *x, const double *b) {
  const int nf = o->nf;
  double *r = (double *)malloc(sizeof(double) * (nf > 0 ? nf : 1));
  double *d = (double *)malloc(sizeof(double) * (nf > 0 ? nf : 1));
  double *q = (double *)malloc(sizeof(double) * (nf > 0 ? nf : 1));
  int iter = 0;
  orc_a_times(o, x, q);
  for (int f = 0; f < nf; f++) r[f] = b[f] - q[f];
  for (int f = 0; f < nf; f++) d[f] = r[f];
  double delta_new = dotf(nf, r, d);
  const double eps = o->cgtol * o->cgtol;
  while (iter < o->cg_max_iter && delta_new > eps) {
    iter = iter + 1;
    orc_a_times(o, d, q);
    const double alpha = delta_new / dotf(nf, d, q);
    for (int f = 0; f < nf; f++) x[f] = x[f] + alpha * d[f];
    if (iter % 50 == 0) {
      orc_a_times(o, x, q);
      for (int f = 0; f < nf; f++) r[f] = b[f] - q[f];
    } else {
      for (int f = 0; f < nf; f++) r[f] = r[f] - alpha * q[f];
    }
    const double delta_old = delta_new;
    delta_new = dotf(nf, r, r);
    const double beta = delta_new / delta_old;
    for (int f = 0; f < nf; f++) d[f] = r[f] + beta * d[f];
  }
  free(r); free(d); free(q);
  o->cg_iters = iter;
  return iter;
}

/* the force solve of advance: Cholesky (stationary) or cjgr
 * (operators.f90:166-171), with cgtol and cg_max_iter of the namelist */
void orc_set_solver(orc_t *o, int stationary, double cgtol, int cg_max_iter) {
  o->stationary = stationary ? 1 : 0;
  o->cgtol = cgtol;
  o->cg_max_iter = cg_max_iter;
}
int orc_cg_iters(const orc_t *o) { return o->cg_iters; }

/* New positions of the body points, then setup_reg: what advance does after
 * move_body when a body moves (operators.f90:142-144, grid.f90:523-576) */
void orc_move(orc_t *o, const double *xb, const double *yb) {
  for (int k = 0; k < o->nb; k++) {
    o->xb[k] = xb[k];
    o->yb[k] = yb[k];
  }
  setup_reg(o);
}

/* preprocess: operators.f90:29-59, without the ibfs.chd cache (the caller
 * owns files).  Builds A column by column and factors it. */
int orc_preprocess(orc_t *o) {
  const int nf = o->nf;
  double *z = (double *)calloc((size_t)nf, sizeof(double));
  for (int i = 0; i < nf; i++) {
    memset(z, 0, sizeof(double) * nf);
    z[i] = 1.0;
    orc_a_times(o, z, o->cholmat + (size_t)i * nf);
  }
  free(z);
  return orc_choldc(o);
}

/* build only column i (0-based) of A: used to spot-check the device factor */
void orc_a_column(orc_t *o, int i, double *col) {
  double *z = (double *)calloc((size_t)o->nf, sizeof(double));
  z[i] = 1.0;
  orc_a_times(o, z, col);
  free(z);
}

/* force values of write_force: operators.f90:583-659; out = [fx(bmin..bmax), fy(bmin..bmax)] */
void orc_force(orc_t *o, const double *xg, double *out) {
  const int m = o->m, n = o->n, nb = o->nb, nbody = o->bmax - o->bmin + 1;
  double *g = (double *)malloc(sizeof(double) * (size_t)o->nf);
  double *qt = (double *)malloc(sizeof(double) * (size_t)o->nq);
  for (int k = o->bmin; k <= o->bmax; k++) {
    for (int f = 0; f < o->nf; f++) g[f] = 0.0;
    for (int i = 0; i < nb; i++)
      if (o->bdyno[i] == k) { g[i] = xg[i]; g[i + nb] = xg[i + nb]; }
    orc_reg(o, g, qt);
    double fx = 0.0, fy = 0.0;
    for (int i = 1; i <= m + 1; i++)
      for (int j = 1; j <= n; j++) fx = fx + qt[UIX(i, j)];
    for (int j = 1; j <= n + 1; j++)
      for (int i = 1; i <= m; i++) fy = fy + qt[VIX(i, j)];
    out[k - o->bmin] = fx * 2.0 * o->delta;
    out[nbody + k - o->bmin] = fy * 2.0 * o->delta;
  }
  free(g); free(qt);
}

/* circulation values of write_circ: operators.f90:662-699 */
void orc_circ(orc_t *o, double *circ) {
  const int m = o->m, n = o->n;
  for (int k = 1; k <= o->L; k++) {
    double s = 0.0;
    const double *w = o->omega + (size_t)o->nw * (k - 1);
    for (int i = 2; i <= m; i++)
      for (int j = 2; j <= n; j++) s = s + OM(w, i, j);
    circ[k - 1] = s;
  }
}

/* STEP 1 of advance (src/operators.f90:94-134): the intermediate vorticity
 * on every level, coarse to fine.  Exposed so that tests can check it
 * against the Crank-Nicolson/AB2 stencil equation. */
void orc_step1(orc_t *o, int first) {
  const int L = o->L;
  const size_t nw = (size_t)o->nw, nq = (size_t)o->nq;
  double *rhs = (double *)malloc(sizeof(double) * nw);
  double *rhsbc = (double *)malloc(sizeof(double) * nw);
  double *t1 = (double *)malloc(sizeof(double) * nw);
  double *t2 = (double *)malloc(sizeof(double) * nw);
  double *fq = (double *)malloc(sizeof(double) * nq);
  double *omega_bc = (double *)calloc((size_t)o->nbc, sizeof(double));
  double *lastbc = (double *)calloc((size_t)o->nbc * L, sizeof(double));
  for (int k = 1; k <= L - 1; k++) orc_get_bc(o, o->omega + nw * k, lastbc + (size_t)o->nbc * (k - 1), 0.25);
  for (int k = L; k >= 1; k--) {
    double *om = o->omega + nw * (k - 1), *ro = o->rhs_old + nw * (k - 1);
    if (k < L) orc_get_bc(o, o->omega + nw * k, omega_bc, 0.25);
    else memset(omega_bc, 0, sizeof(double) * o->nbc);
    orc_nonlinear(o, om, o->q + nq * (k - 1), o->q0 + nq * (k - 1), lastbc + (size_t)o->nbc * (k - 1), fq);
    orc_rot(o, fq, rhs);
    if (first) memcpy(ro, rhs, sizeof(double) * nw);
    memset(rhsbc, 0, sizeof(double) * nw);
    orc_apply_bc(o, rhsbc, lastbc + (size_t)o->nbc * (k - 1), o->vfac[k - 1]);
    orc_apply_bc(o, rhsbc, omega_bc, o->vfac[k - 1]);
    /* omega = dst(lam1i * (dst(con1 rhs + con2 rhs_old + rhsbc) + lam1 * dst(omega))) */
    for (size_t e = 0; e < nw; e++) t1[e] = o->con1[k - 1] * rhs[e] + o->con2[k - 1] * ro[e] + rhsbc[e];
    orc_dst(o, t1, t1);
    orc_dst(o, om, t2);
    const double *l1 = o->lam1 + nw * (k - 1), *l1i = o->lam1i + nw * (k - 1);
    for (size_t e = 0; e < nw; e++) t1[e] = l1i[e] * (t1[e] + l1[e] * t2[e]);
    orc_dst(o, t1, om);
    memcpy(ro, rhs, sizeof(double) * nw);
  }
  free(rhs); free(rhsbc); free(t1); free(t2); free(fq); free(omega_bc); free(lastbc);
}

/* advance: operators.f90:62-188 for stationary bodies.  motion_in is the
 * vector move_body/actuation would return (already scaled by delta, as in
 * grid.f90:573-574); NULL means zero (bodies with bdyno 0 and no
 * actuators).  Outputs (any may be NULL): fb/dt force values as in
 * write_force, circ (L values) and slip.  Returns 0, or the choldc failure
 * row if the itime==1 preprocess failed. */
int orc_advance(orc_t *o, int *itime, const double *motion_in, int do_preprocess,
                double *force, double *circ, double *slip) {
  const int L = o->L, nf = o->nf;
  const size_t nw = (size_t)o->nw, nq = (size_t)o->nq;
  double *t1 = (double *)malloc(sizeof(double) * nw);
  double *t2 = (double *)malloc(sizeof(double) * nw);
  double *fq = (double *)malloc(sizeof(double) * nq);
  double *motion = (double *)malloc(sizeof(double) * (nf > 0 ? nf : 1));
  double *rhsf = (double *)malloc(sizeof(double) * (nf > 0 ? nf : 1));
  double *tmpf = (double *)malloc(sizeof(double) * (nf > 0 ? nf : 1));
  int rc = 0;

  *itime = *itime + 1;
  if (*itime == 1 && do_preprocess && o->stationary) {
    rc = orc_preprocess(o);
    if (rc) goto done;
  }

  /* STEP 1 */
  orc_step1(o, *itime == 1);

  /* STEP 2 */
  for (int f = 0; f < nf; f++) motion[f] = motion_in ? motion_in[f] : 0.0;
  orc_regT(o, o->q0, tmpf);
  for (int f = 0; f < nf; f++) motion[f] = motion[f] - tmpf[f];
  orc_vort2flux(o, o->omega, o->q, 1);
  orc_regT(o, o->q, rhsf);
  for (int f = 0; f < nf; f++) rhsf[f] = rhsf[f] - motion[f];
  if (o->stationary) orc_cholsl(o, rhsf, o->fb);
  else orc_cjgr(o, o->fb, rhsf);

  /* STEP 3 */
  orc_reg(o, o->fb, fq);
  orc_rot(o, fq, t1);
  orc_ainv(o, t1, t2);
  for (size_t e = 0; e < nw; e++) o->omega[e] = o->omega[e] - t2[e];

  /* STEP 4 */
  orc_vort2flux(o, o->omega, o->q, L);

  /* diagnostics */
  if (force) {
    for (int f = 0; f < nf; f++) tmpf[f] = o->fb[f] / o->dt;
    orc_force(o, tmpf, force);
  }
  if (circ) orc_circ(o, circ);
  if (slip) {
    orc_regT(o, o->q, tmpf);
    double s = 0.0;
    for (int f = 0; f < nf; f++) {
      double d = fabs(tmpf[f] - motion[f]);
      if (d > s) s = d;
    }
    *slip = s;
  }
done:
  free(t1); free(t2); free(fq); free(motion); free(rhsf); free(tmpf);
  return rc;
}

/* plain 1-D DST-I of one vector (for pinning against brute force) */
void orc_dst1(const double *x, double *y, int len) {
  fft_plan p;
  fft_plan_init(&p, 2 * (len + 1));
  double *w = (double *)malloc(sizeof(double) * 8 * (size_t)(len + 1) * 2);
  memcpy(y, x, sizeof(double) * len);
  dst_rows(y, 1, len, &p, w);
  free(w);
  fft_plan_free(&p);
}

/* complex FFT of one vector (for pinning against numpy.fft) */
void orc_fft(const double *x, double *y, int len) {
  fft_plan p;
  fft_plan_init(&p, len);
  double *a = (double *)malloc(sizeof(double) * 2 * (size_t)len);
  double *b = (double *)malloc(sizeof(double) * 2 * (size_t)len);
  memcpy(a, x, sizeof(double) * 2 * len);
  double *r = fft_exec(&p, a, b);
  memcpy(y, r, sizeof(double) * 2 * len);
  free(a); free(b);
  fft_plan_free(&p);
}
