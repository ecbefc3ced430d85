/*
 * lmie_oracle.c -- CPU restatement of the size-integrated Mie path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (lmie_amd/) links, loads
 * or calls this file; it is built into oracle/_build/liboracle.so and used by
 * tests/ and by __graft_entry__.smoke() as the checker of the GPU library.
 *
 * It exports the same C ABI as include/lmie_interface.h, so the one Python
 * harness (lmie_amd/abi.py) drives the reference build, this oracle and the
 * product alike, plus a few oracle_* probes of intermediate quantities.
 *
 * Each function restates, in plain double / double complex C, the algorithm of
 * the reference file cited beside it.  Pins (tests/test_oracle.py):
 *   - every output against the reference's own printed example
 *     (examples/example.out) and against full-precision fixtures produced by
 *     the reference build (the .npz files under tests/golden);
 *   - a_n, b_n against arbitrary-precision Bessel-function Lorenz-Mie
 *     coefficients (mpmath), pi_n / tau_n against Legendre derivatives,
 *     Gauss nodes against numpy's leggauss, the generalized spherical
 *     functions against their closed forms at low order, and the invariants
 *     beta_0 = 1, g = beta_1 / 3, C_bak = F11(180 deg) C_sca / (4 pi).
 * The per-size loop runs on n_threads POSIX threads (as the reference
 * partitions it); the result does not depend on the thread count beyond the
 * order of the final sums.
 */
#include <complex.h>
#include <float.h>
#include <pthread.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/lmie_interface.h"

typedef double complex cplx;

#define ORACLE_PI 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------------- */
/* Pointer-table arrays, the layout of misc/gutil_alloc.c:19-75: the row
   pointers, then the contiguous row-major data, in one block. */
static double **blk2(int n0, int n1) {
     size_t head = (size_t) n0 * sizeof(double *);
     head = (head + 15) & ~(size_t) 15;
     char *m = (char *) calloc(1, head + (size_t) n0 * n1 * sizeof(double) + 16);
     double **p = (double **) m;
     double *d = (double *) (m + head);
     for (int i = 0; i < n0; ++i) p[i] = d + (size_t) i * n1;
     return p;
}

static double ***blk3(int n0, int n1, int n2) {
     size_t h0 = (size_t) n0 * sizeof(double **);
     size_t h1 = (size_t) n0 * n1 * sizeof(double *);
     size_t head = (h0 + h1 + 15) & ~(size_t) 15;
     char *m = (char *) calloc(1, head + (size_t) n0 * n1 * n2 * sizeof(double) + 16);
     double ***p = (double ***) m;
     double **q = (double **) (m + h0);
     double *d = (double *) (m + head);
     for (int i = 0; i < n0; ++i) {
          p[i] = q + (size_t) i * n1;
          for (int j = 0; j < n1; ++j) p[i][j] = d + ((size_t) i * n1 + j) * n2;
     }
     return p;
}

/* ------------------------------------------------------------------------- */
/* Truncation orders, src/lmie_core.c:18-38 (Mishchenko's formulas). */
static int order_n1(double x) { return (int) (x + 4.05 * pow(x, 1. / 3.) + 8); }

static int order_n2(int n1, cplx z) {
     double m = cabs(z);
     if (n1 > m) m = n1;
     return (int) (m + 6.40 * pow(m, 1. / 3.) + 8);
}

/* ------------------------------------------------------------------------- */
/* Gauss-Legendre nodes and weights on [x1, x2], ascending, with tangents of a
   moving interval.  Newton iteration with a second-order correction from the
   asymptotic initial guess cos(t + (n-1)/(8n^3)/tan t), t = (4i+3)pi/(4n+2),
   stopped at a step <= DBL_EPSILON (misc/rtutil_math.c:73-129, :231-316). */
static void gl_root(int n, int i, double *root, double *pnm1) {
     double t = (4. * i + 3.) / (4 * n + 2) * ORACLE_PI;
     double y = cos(t + (n - 1) / (8. * pow((double) n, 3)) / tan(t));
     double yprev, p0, p1, p2;
     do {
          p1 = 0.;
          p0 = 1.;
          for (int j = 0; j < n; ++j) {         /* p0 <- P_{j+1}(y) */
               p2 = p1;
               p1 = p0;
               double yp = y * p1;
               p0 = j / (j + 1.) * (yp - p2) + yp;
          }
          double s = 1. - y * y;
          double dp = n * (p1 - y * p0) / s;                /* P_n'  */
          double ddp = (2. * y * dp - n * (n + 1) * p0) / s;  /* P_n'' */
          yprev = y;
          double h = p0 / dp;
          y = yprev - h * (1. + h * ddp / (2. * dp));
     } while (fabs(y - yprev) > DBL_EPSILON);
     *root = y;
     *pnm1 = p1;
}

void oracle_gauss_leg(int n, int nd, double x1, double x2, const double *x1_l,
                      const double *x2_l, double *x, double *w, double *x_l, double *w_l) {
     double mid = (x1 + x2) / 2., half = (x2 - x1) / 2.;
     for (int i = 0, k = n - 1; i < (n + 1) / 2; ++i, --k) {
          double y, pn1;
          gl_root(n, i, &y, &pn1);
          double off = half * y;
          x[i] = mid - off;
          x[k] = mid + off;
          double c = n * pn1;
          c = 2. / (c * c);
          w[i] = (half - off * y) * c;
          w[k] = w[i];
          for (int d = 0; d < nd; ++d) {
               double mid_l = ((x1_l ? x1_l[d] : 0.) + (x2_l ? x2_l[d] : 0.)) / 2.;
               double half_l = ((x2_l ? x2_l[d] : 0.) - (x1_l ? x1_l[d] : 0.)) / 2.;
               double off_l = half_l * y;
               x_l[i * nd + d] = mid_l - off_l;
               x_l[k * nd + d] = mid_l + off_l;
               w_l[i * nd + d] = (half_l - off_l * y) * c;
               w_l[k * nd + d] = w_l[i * nd + d];
          }
     }
}

/* n_int equal sub-intervals of [r1, r2], n_quad Gauss nodes in each, with the
   tangents of moving ends (misc/rtutil_size_dist.c:128-193 and
   misc/rtutil_math.c:329-387: the interval map is applied to the mirrored
   node of the [-1, 1] rule). */
static void size_grid(int n_int, int n_quad, int nd, double r1, double r2,
                      const double *r1_l, const double *r2_l,
                      double *qx, double *qw, double *qx_l, double *qw_l) {
     double *tx = malloc(sizeof(double) * n_quad), *tw = malloc(sizeof(double) * n_quad);
     double *tx_l = calloc((size_t) n_quad * (nd + 1), sizeof(double));
     double *tw_l = calloc((size_t) n_quad * (nd + 1), sizeof(double));
     oracle_gauss_leg(n_quad, nd, -1., 1., NULL, NULL, tx, tw, tx_l, tw_l);
     double h = (r2 - r1) / n_int;
     for (int s = 0; s < n_int; ++s) {
          double lo = r1 + s * h, hi = lo + h;
          double mid = (lo + hi) / 2., half = (hi - lo) / 2.;
          int o = s * n_quad;
          for (int i = 0, k = n_quad - 1; i < (n_quad + 1) / 2; ++i, --k) {
               double off = half * tx[k];
               qx[o + i] = mid - off;
               qx[o + k] = mid + off;
               qw[o + i] = tw[i] * half;
               qw[o + k] = qw[o + i];
               for (int d = 0; d < nd; ++d) {
                    double hl = ((r2_l ? r2_l[d] : 0.) - (r1_l ? r1_l[d] : 0.)) / n_int;
                    double lo_l = (r1_l ? r1_l[d] : 0.) + s * hl, hi_l = lo_l + hl;
                    double mid_l = (lo_l + hi_l) / 2., half_l = (hi_l - lo_l) / 2.;
                    double off_l = half_l * tx[k];
                    qx_l[(o + i) * nd + d] = mid_l - off_l;
                    qx_l[(o + k) * nd + d] = mid_l + off_l;
                    qw_l[(o + i) * nd + d] = tw[i] * half_l;
                    qw_l[(o + k) * nd + d] = qw_l[(o + i) * nd + d];
               }
          }
     }
     free(tx); free(tw); free(tx_l); free(tw_l);
}

/* Brent's root finder on [a, b] (the role of misc/zeroin.c), used for the
   power-law range. */
static double brent(double a, double b, double (*f)(double, const double *), const double *par) {
     double fa = f(a, par), fb = f(b, par), c = a, fc = fa, d = b - a, e = d;
     for (int it = 0; it < 500; ++it) {
          if ((fb > 0 && fc > 0) || (fb < 0 && fc < 0)) { c = a; fc = fa; d = e = b - a; }
          if (fabs(fc) < fabs(fb)) { a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
          double tol = 2. * DBL_EPSILON * fabs(b), m = .5 * (c - b);
          if (fabs(m) <= tol || fb == 0.) break;
          if (fabs(e) >= tol && fabs(fa) > fabs(fb)) {
               double s = fb / fa, p, q, r;
               if (a == c) { p = 2. * m * s; q = 1. - s; }
               else {
                    q = fa / fc; r = fb / fc;
                    p = s * (2. * m * q * (q - r) - (b - a) * (r - 1.));
                    q = (q - 1.) * (r - 1.) * (s - 1.);
               }
               if (p > 0) q = -q; else p = -p;
               if (2. * p < 3. * m * q - fabs(tol * q) && p < fabs(.5 * e * q)) { e = d; d = p / q; }
               else { d = m; e = d; }
          } else { d = m; e = d; }
          a = b; fa = fb;
          b += fabs(d) > tol ? d : (m > 0 ? tol : -tol);
          fb = f(b, par);
     }
     return b;
}

/* Power-law range: r1 solving (r2 - r1)/ln(r2/r1) = a1 with r2 = 2 a1 (1 + a2) - r1
   (misc/rtutil_size_dist.c:95-121). */
static double powerlaw_f(double r1, const double *par) {
     double r2 = (1. + par[1]) * 2. * par[0] - r1;
     return (r2 - r1) / log(r2 / r1) - par[0];
}

static void powerlaw_range(double a1, double a2, double *r1, double *r2) {
     double par[2] = {a1, a2};
     *r1 = brent(DBL_EPSILON, a1, powerlaw_f, par);
     *r2 = (1. + a2) * 2. * a1 - *r1;
}

/* Number density on the grid, normalised to unit integral, with tangents
   (misc/rtutil_size_dist.c:200-646).  Arrays *_l are [N][nd] row-major.
   Returns N or -1. */
typedef struct {
     int dist, n_int1, n_int2, n_quad, nd;
     double a[5], r1, r2;
     const double *a_l[5], *r1_l, *r2_l;
} dist_in;

static int grid_size(int dist, int n_int1, int n_int2, int n_quad) {
     switch (dist) {
     case SIZE_DIST_MONO: return 1;
     case SIZE_DIST_MODIFIED_POWER_LAW: return n_int1 * n_quad + n_int2 * n_quad;
     case SIZE_DIST_MODIFIED_GAMMA: case SIZE_DIST_LOG_NORMAL: case SIZE_DIST_POWER_LAW:
     case SIZE_DIST_GAMMA: case SIZE_DIST_MODIFIED_BIMODAL_LOG_NORMAL: case SIZE_DIST_EXPONENTIAL:
          return n_int2 * n_quad;
     default: return -1;
     }
}

static double seed(const double *v, int d) { return v ? v[d] : 0.; }

static int size_dist(const dist_in *s, double *qx, double *qw, double *nr, double *qx_l,
                     double *qw_l, double *nr_l, double *r1o, double *r2o, double *r1o_l,
                     double *r2o_l, double *norm, double *norm_l) {
     int nd = s->nd, N = grid_size(s->dist, s->n_int1, s->n_int2, s->n_quad);
     double a1 = s->a[0], a2 = s->a[1], a3 = s->a[2], a5 = s->a[4];
     double tot = 0., *tot_l = calloc(nd + 1, sizeof(double));
     if (N < 0) { free(tot_l); return -1; }
     *r1o = s->r1; *r2o = s->r2;
     for (int d = 0; d < nd; ++d) { r1o_l[d] = seed(s->r1_l, d); r2o_l[d] = seed(s->r2_l, d); }

     if (s->dist == SIZE_DIST_MONO) {
          *r1o = *r2o = a1;
          for (int d = 0; d < nd; ++d) { r1o_l[d] = r2o_l[d] = 0.; qx_l[d] = seed(s->a_l[0], d);
                                         qw_l[d] = nr_l[d] = 0.; }
          qx[0] = a1; qw[0] = 1.; nr[0] = 1.;
          tot = 1.;
     }
     else if (s->dist == SIZE_DIST_MODIFIED_POWER_LAW) {
          int N1 = s->n_int1 * s->n_quad;
          size_grid(s->n_int1, s->n_quad, nd, 0., *r1o, NULL, r1o_l, qx, qw, qx_l, qw_l);
          size_grid(s->n_int2, s->n_quad, nd, *r1o, *r2o, r1o_l, r2o_l, qx + N1, qw + N1,
                    qx_l + (size_t) N1 * nd, qw_l + (size_t) N1 * nd);
          for (int i = 0; i < N; ++i) {
               if (i < N1) {
                    nr[i] = 1.;
                    for (int d = 0; d < nd; ++d) nr_l[i * nd + d] = 0.;
               } else {
                    double u = qx[i] / *r1o;
                    nr[i] = pow(u, a1);
                    for (int d = 0; d < nd; ++d)
                         nr_l[i * nd + d] = nr[i] * ((qx_l[i * nd + d] - u * r1o_l[d]) / *r1o * a1 / u
                                                     + seed(s->a_l[0], d) * log(u));
               }
          }
     }
     else {
          if (s->dist == SIZE_DIST_POWER_LAW) {
               powerlaw_range(a1, a2, r1o, r2o);
               for (int d = 0; d < nd; ++d) {
                    /* implicit differentiation of the two range conditions */
                    double L = log(*r2o) - log(*r1o);
                    double p = -L + (*r2o - *r1o) / *r1o, q = L - (*r2o - *r1o) / *r2o;
                    double rhs1 = seed(s->a_l[0], d) * L * L;
                    double rhs2 = seed(s->a_l[1], d) * 2. * a1 + (1. + a2) * 2. * seed(s->a_l[0], d);
                    r1o_l[d] = (q * rhs2 - rhs1) / (q - p);
                    r2o_l[d] = (p * rhs2 - rhs1) / (p - q);
               }
          }
          size_grid(s->n_int2, s->n_quad, nd, *r1o, *r2o, r1o_l, r2o_l, qx, qw, qx_l, qw_l);
          for (int i = 0; i < N; ++i) {
               double r = qx[i], lr = log(r);
               double *nl = nr_l + (size_t) i * nd;
               const double *xl = qx_l + (size_t) i * nd;
               switch (s->dist) {
               case SIZE_DIST_MODIFIED_GAMMA: {
                    /* n = r^a1 exp(-(a1/(a3 a2^a3)) r^a3) */
                    double b = pow(a2, a3), c = a3 * b, gcoef = a1 / c, h = pow(r, a3);
                    nr[i] = pow(r, a1) * exp(-gcoef * h);
                    for (int d = 0; d < nd; ++d) {
                         double a1d = seed(s->a_l[0], d), a2d = seed(s->a_l[1], d), a3d = seed(s->a_l[2], d);
                         double cl = a3d * (b / c) + (a2d * (a3 / a2) + a3d * log(a2));
                         double pl = xl[d] / r;
                         nl[d] = nr[i] * ((pl * a1 + a1d * lr) + h * ((-a1d + a1 * (-(pl * a3 + a3d * lr) + cl)) / c));
                    }
                    break;
               }
               case SIZE_DIST_LOG_NORMAL: {
                    double mu = log(a1), two_s2 = 2. * a2, dv = lr - mu, e = dv * dv / two_s2;
                    nr[i] = exp(-e) / r;
                    for (int d = 0; d < nd; ++d) {
                         double fl = xl[d] / r, mul = seed(s->a_l[0], d) / a1, tl = 2. * seed(s->a_l[1], d);
                         nl[d] = nr[i] * (-fl - ((2. * dv * (fl - mul) - e * tl) / two_s2));
                    }
                    break;
               }
               case SIZE_DIST_POWER_LAW:
                    nr[i] = 1. / pow(r, 3);
                    for (int d = 0; d < nd; ++d) nl[d] = -nr[i] * 3. * xl[d] / r;
                    break;
               case SIZE_DIST_GAMMA: {
                    /* n = r^((1-3b)/b) exp(-r/(a b)) */
                    double ab = a1 * a2, pw = (1. - 3. * a2) / a2, rp = pow(r, pw), ex = exp(-r / ab);
                    nr[i] = rp * ex;
                    for (int d = 0; d < nd; ++d) {
                         double abl = seed(s->a_l[0], d) * a2 + a1 * seed(s->a_l[1], d);
                         double pwl = (-3. * seed(s->a_l[1], d) - pw * seed(s->a_l[1], d)) / a2;
                         nl[d] = rp * ((xl[d] * (pw / r) + pwl * lr) + (-xl[d] + (r / ab) * abl) / ab) * ex;
                    }
                    break;
               }
               case SIZE_DIST_MODIFIED_BIMODAL_LOG_NORMAL: {
                    double m1 = log(a1), t1 = 2. * a2, m2 = log(a3), t2 = 2. * s->a[3];
                    double d1 = lr - m1, e1 = d1 * d1 / t1, d2 = lr - m2, e2 = d2 * d2 / t2;
                    double o1 = exp(-e1), o2 = exp(-e2), r2v = r * r, r4 = r2v * r2v;
                    nr[i] = (o1 + a5 * o2) / r4;
                    for (int d = 0; d < nd; ++d) {
                         double gl = xl[d] / r;
                         double h1 = ((-gl + seed(s->a_l[0], d) / a1) * 2. * d1 + e1 * 2. * seed(s->a_l[1], d)) * o1 / t1;
                         double h2 = ((-gl + seed(s->a_l[2], d) / a3) * 2. * d2 + e2 * 2. * seed(s->a_l[3], d)) * o2 / t2;
                         nl[d] = (h1 + seed(s->a_l[4], d) * o2 + a5 * h2 - nr[i] * 4. * r * r2v * xl[d]) / r4;
                    }
                    break;
               }
               case SIZE_DIST_EXPONENTIAL:
                    /* no tangents in the reference for this kind (misc/rtutil_size_dist.c:607-613) */
                    nr[i] = 2. * exp(-a1 * 2. * r);
                    for (int d = 0; d < nd; ++d) nl[d] = 0.;
                    break;
               }
          }
     }
     if (s->dist != SIZE_DIST_MONO) {
          for (int i = 0; i < N; ++i) {
               tot += nr[i] * qw[i];
               for (int d = 0; d < nd; ++d)
                    tot_l[d] += nr_l[i * nd + d] * qw[i] + nr[i] * qw_l[i * nd + d];
          }
          if (s->dist == SIZE_DIST_EXPONENTIAL)
               for (int d = 0; d < nd; ++d) tot_l[d] = 0.;
     }
     *norm = tot;
     for (int d = 0; d < nd; ++d) norm_l[d] = tot_l[d];
     for (int i = 0; i < N; ++i) {
          nr[i] /= tot;
          for (int d = 0; d < nd; ++d)
               nr_l[i * nd + d] = (nr_l[i * nd + d] - nr[i] * tot_l[d]) / tot;
     }
     free(tot_l);
     return N;
}

/* Moments of the distribution (misc/rtutil_size_dist.c:653-794). */
static void dist_stats(int N, int nd, const double *qx, const double *qw, const double *nr,
                       const double *qx_l, const double *qw_l, const double *nr_l,
                       double *reff, double *veff, double *gavg, double *vavg,
                       double *ravg, double *rvw, double *reff_l, double *veff_l,
                       double *gavg_l, double *vavg_l, double *ravg_l, double *rvw_l) {
     double m[4] = {0, 0, 0, 0};
     double *ml = calloc((size_t) 4 * (nd + 1), sizeof(double));
     for (int i = 0; i < N; ++i) {
          double t = nr[i] * qx[i] * qw[i];
          for (int k = 0; k < 4; ++k) { m[k] += t; t *= qx[i]; }
          for (int d = 0; d < nd; ++d) {
               double h = qx[i] * qw[i] * nr_l[i * nd + d];
               double p = nr[i] * qw[i] * qx_l[i * nd + d];
               double q = nr[i] * qx[i] * qw_l[i * nd + d];
               double rp = 1.;
               for (int k = 0; k < 4; ++k) {
                    ml[k * (nd + 1) + d] += (h + (k + 1) * p + q) * rp;
                    rp *= qx[i];
               }
          }
     }
     *gavg = m[1] * ORACLE_PI;
     *vavg = 4. / 3. * ORACLE_PI * m[2];
     *ravg = m[0];
     *rvw = 4. / 3. * ORACLE_PI * m[3] / *vavg;
     *reff = m[2] / m[1];
     for (int d = 0; d < nd; ++d) {
          gavg_l[d] = ml[1 * (nd + 1) + d] * ORACLE_PI;
          vavg_l[d] = 4. / 3. * ORACLE_PI * ml[2 * (nd + 1) + d];
          ravg_l[d] = ml[d];
          rvw_l[d] = 4. / 3. * ORACLE_PI * (ml[3 * (nd + 1) + d] / *vavg - m[3] * vavg_l[d] / (*vavg * *vavg));
          reff_l[d] = ml[2 * (nd + 1) + d] / m[1] - m[2] * ml[1 * (nd + 1) + d] / (m[1] * m[1]);
     }
     double v = 0., *vl = calloc(nd + 1, sizeof(double));
     for (int i = 0; i < N; ++i) {
          double dr = qx[i] - *reff, r2w = qx[i] * qx[i] * qw[i];
          v += nr[i] * dr * dr * r2w;
          for (int d = 0; d < nd; ++d)
               vl[d] += nr_l[i * nd + d] * dr * dr * r2w + nr[i] *
                        (2. * dr * r2w * (qx_l[i * nd + d] - reff_l[d]) +
                         dr * dr * (2. * qx[i] * qw[i] * qx_l[i * nd + d] + qx[i] * qx[i] * qw_l[i * nd + d]));
     }
     double b = *gavg * *reff * *reff;
     *veff = v * ORACLE_PI / b;
     for (int d = 0; d < nd; ++d)
          veff_l[d] = vl[d] * ORACLE_PI / b - (gavg_l[d] * (*veff * *reff / b) * *reff +
                                               (*veff * *reff / b) * *gavg * 2. * reff_l[d]);
     free(ml); free(vl);
}

/* ------------------------------------------------------------------------- */
/* Per-size Lorenz-Mie coefficients (src/lmie_core.c:263-498) with tangents
   along nt directions given by dx[t] (real) and dm[t] (complex).  Outputs
   a[1..n1], b[1..n1] and a_t/b_t as [n][nt].  Returns n1. */
typedef struct {
     int n1, n2;
     cplx *a, *b, *a_t, *b_t;
} mie_ab;

static int mie_coefs(double x, cplx m, int nt, const double *dx, const cplx *dm, mie_ab *o) {
     int n1 = order_n1(x);
     cplx z = m * x;
     int n2 = order_n2(n1, z);
     o->n1 = n1; o->n2 = n2;
     int T = nt > 0 ? nt : 1;
     double *r = malloc(sizeof(double) * (n1 + 1)), *psi = malloc(sizeof(double) * (n1 + 1));
     double *chi = malloc(sizeof(double) * (n1 + 1));
     cplx *D = malloc(sizeof(cplx) * (n2 + 1));
     double *r_t = calloc((size_t) (n1 + 1) * T, sizeof(double));
     double *psi_t = calloc((size_t) (n1 + 1) * T, sizeof(double));
     double *chi_t = calloc((size_t) (n1 + 1) * T, sizeof(double));
     cplx *D_t = calloc((size_t) (n2 + 1) * T, sizeof(cplx));
     cplx *dz = calloc(T, sizeof(cplx));
     for (int t = 0; t < nt; ++t) dz[t] = dm[t] * x + m * dx[t];

     /* downward ratio r_j = psi_j / psi_{j-1}, r_{n1} = 0 */
     r[n1] = 0.;
     for (int j = n1 - 1; j >= 0; --j) {
          double c = (2. * j + 1.) / x;
          r[j] = 1. / (c - r[j + 1]);
          for (int t = 0; t < nt; ++t)
               r_t[j * T + t] = (c / x * dx[t] + r_t[(j + 1) * T + t]) * r[j] * r[j];
     }
     double sx = sin(x), cx = cos(x);
     psi[0] = sx;
     for (int t = 0; t < nt; ++t) psi_t[t] = cx * dx[t];
     for (int j = 1; j <= n1; ++j) {
          psi[j] = r[j] * psi[j - 1];
          for (int t = 0; t < nt; ++t)
               psi_t[j * T + t] = r_t[j * T + t] * psi[j - 1] + r[j] * psi_t[(j - 1) * T + t];
     }
     /* upward chi: chi_0 = cos x, chi_1 = cos x / x + sin x */
     chi[0] = cx;
     chi[1] = cx / x + sx;
     for (int t = 0; t < nt; ++t) {
          chi_t[t] = -sx * dx[t];
          chi_t[T + t] = -dx[t] / (x * x) * cx + chi_t[t] / x + cx * dx[t];
     }
     for (int j = 1; j < n1; ++j) {
          double c = (2. * j + 1.) / x;
          chi[j + 1] = c * chi[j] - chi[j - 1];
          for (int t = 0; t < nt; ++t)
               chi_t[(j + 1) * T + t] = -c / x * dx[t] * chi[j] + c * chi_t[j * T + t] - chi_t[(j - 1) * T + t];
     }
     /* downward logarithmic derivative D_j(z), D_{n2} = 0 */
     D[n2] = 0.;
     for (int j = n2 - 1; j >= 0; --j) {
          cplx c1 = (j + 1.) / z;
          cplx c2 = D[j + 1] + c1;
          D[j] = c1 - 1. / c2;
          for (int t = 0; t < nt; ++t) {
               cplx dc1 = -c1 / z * dz[t];
               D_t[j * T + t] = dc1 + (D_t[(j + 1) * T + t] + dc1) / (c2 * c2);
          }
     }
     o->a = calloc(n1 + 2, sizeof(cplx));
     o->b = calloc(n1 + 2, sizeof(cplx));
     o->a_t = calloc((size_t) (n1 + 2) * T, sizeof(cplx));
     o->b_t = calloc((size_t) (n1 + 2) * T, sizeof(cplx));
     for (int j = 1; j <= n1; ++j) {
          double jx = j / x;
          cplx zj = psi[j] + I * chi[j], zjm = psi[j - 1] + I * chi[j - 1];
          for (int ab = 0; ab < 2; ++ab) {
               cplx u = ab == 0 ? D[j] / m + jx : m * D[j] + jx;
               cplx num = u * psi[j] - psi[j - 1];
               cplx den = u * zj - zjm;
               cplx v = num / den;
               (ab == 0 ? o->a : o->b)[j] = v;
               for (int t = 0; t < nt; ++t) {
                    cplx dD = D_t[j * T + t];
                    double djx = -jx / x * dx[t];
                    cplx du = ab == 0 ? dD / m - D[j] / (m * m) * dm[t] + djx
                                      : dm[t] * D[j] + m * dD + djx;
                    cplx dzj = psi_t[j * T + t] + I * chi_t[j * T + t];
                    cplx dzjm = psi_t[(j - 1) * T + t] + I * chi_t[(j - 1) * T + t];
                    cplx dnum = du * psi[j] + u * psi_t[j * T + t] - psi_t[(j - 1) * T + t];
                    cplx dden = du * zj + u * dzj - dzjm;
                    (ab == 0 ? o->a_t : o->b_t)[j * T + t] = dnum / den - num * dden / (den * den);
               }
          }
     }
     free(r); free(psi); free(chi); free(D); free(r_t); free(psi_t); free(chi_t); free(D_t); free(dz);
     return n1;
}

static void mie_free(mie_ab *o) { free(o->a); free(o->b); free(o->a_t); free(o->b_t); }

/* a_n, b_n of one sphere (x, m = mr - i mi), interleaved re/im for n = 0..n1
   (index 0 unused).  ab must hold 2 (n1 + 1) doubles each.  Returns n1. */
int oracle_mie_ab(double x, double mr, double mi, int cap, double *a, double *b) {
     mie_ab o;
     int n1 = mie_coefs(x, mr - I * mi, 0, NULL, NULL, &o);
     if (n1 + 1 > cap) { mie_free(&o); return -1; }
     for (int n = 0; n <= n1; ++n) {
          a[2 * n] = creal(o.a[n]); a[2 * n + 1] = cimag(o.a[n]);
          b[2 * n] = creal(o.b[n]); b[2 * n + 1] = cimag(o.b[n]);
     }
     mie_free(&o);
     return n1;
}

int oracle_order_n1(double x) { return order_n1(x); }
int oracle_order_n2(double x, double mr, double mi) { return order_n2(order_n1(x), (mr - I * mi) * x); }

/* pi_n(mu), tau_n(mu) for n = 1..n (src/lmie_core.c:669-691). */
void oracle_pitau(double mu, int n, double *pi, double *tau) {
     double p0 = 0., p1 = 1.;
     for (int k = 1; k <= n; ++k) {
          double a = mu * p1, d = a - p0;
          pi[k - 1] = p1;
          tau[k - 1] = k * d - p0;
          p0 = p1;
          p1 = a + (k + 1.) / k * d;
     }
}

/* ------------------------------------------------------------------------- */
/* Generalized spherical functions P^{0,0}, P^{0,2}, P^{2,2}, P^{2,-2} for
   l = 0..n_l-1 by their three-term recurrences in l
   (misc/rtutil_math.c:867-1017). */
void oracle_gsf(int n_l, double mu, double *p00, double *p02, double *p22, double *p2m2) {
     double mu2 = mu * mu;
     for (int l = 0; l < n_l && l < 2; ++l) {
          p00[l] = l == 0 ? 1. : mu;
          p02[l] = p22[l] = p2m2[l] = 0.;
     }
     if (n_l > 2) {
          p00[2] = (3. * mu2 - 1.) / 2.;
          p02[2] = .25 * sqrt(6.) * (mu2 - 1.);
          p22[2] = .25 * (1. + mu) * (1. + mu);
          p2m2[2] = .25 * (1. - mu) * (1. - mu);
     }
     for (int l = 2; l < n_l - 1; ++l) {
          double L = l, L1 = l + 1., t = 2. * l + 1.;
          double sa = sqrt(L * L - 4.), sb = sqrt(L1 * L1 - 4.);
          p00[l + 1] = (t * mu * p00[l] - L * p00[l - 1]) / L1;
          p02[l + 1] = (t * mu * p02[l] - sa * p02[l - 1]) / sb;
          double c = t * L * L1 * mu, k4 = 4. * t, kb = L1 * (L * L - 4.), kd = L * (L1 * L1 - 4.);
          p22[l + 1] = ((c - k4) * p22[l] - kb * p22[l - 1]) / kd;
          p2m2[l + 1] = ((c + k4) * p2m2[l] - kb * p2m2[l - 1]) / kd;
     }
}

/* Legendre P_l(mu), l = 0..n_l-1 (misc/rtutil_math.c:477-538). */
static void legendre(int n_l, double mu, double *p) {
     if (n_l > 0) p[0] = 1.;
     if (n_l > 1) p[1] = mu;
     for (int l = 1; l < n_l - 1; ++l) {
          double u = mu * p[l];
          p[l + 1] = l / (l + 1.) * (u - p[l - 1]) + u;
     }
}

/* Expansion of the 4 distinct phase-matrix elements pf4 = [F11, F33, F12, F34]
   at the n_q Gauss nodes into generalized-spherical-function coefficients
   [beta, alpha, zeta, delta, gamma, epsilon]: coefs[6][n_coef].  Uses the
   +/- mu symmetry of the nodes (misc/rtutil_scat_int.c:20-230 with
   src/lmie_schedule.c:927-935).  Returns the truncated count. */
static int expand_gc(int n_coef, int n_q, double **pf4, const double *mu, const double *w,
                     double **coefs, double accuracy) {
     double *P = malloc(sizeof(double) * 4 * n_coef);
     for (int k = 0; k < 6; ++k) memset(coefs[k], 0, sizeof(double) * n_coef);
     int h = n_q / 2;
     for (int i = 0; i < h; ++i) {
          int ip = h + i, im = h - 1 - i;   /* +mu node and its mirror */
          if (n_q % 2) { ip = h + 1 + i; im = h - 1 - i; }
          oracle_gsf(n_coef, mu[ip], P, P + n_coef, P + 2 * n_coef, P + 3 * n_coef);
          double wp = w[ip];
          /* six element rows in the order F11, F11, F33, F33, F12, F34 */
          double em[6], ep[6];
          int src[6] = {0, 0, 1, 1, 2, 3};
          for (int k = 0; k < 6; ++k) { em[k] = pf4[src[k]][im] * wp; ep[k] = pf4[src[k]][ip] * wp; }
          double t;
          t = em[1]; em[1] = t + em[2]; em[2] = t - em[2];
          t = ep[1]; ep[1] = t + ep[2]; ep[2] = t - ep[2];
          double s = 1.;
          for (int l = 0; l < n_coef; ++l) {
               double p00 = P[l], p02 = P[n_coef + l], p22 = P[2 * n_coef + l], p2m2 = P[3 * n_coef + l];
               coefs[0][l] += p00 * (em[0] * s + ep[0]);
               coefs[1][l] += p2m2 * em[1] * s + p22 * ep[1];
               coefs[2][l] += p22 * em[2] * s + p2m2 * ep[2];
               coefs[3][l] += p00 * (em[3] * s + ep[3]);
               coefs[4][l] += p02 * (em[4] * s + ep[4]);
               coefs[5][l] += p02 * (em[5] * s + ep[5]);
               s = -s;
          }
     }
     if (n_q % 2) {
          int ic = h;
          oracle_gsf(n_coef, mu[ic], P, P + n_coef, P + 2 * n_coef, P + 3 * n_coef);
          double e[6];
          int src[6] = {0, 0, 1, 1, 2, 3};
          for (int k = 0; k < 6; ++k) e[k] = pf4[src[k]][ic] * w[ic];
          double t = e[1]; e[1] = t + e[2]; e[2] = t - e[2];
          for (int l = 0; l < n_coef; ++l) {
               coefs[0][l] += P[l] * e[0];
               coefs[1][l] += P[2 * n_coef + l] * e[1];
               coefs[2][l] += P[3 * n_coef + l] * e[2];
               coefs[3][l] += P[l] * e[3];
               coefs[4][l] += P[n_coef + l] * e[4];
               coefs[5][l] += P[n_coef + l] * e[5];
          }
     }
     for (int l = 0; l < n_coef; ++l) {
          double t = coefs[1][l];
          coefs[1][l] = (t + coefs[2][l]) / 2.;
          coefs[2][l] = (t - coefs[2][l]) / 2.;
          double f = (2 * l + 1) / 2.;
          for (int k = 0; k < 6; ++k) coefs[k][l] *= f;
     }
     free(P);
     int l;
     for (l = 0; l < n_coef; ++l)
          if (fabs(coefs[0][l]) <= accuracy) return l + 1;
     return l;
}

/* Legendre-series variant (misc/rtutil_scat_int.c:353-480). */
static int expand_lc(int n_coef, int n_q, double **pf4, const double *mu, const double *w,
                     double **coefs, double accuracy) {
     double *P = malloc(sizeof(double) * n_coef);
     int src[6] = {0, 0, 1, 1, 2, 3};
     for (int k = 0; k < 6; ++k) memset(coefs[k], 0, sizeof(double) * n_coef);
     int h = n_q / 2;
     if (n_q % 2) {
          legendre(n_coef, mu[h], P);
          for (int l = 0; l < n_coef; ++l)
               for (int k = 0; k < 6; ++k) coefs[k][l] += P[l] * pf4[src[k]][h] * w[h];
     }
     for (int i = 0; i < h; ++i) {
          int ip = (n_q % 2 ? h + 1 : h) + i, im = h - 1 - i;
          legendre(n_coef, mu[ip], P);
          double s = 1.;
          for (int l = 0; l < n_coef; ++l) {
               for (int k = 0; k < 6; ++k)
                    coefs[k][l] += P[l] * (pf4[src[k]][im] * w[ip] * s + pf4[src[k]][ip] * w[ip]);
               s = -s;
          }
     }
     for (int l = 0; l < n_coef; ++l) {
          coefs[5][l] = -coefs[5][l];
          double f = (2. * l + 1.) / 2.;
          for (int k = 0; k < 6; ++k) coefs[k][l] *= f;
     }
     free(P);
     int l;
     for (l = 0; l < n_coef; ++l)
          if (fabs(coefs[0][l]) <= accuracy) return l + 1;
     return l;
}

/* Scattering matrix at n_theta equally spaced angles from the coefficients
   (misc/rtutil_scat_int.c:487-797). */
static void synth_pf(int n_coef, int n_theta, double **xc, double *theta, double **pf, int gsf) {
     double dth = 180. / (n_theta - 1);
     double *P = malloc(sizeof(double) * 4 * (n_coef > 0 ? n_coef : 1));
     for (int i = 0; i < n_theta; ++i) {
          theta[i] = i * dth;
          double mu = cos(theta[i] * (ORACLE_PI / 180.));
          double s[6] = {0, 0, 0, 0, 0, 0};
          if (gsf) {
               oracle_gsf(n_coef, mu, P, P + n_coef, P + 2 * n_coef, P + 3 * n_coef);
               for (int l = 0; l < n_coef; ++l) {
                    s[0] += xc[0][l] * P[l];
                    s[3] += xc[3][l] * P[l];
                    if (l >= 2) {
                         s[4] += xc[4][l] * P[n_coef + l];
                         s[5] += xc[5][l] * P[n_coef + l];
                         s[1] += (xc[1][l] + xc[2][l]) * P[2 * n_coef + l];
                         s[2] += (xc[1][l] - xc[2][l]) * P[3 * n_coef + l];
                    }
               }
               pf[0][i] = s[0]; pf[3][i] = s[3]; pf[4][i] = s[4]; pf[5][i] = s[5];
               pf[1][i] = (s[1] + s[2]) / 2.;
               pf[2][i] = (s[1] - s[2]) / 2.;
          } else {
               legendre(n_coef, mu, P);
               for (int k = 0; k < 6; ++k) {
                    double a = 0.;
                    for (int l = 0; l < n_coef; ++l) a += xc[k][l] * P[l];
                    pf[k][i] = a;
               }
               pf[5][i] = -pf[5][i];
          }
     }
     free(P);
}

/* ------------------------------------------------------------------------- */
/* The size-integrated solution (src/lmie_schedule.c:942-1268 with the per-size
   kernel src/lmie_core.c:95-842).  Raw sums, normalisation, expansion. */
typedef struct core_in_s {
     int N, nd, nd1, nd2, *idx1, *idx2, save;
     double lambda, *lambda_l, mr, mi, *mr_l, *mi_l;
     const double *qx, *qw, *nr, *qx_l, *qw_l, *nr_l;
     double r2;
} core_in;

typedef struct {
     const struct core_in_s *c;
     int t, P, nq, n1;
     const double *fv1, *fv2, *fv3, *fv4, *mu;
     cplx m;
     const cplx *m_l;
     double ***save1, ****save2;
     double cx[4], *cx_l, **pf, ***pf_l;
} part_job;

static void *core_part(void *arg) {
     part_job *J = (part_job *) arg;
     const core_in *c = J->c;
     int nd = c->nd, nd1 = c->nd1, nd2 = c->nd2, nq = J->nq;
     const double *fv1 = J->fv1, *fv2 = J->fv2, *fv3 = J->fv3, *fv4 = J->fv4, *mu = J->mu;
     cplx m = J->m;
     const cplx *m_l = J->m_l;
     double ***save1 = J->save1;
     double ****save2 = J->save2;
     (void) nd;
          double *dx = calloc(nd1 + 1, sizeof(double)), *fl1 = calloc(nd1 + 1, sizeof(double));
          double *fl2 = calloc(nd2 + 1, sizeof(double));
          cplx *dm = calloc(nd1 + 1, sizeof(cplx));
          int T = nd1 > 0 ? nd1 : 1;
          for (int i = J->t; i < c->N; i += J->P) {
               double r = c->qx[i], x = 2. * ORACLE_PI * r / c->lambda;
               double f = c->nr[i] * c->qw[i];
               for (int e = 0; e < nd1; ++e) {
                    int d = c->idx1[e];
                    dx[e] = 2. * ORACLE_PI * (c->qx_l[i * nd + d] - r * c->lambda_l[d] / c->lambda) / c->lambda;
                    dm[e] = m_l[d];
                    fl1[e] = c->nr_l[i * nd + d] * c->qw[i] + c->nr[i] * c->qw_l[i * nd + d];
               }
               for (int e = 0; e < nd2; ++e) {
                    int d = c->idx2[e];
                    fl2[e] = c->nr_l[i * nd + d] * c->qw[i];
               }
               mie_ab ab;
               int n = mie_coefs(x, m, nd1, dx, dm, &ab);
               cplx *a = ab.a, *b = ab.b;
               /* cross-section series (src/lmie_core.c:504-611) */
               double se = 0, ss = 0, sg = 0;
               cplx cb = 0;
               double *se_t = calloc(T, sizeof(double)), *ss_t = calloc(T, sizeof(double));
               double *sg_t = calloc(T, sizeof(double));
               cplx *cb_t = calloc(T, sizeof(cplx));
               double sgn = -1.;
               for (int j = 1; j <= n; ++j) {
                    se += fv1[j] * creal(a[j] + b[j]);
                    ss += fv1[j] * (creal(a[j] * conj(a[j])) + creal(b[j] * conj(b[j])));
                    cb += fv1[j] * (a[j] - b[j]) * sgn;
                    for (int t = 0; t < nd1; ++t) {
                         cplx at = ab.a_t[j * T + t], bt = ab.b_t[j * T + t];
                         se_t[t] += fv1[j] * creal(at + bt);
                         ss_t[t] += fv1[j] * 2. * (creal(a[j] * conj(at)) + creal(b[j] * conj(bt)));
                         cb_t[t] += fv1[j] * (at - bt) * sgn;
                    }
                    sgn = -sgn;
               }
               for (int j = 1; j < n; ++j) {
                    sg += fv3[j] * creal(a[j] * conj(a[j + 1]) + b[j] * conj(b[j + 1])) +
                          fv4[j] * creal(a[j] * conj(b[j]));
                    for (int t = 0; t < nd1; ++t) {
                         cplx at = ab.a_t[j * T + t], bt = ab.b_t[j * T + t];
                         cplx at1 = ab.a_t[(j + 1) * T + t], bt1 = ab.b_t[(j + 1) * T + t];
                         sg_t[t] += fv3[j] * creal(at * conj(a[j + 1]) + a[j] * conj(at1) +
                                                   bt * conj(b[j + 1]) + b[j] * conj(bt1)) +
                                    fv4[j] * creal(at * conj(b[j]) + a[j] * conj(bt));
                    }
               }
               double cb2 = creal(cb * conj(cb));
               J->cx[0] += f * se; J->cx[1] += f * ss; J->cx[2] += f * cb2; J->cx[3] += f * sg;
               if (c->save == 1) {
                    (*save1)[0][i] = se; (*save1)[1][i] = ss; (*save1)[2][i] = cb2; (*save1)[3][i] = sg;
               }
               for (int t = 0; t < nd1; ++t) {
                    int d = c->idx1[t];
                    J->cx_l[4 * d] += fl1[t] * se + f * se_t[t];
                    J->cx_l[4 * d + 1] += fl1[t] * ss + f * ss_t[t];
                    J->cx_l[4 * d + 2] += fl1[t] * cb2 + f * 2. * creal(cb * conj(cb_t[t]));
                    J->cx_l[4 * d + 3] += fl1[t] * sg + f * sg_t[t];
               }
               for (int t = 0; t < nd2; ++t) {
                    int d = c->idx2[t];
                    J->cx_l[4 * d] += fl2[t] * se; J->cx_l[4 * d + 1] += fl2[t] * ss;
                    J->cx_l[4 * d + 2] += fl2[t] * cb2; J->cx_l[4 * d + 3] += fl2[t] * sg;
               }
               /* angular sums over the Gauss nodes (src/lmie_core.c:617-791) */
               cplx *sm_ = malloc(sizeof(cplx) * (n + 1)), *df_ = malloc(sizeof(cplx) * (n + 1));
               cplx *sm_t = malloc(sizeof(cplx) * (n + 1) * T), *df_t = malloc(sizeof(cplx) * (n + 1) * T);
               for (int k = 1; k <= n; ++k) {
                    sm_[k] = fv4[k] * (a[k] + b[k]);
                    df_[k] = fv4[k] * (a[k] - b[k]);
                    for (int t = 0; t < nd1; ++t) {
                         sm_t[k * T + t] = fv4[k] * (ab.a_t[k * T + t] + ab.b_t[k * T + t]);
                         df_t[k * T + t] = fv4[k] * (ab.a_t[k * T + t] - ab.b_t[k * T + t]);
                    }
               }
               cplx *spt = malloc(sizeof(cplx) * 2 * T), *smt = malloc(sizeof(cplx) * 2 * T);
               for (int j = 0; j < nq / 2; ++j) {
                    cplx sp[2] = {0, 0}, sm[2] = {0, 0};
                    for (int t = 0; t < 2 * T; ++t) spt[t] = smt[t] = 0;
                    double p0 = 0., p1 = 1., sg2 = 1.;
                    for (int k = 1; k <= n; ++k) {
                         double u = mu[j] * p1, dd = u - p0;
                         double tau = k * dd - p0;
                         double pp = p1 + tau, pm = p1 - tau;
                         sp[0] += pp * sm_[k]; sm[0] += pm * df_[k];
                         sp[1] += pm * sm_[k] * sg2; sm[1] += pp * df_[k] * sg2;
                         for (int t = 0; t < nd1; ++t) {
                              spt[t] += pp * sm_t[k * T + t]; smt[t] += pm * df_t[k * T + t];
                              spt[T + t] += pm * sm_t[k * T + t] * sg2;
                              smt[T + t] += pp * df_t[k * T + t] * sg2;
                         }
                         p0 = p1;
                         p1 = u + fv2[k] * dd;
                         sg2 = -sg2;
                    }
                    for (int h = 0; h < 2; ++h) {
                         int jj = h == 0 ? j : nq - j - 1;
                         cplx s1 = .5 * (sp[h] + sm[h]), s2 = .5 * (sp[h] - sm[h]);
                         double q1 = creal(s1 * conj(s1)), q2 = creal(s2 * conj(s2));
                         cplx x12 = s1 * conj(s2);
                         double P[4] = {(q1 + q2) * .5, creal(x12), -(q1 - q2) * .5, cimag(x12)};
                         for (int k = 0; k < 4; ++k) J->pf[k][jj] += f * P[k];
                         if (c->save == 1)
                              for (int k = 0; k < 4; ++k) (*save2)[k][i][jj] = P[k];
                         for (int t = 0; t < nd1; ++t) {
                              cplx s1t = .5 * (spt[h * T + t] + smt[h * T + t]);
                              cplx s2t = .5 * (spt[h * T + t] - smt[h * T + t]);
                              double dq1 = 2. * creal(s1t * conj(s1)), dq2 = 2. * creal(s2t * conj(s2));
                              cplx dx12 = s1t * conj(s2) + s1 * conj(s2t);
                              double dP[4] = {(dq1 + dq2) * .5, creal(dx12), -(dq1 - dq2) * .5, cimag(dx12)};
                              int d = c->idx1[t];
                              for (int k = 0; k < 4; ++k) J->pf_l[d][k][jj] += fl1[t] * P[k] + f * dP[k];
                         }
                         for (int t = 0; t < nd2; ++t) {
                              int d = c->idx2[t];
                              for (int k = 0; k < 4; ++k) J->pf_l[d][k][jj] += fl2[t] * P[k];
                         }
                    }
               }
               free(spt); free(smt); free(sm_); free(df_); free(sm_t); free(df_t);
               free(se_t); free(ss_t); free(sg_t); free(cb_t);
               mie_free(&ab);
          }
          free(dx); free(fl1); free(fl2); free(dm);
     return arg;
}

static int core_solution(const core_in *c, double *cext, double *csca, double *cbak, double *g,
                         double **gc, double **lc, double *cext_l, double *csca_l,
                         double *cbak_l, double *g_l, double ***gc_l, double ***lc_l,
                         double ***save1, double ****save2, double accuracy, int n_threads) {
     int nd = c->nd, nd2 = c->nd2;
     double x2 = 2. * ORACLE_PI * c->r2 / c->lambda;
     cplx m = c->mr - I * c->mi;
     int n1 = order_n1(x2);
     int nq = (!gc && !lc) ? 0 : 2 * n1;
     int n_coef = 2 * n1;
     double *fv1 = calloc(n1 + 2, sizeof(double)), *fv2 = calloc(n1 + 2, sizeof(double));
     double *fv3 = calloc(n1 + 2, sizeof(double)), *fv4 = calloc(n1 + 2, sizeof(double));
     for (int n = 1; n <= n1; ++n) {
          fv1[n] = 2. * n + 1.;
          fv2[n] = (n + 1.) / n;
          fv3[n] = (n + 2.) / fv2[n];
          fv4[n] = fv1[n] / (n * (n + 1.));
     }
     double *mu = calloc(nq + 1, sizeof(double)), *wq = calloc(nq + 1, sizeof(double));
     if (nq > 0) oracle_gauss_leg(nq, 0, -1., 1., NULL, NULL, mu, wq, NULL, NULL);
     double **pf = blk2(4, nq > 0 ? nq : 1);
     double ***pf_l = blk3(nd > 0 ? nd : 1, 4, nq > 0 ? nq : 1);
     cplx *m_l = calloc(nd + 1, sizeof(cplx));
     for (int d = 0; d < nd; ++d) m_l[d] = c->mr_l[d] - I * c->mi_l[d];
     if (c->save == 1) {
          *save1 = blk2(4, c->N);
          *save2 = blk3(4, c->N, nq > 0 ? nq : 1);
     }
     double Cext = 0, Csca = 0, Cbak = 0, G = 0;
     for (int d = 0; d < nd; ++d) cext_l[d] = csca_l[d] = cbak_l[d] = g_l[d] = 0.;

     if (c->save == 2) {
          /* re-weight the cached per-size sums (src/lmie_schedule.c:1087-1124) */
          for (int i = 0; i < c->N; ++i) {
               double f = c->nr[i] * c->qw[i];
               Cext += (*save1)[0][i] * f; Csca += (*save1)[1][i] * f;
               Cbak += (*save1)[2][i] * f; G += (*save1)[3][i] * f;
               for (int j = 0; j < nq; ++j)
                    for (int k = 0; k < 4; ++k) pf[k][j] += (*save2)[k][i][j] * f;
               for (int e = 0; e < nd2; ++e) {
                    int d = c->idx2[e];
                    double fl = c->nr_l[i * nd + d] * c->qw[i] + c->nr[i] * c->qw_l[i * nd + d];
                    cext_l[d] += (*save1)[0][i] * fl; csca_l[d] += (*save1)[1][i] * fl;
                    cbak_l[d] += (*save1)[2][i] * fl; g_l[d] += (*save1)[3][i] * fl;
                    for (int j = 0; j < nq; ++j)
                         for (int k = 0; k < 4; ++k) pf_l[d][k][j] += (*save2)[k][i][j] * fl;
               }
          }
     }
     else {
          /* round-robin partition of the sizes over threads with private
             accumulators, reduced in thread order (src/lmie_schedule.c:33-61,
             :217-366) */
          int P = n_threads > 0 ? n_threads : 1;
          if (P > c->N) P = c->N;
          part_job *jobs = calloc(P, sizeof(part_job));
          pthread_t *th = calloc(P, sizeof(pthread_t));
          for (int t = 0; t < P; ++t) {
               jobs[t].c = c; jobs[t].t = t; jobs[t].P = P; jobs[t].nq = nq; jobs[t].n1 = n1;
               jobs[t].fv1 = fv1; jobs[t].fv2 = fv2; jobs[t].fv3 = fv3; jobs[t].fv4 = fv4;
               jobs[t].mu = mu; jobs[t].m = m; jobs[t].m_l = m_l;
               jobs[t].save1 = save1; jobs[t].save2 = save2;
               jobs[t].cx_l = calloc(4 * (size_t) (nd + 1), sizeof(double));
               jobs[t].pf = blk2(4, nq > 0 ? nq : 1);
               jobs[t].pf_l = blk3(nd > 0 ? nd : 1, 4, nq > 0 ? nq : 1);
               if (P > 1) pthread_create(&th[t], NULL, core_part, &jobs[t]);
               else core_part(&jobs[t]);
          }
          for (int t = 0; t < P; ++t) {
               if (P > 1) pthread_join(th[t], NULL);
               Cext += jobs[t].cx[0]; Csca += jobs[t].cx[1]; Cbak += jobs[t].cx[2]; G += jobs[t].cx[3];
               for (int d = 0; d < nd; ++d) {
                    cext_l[d] += jobs[t].cx_l[4 * d]; csca_l[d] += jobs[t].cx_l[4 * d + 1];
                    cbak_l[d] += jobs[t].cx_l[4 * d + 2]; g_l[d] += jobs[t].cx_l[4 * d + 3];
               }
               for (int j = 0; j < nq; ++j)
                    for (int k = 0; k < 4; ++k) {
                         pf[k][j] += jobs[t].pf[k][j];
                         for (int d = 0; d < nd; ++d) pf_l[d][k][j] += jobs[t].pf_l[d][k][j];
                    }
               free(jobs[t].cx_l); free(jobs[t].pf); free(jobs[t].pf_l);
          }
          free(jobs); free(th);
     }
     /* normalisation (src/lmie_schedule.c:1158-1199) */
     double a = 2. / Csca;
     for (int d = 0; d < nd; ++d) {
          double al = -a * csca_l[d] / Csca;
          g_l[d] = g_l[d] * a + G * al;
          for (int j = 0; j < nq; ++j)
               for (int k = 0; k < 4; ++k) pf_l[d][k][j] = pf_l[d][k][j] * a + pf[k][j] * al;
     }
     G = G / Csca * 2.;
     for (int j = 0; j < nq; ++j)
          for (int k = 0; k < 4; ++k) pf[k][j] *= a;
     double s = c->lambda * c->lambda / (2. * ORACLE_PI);
     for (int d = 0; d < nd; ++d) {
          double sl = 2. * c->lambda_l[d] * c->lambda / (2. * ORACLE_PI);
          cext_l[d] = cext_l[d] * s + Cext * sl;
          csca_l[d] = csca_l[d] * s + Csca * sl;
          cbak_l[d] = (cbak_l[d] * s + Cbak * sl) / (2. * 4. * ORACLE_PI);
     }
     *cext = Cext * s;
     *csca = Csca * s;
     *cbak = Cbak * s / (2. * 4. * ORACLE_PI);
     *g = G;
     int n_coef2 = 0;
     if (gc) {
          n_coef2 = expand_gc(n_coef, nq, pf, mu, wq, gc, accuracy);
          for (int d = 0; d < nd; ++d) {
               expand_gc(n_coef, nq, pf_l[d], mu, wq, gc_l[d], accuracy);
               gc_l[d][0][0] = 0.;
          }
     }
     if (lc) {
          n_coef2 = expand_lc(n_coef, nq, pf, mu, wq, lc, accuracy);
          for (int d = 0; d < nd; ++d) {
               expand_lc(n_coef, nq, pf_l[d], mu, wq, lc_l[d], accuracy);
               lc_l[d][0][0] = 0.;
          }
     }
     free(fv1); free(fv2); free(fv3); free(fv4); free(mu); free(wq); free(pf); free(pf_l); free(m_l);
     return n_coef2;
}

/* ------------------------------------------------------------------------- */
/* The ABI (src/lmie_interface.c). */
int lmie_calc_max_coef(double lambda, enum size_dist_type dist_type,
                       double a1, double a2, double a3, double r1, double r2) {
     (void) a3;
     if (dist_type == SIZE_DIST_POWER_LAW) powerlaw_range(a1, a2, &r1, &r2);
     return 2 * order_n1(2. * ORACLE_PI * r2 / lambda);
}

int lmie_in_alloc(lmie_in_data *in, int n) {
     if (n > 0) {
          double **v[10] = {&in->lambda_l, &in->mr_l, &in->mi_l, &in->a1_l, &in->a2_l,
                            &in->a3_l, &in->a4_l, &in->a5_l, &in->r1_l, &in->r2_l};
          for (int k = 0; k < 10; ++k) *v[k] = calloc(n, sizeof(double));
     }
     return 0;
}

void lmie_in_free(lmie_in_data *in, int flag) {
     if (!flag) return;
     free(in->lambda_l); free(in->mr_l); free(in->mi_l); free(in->a1_l); free(in->a2_l);
     free(in->a3_l); free(in->a4_l); free(in->a5_l); free(in->r1_l); free(in->r2_l);
}

int lmie_out_alloc(lmie_in_data *in, lmie_out_data *out, int max_coef) {
     int n = in->n_derivs;
     if (n > 0) {
          double **v[13] = {&out->r1_l, &out->r2_l, &out->norm_l, &out->reff_l, &out->veff_l,
                            &out->gavg_l, &out->vavg_l, &out->ravg_l, &out->rvw_l, &out->cext_l,
                            &out->csca_l, &out->cbak_l, &out->g_l};
          for (int k = 0; k < 13; ++k) *v[k] = calloc(n, sizeof(double));
     }
     out->gc = NULL;
     if (in->calc_gc) { out->gc = blk2(6, max_coef); if (n > 0) out->gc_l = blk3(n, 6, max_coef); }
     out->lc = NULL;
     if (in->calc_lc) { out->lc = blk2(6, max_coef); if (n > 0) out->lc_l = blk3(n, 6, max_coef); }
     out->pf = NULL;
     if (in->calc_pf) {
          out->theta = calloc(in->n_angles, sizeof(double));
          out->pf = blk2(6, in->n_angles);
          if (n > 0) out->pf_l = blk3(n, 6, in->n_angles);
     }
     if (in->save_control != 2) { out->save1 = NULL; out->save2 = NULL; }
     return 0;
}

void lmie_out_free(lmie_out_data *out, int flag) {
     if (flag) {
          free(out->r1_l); free(out->r2_l); free(out->norm_l); free(out->reff_l); free(out->veff_l);
          free(out->gavg_l); free(out->vavg_l); free(out->ravg_l); free(out->rvw_l);
          free(out->cext_l); free(out->csca_l); free(out->cbak_l); free(out->g_l);
     }
     if (out->gc) { free(out->gc); if (flag) free(out->gc_l); }
     if (out->lc) { free(out->lc); if (flag) free(out->lc_l); }
     if (out->pf) { free(out->theta); free(out->pf); if (flag) free(out->pf_l); }
}

void lmie_in_zero_derivs(lmie_in_data *in, int n) {
     for (int i = 0; i < n; ++i)
          in->lambda_l[i] = in->mr_l[i] = in->mi_l[i] = in->a1_l[i] = in->a2_l[i] =
          in->a3_l[i] = in->a4_l[i] = in->a5_l[i] = in->r1_l[i] = in->r2_l[i] = 0.;
}

int lmie_solution2_l(int calc_gc, int calc_lc, int calc_pf, enum size_dist_type dist_type,
                     int n_int1, int n_int2, int n_quad, int n_angles, int n_derivs,
                     int save_control, double lambda, double mr, double mi,
                     double a1, double a2, double a3, double a4, double a5, double r1, double r2,
                     double *lambda_l, double *mr_l, double *mi_l, double *a1_l, double *a2_l,
                     double *a3_l, double *a4_l, double *a5_l, double *r1_l, double *r2_l,
                     double accuracy, int *n_coef, double *r21, double *r22, double *norm,
                     double *reff, double *veff, double *gavg, double *vavg, double *ravg,
                     double *rvw, double *cext, double *csca, double *cbak, double *g,
                     double **gc, double **lc, double *theta, double **pf, double *r21_l,
                     double *r22_l, double *norm_l, double *reff_l, double *veff_l,
                     double *gavg_l, double *vavg_l, double *ravg_l, double *rvw_l,
                     double *cext_l, double *csca_l, double *cbak_l, double *g_l,
                     double ***gc_l, double ***lc_l, double ***pf_l, double ***save1,
                     double ****save2, int verbose, int n_threads, int use_mpi) {
     (void) verbose; (void) use_mpi;
     if (calc_pf && !calc_gc && !calc_lc) {
          fprintf(stderr, "ERROR: either calc_gc or calc_lc must be set to true when calc_pf is true\n");
          return -1;
     }
     int nd = n_derivs > 0 ? n_derivs : 0;
     int N = grid_size(dist_type, n_int1, n_int2, n_quad);
     if (N <= 0) { fprintf(stderr, "ERROR: invalid distribution type: %d\n", dist_type); return -1; }
     int D = nd > 0 ? nd : 1;
     double *qx = calloc(N, sizeof(double)), *qw = calloc(N, sizeof(double)), *nr = calloc(N, sizeof(double));
     double *qx_l = calloc((size_t) N * D, sizeof(double)), *qw_l = calloc((size_t) N * D, sizeof(double));
     double *nr_l = calloc((size_t) N * D, sizeof(double));
     double *r1o_l = calloc(D, sizeof(double)), *r2o_l = calloc(D, sizeof(double));
     double *nrm_l = calloc(D, sizeof(double));
     dist_in s = {dist_type, n_int1, n_int2, n_quad, nd, {a1, a2, a3, a4, a5}, r1, r2,
                  {a1_l, a2_l, a3_l, a4_l, a5_l}, r1_l, r2_l};
     double r1o = 0., r2o = 0., nrm = 0.;
     size_dist(&s, qx, qw, nr, qx_l, qw_l, nr_l, &r1o, &r2o, r1o_l, r2o_l, &nrm, nrm_l);
     *r21 = r1o; *r22 = r2o; *norm = nrm;
     for (int d = 0; d < nd; ++d) { r21_l[d] = r1o_l[d]; r22_l[d] = r2o_l[d]; norm_l[d] = nrm_l[d]; }
     double dummy[8];
     dist_stats(N, nd, qx, qw, nr, qx_l, qw_l, nr_l, reff, veff, gavg, vavg, ravg, rvw,
                nd ? reff_l : dummy, nd ? veff_l : dummy, nd ? gavg_l : dummy,
                nd ? vavg_l : dummy, nd ? ravg_l : dummy, nd ? rvw_l : dummy);
     /* split seeds into Mie tangents and distribution-weight-only ones
        (src/lmie_interface.c:304-353) */
     int *idx1 = calloc(D, sizeof(int)), *idx2 = calloc(D, sizeof(int)), nd1 = 0, nd2 = 0;
     for (int d = 0; d < nd; ++d) {
          if (lambda_l[d] != 0. || mr_l[d] != 0. || mi_l[d] != 0. || r1o_l[d] != 0. || r2o_l[d] != 0.)
               idx1[nd1++] = d;
          else
               idx2[nd2++] = d;
     }
     double zero[1] = {0.};
     core_in c = {N, nd, nd1, nd2, idx1, idx2, save_control, lambda, nd ? lambda_l : zero,
                  mr, mi, nd ? mr_l : zero, nd ? mi_l : zero, qx, qw, nr, qx_l, qw_l, nr_l, r2o};
     double cx_l[1], dd_l[1];
     *n_coef = core_solution(&c, cext, csca, cbak, g, calc_gc ? gc : NULL, calc_lc ? lc : NULL,
                             nd ? cext_l : cx_l, nd ? csca_l : dd_l, nd ? cbak_l : cx_l,
                             nd ? g_l : dd_l, gc_l, lc_l, save1, save2, accuracy, n_threads);
     if (calc_pf) {
          double **xc = calc_gc ? gc : lc;
          double ***xc_l = calc_gc ? gc_l : lc_l;
          synth_pf(*n_coef, n_angles, xc, theta, pf, calc_gc);
          for (int d = 0; d < nd; ++d) synth_pf(*n_coef, n_angles, xc_l[d], theta, pf_l[d], calc_gc);
     }
     free(qx); free(qw); free(nr); free(qx_l); free(qw_l); free(nr_l); free(r1o_l); free(r2o_l);
     free(nrm_l); free(idx1); free(idx2);
     return 0;
}

int lmie_solution2(int calc_gc, int calc_lc, int calc_pf, enum size_dist_type dist_type,
                   int n_int1, int n_int2, int n_quad, int n_angles, int save_control,
                   double lambda, double mr, double mi, double a1, double a2, double a3,
                   double a4, double a5, double r1, double r2, double accuracy, int *n_coef,
                   double *r21, double *r22, double *norm, double *reff, double *veff,
                   double *gavg, double *vavg, double *ravg, double *rvw, double *cext,
                   double *csca, double *cbak, double *g, double **gc, double **lc,
                   double *theta, double **pf, double ***save1, double ****save2,
                   int verbose, int n_threads, int use_mpi) {
     return lmie_solution2_l(calc_gc, calc_lc, calc_pf, dist_type, n_int1, n_int2, n_quad,
                             n_angles, 0, save_control, lambda, mr, mi, a1, a2, a3, a4, a5, r1, r2,
                             NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, accuracy,
                             n_coef, r21, r22, norm, reff, veff, gavg, vavg, ravg, rvw, cext, csca,
                             cbak, g, gc, lc, theta, pf, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
                             NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, save1, save2,
                             verbose, n_threads, use_mpi);
}

int lmie_solution(lmie_in_data *in, lmie_out_data *out, int alloc_out, int verbose,
                  int n_threads, int use_mpi) {
     if (alloc_out) {
          int mc = lmie_calc_max_coef(in->lambda, in->dist_type, in->a1, in->a2, in->a3, in->r1, in->r2);
          lmie_out_alloc(in, out, mc);
     }
     return lmie_solution2_l(in->calc_gc, in->calc_lc, in->calc_pf, in->dist_type, in->n_int1,
                             in->n_int2, in->n_quad, in->n_angles, in->n_derivs, in->save_control,
                             in->lambda, in->mr, in->mi, in->a1, in->a2, in->a3, in->a4, in->a5,
                             in->r1, in->r2, in->lambda_l, in->mr_l, in->mi_l, in->a1_l, in->a2_l,
                             in->a3_l, in->a4_l, in->a5_l, in->r1_l, in->r2_l, in->accuracy,
                             &out->n_coef, &out->r1, &out->r2, &out->norm, &out->reff, &out->veff,
                             &out->gavg, &out->vavg, &out->ravg, &out->rvw, &out->cext, &out->csca,
                             &out->cbak, &out->g, out->gc, out->lc, out->theta, out->pf, out->r1_l,
                             out->r2_l, out->norm_l, out->reff_l, out->veff_l, out->gavg_l,
                             out->vavg_l, out->ravg_l, out->rvw_l, out->cext_l, out->csca_l,
                             out->cbak_l, out->g_l, out->gc_l, out->lc_l, out->pf_l, &out->save1,
                             &out->save2, verbose, n_threads, use_mpi);
}

int lmie_solution_slave(void) { return -1; }
