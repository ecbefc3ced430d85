/*
 * lmie_host.c -- host C layer of liblmie_amd.so.
 *
 * Implements the drop-in API of include/lmie_interface.h and the extensions of
 * include/lmie_hip.h.  Everything that is O(N * n_derivs) or smaller runs
 * here: the size grid and number density with their tangents
 * (misc/rtutil_size_dist.c:20-646), the distribution statistics
 * (misc/rtutil_size_dist.c:653-794), the split of the seeds into Mie tangents
 * and weight-only ones (src/lmie_interface.c:304-353), the per-size orders,
 * and the normalisation of the raw sums and expansions
 * (src/lmie_schedule.c:1155-1241).  Everything that scales with the number of
 * orders or angles runs in the kernels behind include/lmie_gpu.h.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "../../include/lmie_interface.h"
#include "../../include/lmie_hip.h"
#include "../../include/lmie_gpu.h"

/* host wall clock in ms, for the LMIE_T_HOST_* and LMIE_T_CALL timers */
static double wall_ms(void) {
     struct timespec t;
     clock_gettime(CLOCK_MONOTONIC, &t);
     return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

#define LM_PI 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------------- */
/* Per-size host loops on a small persistent thread pool (up to 16 threads,
   one parallel region at a time; a caller that finds the pool busy runs its
   loop alone).  Chunks are static, so every element is computed by the same
   code as a serial loop would: the results do not depend on the threads. */
typedef void (*par_fn)(void *ctx, int i0, int i1);
static struct {
     pthread_mutex_t region, lk;
     pthread_cond_t go, done;
     int started, nthr, gen, pending, n;
     par_fn fn;
     void *ctx;
} g_pool = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER,
            PTHREAD_COND_INITIALIZER, 0, 0, 0, 0, 0, NULL, NULL};

static void pool_chunk(int t, int T, int n, int *i0, int *i1) {
     *i0 = (int) ((long long) n * t / T);
     *i1 = (int) ((long long) n * (t + 1) / T);
}

static void *pool_worker(void *arg) {
     const int t = (int) (intptr_t) arg;
     int seen = 0;
     pthread_mutex_lock(&g_pool.lk);
     for (;;) {
          while (g_pool.gen == seen) pthread_cond_wait(&g_pool.go, &g_pool.lk);
          seen = g_pool.gen;
          par_fn fn = g_pool.fn;
          void *ctx = g_pool.ctx;
          int n = g_pool.n, T = g_pool.nthr + 1, i0, i1;
          pthread_mutex_unlock(&g_pool.lk);
          pool_chunk(t, T, n, &i0, &i1);
          if (i1 > i0) fn(ctx, i0, i1);
          pthread_mutex_lock(&g_pool.lk);
          if (--g_pool.pending == 0) pthread_cond_signal(&g_pool.done);
     }
     return NULL;
}

/* set on the threads of a fused batch: there the cases are the parallelism,
   and each runs its host loops alone */
static __thread int g_batch_thread;

/* fn over [0, n) on the pool when n >= min_n (else, on a batch thread or
   with the pool busy, in the calling thread) */
static void par_run(int n, par_fn fn, void *ctx, int min_n) {
     if (n < min_n || g_batch_thread || pthread_mutex_trylock(&g_pool.region) != 0) {
          fn(ctx, 0, n);
          return;
     }
     pthread_mutex_lock(&g_pool.lk);
     if (!g_pool.started) {
          g_pool.started = 1;
          long nc = sysconf(_SC_NPROCESSORS_ONLN);
          int T = nc > 16 ? 16 : (nc < 1 ? 1 : (int) nc);
          for (int t = 1; t < T; ++t) {
               pthread_t th;
               if (pthread_create(&th, NULL, pool_worker, (void *) (intptr_t) t) != 0) break;
               pthread_detach(th);
               g_pool.nthr++;
          }
     }
     const int T = g_pool.nthr + 1;
     g_pool.fn = fn;
     g_pool.ctx = ctx;
     g_pool.n = n;
     g_pool.pending = g_pool.nthr;
     g_pool.gen++;
     pthread_cond_broadcast(&g_pool.go);
     pthread_mutex_unlock(&g_pool.lk);
     int i0, i1;
     pool_chunk(0, T, n, &i0, &i1);
     fn(ctx, i0, i1);
     pthread_mutex_lock(&g_pool.lk);
     while (g_pool.pending > 0) pthread_cond_wait(&g_pool.done, &g_pool.lk);
     pthread_mutex_unlock(&g_pool.lk);
     pthread_mutex_unlock(&g_pool.region);
}

static void par_for(int n, par_fn fn, void *ctx) { par_run(n, fn, ctx, 2048); }

/* fixed blocks 0 .. nb-1 spread over the pool: each block's result is the
   same whichever thread runs it */
static void par_blocks(int nb, par_fn fn, void *ctx) { par_run(nb, fn, ctx, 2); }

/* ------------------------------------------------------------------------- */
/* errors: one message per thread; -1 returns leave a breadcrumb like the
   reference's fprintf(stderr, "ERROR: ...") chain. */
static __thread char g_err[512];

static int fail(const char *where, const char *why) {
     char tmp[400];
     snprintf(tmp, sizeof tmp, "%s", why ? why : "");     /* why may be g_err itself */
     snprintf(g_err, sizeof g_err, "%s: %s", where, tmp);
     fprintf(stderr, "ERROR: %s\n", g_err);
     return -1;
}

const char *lmie_last_error(void) { return g_err; }

/* arithmetic the angular contraction of this thread's last solution ran in */
static __thread int g_last_precision = LMIE_PRECISION_FP64;
int lmie_last_precision(void) { return g_last_precision; }

/* ------------------------------------------------------------------------- */
/* Pointer-table blocks (misc/gutil_alloc.c:19-75 layout): row pointers, then
   contiguous row-major data, one malloc, released with free(). */
static double **table2(int n0, int n1) {
     size_t head = ((size_t) n0 * sizeof(double *) + 63) & ~(size_t) 63;
     char *m = calloc(1, head + (size_t) n0 * n1 * sizeof(double) + 64);
     if (!m) return NULL;
     double **p = (double **) m;
     for (int i = 0; i < n0; ++i) p[i] = (double *) (m + head) + (size_t) i * n1;
     return p;
}

static double ***table3(int n0, int n1, int n2) {
     size_t h0 = (size_t) n0 * sizeof(double **), h1 = (size_t) n0 * n1 * sizeof(double *);
     size_t head = (h0 + h1 + 63) & ~(size_t) 63;
     char *m = calloc(1, head + (size_t) n0 * n1 * n2 * sizeof(double) + 64);
     if (!m) return NULL;
     double ***p = (double ***) m;
     double **q = (double **) (m + h0);
     for (int i = 0; i < n0; ++i) {
          p[i] = q + (size_t) i * n1;
          for (int j = 0; j < n1; ++j)
               p[i][j] = (double *) (m + head) + ((size_t) i * n1 + j) * n2;
     }
     return p;
}

static double *vec(size_t n) { return calloc(n ? n : 1, sizeof(double)); }

/* ------------------------------------------------------------------------- */
/* Orders (src/lmie_core.c:18-38).  Computed here, once per size, so host and
   kernels agree on every integer. */
static int mie_n1(double x) { return (int) (x + 4.05 * pow(x, 1. / 3.) + 8); }

static int mie_n2(int n1, double x, double mr, double mi) {
     double az = cabs((mr - I * mi) * x);
     double m = n1 > az ? (double) n1 : az;
     return (int) (m + 6.40 * pow(m, 1. / 3.) + 8);
}

/* ------------------------------------------------------------------------- */
/* Gauss-Legendre rule on [-1, 1] for the size grid: Newton iteration with a
   second-order correction from the initial guess of misc/rtutil_math.c:73-129. */
static void gl_rule(int n, double *x, double *w) {
     for (int i = 0, k = n - 1; i < (n + 1) / 2; ++i, --k) {
          double t = (4. * i + 3.) / (4 * n + 2) * LM_PI;
          double y = cos(t + (n - 1) / (8. * pow((double) n, 3)) / tan(t)), y0, pa, pb;
          do {
               double pm = 0.;
               pa = 1.;
               pb = 0.;
               for (int j = 0; j < n; ++j) {
                    double u;
                    pm = pb;
                    pb = pa;
                    u = y * pb;
                    pa = j / (j + 1.) * (u - pm) + u;
               }
               double s = 1. - y * y;
               double d1 = n * (pb - y * pa) / s;
               double d2 = (2. * y * d1 - n * (n + 1) * pa) / s;
               double h = pa / d1;
               y0 = y;
               y = y0 - h * (1. + h * d2 / (2. * d1));
          } while (fabs(y - y0) > DBL_EPSILON);
          x[i] = -y;
          x[k] = y;
          double c = n * pb;
          c = 2. / (c * c);
          w[i] = (1. - y * y) * c;
          w[k] = w[i];
     }
}

/* Seed columns of the size grid: a[p][d] for a1..a5, r1[d], r2[d]. */
typedef struct {
     int nd;
     const double *a[5];
     const double *r1, *r2;
} seeds;

static double sd(const double *v, int d) { return v ? v[d] : 0.; }

/* n_int equal intervals on [r1, r2] with n_quad nodes each; the map of the
   [-1, 1] rule onto each interval follows misc/rtutil_math.c:329-387 (the
   mirrored node sets the offset). */
/* gl_rule of the last few orders asked for, kept for the next calls (a
   batch of cases builds the same rule once per case). */
static void gl_rule_cached(int n, double *x, double *w) {
     static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
     static struct { int n; double *x, *w; } memo[4];
     static int next;
     pthread_mutex_lock(&mu);
     for (int k = 0; k < 4; ++k)
          if (memo[k].n == n && memo[k].x) {
               memcpy(x, memo[k].x, sizeof(double) * n);
               memcpy(w, memo[k].w, sizeof(double) * n);
               pthread_mutex_unlock(&mu);
               return;
          }
     pthread_mutex_unlock(&mu);
     gl_rule(n, x, w);
     double *cx = malloc(sizeof(double) * n), *cw = malloc(sizeof(double) * n);
     if (!cx || !cw) { free(cx); free(cw); return; }
     memcpy(cx, x, sizeof(double) * n);
     memcpy(cw, w, sizeof(double) * n);
     pthread_mutex_lock(&mu);
     const int k = next++ % 4;
     free(memo[k].x); free(memo[k].w);
     memo[k].n = n; memo[k].x = cx; memo[k].w = cw;
     pthread_mutex_unlock(&mu);
}

static void grid_intervals(int n_int, int n_quad, int nd, double r1, double r2,
                           const double *r1_l, const double *r2_l,
                           double *qx, double *qw, double *qx_l, double *qw_l) {
     double *tx = vec(n_quad), *tw = vec(n_quad);
     gl_rule_cached(n_quad, tx, tw);
     /* the reference scales the [-1,1] weights by the half-width of each
        interval with its own weight formula ((half - off*y) * c); on [-1, 1]
        that is (1 - y^2) * c, as gl_rule computes */
     double step = (r2 - r1) / n_int;
     for (int s = 0; s < n_int; ++s) {
          double lo = r1 + s * step, hi = lo + step;
          double mid = (lo + hi) / 2., half = (hi - lo) / 2.;
          for (int i = 0, k = n_quad - 1; i < (n_quad + 1) / 2; ++i, --k) {
               int a = s * n_quad + i, b = s * n_quad + k;
               qx[a] = mid - half * tx[k];
               qx[b] = mid + half * tx[k];
               qw[a] = qw[b] = tw[i] * half;
               for (int d = 0; d < nd && (r1_l || r2_l); ++d) {   /* else the tangents stay zero */
                    double st_l = (sd(r2_l, d) - sd(r1_l, d)) / n_int;
                    double lo_l = sd(r1_l, d) + s * st_l, hi_l = lo_l + st_l;
                    double mid_l = (lo_l + hi_l) / 2., half_l = (hi_l - lo_l) / 2.;
                    qx_l[(size_t) a * nd + d] = mid_l - half_l * tx[k];
                    qx_l[(size_t) b * nd + d] = mid_l + half_l * tx[k];
                    qw_l[(size_t) a * nd + d] = qw_l[(size_t) b * nd + d] = tw[i] * half_l;
               }
          }
     }
     free(tx);
     free(tw);
}

static int grid_points(int dist, int n_int1, int n_int2, int n_quad) {
     switch (dist) {
     case SIZE_DIST_MONO: return 1;
     case SIZE_DIST_MODIFIED_POWER_LAW: return (n_int1 + n_int2) * n_quad;
     case SIZE_DIST_MODIFIED_GAMMA: case SIZE_DIST_LOG_NORMAL: case SIZE_DIST_POWER_LAW:
     case SIZE_DIST_GAMMA: case SIZE_DIST_MODIFIED_BIMODAL_LOG_NORMAL: case SIZE_DIST_EXPONENTIAL:
          return n_int2 * n_quad;
     default: return -1;
     }
}

/* Brent root bracketing for the power-law range (the role of misc/zeroin.c). */
static double pl_f(double r1, double a1, double a2) {
     double r2 = (1. + a2) * 2. * a1 - r1;
     return (r2 - r1) / log(r2 / r1) - a1;
}

static void pl_range(double a1, double a2, double *r1, double *r2) {
     double a = DBL_EPSILON, b = a1, fa = pl_f(a, a1, a2), fb = pl_f(b, a1, a2);
     double c = a, fc = fa, d = b - a, e = d;
     for (int it = 0; it < 400; ++it) {
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
          fb = pl_f(b, a1, a2);
     }
     *r1 = b;
     *r2 = (1. + a2) * 2. * a1 - b;
}

typedef struct {
     int N, nd;
     double *qx, *qw, *nr, *qx_l, *qw_l, *nr_l;    /* [N], [N][nd] */
     double r1, r2, *r1_l, *r2_l, norm, *norm_l;
} grid_t;

static void grid_free(grid_t *g) {
     free(g->qx); free(g->qw); free(g->nr); free(g->qx_l); free(g->qw_l); free(g->nr_l);
     free(g->r1_l); free(g->r2_l); free(g->norm_l);
}

/* Grid, number density normalised to unit integral, and tangents along each
   seed column (misc/rtutil_size_dist.c:200-646). */
/* Number density and its tangents at the nodes i0 .. i1-1
   (misc/rtutil_size_dist.c:20-646). */
typedef struct {
     int dist, nd, N1;
     const seeds *sv;
     grid_t *g;
     double a[5];
} dens_ctx;

static void dens_range(void *vp, int i0, int i1) {
     const dens_ctx *dc = vp;
     const int dist = dc->dist, nd = dc->nd, N1 = dc->N1;
     const seeds *sv = dc->sv;
     grid_t *g = dc->g;
     const double a1 = dc->a[0], a2 = dc->a[1], a3 = dc->a[2], a4 = dc->a[3], a5 = dc->a[4];
     (void) N1;
     for (int i = i0; i < i1; ++i) {
          double r = g->qx[i], lr = log(r), *nl = g->nr_l + (size_t) i * nd;
          const double *xl = g->qx_l + (size_t) i * nd;
          double n = 0.;
          switch (dist) {
          case SIZE_DIST_MODIFIED_GAMMA: {
               double b = pow(a2, a3), c = a3 * b, h = pow(r, a3);
               n = pow(r, a1) * exp(-a1 / c * h);
               for (int d = 0; d < nd; ++d) {
                    double s1 = sd(sv->a[0], d), s2 = sd(sv->a[1], d), s3 = sd(sv->a[2], d);
                    double lc = s3 / a3 + s2 * (a3 / a2) + s3 * log(a2);    /* d ln c */
                    double px = xl[d] / r;
                    nl[d] = n * ((px * a1 + s1 * lr) + h * ((-s1 + a1 * (-(px * a3 + s3 * lr) + lc)) / c));
               }
               break;
          }
          case SIZE_DIST_LOG_NORMAL: {
               double dv = lr - log(a1), e = dv * dv / (2. * a2);
               n = exp(-e) / r;
               for (int d = 0; d < nd; ++d) {
                    double px = xl[d] / r;
                    nl[d] = n * (-px - (2. * dv * (px - sd(sv->a[0], d) / a1) - e * 2. * sd(sv->a[1], d)) / (2. * a2));
               }
               break;
          }
          case SIZE_DIST_POWER_LAW:
               n = 1. / pow(r, 3);
               for (int d = 0; d < nd; ++d) nl[d] = -n * 3. * xl[d] / r;
               break;
          case SIZE_DIST_GAMMA: {
               double ab = a1 * a2, pw = (1. - 3. * a2) / a2, rp = pow(r, pw), ex = exp(-r / ab);
               n = rp * ex;
               for (int d = 0; d < nd; ++d) {
                    double ab_l = sd(sv->a[0], d) * a2 + a1 * sd(sv->a[1], d);
                    double pw_l = (-3. * sd(sv->a[1], d) - pw * sd(sv->a[1], d)) / a2;
                    nl[d] = rp * ((xl[d] * (pw / r) + pw_l * lr) + (-xl[d] + r / ab * ab_l) / ab) * ex;
               }
               break;
          }
          case SIZE_DIST_MODIFIED_POWER_LAW:
               if (i < N1) {
                    n = 1.;
               } else {
                    double u = r / g->r1;
                    n = pow(u, a1);
                    for (int d = 0; d < nd; ++d)
                         nl[d] = n * ((xl[d] - u * g->r1_l[d]) / g->r1 * a1 / u + sd(sv->a[0], d) * log(u));
               }
               break;
          case SIZE_DIST_MODIFIED_BIMODAL_LOG_NORMAL: {
               double d1 = lr - log(a1), e1 = d1 * d1 / (2. * a2);
               double d2 = lr - log(a3), e2 = d2 * d2 / (2. * a4);
               double o1 = exp(-e1), o2 = exp(-e2), r4 = r * r * r * r;
               n = (o1 + a5 * o2) / r4;
               for (int d = 0; d < nd; ++d) {
                    double px = xl[d] / r;
                    double h1 = ((-px + sd(sv->a[0], d) / a1) * 2. * d1 + e1 * 2. * sd(sv->a[1], d)) * o1 / (2. * a2);
                    double h2 = ((-px + sd(sv->a[2], d) / a3) * 2. * d2 + e2 * 2. * sd(sv->a[3], d)) * o2 / (2. * a4);
                    nl[d] = (h1 + sd(sv->a[4], d) * o2 + a5 * h2 - n * 4. * r * r * r * xl[d]) / r4;
               }
               break;
          }
          case SIZE_DIST_EXPONENTIAL:
               /* the reference defines no tangents for this kind */
               n = 2. * exp(-a1 * 2. * r);
               break;
          }
          g->nr[i] = n;
     }
}

static int grid_build(int dist, int n_int1, int n_int2, int n_quad, const double a[5],
                      double r1, double r2, const seeds *sv, grid_t *g) {
     int nd = sv->nd, N = grid_points(dist, n_int1, n_int2, n_quad);
     if (N <= 0) return fail("size distribution", "invalid distribution type");
     size_t L = (size_t) N * (nd > 0 ? nd : 1);
     g->N = N; g->nd = nd;
     g->qx = vec(N); g->qw = vec(N); g->nr = vec(N);
     g->qx_l = vec(L); g->qw_l = vec(L); g->nr_l = vec(L);
     g->r1_l = vec(nd); g->r2_l = vec(nd); g->norm_l = vec(nd);
     double *tot_l = vec(nd), tot = 0.;
     if (!g->qx || !g->qw || !g->nr || !g->qx_l || !g->qw_l || !g->nr_l || !g->r1_l || !g->r2_l ||
         !g->norm_l || !tot_l) {
          free(tot_l);
          return fail("size distribution", "out of memory");   /* the caller frees g */
     }
     g->r1 = r1; g->r2 = r2;
     for (int d = 0; d < nd; ++d) { g->r1_l[d] = sd(sv->r1, d); g->r2_l[d] = sd(sv->r2, d); }
     double a1 = a[0], a2 = a[1], a3 = a[2], a4 = a[3], a5 = a[4];

     if (dist == SIZE_DIST_MONO) {
          g->r1 = g->r2 = a1;
          for (int d = 0; d < nd; ++d) { g->r1_l[d] = g->r2_l[d] = 0.; g->qx_l[d] = sd(sv->a[0], d); }
          g->qx[0] = a1; g->qw[0] = 1.; g->nr[0] = 1.;
          g->norm = 1.;
          free(tot_l);
          return 0;
     }
     if (dist == SIZE_DIST_POWER_LAW) {
          pl_range(a1, a2, &g->r1, &g->r2);
          double Lg = log(g->r2) - log(g->r1);
          double p = -Lg + (g->r2 - g->r1) / g->r1, q = Lg - (g->r2 - g->r1) / g->r2;
          for (int d = 0; d < nd; ++d) {
               double u = sd(sv->a[0], d) * Lg * Lg;
               double v = sd(sv->a[1], d) * 2. * a1 + (1. + a2) * 2. * sd(sv->a[0], d);
               g->r1_l[d] = (q * v - u) / (q - p);
               g->r2_l[d] = (p * v - u) / (p - q);
          }
     }
     int N1 = 0;
     if (dist == SIZE_DIST_MODIFIED_POWER_LAW) {
          N1 = n_int1 * n_quad;
          grid_intervals(n_int1, n_quad, nd, 0., g->r1, NULL, g->r1_l, g->qx, g->qw, g->qx_l, g->qw_l);
          grid_intervals(n_int2, n_quad, nd, g->r1, g->r2, g->r1_l, g->r2_l, g->qx + N1, g->qw + N1,
                         g->qx_l + (size_t) N1 * nd, g->qw_l + (size_t) N1 * nd);
     } else
          grid_intervals(n_int2, n_quad, nd, g->r1, g->r2, g->r1_l, g->r2_l, g->qx, g->qw, g->qx_l, g->qw_l);

     {
          dens_ctx dc = {dist, nd, N1, sv, g, {a1, a2, a3, a4, a5}};
          par_for(N, dens_range, &dc);
     }
     for (int i = 0; i < N; ++i) {
          tot += g->nr[i] * g->qw[i];
          if (dist != SIZE_DIST_EXPONENTIAL)
               for (int d = 0; d < nd; ++d)
                    tot_l[d] += g->nr_l[(size_t) i * nd + d] * g->qw[i] + g->nr[i] * g->qw_l[(size_t) i * nd + d];
     }
     g->norm = tot;
     for (int d = 0; d < nd; ++d) g->norm_l[d] = tot_l[d];
     for (int i = 0; i < N; ++i) {
          g->nr[i] /= tot;
          for (int d = 0; d < nd; ++d)
               g->nr_l[(size_t) i * nd + d] = (g->nr_l[(size_t) i * nd + d] - g->nr[i] * tot_l[d]) / tot;
     }
     free(tot_l);
     return 0;
}

/* Distribution statistics and tangents (misc/rtutil_size_dist.c:653-794). */
typedef struct { double reff, veff, gavg, vavg, ravg, rvw; } dstats;

static void grid_stats(const grid_t *g, dstats *s, dstats *sl /* [nd] or NULL */) {
     int N = g->N, nd = g->nd;
     double m[4] = {0, 0, 0, 0};
     double *ml = vec((size_t) 4 * (nd ? nd : 1));
     for (int i = 0; i < N; ++i) {
          double t = g->nr[i] * g->qx[i] * g->qw[i];
          for (int k = 0; k < 4; ++k) { m[k] += t; t *= g->qx[i]; }
          for (int d = 0; d < nd; ++d) {
               size_t o = (size_t) i * nd + d;
               double h = g->qx[i] * g->qw[i] * g->nr_l[o], p = g->nr[i] * g->qw[i] * g->qx_l[o];
               double q = g->nr[i] * g->qx[i] * g->qw_l[o], rp = 1.;
               for (int k = 0; k < 4; ++k) { ml[k * nd + d] += (h + (k + 1) * p + q) * rp; rp *= g->qx[i]; }
          }
     }
     s->gavg = m[1] * LM_PI;
     s->vavg = 4. / 3. * LM_PI * m[2];
     s->ravg = m[0];
     s->rvw = 4. / 3. * LM_PI * m[3] / s->vavg;
     s->reff = m[2] / m[1];
     for (int d = 0; d < nd && sl; ++d) {
          sl[d].gavg = ml[nd + d] * LM_PI;
          sl[d].vavg = 4. / 3. * LM_PI * ml[2 * nd + d];
          sl[d].ravg = ml[d];
          sl[d].rvw = 4. / 3. * LM_PI * (ml[3 * nd + d] / s->vavg - m[3] * sl[d].vavg / (s->vavg * s->vavg));
          sl[d].reff = ml[2 * nd + d] / m[1] - m[2] * ml[nd + d] / (m[1] * m[1]);
     }
     double v = 0., *vl = vec(nd);
     for (int i = 0; i < N; ++i) {
          double dr = g->qx[i] - s->reff, r2w = g->qx[i] * g->qx[i] * g->qw[i];
          v += g->nr[i] * dr * dr * r2w;
          for (int d = 0; d < nd && sl; ++d) {
               size_t o = (size_t) i * nd + d;
               vl[d] += g->nr_l[o] * dr * dr * r2w + g->nr[i] *
                        (2. * dr * r2w * (g->qx_l[o] - sl[d].reff) +
                         dr * dr * (2. * g->qx[i] * g->qw[i] * g->qx_l[o] + g->qx[i] * g->qx[i] * g->qw_l[o]));
          }
     }
     double b = s->gavg * s->reff * s->reff;
     s->veff = v * LM_PI / b;
     for (int d = 0; d < nd && sl; ++d) {
          double q = s->veff * s->reff / b;
          sl[d].veff = vl[d] * LM_PI / b - (sl[d].gavg * q * s->reff + q * s->gavg * 2. * sl[d].reff);
     }
     free(ml);
     free(vl);
}

/* ------------------------------------------------------------------------- */
/* Core: raw sums on the GPU, normalisation and expansion bookkeeping here
   (src/lmie_schedule.c:942-1268).  Seeds as columns d of [N][nd] arrays. */
typedef struct {
     int N, nd, nd1, nd2, *idx1, *idx2, save_control;
     double lambda, mr, mi;
     const double *lambda_l, *mr_l, *mi_l;      /* [nd] */
     const double *qx, *qw, *nr;                /* [N]  */
     const double *qx_l, *qw_l, *nr_l;          /* [N][nd] */
     double r2;
     int calc_gc, calc_lc;
     double accuracy;
     const lmie_opts *opts;
     int keep;                      /* keep the device state for a reverse pass */
     void (*on_device)(void *);     /* called once the host inputs are ready, just */
     void *on_device_arg;           /* before the device pipeline (or NULL)        */
} core_args;

typedef struct {
     double *cext, *csca, *cbak, *g;
     double **gc, **lc;
     double *cext_l, *csca_l, *cbak_l, *g_l;
     double ***gc_l, ***lc_l;
     double ***save1, ****save2;
     int n_qang, n_coef;
     double sums[4];                /* raw sums, kept for the adjoint */
     void *state;                   /* device state when core_args.keep */
} core_out;

typedef struct {
     const core_args *c;
     double *x, *f, *dx, *dft, *dfw, *sx, *cx;
     int *n1, *n2;
} job_ctx;

/* Per-size inputs of the kernels at the nodes i0 .. i1-1: size parameter,
   weight, orders (src/lmie_core.c:18-38), sin x and cos x from the C library
   (as the reference's recurrences start from them), and the seed columns. */
static void job_range(void *vp, int i0, int i1) {
     const job_ctx *jc = vp;
     const core_args *c = jc->c;
     for (int i = i0; i < i1; ++i) {
          double *x = jc->x;
          x[i] = 2. * LM_PI * c->qx[i] / c->lambda;
          jc->f[i] = c->nr[i] * c->qw[i];
          jc->n1[i] = mie_n1(x[i]);
          jc->n2[i] = mie_n2(jc->n1[i], x[i], c->mr, c->mi);
          jc->sx[i] = sin(x[i]);
          jc->cx[i] = cos(x[i]);
          for (int t = 0; t < c->nd1; ++t) {
               int d = c->idx1[t];
               size_t o = (size_t) i * c->nd + d;
               jc->dx[(size_t) t * c->N + i] = 2. * LM_PI * (c->qx_l[o] - c->qx[i] * c->lambda_l[d] / c->lambda) / c->lambda;
               jc->dft[(size_t) t * c->N + i] = c->nr_l[o] * c->qw[i] + c->nr[i] * c->qw_l[o];
          }
          for (int t = 0; t < c->nd2; ++t) {
               size_t o = (size_t) i * c->nd + c->idx2[t];
               jc->dfw[(size_t) t * c->N + i] = c->nr_l[o] * c->qw[i];
          }
     }
}

static void fill_job(const core_args *c, lmie_gpu_job *j, double *x, double *f, int *n1, int *n2,
                     double *dx, double *dm, double *dft, double *dfw, double *sx, double *cx) {
     double x2 = 2. * LM_PI * c->r2 / c->lambda;
     int n1m = mie_n1(x2);
     memset(j, 0, sizeof *j);
     job_ctx jc = {c, x, f, dx, dft, dfw, sx, cx, n1, n2};
     par_for(c->N, job_range, &jc);
     for (int t = 0; t < c->nd1; ++t) {
          dm[2 * t] = c->mr_l[c->idx1[t]];
          dm[2 * t + 1] = -c->mi_l[c->idx1[t]];
     }
     j->n_size = c->N; j->x = x; j->f = f; j->n1 = n1; j->n2 = n2; j->sx = sx; j->cx = cx;
     j->mr = c->mr; j->mi = c->mi;
     j->n1_max = n1m;
     j->n_qang = (c->calc_gc || c->calc_lc) ? 2 * n1m : 0;
     j->n_coef = 2 * n1m;
     j->calc_gc = c->calc_gc; j->calc_lc = c->calc_lc;
     j->n_tan = c->nd1; j->dx = dx; j->dm = dm; j->df_tan = dft;
     j->n_wgt = c->nd2; j->df_wgt = dfw;
     j->save_control = c->save_control;
     j->precision = c->opts ? c->opts->precision : LMIE_PRECISION_FP64;
     j->device = c->opts ? c->opts->device : 0;
     j->keep = c->keep;
}

typedef struct {
     double *x, *f, *dx, *dm, *dft, *dfw, *sx, *cx;
     int *n1, *n2;
} job_mem;

static void job_mem_alloc(job_mem *m, int N, int nd1, int nd2) {
     m->x = vec(N); m->f = vec(N); m->sx = vec(N); m->cx = vec(N);
     m->n1 = calloc(N, sizeof(int)); m->n2 = calloc(N, sizeof(int));
     m->dx = vec((size_t) nd1 * N); m->dm = vec(2 * (size_t) nd1); m->dft = vec((size_t) nd1 * N);
     m->dfw = vec((size_t) nd2 * N);
}

static void job_mem_free(job_mem *m) {
     free(m->x); free(m->f); free(m->n1); free(m->n2); free(m->dx); free(m->dm); free(m->dft); free(m->dfw);
     free(m->sx); free(m->cx);
}

/* Normalise the expansion of channel c of raw into out (one 6 x n_coef set):
   value a * raw_c (+ a_l * raw_0 for tangent channels). */
static void norm_set(const double *raw, int ch, int n_coef, double a, double a_l, double **dst) {
     const double *r = raw + (size_t) ch * 6 * n_coef, *r0 = raw;
     for (int k = 0; k < 6; ++k)
          for (int l = 0; l < n_coef; ++l)
               dst[k][l] = ch == 0 ? r[k * n_coef + l] * a : r[k * n_coef + l] * a + r0[k * n_coef + l] * a_l;
}

static int truncated_count(double **coef, int n_coef, double accuracy) {
     for (int l = 0; l < n_coef; ++l)
          if (fabs(coef[0][l]) <= accuracy) return l + 1;
     return n_coef;
}

static int core_run(const core_args *c, core_out *o) {
     int nd = c->nd, nch = 1 + c->nd1 + c->nd2;
     job_mem jm;
     lmie_gpu_job job;
     lmie_gpu_result res;
     job_mem_alloc(&jm, c->N, c->nd1, c->nd2);
     fill_job(c, &job, jm.x, jm.f, jm.n1, jm.n2, jm.dx, jm.dm, jm.dft, jm.dfw, jm.sx, jm.cx);
     int C = job.n_coef, Q = job.n_qang, rc = -1;
     o->n_qang = Q; o->n_coef = C;
     memset(&res, 0, sizeof res);
     res.sums_ch = vec((size_t) 4 * (nch - 1));
     res.gc_raw = c->calc_gc ? vec((size_t) nch * 6 * C) : NULL;
     res.lc_raw = c->calc_lc ? vec((size_t) nch * 6 * C) : NULL;
     if (c->save_control == 1) {
          *o->save1 = table2(4, c->N);
          *o->save2 = table3(4, c->N, Q > 0 ? Q : 1);
          res.save1 = (*o->save1)[0];
          res.save2 = (*o->save2)[0][0];
     }
     if (c->save_control == 2) {
          /* the per-size arrays of an earlier save_control = 1 call on the same
             grid (src/lmie_schedule.c:1087-1124) */
          if (!o->save1 || !o->save2 || !*o->save1 || !**o->save1 || !*o->save2 || !**o->save2 ||
              (Q > 0 && !***o->save2)) {
               fail("lmie_core_solution", "save_control = 2 needs save1 and save2 of an earlier call");
               goto done;
          }
          job.saved1 = (*o->save1)[0];
          job.saved2 = Q > 0 ? (*o->save2)[0][0] : NULL;
     }
     if (c->on_device) c->on_device(c->on_device_arg);
     rc = lmie_gpu_forward(&job, &res);
     if (rc) {
          fail("lmie_gpu_forward", lmie_gpu_last_error());
          goto done;
     }
     o->state = res.state;
     g_last_precision = res.precision;
     /* per-channel raw tangent sums -> seed order */
     double *rs_l = vec((size_t) 4 * (nd ? nd : 1));
     for (int t = 0; t < c->nd1; ++t)
          for (int q = 0; q < 4; ++q) rs_l[4 * c->idx1[t] + q] = res.sums_ch[4 * t + q];
     for (int t = 0; t < c->nd2; ++t)
          for (int q = 0; q < 4; ++q) rs_l[4 * c->idx2[t] + q] = res.sums_ch[4 * (c->nd1 + t) + q];
     double Cext = res.sums[0], Csca = res.sums[1], Cbak = res.sums[2], G = res.sums[3];
     memcpy(o->sums, res.sums, sizeof o->sums);
     double a = 2. / Csca, s = c->lambda * c->lambda / (2. * LM_PI);
     for (int d = 0; d < nd; ++d) {
          double al = -a * rs_l[4 * d + 1] / Csca;
          double sl = 2. * c->lambda_l[d] * c->lambda / (2. * LM_PI);
          o->g_l[d] = rs_l[4 * d + 3] * a + G * al;
          o->cext_l[d] = rs_l[4 * d + 0] * s + Cext * sl;
          o->csca_l[d] = rs_l[4 * d + 1] * s + Csca * sl;
          o->cbak_l[d] = (rs_l[4 * d + 2] * s + Cbak * sl) / (2. * 4. * LM_PI);
          /* channel of seed d in the raw arrays */
          int ch = -1;
          for (int t = 0; t < c->nd1; ++t) if (c->idx1[t] == d) ch = 1 + t;
          for (int t = 0; t < c->nd2; ++t) if (c->idx2[t] == d) ch = 1 + c->nd1 + t;
          if (c->calc_gc) { norm_set(res.gc_raw, ch, C, a, al, o->gc_l[d]); o->gc_l[d][0][0] = 0.; }
          if (c->calc_lc) { norm_set(res.lc_raw, ch, C, a, al, o->lc_l[d]); o->lc_l[d][0][0] = 0.; }
     }
     free(rs_l);
     *o->g = G / Csca * 2.;
     *o->cext = Cext * s;
     *o->csca = Csca * s;
     *o->cbak = Cbak * s / (2. * 4. * LM_PI);
     rc = 0;
     if (c->calc_gc) { norm_set(res.gc_raw, 0, C, a, 0., o->gc); rc = truncated_count(o->gc, C, c->accuracy); }
     if (c->calc_lc) { norm_set(res.lc_raw, 0, C, a, 0., o->lc); rc = truncated_count(o->lc, C, c->accuracy); }
done:
     free(res.sums_ch); free(res.gc_raw); free(res.lc_raw);
     job_mem_free(&jm);
     return rc;
}

/* ------------------------------------------------------------------------- */
/* Reference seam (src/lmie_schedule.h:18-30). */
int lmie_core_solution_ex(int n_qsiz, int n_derivs1, int n_derivs2, int save_control,
                          int *index1, int *index2, double lambda, double *lambda_l,
                          double mr, double mi, double *mr_l, double *mi_l,
                          double *qx, double *qw, double *nr, double r1, double r2,
                          double *cext_, double *csca_, double *pbak_, double *g,
                          double **gc, double **lc, double **qx_l, double **qw_l, double **nr_l,
                          double *cext_l_, double *csca_l_, double *pbak_l_, double *g_l,
                          double ***gc_l, double ***lc_l, double ***save1, double ****save2,
                          double accuracy, const lmie_opts *opts) {
     (void) r1;
     int nd = n_derivs1 + n_derivs2;
     core_args c;
     memset(&c, 0, sizeof c);
     c.N = n_qsiz; c.nd = nd; c.nd1 = n_derivs1; c.nd2 = n_derivs2;
     c.idx1 = index1; c.idx2 = index2; c.save_control = save_control;
     c.lambda = lambda; c.mr = mr; c.mi = mi;
     c.lambda_l = lambda_l; c.mr_l = mr_l; c.mi_l = mi_l;
     c.qx = qx; c.qw = qw; c.nr = nr;
     /* pointer tables from the reference allocator are contiguous row-major */
     c.qx_l = nd ? qx_l[0] : NULL; c.qw_l = nd ? qw_l[0] : NULL; c.nr_l = nd ? nr_l[0] : NULL;
     c.r2 = r2; c.calc_gc = gc != NULL; c.calc_lc = lc != NULL; c.accuracy = accuracy;
     c.opts = opts;
     double dummy[4];
     core_out o = {cext_, csca_, pbak_, g, gc, lc, nd ? cext_l_ : dummy, nd ? csca_l_ : dummy,
                   nd ? pbak_l_ : dummy, nd ? g_l : dummy, gc_l, lc_l, save1, save2, 0, 0, {0, 0, 0, 0}, NULL};
     return core_run(&c, &o);
}

int lmie_core_solution(int n_qsiz, int n_derivs1, int n_derivs2, int save_control,
                       int *index1, int *index2, double lambda, double *lambda_l,
                       double mr, double mi, double *mr_l, double *mi_l,
                       double *qx, double *qw, double *nr, double r1, double r2,
                       double *cext_, double *csca_, double *pbak_, double *g,
                       double **gc, double **lc, double **qx_l, double **qw_l, double **nr_l,
                       double *cext_l_, double *csca_l_, double *pbak_l_, double *g_l,
                       double ***gc_l, double ***lc_l, double ***save1, double ****save2,
                       double accuracy, int verbose, int n_threads, int use_mpi) {
     (void) verbose; (void) n_threads;
     if (use_mpi) return fail("lmie_core_solution", "MPI is not supported (one GPU)");
     return lmie_core_solution_ex(n_qsiz, n_derivs1, n_derivs2, save_control, index1, index2, lambda,
                                  lambda_l, mr, mi, mr_l, mi_l, qx, qw, nr, r1, r2, cext_, csca_,
                                  pbak_, g, gc, lc, qx_l, qw_l, nr_l, cext_l_, csca_l_, pbak_l_, g_l,
                                  gc_l, lc_l, save1, save2, accuracy, NULL);
}

/* ------------------------------------------------------------------------- */
/* Public API (src/lmie_interface.c). */
int lmie_calc_max_coef(double lambda, enum size_dist_type dist_type,
                       double a1, double a2, double a3, double r1, double r2) {
     (void) a3;
     if (dist_type == SIZE_DIST_POWER_LAW) pl_range(a1, a2, &r1, &r2);
     return 2 * mie_n1(2. * LM_PI * r2 / lambda);
}

int lmie_in_alloc(lmie_in_data *in, int n) {
     if (n > 0) {
          double **v[10] = {&in->lambda_l, &in->mr_l, &in->mi_l, &in->a1_l, &in->a2_l,
                            &in->a3_l, &in->a4_l, &in->a5_l, &in->r1_l, &in->r2_l};
          for (int k = 0; k < 10; ++k)
               if (!(*v[k] = calloc(n, sizeof(double)))) return fail("lmie_in_alloc", "out of memory");
     }
     return 0;
}

void lmie_in_free(lmie_in_data *in, int flag) {
     if (!flag) return;
     free(in->lambda_l); free(in->mr_l); free(in->mi_l); free(in->a1_l); free(in->a2_l);
     free(in->a3_l); free(in->a4_l); free(in->a5_l); free(in->r1_l); free(in->r2_l);
}

/* Every block is checked; on a failure everything allocated here is freed,
   the pointers are left NULL and -1 is returned (src/lmie_interface.c:79-125
   returns -1 from the first failing allocation).  Arrays of outputs that are
   not requested are set to NULL, so writers and lmie_out_free can test them. */
int lmie_out_alloc(lmie_in_data *in, lmie_out_data *out, int max_coef) {
     if (!in || !out) return fail("lmie_out_alloc", "missing argument");
     const int n = in->n_derivs > 0 ? in->n_derivs : 0;
     if (max_coef < 0 || (in->calc_pf && in->n_angles < 1))
          return fail("lmie_out_alloc", "invalid array size");
     double **v[13] = {&out->r1_l, &out->r2_l, &out->norm_l, &out->reff_l, &out->veff_l,
                       &out->gavg_l, &out->vavg_l, &out->ravg_l, &out->rvw_l, &out->cext_l,
                       &out->csca_l, &out->cbak_l, &out->g_l};
     int bad = 0;
     for (int k = 0; k < 13; ++k) {
          *v[k] = n > 0 ? vec(n) : NULL;
          bad |= n > 0 && !*v[k];
     }
     out->gc = NULL; out->gc_l = NULL;
     out->lc = NULL; out->lc_l = NULL;
     out->pf = NULL; out->pf_l = NULL; out->theta = NULL;
     if (in->calc_gc) {
          bad |= !(out->gc = table2(6, max_coef));
          if (n > 0) bad |= !(out->gc_l = table3(n, 6, max_coef));
     }
     if (in->calc_lc) {
          bad |= !(out->lc = table2(6, max_coef));
          if (n > 0) bad |= !(out->lc_l = table3(n, 6, max_coef));
     }
     if (in->calc_pf) {
          bad |= !(out->theta = vec(in->n_angles));
          bad |= !(out->pf = table2(6, in->n_angles));
          if (n > 0) bad |= !(out->pf_l = table3(n, 6, in->n_angles));
     }
     if (in->save_control != 2) { out->save1 = NULL; out->save2 = NULL; }
     if (bad) {
          for (int k = 0; k < 13; ++k) { free(*v[k]); *v[k] = NULL; }
          free(out->gc); free(out->gc_l); free(out->lc); free(out->lc_l);
          free(out->pf); free(out->pf_l); free(out->theta);
          out->gc = NULL; out->gc_l = NULL; out->lc = NULL; out->lc_l = NULL;
          out->pf = NULL; out->pf_l = NULL; out->theta = NULL;
          return fail("lmie_out_alloc", "out of memory");
     }
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

/* Scattering matrix at user angles theta_i = i * 180 / (n - 1)
   (misc/rtutil_scat_int.c:487-503, :771-797), evaluated on the GPU. */
static int user_angles(int n_coef, int n_angles, int nd, int ld, double **xc, double ***xc_l,
                       double *theta, double **pf, double ***pf_l, int gsf, int device) {
     double dth = 180. / (n_angles - 1);
     for (int i = 0; i < n_angles; ++i) theta[i] = i * dth;
     if (lmie_gpu_phase_func(1, n_coef, ld, xc[0], n_angles, theta, pf[0], gsf, device))
          return fail("lmie_gpu_phase_func", lmie_gpu_last_error());
     if (nd > 0 && lmie_gpu_phase_func(nd, n_coef, ld, xc_l[0][0], n_angles, theta, pf_l[0][0], gsf, device))
          return fail("lmie_gpu_phase_func", lmie_gpu_last_error());
     return 0;
}

/* Everything of lmie_solution2_l for one or several distributions on one grid. */
typedef struct {
     int calc_gc, calc_lc, calc_pf, n_angles, save_control;
     double lambda, mr, mi, accuracy;
     int keep;
     void (*on_device)(void *);     /* core_args.on_device */
     void *on_device_arg;
} run_flags;

static int solve_grid(const run_flags *rf, grid_t *g, const double *lambda_l, const double *mr_l,
                      const double *mi_l, const lmie_opts *opts, int *n_coef,
                      double *r21, double *r22, double *norm, double *reff, double *veff,
                      double *gavg, double *vavg, double *ravg, double *rvw,
                      double *cext, double *csca, double *cbak, double *g_,
                      double **gc, double **lc, double *theta, double **pf,
                      double *r21_l, double *r22_l, double *norm_l, double *reff_l, double *veff_l,
                      double *gavg_l, double *vavg_l, double *ravg_l, double *rvw_l,
                      double *cext_l, double *csca_l, double *cbak_l, double *g_l,
                      double ***gc_l, double ***lc_l, double ***pf_l,
                      double ***save1, double ****save2, core_out *keep) {
     int nd = g->nd;
     dstats st, *stl = calloc(nd ? nd : 1, sizeof(dstats));
     grid_stats(g, &st, stl);
     *r21 = g->r1; *r22 = g->r2; *norm = g->norm;
     *reff = st.reff; *veff = st.veff; *gavg = st.gavg; *vavg = st.vavg; *ravg = st.ravg; *rvw = st.rvw;
     for (int d = 0; d < nd; ++d) {
          r21_l[d] = g->r1_l[d]; r22_l[d] = g->r2_l[d]; norm_l[d] = g->norm_l[d];
          reff_l[d] = stl[d].reff; veff_l[d] = stl[d].veff; gavg_l[d] = stl[d].gavg;
          vavg_l[d] = stl[d].vavg; ravg_l[d] = stl[d].ravg; rvw_l[d] = stl[d].rvw;
     }
     free(stl);
     int *idx1 = calloc(nd + 1, sizeof(int)), *idx2 = calloc(nd + 1, sizeof(int)), nd1 = 0, nd2 = 0;
     for (int d = 0; d < nd; ++d) {
          if (lambda_l[d] != 0. || mr_l[d] != 0. || mi_l[d] != 0. || g->r1_l[d] != 0. || g->r2_l[d] != 0.)
               idx1[nd1++] = d;
          else
               idx2[nd2++] = d;
     }
     core_args c;
     memset(&c, 0, sizeof c);
     double zero = 0.;
     c.N = g->N; c.nd = nd; c.nd1 = nd1; c.nd2 = nd2; c.idx1 = idx1; c.idx2 = idx2;
     c.save_control = rf->save_control;
     c.lambda = rf->lambda; c.mr = rf->mr; c.mi = rf->mi;
     c.lambda_l = nd ? lambda_l : &zero; c.mr_l = nd ? mr_l : &zero; c.mi_l = nd ? mi_l : &zero;
     c.qx = g->qx; c.qw = g->qw; c.nr = g->nr; c.qx_l = g->qx_l; c.qw_l = g->qw_l; c.nr_l = g->nr_l;
     c.r2 = g->r2; c.calc_gc = rf->calc_gc; c.calc_lc = rf->calc_lc; c.accuracy = rf->accuracy;
     c.opts = opts;
     c.keep = rf->keep;
     c.on_device = rf->on_device;
     c.on_device_arg = rf->on_device_arg;
     double dummy[4];
     core_out o = {cext, csca, cbak, g_, gc, lc, nd ? cext_l : dummy, nd ? csca_l : dummy,
                   nd ? cbak_l : dummy, nd ? g_l : dummy, gc_l, lc_l, save1, save2, 0, 0, {0, 0, 0, 0}, NULL};
     int rc = core_run(&c, &o);
     free(idx1); free(idx2);
     if (rc < 0) return fail("lmie_core_solution", lmie_last_error());
     *n_coef = rc;
     if (keep) *keep = o;
     if (rf->calc_pf) {
          int dev = opts ? opts->device : 0;
          if (rf->calc_gc) rc = user_angles(*n_coef, rf->n_angles, nd, o.n_coef, gc, gc_l, theta, pf, pf_l, 1, dev);
          else rc = user_angles(*n_coef, rf->n_angles, nd, o.n_coef, lc, lc_l, theta, pf, pf_l, 0, dev);
          if (rc) return -1;
     }
     return 0;
}

static int solution2_l_opts(int calc_gc, int calc_lc, int calc_pf, enum size_dist_type dist_type,
                            int n_int1, int n_int2, int n_quad, int n_angles, int n_derivs,
                            int save_control, double lambda, double mr, double mi,
                            double a1, double a2, double a3, double a4, double a5,
                            double r1, double r2, double *lambda_l, double *mr_l, double *mi_l,
                            double *a1_l, double *a2_l, double *a3_l, double *a4_l, double *a5_l,
                            double *r1_l, double *r2_l, double accuracy, int *n_coef,
                            double *r21, double *r22, double *norm, double *reff, double *veff,
                            double *gavg, double *vavg, double *ravg, double *rvw,
                            double *cext, double *csca, double *cbak, double *g,
                            double **gc, double **lc, double *theta, double **pf,
                            double *r21_l, double *r22_l, double *norm_l, double *reff_l,
                            double *veff_l, double *gavg_l, double *vavg_l, double *ravg_l,
                            double *rvw_l, double *cext_l, double *csca_l, double *cbak_l,
                            double *g_l, double ***gc_l, double ***lc_l, double ***pf_l,
                            double ***save1, double ****save2, int use_mpi, const lmie_opts *opts) {
     if (use_mpi) return fail("lmie_solution", "MPI is not supported (one GPU)");
     if (calc_pf && !calc_gc && !calc_lc)
          return fail("lmie_solution", "either calc_gc or calc_lc must be set to true when calc_pf is true");
     int nd = n_derivs > 0 ? n_derivs : 0;
     double a[5] = {a1, a2, a3, a4, a5};
     seeds sv = {nd, {a1_l, a2_l, a3_l, a4_l, a5_l}, r1_l, r2_l};
     grid_t gr;
     memset(&gr, 0, sizeof gr);
     const double tg = wall_ms();
     if (grid_build(dist_type, n_int1, n_int2, n_quad, a, r1, r2, &sv, &gr)) {
          grid_free(&gr);
          return fail("get_particle_dist", lmie_last_error());
     }
     lmie_gpu_host_time(LMIE_T_HOST_GRID, wall_ms() - tg);
     run_flags rf = {calc_gc, calc_lc, calc_pf, n_angles, save_control, lambda, mr, mi, accuracy, 0, NULL, NULL};
     int rc = solve_grid(&rf, &gr, lambda_l, mr_l, mi_l, opts, n_coef, r21, r22, norm, reff, veff,
                         gavg, vavg, ravg, rvw, cext, csca, cbak, g, gc, lc, theta, pf, r21_l, r22_l,
                         norm_l, reff_l, veff_l, gavg_l, vavg_l, ravg_l, rvw_l, cext_l, csca_l,
                         cbak_l, g_l, gc_l, lc_l, pf_l, save1, save2, NULL);
     grid_free(&gr);
     return rc;
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
     (void) verbose; (void) n_threads;
     return solution2_l_opts(calc_gc, calc_lc, calc_pf, dist_type, n_int1, n_int2, n_quad, n_angles,
                             n_derivs, save_control, lambda, mr, mi, a1, a2, a3, a4, a5, r1, r2,
                             lambda_l, mr_l, mi_l, a1_l, a2_l, a3_l, a4_l, a5_l, r1_l, r2_l, accuracy,
                             n_coef, r21, r22, norm, reff, veff, gavg, vavg, ravg, rvw, cext, csca,
                             cbak, g, gc, lc, theta, pf, r21_l, r22_l, norm_l, reff_l, veff_l, gavg_l,
                             vavg_l, ravg_l, rvw_l, cext_l, csca_l, cbak_l, g_l, gc_l, lc_l, pf_l,
                             save1, save2, use_mpi, NULL);
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
     return lmie_solution2_l(calc_gc, calc_lc, calc_pf, dist_type, n_int1, n_int2, n_quad, n_angles,
                             0, save_control, lambda, mr, mi, a1, a2, a3, a4, a5, r1, r2, NULL, NULL,
                             NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, accuracy, n_coef, r21,
                             r22, norm, reff, veff, gavg, vavg, ravg, rvw, cext, csca, cbak, g, gc,
                             lc, theta, pf, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
                             NULL, NULL, NULL, NULL, NULL, NULL, NULL, save1, save2, verbose,
                             n_threads, use_mpi);
}

static int solution_opts_run(lmie_in_data *in, lmie_out_data *out, int alloc_out, int use_mpi,
                             const lmie_opts *opts);

static int solution_opts(lmie_in_data *in, lmie_out_data *out, int alloc_out, int use_mpi,
                         const lmie_opts *opts) {
     const double t0 = wall_ms();
     lmie_gpu_host_times_clear();
     const int rc = solution_opts_run(in, out, alloc_out, use_mpi, opts);
     lmie_gpu_host_time(LMIE_T_CALL, wall_ms() - t0);
     return rc;
}

static int solution_opts_run(lmie_in_data *in, lmie_out_data *out, int alloc_out, int use_mpi,
                             const lmie_opts *opts) {
     if (alloc_out) {
          int mc = lmie_calc_max_coef(in->lambda, in->dist_type, in->a1, in->a2, in->a3, in->r1, in->r2);
          if (lmie_out_alloc(in, out, mc)) return fail("lmie_out_alloc", "failed");
     }
     if (solution2_l_opts(in->calc_gc, in->calc_lc, in->calc_pf, in->dist_type, in->n_int1, in->n_int2,
                          in->n_quad, in->n_angles, in->n_derivs, in->save_control, in->lambda,
                          in->mr, in->mi, in->a1, in->a2, in->a3, in->a4, in->a5, in->r1, in->r2,
                          in->lambda_l, in->mr_l, in->mi_l, in->a1_l, in->a2_l, in->a3_l, in->a4_l,
                          in->a5_l, in->r1_l, in->r2_l, in->accuracy, &out->n_coef, &out->r1,
                          &out->r2, &out->norm, &out->reff, &out->veff, &out->gavg, &out->vavg,
                          &out->ravg, &out->rvw, &out->cext, &out->csca, &out->cbak, &out->g,
                          out->gc, out->lc, out->theta, out->pf, out->r1_l, out->r2_l, out->norm_l,
                          out->reff_l, out->veff_l, out->gavg_l, out->vavg_l, out->ravg_l,
                          out->rvw_l, out->cext_l, out->csca_l, out->cbak_l, out->g_l, out->gc_l,
                          out->lc_l, out->pf_l, &out->save1, &out->save2, use_mpi, opts))
          return fail("lmie_solution2_l", lmie_last_error());
     return 0;
}

int lmie_solution(lmie_in_data *in, lmie_out_data *out, int alloc_out, int verbose,
                  int n_threads, int use_mpi) {
     (void) verbose; (void) n_threads;
     return solution_opts(in, out, alloc_out, use_mpi, NULL);
}

int lmie_solution_ex(lmie_in_data *in, lmie_out_data *out, int alloc_out, const lmie_opts *opts) {
     return solution_opts(in, out, alloc_out, 0, opts);
}

int lmie_solution_slave(void) { return fail("lmie_solution_slave", "MPI is not supported (one GPU)"); }

/* Frees what lmie_out_alloc gave an output (with nd tangent columns) and
   clears the pointers, for the error paths of calls that allocated it. */
static void out_release_nd(lmie_out_data *out, int nd) {
     lmie_out_free(out, nd > 0);
     out->gc = NULL; out->lc = NULL; out->pf = NULL; out->theta = NULL;
     out->gc_l = NULL; out->lc_l = NULL; out->pf_l = NULL;
}

/* ------------------------------------------------------------------------- */
/* Mixtures of distributions on one grid. */

/* Grid of the mixture: sum_k w_k n_k(r), each n_k normalised.  Seed column d
   belongs to mode owner[d]; its a1..a5 seeds act on that mode only, its
   r1/r2 seeds move the shared grid for all modes. */
static int mix_grid(lmie_in_data *const *modes, const double *weight, int n_modes, int nd,
                    grid_t *mix) {
     const lmie_in_data *m0 = modes[0];
     /* the power law and mono place their nodes from their own parameters, so
        such a mode has a grid of its own and cannot share one */
     for (int k = 0; k < n_modes && n_modes > 1; ++k)
          if (modes[k]->dist_type == SIZE_DIST_POWER_LAW || modes[k]->dist_type == SIZE_DIST_MONO)
               return fail("lmie_solution_mix", "power-law and mono modes cannot be mixed");
     int D = nd > 0 ? nd : 1;
     double **col = calloc(12, sizeof(double *));
     for (int q = 0; q < 7; ++q) col[q] = vec(D);    /* a1..a5, r1, r2 */
     memset(mix, 0, sizeof *mix);
     int off = 0, rc = 0;
     for (int k = 0; k < n_modes && !rc; ++k) {
          const lmie_in_data *m = modes[k];
          for (int q = 0; q < 7; ++q) memset(col[q], 0, sizeof(double) * D);
          for (int kk = 0, o2 = 0; kk < n_modes; o2 += modes[kk]->n_derivs, ++kk) {
               const lmie_in_data *mm = modes[kk];
               for (int d = 0; d < mm->n_derivs; ++d) {
                    if (kk == k) {
                         const double *src[5] = {mm->a1_l, mm->a2_l, mm->a3_l, mm->a4_l, mm->a5_l};
                         for (int q = 0; q < 5; ++q) col[q][o2 + d] = src[q][d];
                    }
                    col[5][o2 + d] = mm->r1_l[d];
                    col[6][o2 + d] = mm->r2_l[d];
               }
          }
          double a[5] = {m->a1, m->a2, m->a3, m->a4, m->a5};
          seeds sv = {nd, {col[0], col[1], col[2], col[3], col[4]}, col[5], col[6]};
          grid_t gk;
          memset(&gk, 0, sizeof gk);
          rc = grid_build(m->dist_type, m0->n_int1, m0->n_int2, m0->n_quad, a, m0->r1, m0->r2, &sv, &gk);
          if (rc) grid_free(&gk);
          if (!rc) {
               double w = weight ? weight[k] : 1.;
               if (k == 0) {
                    *mix = gk;
                    for (int i = 0; i < gk.N; ++i) mix->nr[i] *= w;
                    for (size_t i = 0; i < (size_t) gk.N * nd; ++i) mix->nr_l[i] *= w;
                    mix->norm *= w;
                    for (int d = 0; d < nd; ++d) mix->norm_l[d] *= w;
               } else {
                    if (gk.N != mix->N || memcmp(gk.qx, mix->qx, sizeof(double) * gk.N))
                         rc = fail("lmie_solution_mix", "modes have different grids");
                    for (int i = 0; i < gk.N && !rc; ++i) mix->nr[i] += w * gk.nr[i];
                    for (size_t i = 0; i < (size_t) gk.N * nd && !rc; ++i) mix->nr_l[i] += w * gk.nr_l[i];
                    mix->norm += w * gk.norm;
                    for (int d = 0; d < nd; ++d) mix->norm_l[d] += w * gk.norm_l[d];
                    grid_free(&gk);
               }
          }
          off += m->n_derivs;
     }
     (void) off;
     for (int q = 0; q < 7; ++q) free(col[q]);
     free(col);
     if (rc) {
          grid_free(mix);
          memset(mix, 0, sizeof *mix);
     }
     return rc;
}

static void mix_seed_column(lmie_in_data *const *modes, int n_modes, int which, double *dst) {
     for (int k = 0, o = 0; k < n_modes; o += modes[k]->n_derivs, ++k) {
          const double *src = which == 0 ? modes[k]->lambda_l : which == 1 ? modes[k]->mr_l : modes[k]->mi_l;
          for (int d = 0; d < modes[k]->n_derivs; ++d) dst[o + d] = src[d];
     }
}

static int solution_mix_run(lmie_in_data *const *modes, const double *weight, int n_modes,
                            lmie_out_data *out, int alloc_out, const lmie_opts *opts);

int lmie_solution_mix(lmie_in_data *const *modes, const double *weight, int n_modes,
                      lmie_out_data *out, int alloc_out, const lmie_opts *opts) {
     const double t0 = wall_ms();
     lmie_gpu_host_times_clear();
     const int rc = solution_mix_run(modes, weight, n_modes, out, alloc_out, opts);
     lmie_gpu_host_time(LMIE_T_CALL, wall_ms() - t0);
     return rc;
}

static int solution_mix_run(lmie_in_data *const *modes, const double *weight, int n_modes,
                            lmie_out_data *out, int alloc_out, const lmie_opts *opts) {
     if (n_modes < 1 || !modes || !modes[0]) return fail("lmie_solution_mix", "no modes");
     const lmie_in_data *m0 = modes[0];
     int nd = 0;
     for (int k = 0; k < n_modes; ++k) nd += modes[k]->n_derivs > 0 ? modes[k]->n_derivs : 0;
     if (m0->calc_pf && !m0->calc_gc && !m0->calc_lc)
          return fail("lmie_solution_mix", "either calc_gc or calc_lc must be set to true when calc_pf is true");
     if (alloc_out) {
          lmie_in_data tmp = *m0;
          tmp.n_derivs = nd;
          int mc = lmie_calc_max_coef(m0->lambda, m0->dist_type, m0->a1, m0->a2, m0->a3, m0->r1, m0->r2);
          if (lmie_out_alloc(&tmp, out, mc)) return fail("lmie_out_alloc", "failed");
     }
     grid_t g;
     /* as lmie_solution (src/lmie_interface.c:208-248): on a failure after
        alloc_out the output stays allocated and the caller frees it */
     const double tg = wall_ms();
     if (mix_grid(modes, weight, n_modes, nd, &g)) return -1;
     lmie_gpu_host_time(LMIE_T_HOST_GRID, wall_ms() - tg);
     double *lam_l = vec(nd), *mr_l = vec(nd), *mi_l = vec(nd);
     mix_seed_column(modes, n_modes, 0, lam_l);
     mix_seed_column(modes, n_modes, 1, mr_l);
     mix_seed_column(modes, n_modes, 2, mi_l);
     run_flags rf = {m0->calc_gc, m0->calc_lc, m0->calc_pf, m0->n_angles, m0->save_control,
                     m0->lambda, m0->mr, m0->mi, m0->accuracy, 0, NULL, NULL};
     int rc = solve_grid(&rf, &g, lam_l, mr_l, mi_l, opts, &out->n_coef, &out->r1, &out->r2,
                         &out->norm, &out->reff, &out->veff, &out->gavg, &out->vavg, &out->ravg,
                         &out->rvw, &out->cext, &out->csca, &out->cbak, &out->g, out->gc, out->lc,
                         out->theta, out->pf, out->r1_l, out->r2_l, out->norm_l, out->reff_l,
                         out->veff_l, out->gavg_l, out->vavg_l, out->ravg_l, out->rvw_l, out->cext_l,
                         out->csca_l, out->cbak_l, out->g_l, out->gc_l, out->lc_l, out->pf_l,
                         &out->save1, &out->save2, NULL);
     free(lam_l); free(mr_l); free(mi_l);
     grid_free(&g);
     return rc;
}

/* ------------------------------------------------------------------------- */
/* Reverse mode.  The forward keeps its device state; the reverse pass pulls
   the output cotangents back through the normalisation here
   (src/lmie_schedule.c:1158-1199 read backwards), through the kernels to
   f_i, x_i and m, and from f_i = qw_i sum_k w_k n_k(r_i) and
   x_i = 2 pi r_i / lambda to the distribution parameters, r1 and r2 with the
   grid tangents of each mode. */
#define ADJ_SEEDS 7                 /* a1 .. a5, r1, r2 */
struct lmie_adj_ctx {
     void *state;
     int n_modes, N, C, calc_gc, calc_lc;
     size_t ld_gc, ld_lc;           /* row strides of the caller's gc / lc blocks (and cotangents) */
     double lambda, Csca, g, cext, csca, cbak;
     double *gc, *lc;               /* normalised outputs [6][C] */
     double *w;                     /* mode weights */
     double **df;                   /* per mode: d (w_k qw_i n_k(r_i)) / d seed, [N][ADJ_SEEDS] */
     double **dx;                   /* per mode: d x_i / d seed, [N][ADJ_SEEDS] */
     double *xl;                    /* [N] x_i / lambda */
};

static void adj_tangents_free(double **v, int n) {
     if (v)
          for (int k = 0; k < n; ++k) free(v[k]);
     free(v);
}

void lmie_adjoint_free(lmie_adj_ctx *ctx) {
     if (!ctx) return;
     lmie_gpu_release(ctx->state);
     free(ctx->gc); free(ctx->lc); free(ctx->w); free(ctx->xl);
     adj_tangents_free(ctx->df, ctx->n_modes);
     adj_tangents_free(ctx->dx, ctx->n_modes);
     free(ctx);
}

/* Per mode k, unit seeds on a1..a5, r1, r2: the tangents of the size weight
   w_k qw_i n_k(r_i) (node weight and normalised density both move) and of
   x_i = 2 pi r_i / lambda. */
typedef struct {
     lmie_in_data *const *modes;
     const double *weight;
     int n_modes, N;
     double lambda;
     double **df, **dx;
} dn_job;

static void *dn_build(void *vp) {
     dn_job *j = vp;
     const lmie_in_data *m0 = j->modes[0];
     const int S = ADJ_SEEDS;
     j->df = calloc(j->n_modes, sizeof(double *));
     j->dx = calloc(j->n_modes, sizeof(double *));
     for (int k = 0; k < j->n_modes; ++k) {
          const lmie_in_data *m = j->modes[k];
          const double w = j->weight ? j->weight[k] : 1.;
          double e[ADJ_SEEDS][ADJ_SEEDS];
          memset(e, 0, sizeof e);
          for (int q = 0; q < S; ++q) e[q][q] = 1.;
          seeds sv = {S, {e[0], e[1], e[2], e[3], e[4]}, e[5], e[6]};
          double a[5] = {m->a1, m->a2, m->a3, m->a4, m->a5};
          grid_t gk;
          memset(&gk, 0, sizeof gk);
          j->df[k] = vec((size_t) S * j->N);
          j->dx[k] = vec((size_t) S * j->N);
          /* a seed moves the size parameters only where the forward makes it a
             Mie tangent channel: where it moves the radius range
             (src/lmie_interface.c:327-331), so the mono distribution's a1,
             whose node moves with r1, r2 held (misc/rtutil_size_dist.c:267-290),
             acts on the weights only, as in the reference's tangents */
          if (grid_build(m->dist_type, m0->n_int1, m0->n_int2, m0->n_quad, a, m0->r1, m0->r2, &sv, &gk) == 0 &&
              gk.N == j->N)
               for (int i = 0; i < gk.N; ++i)
                    for (int q = 0; q < S; ++q) {
                         const size_t o = (size_t) i * S + q;
                         const int moves = gk.r1_l[q] != 0. || gk.r2_l[q] != 0.;
                         j->df[k][o] = w * (gk.nr_l[o] * gk.qw[i] + gk.nr[i] * gk.qw_l[o]);
                         j->dx[k][o] = moves ? 2. * LM_PI * gk.qx_l[o] / j->lambda : 0.;
                    }
          grid_free(&gk);
     }
     return NULL;
}

/* One persistent helper thread for dn_build, created on first use: a call
   hands it its job instead of creating a thread (thread creation sat on the
   forward's path).  A second concurrent caller finds it busy and creates a
   thread of its own. */
static struct {
     pthread_mutex_t mu;
     pthread_cond_t go, done;
     int started, busy, has_job;
     dn_job *job;
} g_helper = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, PTHREAD_COND_INITIALIZER, 0, 0, 0, NULL};

static void *helper_main(void *vp) {
     (void) vp;
     pthread_mutex_lock(&g_helper.mu);
     for (;;) {
          while (!g_helper.has_job) pthread_cond_wait(&g_helper.go, &g_helper.mu);
          dn_job *j = g_helper.job;
          pthread_mutex_unlock(&g_helper.mu);
          dn_build(j);
          pthread_mutex_lock(&g_helper.mu);
          g_helper.has_job = 0;
          pthread_cond_broadcast(&g_helper.done);
     }
     return NULL;
}

/* hand j to the helper: 0 if it took it (then helper_join), -1 if busy */
static int helper_post(dn_job *j) {
     pthread_mutex_lock(&g_helper.mu);
     if (g_helper.busy) {
          pthread_mutex_unlock(&g_helper.mu);
          return -1;
     }
     if (!g_helper.started) {
          pthread_t th;
          if (pthread_create(&th, NULL, helper_main, NULL) != 0) {
               pthread_mutex_unlock(&g_helper.mu);
               return -1;
          }
          pthread_detach(th);
          g_helper.started = 1;
     }
     g_helper.busy = 1;
     g_helper.job = j;
     g_helper.has_job = 1;
     pthread_cond_signal(&g_helper.go);
     pthread_mutex_unlock(&g_helper.mu);
     return 0;
}

static void helper_join(void) {
     pthread_mutex_lock(&g_helper.mu);
     while (g_helper.has_job) pthread_cond_wait(&g_helper.done, &g_helper.mu);
     g_helper.busy = 0;
     pthread_mutex_unlock(&g_helper.mu);
}

/* dn_build on the helper thread (or a thread of its own), started by
   core_run's on_device hook; inline where no thread can be had, and on the
   threads of a fused batch */
typedef struct {
     dn_job *job;
     int running, pooled, done;
     pthread_t th;
} dn_start;

static void dn_launch(void *vp) {
     dn_start *st = vp;
     if (st->running || st->pooled || st->done) return;
     if (!g_batch_thread) {
          if (helper_post(st->job) == 0) {
               st->pooled = 1;
               return;
          }
          if (pthread_create(&st->th, NULL, dn_build, st->job) == 0) {
               st->running = 1;
               return;
          }
     }
     dn_build(st->job);
     st->done = 1;
}

static void out_release(lmie_out_data *out) { out_release_nd(out, 0); }

static int adjoint_forward_run(lmie_in_data *const *modes, const double *weight, int n_modes,
                               lmie_out_data *out, int alloc_out, const lmie_opts *opts,
                               lmie_adj_ctx **ctxp);

int lmie_adjoint_forward(lmie_in_data *const *modes, const double *weight, int n_modes,
                         lmie_out_data *out, int alloc_out, const lmie_opts *opts,
                         lmie_adj_ctx **ctxp) {
     const double t0 = wall_ms();
     lmie_gpu_host_times_clear();
     const int rc = adjoint_forward_run(modes, weight, n_modes, out, alloc_out, opts, ctxp);
     lmie_gpu_host_time(LMIE_T_CALL, wall_ms() - t0);
     return rc;
}

static int adjoint_forward_run(lmie_in_data *const *modes, const double *weight, int n_modes,
                               lmie_out_data *out, int alloc_out, const lmie_opts *opts,
                               lmie_adj_ctx **ctxp) {
     if (n_modes < 1 || !modes || !modes[0] || !ctxp) return fail("lmie_adjoint_forward", "missing argument");
     *ctxp = NULL;
     const lmie_in_data *m0 = modes[0];
     if (m0->save_control != 0) return fail("lmie_adjoint_forward", "save_control must be 0");
     if (alloc_out) {
          lmie_in_data tmp = *m0;
          tmp.n_derivs = 0;
          int mc = lmie_calc_max_coef(m0->lambda, m0->dist_type, m0->a1, m0->a2, m0->a3, m0->r1, m0->r2);
          if (lmie_out_alloc(&tmp, out, mc)) return fail("lmie_out_alloc", "failed");
     }
     lmie_in_data *copies = calloc(n_modes, sizeof(*copies));
     lmie_in_data **nm = calloc(n_modes, sizeof(*nm));
     for (int k = 0; k < n_modes; ++k) { copies[k] = *modes[k]; copies[k].n_derivs = 0; nm[k] = &copies[k]; }
     grid_t g;
     const double tg = wall_ms();
     int rc = mix_grid(nm, weight, n_modes, 0, &g);
     lmie_gpu_host_time(LMIE_T_HOST_GRID, wall_ms() - tg);
     free(nm); free(copies);
     if (rc) {
          if (alloc_out) out_release(out);
          return -1;
     }
     /* the number density tangents of the modes (for the reverse pass) on a
        helper thread while the forward runs on the GPU: started once the
        host's own per-size loops are done, so the two do not share the cores */
     dn_job dj = {modes, weight, n_modes, g.N, m0->lambda, NULL, NULL};
     dn_start st;
     memset(&st, 0, sizeof st);
     st.job = &dj;
     run_flags rf = {m0->calc_gc, m0->calc_lc, 0, 0, 0, m0->lambda, m0->mr, m0->mi, m0->accuracy, 1,
                     dn_launch, &st};
     core_out keep;
     double zero = 0.;
     rc = solve_grid(&rf, &g, &zero, &zero, &zero, opts, &out->n_coef, &out->r1, &out->r2,
                     &out->norm, &out->reff, &out->veff, &out->gavg, &out->vavg, &out->ravg,
                     &out->rvw, &out->cext, &out->csca, &out->cbak, &out->g, out->gc, out->lc,
                     out->theta, out->pf, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
                     NULL, NULL, NULL, NULL, NULL, NULL, NULL, &out->save1, &out->save2, &keep);
     const double tw = wall_ms();
     if (st.pooled) helper_join();
     else if (st.running) pthread_join(st.th, NULL);
     else if (!st.done && rc == 0) dn_build(&dj);     /* not started (cannot happen on success) */
     lmie_gpu_host_time(LMIE_T_HOST_WAIT, wall_ms() - tw);
     if (rc) {
          adj_tangents_free(dj.df, n_modes);
          adj_tangents_free(dj.dx, n_modes);
          grid_free(&g);
          if (alloc_out) out_release(out);
          return -1;
     }
     lmie_adj_ctx *c = calloc(1, sizeof *c);
     c->state = keep.state;
     c->n_modes = n_modes; c->N = g.N; c->C = keep.n_coef;
     c->calc_gc = m0->calc_gc; c->calc_lc = m0->calc_lc;
     c->lambda = m0->lambda; c->Csca = keep.sums[1]; c->g = out->g;
     c->cext = out->cext; c->csca = out->csca; c->cbak = out->cbak;
     /* the caller's blocks may be wider than the coefficient count (alloc_out
        = 0 with a larger max_coef): rows are copied with their own stride, and
        the cotangents are read with the same stride */
     if (c->calc_gc) {
          c->ld_gc = (size_t) (out->gc[1] - out->gc[0]);
          c->gc = vec((size_t) 6 * c->C);
          for (int k = 0; k < 6; ++k) memcpy(c->gc + (size_t) k * c->C, out->gc[k], sizeof(double) * c->C);
     }
     if (c->calc_lc) {
          c->ld_lc = (size_t) (out->lc[1] - out->lc[0]);
          c->lc = vec((size_t) 6 * c->C);
          for (int k = 0; k < 6; ++k) memcpy(c->lc + (size_t) k * c->C, out->lc[k], sizeof(double) * c->C);
     }
     c->w = vec(n_modes);
     for (int k = 0; k < n_modes; ++k) c->w[k] = weight ? weight[k] : 1.;
     c->df = dj.df; c->dx = dj.dx;
     c->xl = vec(g.N);
     for (int i = 0; i < g.N; ++i) c->xl[i] = 2. * LM_PI * g.qx[i] / (m0->lambda * m0->lambda);
     grid_free(&g);
     *ctxp = c;
     return 0;
}

static int adjoint_reverse_run(lmie_adj_ctx *c, const lmie_adj_in *adj, lmie_grad *grad, double *grad_modes);

int lmie_adjoint_reverse(lmie_adj_ctx *c, const lmie_adj_in *adj, lmie_grad *grad, double *grad_modes) {
     const double t0 = wall_ms();
     lmie_gpu_host_times_clear();
     const int rc = adjoint_reverse_run(c, adj, grad, grad_modes);
     lmie_gpu_host_time(LMIE_T_CALL, wall_ms() - t0);
     return rc;
}

/* The reverse pass's sums over sizes of the mode-parameter gradients, in
   GRAD_BLOCKS fixed blocks: part[b] = (5 per mode, r1, r2, lambda) */
#define GRAD_BLOCKS 32
typedef struct {
     const lmie_adj_ctx *c;
     const double *f_bar, *x_bar;
     double *part;
} grad_ctx;

static void grad_block(void *vp, int b0, int b1) {
     const grad_ctx *g = vp;
     const lmie_adj_ctx *c = g->c;
     const int S = ADJ_SEEDS, W = 5 * c->n_modes + 3;
     for (int b = b0; b < b1; ++b) {
          double *pb = g->part + (size_t) b * W;
          const int i0 = (int) ((long long) c->N * b / GRAD_BLOCKS), i1 = (int) ((long long) c->N * (b + 1) / GRAD_BLOCKS);
          for (int q = 0; q < W; ++q) pb[q] = 0.;
          for (int k = 0; k < c->n_modes; ++k)
               for (int i = i0; i < i1; ++i) {
                    const double *df = c->df[k] + (size_t) i * S, *dx = c->dx[k] + (size_t) i * S;
                    for (int q = 0; q < 5; ++q) pb[5 * k + q] += g->f_bar[i] * df[q] + g->x_bar[i] * dx[q];
                    pb[W - 3] += g->f_bar[i] * df[5];
                    pb[W - 2] += g->f_bar[i] * df[6];
               }
          for (int i = i0; i < i1; ++i) {
               pb[W - 3] += g->x_bar[i] * c->dx[0][(size_t) i * S + 5];
               pb[W - 2] += g->x_bar[i] * c->dx[0][(size_t) i * S + 6];
               pb[W - 1] += g->x_bar[i] * c->xl[i];
          }
     }
}

static int adjoint_reverse_run(lmie_adj_ctx *c, const lmie_adj_in *adj, lmie_grad *grad, double *grad_modes) {
     if (!c || !adj || !grad) return fail("lmie_adjoint_reverse", "missing argument");
     int C = c->C;
     double Csca = c->Csca, s = c->lambda * c->lambda / (2. * LM_PI);
     double sb[4];
     sb[0] = s * adj->cext;
     sb[2] = s * adj->cbak / (2. * 4. * LM_PI);
     sb[3] = 2. * adj->g / Csca;
     double dot = adj->g * c->g;
     double *gcb = NULL, *lcb = NULL;
     if (adj->gc && c->calc_gc) {
          gcb = vec((size_t) 6 * C);
          for (int k = 0; k < 6; ++k)
               for (int l = 0; l < C; ++l) {
                    const double v = adj->gc[k * c->ld_gc + l];
                    gcb[k * C + l] = v * 2. / Csca;
                    dot += v * c->gc[k * C + l];
               }
     }
     if (adj->lc && c->calc_lc) {
          lcb = vec((size_t) 6 * C);
          for (int k = 0; k < 6; ++k)
               for (int l = 0; l < C; ++l) {
                    const double v = adj->lc[k * c->ld_lc + l];
                    lcb[k * C + l] = v * 2. / Csca;
                    dot += v * c->lc[k * C + l];
               }
     }
     sb[1] = s * adj->csca - dot / Csca;
     lmie_gpu_adj_in ain = {sb, gcb, lcb};
     lmie_gpu_adj_out aout;
     aout.f_bar = vec(c->N);
     aout.x_bar = vec(c->N);
     int rc = lmie_gpu_adjoint(c->state, &ain, &aout);
     if (rc) {
          fail("lmie_gpu_adjoint", lmie_gpu_last_error());
     } else {
          grad->mr = aout.m_bar[0];
          grad->mi = aout.m_bar[1];
          /* a_p of mode k moves f_i through n_k (and, for grids set by the
             parameters, the nodes and weights); r1 and r2 move the shared grid
             of every mode: its weights and densities per mode, its nodes once */
          double *gp = vec((size_t) 5 * c->n_modes), gr[2] = {0., 0.}, lx = 0.;
          /* per fixed block of sizes on the pool, then the blocks in order:
             the sums do not depend on the threads */
          grad_ctx gx = {c, aout.f_bar, aout.x_bar, vec((size_t) GRAD_BLOCKS * (5 * c->n_modes + 3))};
          par_blocks(GRAD_BLOCKS, grad_block, &gx);
          for (int b = 0; b < GRAD_BLOCKS; ++b) {
               const double *pb = gx.part + (size_t) b * (5 * c->n_modes + 3);
               for (int q = 0; q < 5 * c->n_modes; ++q) gp[q] += pb[q];
               gr[0] += pb[5 * c->n_modes];
               gr[1] += pb[5 * c->n_modes + 1];
               lx += pb[5 * c->n_modes + 2];
          }
          free(gx.part);
          /* lambda: x_i = 2 pi r_i / lambda, and cext, csca, cbak carry
             lambda^2 / (2 pi) (src/lmie_schedule.c:1158-1199) */
          grad->lambda = 2. / c->lambda * (adj->cext * c->cext + adj->csca * c->csca + adj->cbak * c->cbak) - lx;
          grad->r1 = gr[0];
          grad->r2 = gr[1];
          grad->a1 = gp[0]; grad->a2 = gp[1]; grad->a3 = gp[2]; grad->a4 = gp[3]; grad->a5 = gp[4];
          if (grad_modes)
               for (int q = 0; q < 5 * c->n_modes; ++q) grad_modes[q] = gp[q];
          free(gp);
     }
     free(aout.f_bar); free(aout.x_bar); free(gcb); free(lcb);
     return rc ? -1 : 0;
}

int lmie_solution_adjoint(lmie_in_data *const *modes, const double *weight, int n_modes,
                          lmie_out_data *out, int alloc_out, const lmie_opts *opts,
                          lmie_cotangent_fn cot, void *user, lmie_grad *grad, double *grad_modes) {
     if (!cot) return fail("lmie_solution_adjoint", "no cotangent function");
     lmie_adj_ctx *c = NULL;
     if (lmie_adjoint_forward(modes, weight, n_modes, out, alloc_out, opts, &c)) return -1;
     lmie_adj_in adj;
     memset(&adj, 0, sizeof adj);
     int rc = cot(out, &adj, user);
     if (rc) rc = fail("lmie_solution_adjoint", "the cotangent function failed");
     else rc = lmie_adjoint_reverse(c, &adj, grad, grad_modes);
     lmie_adjoint_free(c);
     return rc;
}

/* ------------------------------------------------------------------------- */
/* Independent cases in one call: worker threads take the next case and run
   forward, cotangent and reverse on it.  Each worker holds its own device
   workspace (streams and buffers) for the duration of a case, so the device
   sees up to n_workers independent pipelines at once. */
typedef struct {
     int n_case, next, first_failed;
     lmie_in_data *const *cases;
     lmie_out_data *const *outs;
     int alloc_out;
     const lmie_opts *opts;
     lmie_cotangent_fn cot;
     void *const *users;
     lmie_grad *grads;
     double *grad_modes;
     int *status;
     pthread_mutex_t mu;
     char err[512];
} batch_job;

static int batch_case(batch_job *b, int c) {
     lmie_in_data *one[1] = {b->cases[c]};
     lmie_out_data *out = b->outs[c];
     if (!one[0] || !out) return fail("lmie_solution_adjoint_batch", "missing case");
     lmie_adj_ctx *ctx = NULL;
     if (lmie_adjoint_forward(one, NULL, 1, out, b->alloc_out, b->opts, &ctx)) return -1;
     lmie_adj_in adj;
     memset(&adj, 0, sizeof adj);
     int rc = b->cot(out, &adj, b->users ? b->users[c] : NULL);
     if (rc) rc = fail("lmie_solution_adjoint_batch", "the cotangent function failed");
     else rc = lmie_adjoint_reverse(ctx, &adj, &b->grads[c], b->grad_modes ? b->grad_modes + 5 * c : NULL);
     lmie_adjoint_free(ctx);
     if (rc && b->alloc_out) out_release(out);
     return rc ? -1 : 0;
}

/* Fused mode: one host thread per case of a device batch; the device passes
   of all cases run as single launches (include/lmie_gpu.h). */
typedef struct {
     batch_job *b;
     void *gpu_batch;
     int first, slot;
} batch_member;

static void batch_record(batch_job *b, int c, int rc) {
     if (b->status) b->status[c] = rc;
     if (rc) {
          pthread_mutex_lock(&b->mu);
          if (b->first_failed < 0 || c < b->first_failed) {
               b->first_failed = c;
               snprintf(b->err, sizeof b->err, "case %d: %s", c, lmie_last_error());
          }
          pthread_mutex_unlock(&b->mu);
     }
}

static void *batch_member_run(void *vp) {
     batch_member *m = vp;
     const int c = m->first + m->slot;
     g_batch_thread = 1;
     /* a fresh thread holding a valid slot: entering cannot fail */
     int rc = lmie_gpu_batch_enter(m->gpu_batch, m->slot);
     if (rc) {
          rc = fail("lmie_gpu_batch_enter", lmie_gpu_last_error());
     } else {
          rc = batch_case(m->b, c);
          lmie_gpu_batch_leave();
     }
     batch_record(m->b, c, rc);
     return NULL;
}

#define LMIE_BATCH_GROUP 64

static __thread int g_batch_fused[2];

int lmie_last_batch_fused(int pass) { return pass >= 0 && pass < 2 ? g_batch_fused[pass] : 0; }

/* Persistent threads for the cases of fused batches: a batch call hands each
   case to one of them instead of creating a thread per case and call (thread
   creation costs about as much as a small case's host work).  One batch call
   uses them at a time (region); a call that finds them busy, or cases beyond
   the threads that could be started, get threads of their own as before. */
static struct {
     pthread_mutex_t region;
     pthread_mutex_t mu;
     pthread_cond_t go, done;
     int n;                                  /* threads started */
     unsigned gen;                           /* generation of the current work */
     int m, pending;                         /* members of the generation, not yet done */
     batch_member *mem[LMIE_BATCH_GROUP];
} g_bpool = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER,
             PTHREAD_COND_INITIALIZER, 0, 0, 0, 0, {0}};

static void *bpool_worker(void *vp) {
     const int k = (int) (intptr_t) vp;
     unsigned seen = 0;
     pthread_mutex_lock(&g_bpool.mu);
     for (;;) {
          while (g_bpool.gen == seen) pthread_cond_wait(&g_bpool.go, &g_bpool.mu);
          seen = g_bpool.gen;
          if (k < g_bpool.m) {
               batch_member *mb = g_bpool.mem[k];
               pthread_mutex_unlock(&g_bpool.mu);
               batch_member_run(mb);
               pthread_mutex_lock(&g_bpool.mu);
               if (--g_bpool.pending == 0) pthread_cond_signal(&g_bpool.done);
          }
     }
     return NULL;
}

/* Runs members 0 .. m-1 on the persistent threads; returns how many it took
   (0 when the threads are in use by another call), the caller starts the rest. */
static int bpool_start(batch_member *mem, int m) {
     if (pthread_mutex_trylock(&g_bpool.region) != 0) return 0;
     pthread_mutex_lock(&g_bpool.mu);
     while (g_bpool.n < m) {
          pthread_t th;
          if (pthread_create(&th, NULL, bpool_worker, (void *) (intptr_t) g_bpool.n) != 0) break;
          pthread_detach(th);
          g_bpool.n++;
     }
     const int take = m < g_bpool.n ? m : g_bpool.n;
     for (int k = 0; k < take; ++k) g_bpool.mem[k] = &mem[k];
     g_bpool.m = take;
     g_bpool.pending = take;
     g_bpool.gen++;
     pthread_cond_broadcast(&g_bpool.go);
     pthread_mutex_unlock(&g_bpool.mu);
     if (take == 0) pthread_mutex_unlock(&g_bpool.region);
     return take;
}

static void bpool_wait(void) {
     pthread_mutex_lock(&g_bpool.mu);
     while (g_bpool.pending > 0) pthread_cond_wait(&g_bpool.done, &g_bpool.mu);
     pthread_mutex_unlock(&g_bpool.mu);
     pthread_mutex_unlock(&g_bpool.region);
}

static void batch_fused(batch_job *b) {
     g_batch_fused[0] = g_batch_fused[1] = 0;
     for (int g0 = 0; g0 < b->n_case; g0 += LMIE_BATCH_GROUP) {
          const int m = b->n_case - g0 < LMIE_BATCH_GROUP ? b->n_case - g0 : LMIE_BATCH_GROUP;
          void *gb = lmie_gpu_batch_open(m);
          if (!gb) {
               fail("lmie_gpu_batch_open", lmie_gpu_last_error());
               for (int k = 0; k < m; ++k) batch_record(b, g0 + k, -1);
               continue;
          }
          batch_member mem[LMIE_BATCH_GROUP];
          pthread_t th[LMIE_BATCH_GROUP];
          int started[LMIE_BATCH_GROUP];
          for (int k = 0; k < m; ++k) {
               mem[k].b = b; mem[k].gpu_batch = gb; mem[k].first = g0; mem[k].slot = k;
          }
          const int pooled = bpool_start(mem, m);
          for (int k = 0; k < m; ++k)
               started[k] = k < pooled ? 1 : pthread_create(&th[k], NULL, batch_member_run, &mem[k]) == 0;
          /* a case whose thread did not start leaves the device batch from here */
          for (int k = 0; k < m; ++k)
               if (!started[k]) {
                    if (lmie_gpu_batch_enter(gb, k) == 0) lmie_gpu_batch_leave();
                    fail("lmie_solution_adjoint_batch", "could not start a thread");
                    batch_record(b, g0 + k, -1);
               }
          for (int k = pooled; k < m; ++k)
               if (started[k]) pthread_join(th[k], NULL);
          if (pooled) bpool_wait();
          g_batch_fused[0] += lmie_gpu_batch_fused(gb, 0);
          g_batch_fused[1] += lmie_gpu_batch_fused(gb, 1);
          lmie_gpu_batch_close(gb);
     }
}

static void *batch_worker(void *vp) {
     batch_job *b = vp;
     for (;;) {
          pthread_mutex_lock(&b->mu);
          const int c = b->next++;
          pthread_mutex_unlock(&b->mu);
          if (c >= b->n_case) return NULL;
          batch_record(b, c, batch_case(b, c));
     }
}

int lmie_solution_adjoint_batch(int n_case, lmie_in_data *const *cases, lmie_out_data *const *outs,
                                int alloc_out, const lmie_opts *opts, lmie_cotangent_fn cot,
                                void *const *users, lmie_grad *grads, double *grad_modes,
                                int *status, int n_workers) {
     if (n_case < 0 || (n_case > 0 && (!cases || !outs || !grads)))
          return fail("lmie_solution_adjoint_batch", "missing argument");
     if (!cot) return fail("lmie_solution_adjoint_batch", "no cotangent function");
     if (n_case == 0) return 0;
     batch_job b;
     memset(&b, 0, sizeof b);
     b.n_case = n_case; b.first_failed = -1;
     b.cases = cases; b.outs = outs; b.alloc_out = alloc_out; b.opts = opts; b.cot = cot;
     b.users = users; b.grads = grads; b.grad_modes = grad_modes; b.status = status;
     pthread_mutex_init(&b.mu, NULL);
     g_batch_fused[0] = g_batch_fused[1] = 0;
     if (n_workers <= 0) {
          batch_fused(&b);
          pthread_mutex_destroy(&b.mu);
          if (b.first_failed >= 0) return fail("lmie_solution_adjoint_batch", b.err);
          return 0;
     }
     int T = n_workers;
     if (T > n_case) T = n_case;
     if (T > 64) T = 64;
     pthread_t th[64];
     int started = 0;
     for (int t = 1; t < T; ++t) {
          if (pthread_create(&th[started], NULL, batch_worker, &b) != 0) break;
          started++;
     }
     batch_worker(&b);
     for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
     pthread_mutex_destroy(&b.mu);
     if (b.first_failed >= 0) return fail("lmie_solution_adjoint_batch", b.err);
     return 0;
}

int lmie_cotangent_gc_l2(const lmie_out_data *out, lmie_adj_in *adj, void *user) {
     lmie_gc_l2 *l = user;
     if (!out || !adj || !l || !l->target || !l->weight || !l->gc_bar || !out->gc)
          return fail("lmie_cotangent_gc_l2", "missing argument");
     const size_t ld = (size_t) (out->gc[1] - out->gc[0]);
     const double *c = out->gc[0];
     double loss = 0.;
     for (size_t q = 0; q < 6 * ld; ++q) {
          const double d = c[q] - l->target[q];
          loss += l->weight[q] * d * d;
          l->gc_bar[q] = 2.0 * l->weight[q] * d;
     }
     l->loss = loss;
     adj->gc = l->gc_bar;
     return 0;
}

/* ------------------------------------------------------------------------- */
/* Error record for the other source files of the library (lmie_scat_io.c). */
void lmie_host_error(const char *where, const char *why) { fail(where, why); }
