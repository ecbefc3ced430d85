/*
 * kernel_arith.c -- CPU checks of the arithmetic identities the gfx950
 * kernels rely on (lmie_amd/csrc/lmie_gpu.hip), in IEEE double with
 * separately rounded operations and the C library's correctly rounded fma:
 *
 *  1. gsf_table's quotients: with y = RN(1/d) and q = RN(x y), the corrected
 *     q + y RN(x - q d) equals the IEEE quotient RN(x/d) bit for bit
 *     (Markstein's theorem), for the divisors of the generalized spherical
 *     function recurrences (l, sqrt(l^2 - 4), (l-1)(l^2 - 4)) and numerators
 *     of every size the recurrences produce, plus random pairs.
 *  2. mie_chains' psi: Miller's downward recurrence U_{n1} = 0, U_{n1-1} = 1,
 *     U_{j-1} = (2j+1)/x U_j - U_{j+1}, normalised by sin x / U_0, is the same
 *     function as the reference's ratio chain r_j = 1/((2j+1)/x - r_{j+1}),
 *     r_{n1} = 0, and product psi_j = sin x r_1 ... r_j (src/lmie_core.c:
 *     298-330): the two agree to rounding over x from 1e-3 to 2000, near
 *     x = k pi to the conditioning of the normalisation, which both share.
 *  3. mie_chains' reciprocal refinement r (1 + e + e^2), e = 1 - q r, from a
 *     starting reciprocal off by up to 2^-23 (the hardware's is good to about
 *     2^-24.6): within one ulp of 1/q.
 *  4. mie_chains' D chain carried as w_j = D_{j+1} + (j+1)/z,
 *     w_{j-1} = (2j+1)/z - conj(w_j) / |w_j|^2, D_j = (j+1)/z - conj(w_j) / |w_j|^2,
 *     against the reference's D_j = (j+1)/z - 1 / (D_{j+1} + (j+1)/z) with
 *     Smith's complex division (src/lmie_core.c:386-405): the same D_j to
 *     1e-12 relative for absorbing spheres, to 1e-10 for m nearly real (there
 *     any two roundings drift apart by about n2 ulp).
 * Prints "kernel arithmetic: ok" on success.
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint64_t rng = 88172645463325252ull;
static double urand(void) {   /* xorshift64*, uniform in [0, 1) */
     rng ^= rng >> 12; rng ^= rng << 25; rng ^= rng >> 27;
     return (double) ((rng * 2685821657736338717ull) >> 11) * 0x1p-53;
}

static int same(double a, double b) { return memcmp(&a, &b, sizeof a) == 0 || (a == 0. && b == 0.); }

static long check_quotient(double x, double d) {
     volatile double vx = x, vd = d;   /* no constant folding */
     const double y = 1. / vd, q = vx * y;
     const double v = fma(fma(-q, vd, vx), y, q);
     return same(v, vx / vd) ? 0 : 1;
}

static int markstein(void) {
     long bad = 0, n = 0;
     for (int l = 3; l <= 8000; ++l) {
          const double L = l - 1, L1 = l;
          const double d[3] = {L1, sqrt(L1 * L1 - 4.), L * (L1 * L1 - 4)};
          for (int k = 0; k < 3; ++k)
               for (int t = 0; t < 64; ++t) {
                    const double x = (urand() - .5) * ldexp(1., (int) (urand() * 60.) - 10) * d[k];
                    bad += check_quotient(x, d[k]);
                    ++n;
               }
     }
     for (int t = 0; t < 2000000; ++t) {
          const double x = (urand() - .5) * ldexp(1., (int) (urand() * 200.) - 100);
          const double d = (urand() + .5) * ldexp(1., (int) (urand() * 200.) - 100);
          bad += check_quotient(x, d);
          ++n;
     }
     if (bad) printf("markstein: %ld of %ld quotients differ from the IEEE division\n", bad, n);
     return bad != 0;
}

static int miller(void) {
     const double xs[] = {1e-3, 0.05, 0.7, 3.14159, 12.5, 99.9, 314.159, 1000.3, 1848.0, 2000.0};
     int fails = 0;
     for (size_t k = 0; k < sizeof xs / sizeof xs[0]; ++k) {
          const double x = xs[k];
          const int n1 = (int) (x + 4.05 * cbrt(x) + 8.);
          double *r = calloc(n1 + 1, sizeof *r), *pr = calloc(n1 + 1, sizeof *pr), *pm = calloc(n1 + 1, sizeof *pm);
          /* the reference's ratio chain and product */
          r[n1] = 0.;
          for (int j = n1 - 1; j >= 1; --j) r[j] = 1. / ((2. * j + 1.) / x - r[j + 1]);
          pr[0] = sin(x);
          for (int j = 1; j <= n1; ++j) pr[j] = r[j] * pr[j - 1];
          /* Miller's form, as mie_chains */
          const double ax = 1. / x;
          double ua = 0., ub = 1.;
          pm[n1] = 0.;
          pm[n1 - 1] = 1.;
          for (int j = n1 - 1; j >= 1; --j) {
               const double un = (2. * j + 1.) * ax * ub - ua;
               pm[j - 1] = un;
               ua = ub;
               ub = un;
          }
          const double s = sin(x) / ub;
          double mx = 0., err = 0.;
          for (int j = 0; j <= n1; ++j) mx = fmax(mx, fabs(pr[j]));
          for (int j = 0; j <= n1; ++j) err = fmax(err, fabs(s * pm[j] - pr[j]));
          /* both forms carry the cancellation of psi_0 / psi_1 (the reference in
             its first ratio r_1, Miller's in U_0), so near x = k pi they agree
             to the conditioning 1 / |sin x| of that step */
          if (!(err <= 1e-12 * mx / fmin(1., fabs(sin(x))))) {
               printf("miller: x %g n1 %d max |psi diff| %.3e of max |psi| %.3e\n", x, n1, err, mx);
               fails = 1;
          }
          free(r); free(pr); free(pm);
     }
     return fails;
}

static int refine(void) {
     long bad = 0;
     for (int t = 0; t < 2000000; ++t) {
          const double q = (urand() + .5) * ldexp(1., (int) (urand() * 120.) - 60);
          const double d = (2. * urand() - 1.) * 0x1p-23;
          volatile double r = (1. / q) * (1. + d);
          const double e = fma(-q, r, 1.);
          const double v = fma(r, fma(e, e, e), r), t1 = 1. / q;
          if (fabs(v - t1) > nextafter(fabs(t1), INFINITY) - fabs(t1)) ++bad;
     }
     if (bad) printf("refine: %ld reciprocals off by more than one ulp\n", bad);
     return bad != 0;
}

typedef struct { double r, i; } cplx;
static cplx cdiv_smith(cplx a, cplx b) {
     cplx o;
     if (fabs(b.r) >= fabs(b.i)) {
          const double t = b.i / b.r, d = b.r + b.i * t;
          o.r = (a.r + a.i * t) / d; o.i = (a.i - a.r * t) / d;
     } else {
          const double t = b.r / b.i, d = b.r * t + b.i;
          o.r = (a.r * t + a.i) / d; o.i = (a.i * t - a.r) / d;
     }
     return o;
}

static int dchain(void) {
     const double ms[][2] = {{1.53, 0.005}, {1.33, 1e-8}, {1.5, 0.5}, {2.0, 1.0}};
     const double xs[] = {0.01, 1.0, 30.0, 500.0, 1848.0};
     int fails = 0;
     for (size_t a = 0; a < 4; ++a)
          for (size_t b = 0; b < 5; ++b) {
               const double x = xs[b];
               const cplx z = {ms[a][0] * x, -ms[a][1] * x};
               const cplx iz = cdiv_smith((cplx) {1., 0.}, z);
               const int n2 = (int) (fmax(x + 4.05 * cbrt(x) + 8., hypot(z.r, z.i)) + 16.);
               cplx D = {0., 0.}, w = {iz.r * n2, iz.i * n2};
               double err = 0.;
               for (int j = n2 - 1; j >= 1; --j) {
                    const cplx c1 = {iz.r * (j + 1.), iz.i * (j + 1.)}, e = {iz.r * (2. * j + 1.), iz.i * (2. * j + 1.)};
                    /* reference form */
                    const cplx rc = cdiv_smith((cplx) {1., 0.}, (cplx) {D.r + c1.r, D.i + c1.i});
                    D.r = c1.r - rc.r; D.i = c1.i - rc.i;
                    /* the kernel's w form */
                    const double q = 1. / fma(w.r, w.r, w.i * w.i);
                    const cplx Dw = {fma(-w.r, q, c1.r), fma(w.i, q, c1.i)};
                    w.r = fma(-w.r, q, e.r); w.i = fma(w.i, q, e.i);
                    err = fmax(err, hypot(Dw.r - D.r, Dw.i - D.i) / hypot(D.r, D.i));
               }
               /* with m nearly real the recurrence neither damps nor grows
                  differences over the oscillatory orders, so any two roundings
                  drift apart by about n2 ulp */
               const double tol = ms[a][1] >= 0.005 ? 1e-12 : 1e-10;
               if (!(err <= tol)) {
                    printf("dchain: m %g-%gi x %g max relative |D diff| %.3e\n", ms[a][0], ms[a][1], x, err);
                    fails = 1;
               }
          }
     return fails;
}

int main(void) {
     const int f = markstein() | miller() | refine() | dchain();
     if (!f) printf("kernel arithmetic: ok\n");
     return f;
}
