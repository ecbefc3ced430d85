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

int main(void) {
     const int f = markstein() | miller();
     if (!f) printf("kernel arithmetic: ok\n");
     return f;
}
