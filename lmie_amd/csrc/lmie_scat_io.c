/*
 * lmie_scat_io.c -- scattering files in the text format of the reference's
 * calllmie (the default build of misc/rtutil_scat_io.c): the header of
 * "name = value" lines (misc/rtutil_scat_io.c:250-294), expansion coefficients
 * (:823-889) and the phase function (:636-702).  Plain C on the host; errors
 * go to stderr and, inside the product library, to lmie_last_error().
 */
#include <stdio.h>

#include "../../include/lmie_interface.h"
#include "../../include/lmie_hip.h"

/* the host layer's error record (lmie_host.c); absent where these writers are
   linked without it (the command line program's CPU check builds) */
void lmie_host_error(const char *where, const char *why) __attribute__((weak));

static int fail(const char *where, const char *why) {
     if (lmie_host_error)
          lmie_host_error(where, why);
     else
          fprintf(stderr, "ERROR: %s: %s\n", where, why);
     return -1;
}

/* Scattering files in the text format the reference's calllmie writes
   (misc/rtutil_scat_io.c:250-294, 636-702, 823-889, default build): a header
   of "name = value" lines, the row count, then one row per coefficient or
   angle with the six elements.  The header carries the inputs and the
   solution's radius range and distribution statistics
   (utils/calllmie.c:758-760). */
static const char *const scat_dist_names[] = {
     "none", "modified_gamma", "log_normal", "power_law", "gamma",
     "modified_power_law", "modified_bimodal_log_normal", "exponential", "mono"};

int lmie_scat_header_write(FILE *fp, const lmie_in_data *in, const lmie_out_data *out) {
     if (!fp || !in || !out) return fail("lmie_scat_header_write", "missing argument");
     if ((int) in->dist_type < 0 || (int) in->dist_type >= (int) (sizeof scat_dist_names / sizeof *scat_dist_names))
          return fail("lmie_scat_header_write", "invalid size distribution type");
     const struct { const char *key; double v; } pre[] = {
          {"lambda   ", in->lambda}, {"mr       ", in->mr}, {"mi       ", in->mi}};
     const struct { const char *key; double v; } mid[] = {
          {"a1       ", in->a1}, {"a2       ", in->a2}, {"a3       ", in->a3}, {"a4       ", in->a4},
          {"a5       ", in->a5}, {"r1       ", out->r1}, {"r2       ", out->r2}};
     const struct { const char *key; double v; } post[] = {
          {"norm     ", out->norm}, {"reff     ", out->reff}, {"veff     ", out->veff},
          {"gavg     ", out->gavg}, {"vavg     ", out->vavg}, {"ravg     ", out->ravg},
          {"rvw      ", out->rvw}, {"cext     ", out->cext}, {"csca     ", out->csca},
          {"cbak     ", out->cbak}, {"g        ", out->g}};
     int bad = 0;
     for (size_t k = 0; k < 3; ++k) bad |= fprintf(fp, "%s = %e\n", pre[k].key, pre[k].v) < 1;
     bad |= fprintf(fp, "dist_type = %s\n", scat_dist_names[in->dist_type]) < 1;
     for (size_t k = 0; k < 7; ++k) bad |= fprintf(fp, "%s = %e\n", mid[k].key, mid[k].v) < 1;
     bad |= fprintf(fp, "n_int1    = %d\nn_int2    = %d\nn_quad    = %d\n", in->n_int1, in->n_int2, in->n_quad) < 1;
     for (size_t k = 0; k < 11; ++k) bad |= fprintf(fp, "%s = %e\n", post[k].key, post[k].v) < 1;
     return bad ? fail("lmie_scat_header_write", "write failed") : 0;
}

/* rows [6][ld] of n rows (coefficients) or, with x non-NULL, one leading
   column x (angles) */
static int scat_rows_write(FILE *fp, const char *who, const double *x, double *const *rows, int n) {
     if (fprintf(fp, "%d\n", n) < 1) return fail(who, "write failed");
     for (int i = 0; i < n; ++i) {
          if (x && fprintf(fp, "%e ", x[i]) < 1) return fail(who, "write failed");
          if (fprintf(fp, "%e %e %e %e %e %e\n", rows[0][i], rows[1][i], rows[2][i], rows[3][i], rows[4][i],
                      rows[5][i]) < 6)
               return fail(who, "write failed");
     }
     return 0;
}

int lmie_scat_coefs_write(FILE *fp, const lmie_in_data *in, const lmie_out_data *out, int legendre, int deriv) {
     if (!fp || !in || !out) return fail("lmie_scat_coefs_write", "missing argument");
     if (deriv >= in->n_derivs) return fail("lmie_scat_coefs_write", "no such derivative");
     double ***cl = legendre ? out->lc_l : out->gc_l, **c0 = legendre ? out->lc : out->gc;
     double **c = deriv < 0 ? c0 : c0 && cl ? cl[deriv] : NULL;
     if (!c) return fail("lmie_scat_coefs_write", legendre ? "no Legendre coefficients" : "no expansion coefficients");
     if (lmie_scat_header_write(fp, in, out)) return -1;
     return scat_rows_write(fp, "lmie_scat_coefs_write", NULL, c, out->n_coef);
}

int lmie_phase_func_write(FILE *fp, const lmie_in_data *in, const lmie_out_data *out, int deriv) {
     if (!fp || !in || !out) return fail("lmie_phase_func_write", "missing argument");
     if (deriv >= in->n_derivs) return fail("lmie_phase_func_write", "no such derivative");
     double **p = deriv < 0 ? out->pf : out->pf && out->pf_l ? out->pf_l[deriv] : NULL;
     if (!p || !out->theta) return fail("lmie_phase_func_write", "no phase function");
     if (lmie_scat_header_write(fp, in, out)) return -1;
     return scat_rows_write(fp, "lmie_phase_func_write", out->theta, p, in->n_angles);
}
