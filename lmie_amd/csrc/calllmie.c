/*
 * calllmie.c -- command line front end of liblmie_amd.so with the options and
 * output files of the reference's calllmie (utils/calllmie.c:279-966): one
 * size-integrated solution from the command line, written as the reference's
 * scattering text files (header, expansion coefficients, phase function) and
 * their derivative files.
 *
 *   calllmie -dist <name> <values> -lambda <l> -m <mr> <mi> -n_int2 <n>
 *            -n_quad <n> [-n_int1 <n>] [-n_angles <n>] [-accuracy <a>]
 *            [-derivs <n> [-lambda_l <list>] [-m_l <mr list> <mi list>]
 *             [-dist_l <lists>]]
 *            [-total <file> [<list>]] [-gc <file> [<list>]] [-lc <file> [<list>]]
 *            [-pf <file> [<list>]] [-check_cbak <tol>] [-check_g <tol>]
 *            [-save_dist] [-timing <n>] [-n_threads <n>] [-verbose]
 *            [-precision fp64|mixed] [-help] [-version]
 *
 * <list> is comma separated, one entry per derivative (a file name, or a seed
 * value); a file name "-" is standard output.  Distribution values follow the
 * reference's order per kind (misc/rtutil_support.c:580-668): mono a1;
 * modified_gamma a1 a2 a3 r1 r2; log_normal a1 a2 r1 r2; power_law a1 a2;
 * gamma a1 a2 r1 r2; modified_power_law a1 r1 r2;
 * modified_bimodal_log_normal a1 a2 a3 a4 a5 r1 r2; exponential a1 r1 r2.
 * -dist_l takes one list per value of the kind, in the same order.
 *
 * Differences from the reference: -timing prints wall-clock seconds of the
 * repeated solutions (the reference prints the process CPU time of clock(),
 * which sums over its threads); -n_threads only sizes host-side work;
 * -precision selects the arithmetic of the angular contraction (default
 * fp64, the reference's); -findif is not supported (the reference leaves it
 * unimplemented).  Errors print "ERROR: ..." and exit with status 1.
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/lmie_hip.h"
#include "../../include/lmie_interface.h"

#define NAME "calllmie"

/* the options block is an extension of the product library; a build against
   another implementation of the reference interface has only lmie_solution */
#pragma weak lmie_solution_ex

static int solve(lmie_in_data *in, lmie_out_data *out, int alloc_out, const lmie_opts *opts) {
     if (opts->precision == LMIE_PRECISION_FP64 || !lmie_solution_ex)
          return opts->precision == LMIE_PRECISION_FP64 ? lmie_solution(in, out, alloc_out, 0, 1, 0) : -1;
     return lmie_solution_ex(in, out, alloc_out, opts);
}

static void die(const char *fmt, const char *a) {
     fprintf(stderr, "ERROR: ");
     fprintf(stderr, fmt, a);
     fprintf(stderr, "\n");
     exit(1);
}

static void need(int i, int argc, int n, const char *opt) {
     if (i + n >= argc) die("not enough arguments for %s", opt);
}

static double num(const char *s, const char *opt) {
     char *end;
     errno = 0;
     double v = strtod(s, &end);
     if (errno || end == s || *end) {
          fprintf(stderr, "ERROR: invalid value for %s: %s\n", opt, s);
          exit(1);
     }
     return v;
}

static int inum(const char *s, const char *opt) {
     char *end;
     errno = 0;
     long v = strtol(s, &end, 10);
     if (errno || end == s || *end) {
          fprintf(stderr, "ERROR: invalid value for %s: %s\n", opt, s);
          exit(1);
     }
     return (int) v;
}

/* entries of a comma/space separated list (n_derivs of them) */
static char **list_s(int n_derivs, const char *opt, const char *list) {
     if (n_derivs <= 0) die("Must specify -derivs <n_derivs> option before %s", opt);
     char *tmp = strdup(list), *save = NULL;
     char **v = calloc((size_t) n_derivs, sizeof *v);
     int n = 0;
     for (char *t = strtok_r(tmp, " \n,", &save); t; t = strtok_r(NULL, " \n,", &save)) {
          if (n < n_derivs) v[n] = strdup(t);
          ++n;
     }
     free(tmp);
     if (n != n_derivs) {
          fprintf(stderr, "ERROR: Number of values in %s list is not consistent with the specified "
                  "number of derivs: %d\n", opt, n_derivs);
          exit(1);
     }
     return v;
}

static double *list_d(int n_derivs, const char *opt, const char *list) {
     char **s = list_s(n_derivs, opt, list);
     double *v = calloc((size_t) n_derivs, sizeof *v);
     for (int k = 0; k < n_derivs; ++k) {
          v[k] = num(s[k], opt);
          free(s[k]);
     }
     free(s);
     return v;
}

static void free_list(char **v, int n) {
     if (!v) return;
     for (int k = 0; k < n; ++k) free(v[k]);
     free(v);
}

static const char *const dist_names[] = {"none", "modified_gamma", "log_normal", "power_law", "gamma",
                                         "modified_power_law", "modified_bimodal_log_normal",
                                         "exponential", "mono"};

/* values each distribution kind takes on the command line, by slot:
   0..4 = a1..a5, 5 = r1, 6 = r2 (misc/rtutil_support.c:580-668) */
static int dist_slots(int kind, int slot[7]) {
     static const int tab[9][8] = {
          {0},
          {5, 0, 1, 2, 5, 6},                  /* modified_gamma */
          {4, 0, 1, 5, 6},                     /* log_normal */
          {2, 0, 1},                           /* power_law */
          {4, 0, 1, 5, 6},                     /* gamma */
          {3, 0, 5, 6},                        /* modified_power_law */
          {7, 0, 1, 2, 3, 4, 5, 6},            /* modified_bimodal_log_normal */
          {3, 0, 5, 6},                        /* exponential */
          {1, 0},                              /* mono */
     };
     if (kind < 1 || kind > 8) return -1;
     for (int k = 0; k < tab[kind][0]; ++k) slot[k] = tab[kind][1 + k];
     return tab[kind][0];
}

static void usage(void) {
     printf("usage: %s -dist <name> <values> -lambda <wavelength> -m <mr> <mi> -n_int2 <n> -n_quad <n>\n"
            "       [-n_int1 <n>] [-n_angles <n>] [-accuracy <a>] [-derivs <n> [-lambda_l <list>]\n"
            "       [-m_l <mr list> <mi list>] [-dist_l <lists>]] [-total <file> [<list>]]\n"
            "       [-gc <file> [<list>]] [-lc <file> [<list>]] [-pf <file> [<list>]]\n"
            "       [-check_cbak <tol>] [-check_g <tol>] [-save_dist] [-timing <n>]\n"
            "       [-n_threads <n>] [-precision fp64|mixed] [-verbose] [-help] [-version]\n", NAME);
}

static FILE *open_out(const char *fn) {
     if (fn[0] == '-') return stdout;
     FILE *fp = fopen(fn, "w");
     if (!fp) {
          fprintf(stderr, "ERROR: Error opening file for writing: %s ... %s\n", fn, strerror(errno));
          exit(1);
     }
     return fp;
}

static void close_out(FILE *fp, const char *fn) {
     if (fn[0] != '-') fclose(fp);
}

/* the four output files of one set (values, or the tangents of one seed: the
   header then carries the seeds and the tangents of the statistics, as the
   reference's derivative files, utils/calllmie.c:855-957) */
static void write_set(const lmie_in_data *in, const lmie_out_data *out, const char *fn_total,
                      const char *fn_gc, const char *fn_lc, const char *fn_pf) {
     FILE *fp;
     if (fn_total) {
          fp = open_out(fn_total);
          if (lmie_scat_header_write(fp, in, out)) die("%s", "lmie_scat_header_write()");
          close_out(fp, fn_total);
     }
     if (fn_gc) {
          fp = open_out(fn_gc);
          if (lmie_scat_coefs_write(fp, in, out, 0, -1)) die("%s", "lmie_scat_coefs_write()");
          close_out(fp, fn_gc);
     }
     if (fn_lc) {
          fp = open_out(fn_lc);
          if (lmie_scat_coefs_write(fp, in, out, 1, -1)) die("%s", "lmie_scat_coefs_write()");
          close_out(fp, fn_lc);
     }
     if (fn_pf) {
          fp = open_out(fn_pf);
          if (lmie_phase_func_write(fp, in, out, -1)) die("%s", "lmie_phase_func_write()");
          close_out(fp, fn_pf);
     }
}

int main(int argc, char *argv[]) {
     int o_dist = 0, o_lambda = 0, o_m = 0, o_n_int1 = 0, o_n_int2 = 0, o_n_quad = 0, o_derivs = 0;
     int o_check_cbak = 0, o_check_g = 0, o_save_dist = 0, o_timing = 0, o_verbose = 0;
     const char *fn_total = NULL, *fn_gc = NULL, *fn_lc = NULL, *fn_pf = NULL;
     char **fn_total_l = NULL, **fn_gc_l = NULL, **fn_lc_l = NULL, **fn_pf_l = NULL;
     int dist_type = 0, n_int1 = 0, n_int2 = 0, n_quad = 0, n_angles = 18001, n_derivs = 0;
     int n_timing = 1, n_threads = 1, precision = LMIE_PRECISION_FP64;
     double check_tol_cbak = 1e-5, check_tol_g = 1e-5, accuracy = 1e-7;
     double lambda = 0., mr = 0., mi = 0., a[7] = {0., 0., 0., 0., 0., 0., 0.};
     double *lambda_l = NULL, *mr_l = NULL, *mi_l = NULL, *a_l[7] = {NULL};

     for (int i = 1; i < argc; ++i) {
          const char *opt = argv[i];
          if (opt[0] != '-') {
               printf("Too many arguments, use -help for more information\n");
               return 1;
          }
          if (!strcmp(opt, "-accuracy")) {
               need(i, argc, 1, opt);
               accuracy = num(argv[++i], opt);
          } else if (!strcmp(opt, "-check_cbak")) {
               need(i, argc, 1, opt);
               o_check_cbak = 1;
               check_tol_cbak = num(argv[++i], opt);
          } else if (!strcmp(opt, "-check_g")) {
               need(i, argc, 1, opt);
               o_check_g = 1;
               check_tol_g = num(argv[++i], opt);
          } else if (!strcmp(opt, "-derivs")) {
               need(i, argc, 1, opt);
               o_derivs = 1;
               n_derivs = inum(argv[++i], opt);
          } else if (!strcmp(opt, "-findif")) {
               die("%s is not supported", opt);
          } else if (!strcmp(opt, "-dist")) {
               need(i, argc, 1, opt);
               const char *name = argv[++i];
               dist_type = -1;
               for (int k = 1; k < 9; ++k)
                    if (!strcmp(name, dist_names[k])) dist_type = k;
               if (dist_type < 0) die("(-dist) invalid distribution name: %s", name);
               int slot[7], n = dist_slots(dist_type, slot);
               char label[64];
               snprintf(label, sizeof label, "-dist %s", name);
               need(i, argc, n, label);
               for (int k = 0; k < n; ++k) a[slot[k]] = num(argv[++i], label);
               if (dist_type == SIZE_DIST_MONO) a[5] = a[6] = a[0];
               o_dist = 1;
          } else if (!strcmp(opt, "-dist_l")) {
               if (!o_dist) die("%s", "must specify -dist before -dist_l");
               int slot[7], n = dist_slots(dist_type, slot);
               need(i, argc, n, opt);
               for (int k = 0; k < n; ++k) {
                    free(a_l[slot[k]]);
                    a_l[slot[k]] = list_d(n_derivs, opt, argv[++i]);
               }
          } else if (!strcmp(opt, "-total") || !strcmp(opt, "-gc") || !strcmp(opt, "-lc") || !strcmp(opt, "-pf")) {
               const int nd = o_derivs ? 1 : 0;
               need(i, argc, 1 + nd, opt);
               const char *fn = argv[++i];
               char **fl = nd ? list_s(n_derivs, opt, argv[++i]) : NULL;
               if (!strcmp(opt, "-total")) { fn_total = fn; free_list(fn_total_l, n_derivs); fn_total_l = fl; }
               else if (!strcmp(opt, "-gc")) { fn_gc = fn; free_list(fn_gc_l, n_derivs); fn_gc_l = fl; }
               else if (!strcmp(opt, "-lc")) { fn_lc = fn; free_list(fn_lc_l, n_derivs); fn_lc_l = fl; }
               else { fn_pf = fn; free_list(fn_pf_l, n_derivs); fn_pf_l = fl; }
          } else if (!strcmp(opt, "-help")) {
               usage();
               return 0;
          } else if (!strcmp(opt, "-lambda")) {
               need(i, argc, 1, opt);
               o_lambda = 1;
               lambda = num(argv[++i], opt);
          } else if (!strcmp(opt, "-lambda_l")) {
               need(i, argc, 1, opt);
               free(lambda_l);
               lambda_l = list_d(n_derivs, opt, argv[++i]);
          } else if (!strcmp(opt, "-m")) {
               need(i, argc, 2, opt);
               o_m = 1;
               mr = num(argv[++i], opt);
               mi = num(argv[++i], opt);
          } else if (!strcmp(opt, "-m_l")) {
               need(i, argc, 2, opt);
               free(mr_l);
               free(mi_l);
               mr_l = list_d(n_derivs, opt, argv[++i]);
               mi_l = list_d(n_derivs, opt, argv[++i]);
          } else if (!strcmp(opt, "-n_int1")) {
               need(i, argc, 1, opt);
               o_n_int1 = 1;
               n_int1 = inum(argv[++i], opt);
          } else if (!strcmp(opt, "-n_int2")) {
               need(i, argc, 1, opt);
               o_n_int2 = 1;
               n_int2 = inum(argv[++i], opt);
          } else if (!strcmp(opt, "-n_quad")) {
               need(i, argc, 1, opt);
               o_n_quad = 1;
               n_quad = inum(argv[++i], opt);
          } else if (!strcmp(opt, "-n_angles")) {
               need(i, argc, 1, opt);
               n_angles = inum(argv[++i], opt);
          } else if (!strcmp(opt, "-n_threads")) {
               need(i, argc, 1, opt);
               n_threads = inum(argv[++i], opt);
          } else if (!strcmp(opt, "-timing")) {
               need(i, argc, 1, opt);
               o_timing = 1;
               n_timing = inum(argv[++i], opt);
          } else if (!strcmp(opt, "-precision")) {
               need(i, argc, 1, opt);
               const char *p = argv[++i];
               if (!strcmp(p, "fp64")) precision = LMIE_PRECISION_FP64;
               else if (!strcmp(p, "mixed")) precision = LMIE_PRECISION_MIXED;
               else die("invalid value for -precision: %s", p);
          } else if (!strcmp(opt, "-save_dist")) {
               o_save_dist = 1;
          } else if (!strcmp(opt, "-verbose")) {
               o_verbose = 1;
          } else if (!strcmp(opt, "-version")) {
               printf("%s, version 0.91 (liblmie_amd)\n", NAME);
               return 0;
          } else {
               printf("Invalid option: %s, use -help for more information\n", opt);
               return 1;
          }
     }
     (void) o_verbose;
     (void) n_threads;
     if (!o_dist) die("%s", "must specify a size distribution with -dist");
     if (!o_lambda) die("%s", "must specify the wavelength with -lambda");
     if (!o_m) die("%s", "must specify the index of refraction with -m");
     if (dist_type == SIZE_DIST_MODIFIED_POWER_LAW && !o_n_int1)
          die("%s", "must specify the # of integration subintervals from r0 to r1 with -np (-n01)");
     if (!o_n_int2) die("%s", "must specify the # of integration subintervals from r1 to r2 with -n (-n12)");
     if (!o_n_quad) die("%s", "must specify the # of guassian division points on each interval with -nk");
     if (o_check_cbak && !fn_pf) die("%s", "must use -pf when using -check_cbak");
     if (o_check_g && !fn_pf) die("%s", "must use -gc or -lc when using -check_g");

     lmie_in_data in;
     lmie_out_data out, out2;
     memset(&in, 0, sizeof in);
     memset(&out, 0, sizeof out);
     if (lmie_in_alloc(&in, n_derivs)) die("%s", "lmie_in_alloc()");
     in.save_control = o_save_dist ? 1 : 0;
     in.calc_gc = fn_gc || fn_pf;
     in.calc_lc = fn_lc != NULL;
     in.calc_pf = fn_pf != NULL;
     in.dist_type = dist_type;
     in.n_int1 = n_int1;
     in.n_int2 = n_int2;
     in.n_quad = n_quad;
     in.n_angles = n_angles;
     in.n_derivs = n_derivs;
     in.lambda = lambda;
     in.mr = mr;
     in.mi = mi;
     in.a1 = a[0]; in.a2 = a[1]; in.a3 = a[2]; in.a4 = a[3]; in.a5 = a[4];
     in.r1 = a[5]; in.r2 = a[6];
     for (int d = 0; d < n_derivs; ++d) {
          in.lambda_l[d] = lambda_l ? lambda_l[d] : 0.;
          in.mr_l[d] = mr_l ? mr_l[d] : 0.;
          in.mi_l[d] = mi_l ? mi_l[d] : 0.;
          double *dst[7] = {in.a1_l, in.a2_l, in.a3_l, in.a4_l, in.a5_l, in.r1_l, in.r2_l};
          for (int q = 0; q < 7; ++q) dst[q][d] = a_l[q] ? a_l[q][d] : 0.;
     }
     in.accuracy = accuracy;
     lmie_opts opts;
     memset(&opts, 0, sizeof opts);
     opts.precision = precision;
     if (solve(&in, &out, 1, &opts)) die("%s", "lmie_solution()");

     if (o_save_dist) {
          /* the per-size arrays of the first call, re-weighted by a second
             call into outputs reset to -999 (utils/calllmie.c:604-664) */
          in.save_control = 2;
          const int max_coef = lmie_calc_max_coef(in.lambda, in.dist_type, in.a1, in.a2, in.a3, in.r1, in.r2);
          out.r1 = out.r2 = out.norm = out.reff = out.veff = out.gavg = out.vavg = out.ravg = out.rvw = -999.;
          out.cext = out.csca = out.cbak = out.g = -999.;
          for (int d = 0; d < n_derivs; ++d) {
               double *v[13] = {out.r1_l, out.r2_l, out.norm_l, out.reff_l, out.veff_l, out.gavg_l, out.vavg_l,
                                out.ravg_l, out.rvw_l, out.cext_l, out.csca_l, out.cbak_l, out.g_l};
               for (int q = 0; q < 13; ++q) v[q][d] = -999.;
          }
          for (int k = 0; k < 6; ++k)
               for (int l = 0; l < max_coef; ++l) {
                    if (in.calc_gc) {
                         out.gc[k][l] = -999.;
                         for (int d = 0; d < n_derivs; ++d) out.gc_l[d][k][l] = -999.;
                    }
                    if (in.calc_lc) {
                         out.lc[k][l] = -999.;
                         for (int d = 0; d < n_derivs; ++d) out.lc_l[d][k][l] = -999.;
                    }
               }
          if (in.calc_pf)
               for (int j = 0; j < n_angles; ++j) {
                    out.theta[j] = -999.;
                    for (int k = 0; k < 6; ++k) {
                         out.pf[k][j] = -999.;
                         for (int d = 0; d < n_derivs; ++d) out.pf_l[d][k][j] = -999.;
                    }
               }
          if (solve(&in, &out, 0, &opts)) die("%s", "lmie_solution()");
     }

     double seconds = 0.;
     if (o_timing) {
          struct timespec t0, t1;
          clock_gettime(CLOCK_MONOTONIC, &t0);
          for (int k = 0; k < n_timing; ++k) {
               memset(&out2, 0, sizeof out2);
               if (solve(&in, &out2, 1, &opts)) die("%s", "lmie_solution()");
               lmie_out_free(&out2, n_derivs > 0);
               free(out2.save1);
               free(out2.save2);
          }
          clock_gettime(CLOCK_MONOTONIC, &t1);
          seconds = (double) (t1.tv_sec - t0.tv_sec) + 1e-9 * (double) (t1.tv_nsec - t0.tv_nsec);
     }

     if (o_check_cbak) {
          const double cbak2 = out.pf[0][n_angles - 1] * out.csca / (4. * M_PI);
          if (fabs(out.cbak - cbak2) > check_tol_cbak) {
               fprintf(stderr, "ERROR: check_cbak failed: cbak = %.6e and pf[0][n_angles - 1] * csca / (4. * PI) = "
                       "%.6e\n", out.cbak, cbak2);
               return 1;
          }
     }
     if (o_check_g) {
          const double g2 = (fn_gc || fn_pf) ? out.gc[0][1] / 3. : out.lc[0][1] / 3.;
          if (fabs(out.g - g2) > check_tol_g) {
               fprintf(stderr, "ERROR: check_g failed: g = %.6e and gc[0][1] / 3. = %.6e\n", out.g, g2);
               return 1;
          }
     }

     write_set(&in, &out, fn_total, fn_gc, fn_lc, fn_pf);
     for (int d = 0; d < n_derivs; ++d) {
          lmie_in_data ind = in;
          lmie_out_data od = out;
          ind.lambda = in.lambda_l[d]; ind.mr = in.mr_l[d]; ind.mi = in.mi_l[d];
          ind.a1 = in.a1_l[d]; ind.a2 = in.a2_l[d]; ind.a3 = in.a3_l[d]; ind.a4 = in.a4_l[d]; ind.a5 = in.a5_l[d];
          ind.n_derivs = 0;
          od.r1 = out.r1_l[d]; od.r2 = out.r2_l[d]; od.norm = out.norm_l[d]; od.reff = out.reff_l[d];
          od.veff = out.veff_l[d]; od.gavg = out.gavg_l[d]; od.vavg = out.vavg_l[d]; od.ravg = out.ravg_l[d];
          od.rvw = out.rvw_l[d]; od.cext = out.cext_l[d]; od.csca = out.csca_l[d]; od.cbak = out.cbak_l[d];
          od.g = out.g_l[d];
          od.gc = out.gc_l ? out.gc_l[d] : NULL;
          od.lc = out.lc_l ? out.lc_l[d] : NULL;
          od.pf = out.pf_l ? out.pf_l[d] : NULL;
          write_set(&ind, &od, fn_total_l ? fn_total_l[d] : NULL, fn_gc_l ? fn_gc_l[d] : NULL,
                    fn_lc_l ? fn_lc_l[d] : NULL, fn_pf_l ? fn_pf_l[d] : NULL);
     }
     if (o_timing) printf("%.2f\n", seconds);

     lmie_out_free(&out, n_derivs > 0);
     free(out.save1);
     free(out.save2);
     lmie_in_free(&in, n_derivs > 0);
     free_list(fn_total_l, n_derivs);
     free_list(fn_gc_l, n_derivs);
     free_list(fn_lc_l, n_derivs);
     free_list(fn_pf_l, n_derivs);
     free(lambda_l);
     free(mr_l);
     free(mi_l);
     for (int q = 0; q < 7; ++q) free(a_l[q]);
     return 0;
}
