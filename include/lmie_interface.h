/*
 * lmie_interface.h -- drop-in C ABI of the size-integrated Mie solver.
 *
 * Same type names, field order, field types and function signatures as the
 * reference interface (src/lmie_interface.h:20-157 and the distribution enum
 * misc/rtutil_size_dist.h:22-34), so a program written against the reference
 * recompiles and relinks against liblmie_amd.so unchanged.  The arrays in
 * lmie_out_data are pointer-table blocks: one allocation holding the row
 * pointers followed by the contiguous row-major data, as the reference's
 * allocator lays them out (misc/gutil_alloc.c:19-75), so out->gc[0] points at
 * 6*max_coef contiguous doubles and out->gc_l[0][0] at n_derivs*6*max_coef.
 *
 * Work placement: the distribution grid, its statistics and the derivative
 * split run on the host in C; the per-size Mie coefficients, the angular
 * sums, their reduction over sizes and the phase-matrix expansion run on the
 * GPU (see lmie_hip.h for the precision and adjoint extensions).
 */
#ifndef LMIE_INTERFACE_H
#define LMIE_INTERFACE_H

/* Platform identifiers that the reference's headers give its callers
   (misc/gutil.h:60-77, reached through src/lmie_interface.h:13), for code
   built with the reference's -DPLATFORM=... flag (common.inc:5) that compares
   against them; same values. */
#ifndef LINUX_GCC
#define AIX_CC       1
#define AIX_GCC      2
#define CYGWIN_GCC   3
#define DEC_UNIX_CC  4
#define DEC_UNIX_GCC 5
#define HP_CC        6
#define HP_GCC       7
#define IRIX_CC      8
#define IRIX_GCC     9
#define LINUX_GCC    10
#define LINUX_ICC    11
#define LINUX_PGCC   12
#define MACOSX_GCC   13
#define MACOSX_XLC   14
#define SOLARIS_CC   15
#define SOLARIS_GCC  16
#define WIN32_MSVC   17
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Size distribution kinds; numbering as misc/rtutil_size_dist.h:22-34. */
enum size_dist_type {
     SIZE_DIST_NONE,
     SIZE_DIST_MODIFIED_GAMMA,
     SIZE_DIST_LOG_NORMAL,
     SIZE_DIST_POWER_LAW,
     SIZE_DIST_GAMMA,
     SIZE_DIST_MODIFIED_POWER_LAW,
     SIZE_DIST_MODIFIED_BIMODAL_LOG_NORMAL,
     SIZE_DIST_EXPONENTIAL,
     SIZE_DIST_MONO,

     N_SIZE_DIST_TYPES
};

/* Inputs; layout as src/lmie_interface.h:20-58. */
typedef struct {
     int save_control;

     int calc_gc;
     int calc_lc;
     int calc_pf;

     enum size_dist_type dist_type;

     int n_int1;
     int n_int2;
     int n_quad;
     int n_angles;
     int n_derivs;

     double lambda;
     double mr;
     double mi;
     double a1;
     double a2;
     double a3;
     double a4;
     double a5;
     double r1;
     double r2;

     double *lambda_l;
     double *mr_l;
     double *mi_l;
     double *a1_l;
     double *a2_l;
     double *a3_l;
     double *a4_l;
     double *a5_l;
     double *r1_l;
     double *r2_l;

     double accuracy;
} lmie_in_data;

/* Outputs; layout as src/lmie_interface.h:61-101. */
typedef struct {
     int n_coef;

     double r1;
     double r2;
     double norm;
     double reff;
     double veff;
     double gavg;
     double vavg;
     double ravg;
     double rvw;
     double cext;
     double csca;
     double cbak;
     double g;
     double **gc;
     double **lc;
     double *theta;
     double **pf;

     double *r1_l;
     double *r2_l;
     double *norm_l;
     double *reff_l;
     double *veff_l;
     double *gavg_l;
     double *vavg_l;
     double *ravg_l;
     double *rvw_l;
     double *cext_l;
     double *csca_l;
     double *cbak_l;
     double *g_l;
     double ***gc_l;
     double ***lc_l;
     double ***pf_l;

     double **save1;
     double ***save2;
} lmie_out_data;

/* 2 * n1(2 pi r2 / lambda); src/lmie_interface.c:24-31. */
int lmie_calc_max_coef(double lambda, enum size_dist_type dist_type,
                       double a1, double a2, double a3, double r1, double r2);

/* Seed-vector allocation for n_derivs linearised inputs; src/lmie_interface.c:38-72. */
int lmie_in_alloc(lmie_in_data *in, int n_derivs);
void lmie_in_free(lmie_in_data *in, int flag);

/* Output allocation sized by max_coef and in->n_angles; src/lmie_interface.c:79-169.
   As in the reference, save1/save2 are not released by lmie_out_free. */
int lmie_out_alloc(lmie_in_data *in, lmie_out_data *out, int max_coef);
void lmie_out_free(lmie_out_data *out, int flag);

/* Zero every seed vector; src/lmie_interface.c:176-192. */
void lmie_in_zero_derivs(lmie_in_data *in, int n_derivs);

/* The solution call; src/lmie_interface.c:199-250.  Returns 0 or -1.
   n_threads sizes the host-side work only; use_mpi != 0 is rejected (one GPU). */
int lmie_solution(lmie_in_data *in, lmie_out_data *out, int alloc_out, int verbose,
                  int n_threads, int use_mpi);

/* Flat-argument forms; src/lmie_interface.c:257-297 and :360-507. */
int lmie_solution2(int calc_gc, int calc_lc, int calc_pf,
                   enum size_dist_type dist_type,
                   int n_int1, int n_int2,
                   int n_quad, int n_angles, int save_control,
                   double lambda, double mr, double mi,
                   double a1, double a2, double a3, double a4, double a5,
                   double r1, double r2,
                   double accuracy, int *n_coef,
                   double *r21, double *r22,
                   double *norm, double *reff, double *veff,
                   double *gavg, double *vavg, double *ravg, double *rvw,
                   double *cext, double *csca, double *cbak, double *g,
                   double **gc, double **lc, double *theta, double **pf,
                   double ***save1, double ****save2,
                   int verbose, int n_threads, int use_mpi);

int lmie_solution2_l(int calc_gc, int calc_lc, int calc_pf,
                     enum size_dist_type dist_type,
                     int n_int1, int n_int2, int n_quad,
                     int n_angles, int n_derivs, int save_control,
                     double lambda, double mr, double mi,
                     double a1, double a2, double a3, double a4, double a5,
                     double r1, double r2,
                     double *lambda_l, double *mr_l, double *mi_l,
                     double *a1_l, double *a2_l, double *a3_l, double *a4_l, double *a5_l,
                     double *r1_l, double *r2_l,
                     double accuracy, int *n_coef,
                     double *r21, double *r22,
                     double *norm, double *reff, double *veff,
                     double *gavg, double *vavg, double *ravg, double *rvw,
                     double *cext, double *csca, double *cbak, double *g,
                     double **gc, double **lc, double *theta, double **pf,
                     double *r21_l, double *r22_l,
                     double *norm_l, double *reff_l, double *veff_l,
                     double *gavg_l, double *vavg_l, double *ravg_l, double *rvw_l,
                     double *cext_l, double *csca_l, double *cbak_l, double *g_l,
                     double ***gc_l, double ***lc_l, double ***pf_l,
                     double ***save1, double ****save2,
                     int verbose, int n_threads, int use_mpi);

/* MPI worker entry (src/lmie_interface.c:516-524).  There is no MPI path on
   one GPU: always returns -1. */
int lmie_solution_slave(void);

#ifdef __cplusplus
}
#endif

#endif /* LMIE_INTERFACE_H */
