/*
 * lmie_hip.h -- GPU-side C ABI of liblmie_amd.so (gfx950, one GPU).
 *
 * Plain C types only.  Three groups of entry points:
 *
 *  1. lmie_core_solution: the reference's own scheduler seam
 *     (src/lmie_schedule.h:18-30, called from src/lmie_interface.c:453-462),
 *     same signature.  It replaces the CPU dispatch of lmie_core_shared
 *     (src/lmie_schedule.c:391-452, the hook the reference left for an
 *     accelerator at :443-449) together with the normalisation and the
 *     expansion (src/lmie_schedule.c:1155-1241).  FP64 throughout.
 *
 *  2. Extensions with no counterpart in the reference, for precision choice,
 *     gradients and the two-mode mixture:
 *       lmie_solution_ex       lmie_solution with an options block
 *       lmie_solution_mix      several distributions on one size grid, mixed
 *                              by number weights (one GPU pass)
 *       lmie_solution_adjoint  forward pass plus one reverse-mode pass: the
 *                              gradient of a scalar loss over the outputs with
 *                              respect to every scalar input
 *     (the reference's derivatives are forward-mode only; its adjoint
 *      equations are left empty in doc/lmie_equations.tex:247-249).
 *
 *  3. Diagnostics: lmie_last_error and per-kernel timing.
 *
 * All functions return 0 on success and -1 on error (then lmie_last_error()
 * names the failing step), like the reference's API.
 */
#ifndef LMIE_HIP_H
#define LMIE_HIP_H

#include <stdio.h>

#include "lmie_interface.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the angular contraction.  FP64: every stage in double.
   MIXED: the Mie coefficients, cross sections, reductions, normalisation and
   expansion stay in double; the angular sums S1/S2 (and their adjoint) run on
   the matrix cores with fp32 accumulation, the operands split into 16-bit
   parts (forward: the size operand in two, the angle operand in three, five
   products; reverse: two and two, three products); MIXED uses fp16 parts
   (operands scaled into the fp16 range by exact powers of two), MIXED_BF16
   bf16 parts.  Forward
   tangents are always computed in FP64 (lmie_last_precision reports the
   arithmetic a call ran in). */
enum lmie_precision {
     LMIE_PRECISION_FP64       = 0,
     LMIE_PRECISION_MIXED      = 1,
     LMIE_PRECISION_MIXED_BF16 = 2
};

typedef struct {
     int precision;        /* enum lmie_precision */
     int device;           /* HIP device ordinal */
     int reserved[6];      /* zero */
} lmie_opts;

/* (1) Scheduler seam, src/lmie_schedule.h:18-30.  Returns the truncated
   coefficient count (>= 0) or -1. */
int lmie_core_solution(int n_qsiz, int n_derivs1, int n_derivs2,
                       int save_control, int *index1, int *index2,
                       double lambda, double *lambda_l,
                       double mr, double mi, double *mr_l, double *mi_l,
                       double *qx, double *qw, double *nr,
                       double r1, double r2,
                       double *cext_, double *csca_, double *pbak_,
                       double *g, double **gc, double **lc,
                       double **qx_l, double **qw_l, double **nr_l,
                       double *cext_l_, double *csca_l_, double *pbak_l_,
                       double *g_l, double ***gc_l, double ***lc_l,
                       double ***save1, double ****save2, double accuracy,
                       int verbose, int n_threads, int use_mpi);

/* Same seam with an options block (NULL = FP64 on device 0). */
int lmie_core_solution_ex(int n_qsiz, int n_derivs1, int n_derivs2,
                          int save_control, int *index1, int *index2,
                          double lambda, double *lambda_l,
                          double mr, double mi, double *mr_l, double *mi_l,
                          double *qx, double *qw, double *nr,
                          double r1, double r2,
                          double *cext_, double *csca_, double *pbak_,
                          double *g, double **gc, double **lc,
                          double **qx_l, double **qw_l, double **nr_l,
                          double *cext_l_, double *csca_l_, double *pbak_l_,
                          double *g_l, double ***gc_l, double ***lc_l,
                          double ***save1, double ****save2, double accuracy,
                          const lmie_opts *opts);

/* (2) lmie_solution with options (NULL = FP64 on device 0).  No
   counterpart in the reference. */
int lmie_solution_ex(lmie_in_data *in, lmie_out_data *out, int alloc_out,
                     const lmie_opts *opts);

/* Mixture of n_modes distributions on one grid: every mode shares
   dist grid fields (n_int1, n_int2, n_quad, r1, r2), lambda, mr, mi,
   accuracy and the calc_* flags of modes[0]; mode k contributes number
   density weight[k] * n_k(r) (each n_k normalised to unit integral).
   Outputs are those of the mixed distribution; the distribution statistics
   (reff ...) are those of the mixture.  Seeds: modes[k]->n_derivs columns
   each; the output tangent columns are the modes' columns in order (mode 0
   first), so out must be allocated for sum_k n_derivs.  Seeds of lambda,
   mr, mi are read from each mode's own columns.  No counterpart in the
   reference. */
int lmie_solution_mix(lmie_in_data *const *modes, const double *weight, int n_modes,
                      lmie_out_data *out, int alloc_out, const lmie_opts *opts);

/* Cotangents of the outputs (the loss gradient dL/d output).  NULL arrays
   are zero.  gc / lc are [6][n_coef_max] row-major (the layout of out->gc[0]). */
typedef struct {
     double cext, csca, cbak, g;
     const double *gc;
     const double *lc;
} lmie_adj_in;

/* Gradient of the loss with respect to each scalar input, in the order of
   lmie_in_data: lambda, mr, mi, a1, a2, a3, a4, a5, r1, r2.  For a mixture,
   mode k's distribution parameters a1..a5 go to grad_modes[5 k .. 5 k + 4]
   (grad_modes may be NULL for one mode). */
typedef struct {
     double lambda, mr, mi, a1, a2, a3, a4, a5, r1, r2;
} lmie_grad;

/* Reverse mode in two calls, so a caller (e.g. an autograd wrapper) can form
   the loss gradient from the forward outputs in between:
     lmie_adjoint_forward  forward pass (outputs written to out exactly as
                           lmie_solution_ex with n_derivs = 0) that keeps what
                           the reverse pass needs on the device: a_n, b_n per
                           size point and the angular sums per size and angle;
     lmie_adjoint_reverse  one reverse pass for the given output cotangents;
                           may be called again on the same context;
     lmie_adjoint_free     releases the device state.
   modes / weight / n_modes as lmie_solution_mix (n_modes = 1, weight = NULL
   for a single distribution).  The gradient covers every scalar input:
   lambda, mr, mi, the distribution parameters a1..a5, r1 and r2 (lambda, r1,
   r2 and grids set by a1, a2 move the size parameters x_i = 2 pi r_i / lambda,
   whose cotangents the reverse pass forms per size).  No counterpart in the
   reference. */
typedef struct lmie_adj_ctx lmie_adj_ctx;

int lmie_adjoint_forward(lmie_in_data *const *modes, const double *weight, int n_modes,
                         lmie_out_data *out, int alloc_out, const lmie_opts *opts,
                         lmie_adj_ctx **ctx);
int lmie_adjoint_reverse(lmie_adj_ctx *ctx, const lmie_adj_in *adj, lmie_grad *grad,
                         double *grad_modes);
void lmie_adjoint_free(lmie_adj_ctx *ctx);

/* Both in one call: cot(out, adj, user) is called between the passes to fill
   the cotangents from the forward outputs. */
typedef int (*lmie_cotangent_fn)(const lmie_out_data *out, lmie_adj_in *adj, void *user);
int lmie_solution_adjoint(lmie_in_data *const *modes, const double *weight, int n_modes,
                          lmie_out_data *out, int alloc_out, const lmie_opts *opts,
                          lmie_cotangent_fn cot, void *user, lmie_grad *grad, double *grad_modes);

/* Independent retrieval cases in one call: case c is the single distribution
   cases[c] with its own wavelength, m and size grid; lmie_solution_adjoint is
   applied to every case.  n_workers = 0: the cases run together, up to 64 at a
   time, and each kernel of the device passes is launched once for all of
   them (mixed precision with a gc or lc cotangent; other cases run one
   after another).  n_workers > 0: that many cases at a time, each on its own
   set of device streams.
   cot is called once per case, from the worker threads, with users[c] (users
   may be NULL): it must be thread-safe.  outs[c], grads[c] and
   grad_modes[5 c .. 5 c + 4] (grad_modes may be NULL) receive case c's
   outputs and gradient; status[c] (may be NULL) is 0 for a case that
   succeeded and -1 for one that failed.  With alloc_out the outputs of every
   case with status 0 are allocated (free with lmie_out_free); a failed case
   leaves nothing allocated.  Every case runs even if another fails; the
   return value is -1 if any failed, and lmie_last_error() then names the
   first failing case.  No counterpart in the reference. */
int lmie_solution_adjoint_batch(int n_case, lmie_in_data *const *cases, lmie_out_data *const *outs,
                                int alloc_out, const lmie_opts *opts, lmie_cotangent_fn cot,
                                void *const *users, lmie_grad *grads, double *grad_modes,
                                int *status, int n_workers);

/* Cases of this thread's last lmie_solution_adjoint_batch call whose forward
   (pass 0) or reverse (pass 1) ran in the single launches of the fused mode. */
int lmie_last_batch_fused(int pass);

/* A ready-made cotangent function: the weighted L2 loss of the GSF
   expansion, L = sum_kl weight_kl (gc_kl - target_kl)^2 over the [6][max_coef]
   block of out->gc[0] (the row stride is taken from out->gc).  user points to
   an lmie_gc_l2; gc_bar is caller memory of the same shape that receives
   dL/d gc, and loss receives L. */
typedef struct {
     const double *target;  /* [6][max_coef] */
     const double *weight;  /* [6][max_coef] */
     double *gc_bar;        /* [6][max_coef], written */
     double loss;           /* written */
} lmie_gc_l2;
int lmie_cotangent_gc_l2(const lmie_out_data *out, lmie_adj_in *adj, void *user);

/* Scattering files in the reference's calllmie text format (the default build
   of misc/rtutil_scat_io.c): header (misc/rtutil_scat_io.c:250-294), expansion
   coefficients (:823-889) and phase function (:636-702).  The header holds the
   inputs of in and the radius range and statistics of out, as
   utils/calllmie.c:758-760 fills it.  legendre: 0 = out->gc, 1 = out->lc;
   deriv: -1 = the values, d >= 0 = the tangents of seed d (gc_l / lc_l /
   pf_l).  The coefficient file has out->n_coef rows, the phase function
   in->n_angles. */
int lmie_scat_header_write(FILE *fp, const lmie_in_data *in, const lmie_out_data *out);
int lmie_scat_coefs_write(FILE *fp, const lmie_in_data *in, const lmie_out_data *out, int legendre, int deriv);
int lmie_phase_func_write(FILE *fp, const lmie_in_data *in, const lmie_out_data *out, int deriv);

/* (3) Text of the last error on this thread ("" if none). */
const char *lmie_last_error(void);

/* Arithmetic (enum lmie_precision) the angular contraction of this thread's
   last solution ran in.  A MIXED request runs in FP64 where the split 16-bit
   kernels do not cover it: with forward tangent channels (seeds that move the
   size parameters or m), more than 3 weight-only seed channels, or
   save_control 1 (per-size outputs); the re-weighting of save_control 2 is
   FP64.  No counterpart in the reference. */
int lmie_last_precision(void);

/* Per-kernel device time of the last call on this thread, in milliseconds,
   measured with HIP events on the library's stream when enabled; the
   LMIE_T_HOST_* and LMIE_T_CALL entries are host wall-clock times of the
   same call. */
enum lmie_timer {
     LMIE_T_TABLES = 0,    /* Gauss nodes + angular and expansion tables  */
     LMIE_T_COEFS,         /* per-size recurrences, a_n/b_n, series        */
     LMIE_T_ANGULAR,       /* angular contraction (dominant kernel)        */
     LMIE_T_REDUCE,        /* reductions over sizes                        */
     LMIE_T_EXPAND,        /* expansion into coefficients                  */
     LMIE_T_ADJ_ANGULAR,   /* adjoint angular contraction                  */
     LMIE_T_ADJ_COEFS,     /* adjoint per-size recurrences                 */
     LMIE_T_TOTAL,         /* whole device pipeline of the call            */
     LMIE_T_ADJ_GEMM,      /* reverse contraction kernel alone (mixed)     */
     LMIE_T_HOST_GRID,     /* host: size grid and number densities          */
     LMIE_T_HOST_PREP,     /* host: device-layer set-up before the first launch */
     LMIE_T_HOST_WAIT,     /* host: wait for the reverse pass's density
                              tangents (built beside the device pipeline)  */
     LMIE_T_CALL,          /* host: the whole call, entry to return         */
     LMIE_N_TIMERS
};
int lmie_set_timing(int on);
int lmie_get_timing(double *ms, int n);

#ifdef __cplusplus
}
#endif

#endif /* LMIE_HIP_H */
