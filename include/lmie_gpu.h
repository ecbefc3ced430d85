/*
 * lmie_gpu.h -- thin C ABI between the host C layer of liblmie_amd.so and its
 * gfx950 kernels.  Plain pointers and sizes; all arrays are host memory owned
 * by the caller.  Counterpart of the accelerator hook the reference left in
 * its scheduler (lmie_core_cuda, src2/lmie_cuda.h:19-30, called from
 * src/lmie_schedule.c:443-449), re-cut for this design: the kernels return
 * the raw size-integrated sums and the expansion of the raw scattering matrix;
 * the host applies the normalisation (src/lmie_schedule.c:1155-1199) to the
 * small results.
 */
#ifndef LMIE_GPU_H
#define LMIE_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
     int n_size;              /* size points N                                   */
     const double *x;         /* [N] size parameter 2 pi r / lambda              */
     const double *f;         /* [N] number density x quadrature weight          */
     const int *n1;           /* [N] per-size order n1(x_i), src/lmie_core.c:18-26 */
     const int *n2;           /* [N] per-size D_n start,   src/lmie_core.c:30-38 */
     const double *sx, *cx;   /* [N] sin x_i, cos x_i from the C library, or NULL */
     double mr, mi;           /* m = mr - i mi (src/lmie_schedule.c:1048)         */
     int n1_max;              /* order of the largest radius r2                   */
     int n_qang;              /* Gauss angles (2 n1_max), 0 = no angular work     */
     int n_coef;              /* expansion length (2 n1_max)                      */
     int calc_gc, calc_lc;
     /* Mie tangent channels: seeds that move x or m (src/lmie_interface.c:327-331) */
     int n_tan;
     const double *dx;        /* [n_tan][N]                                       */
     const double *dm;        /* [n_tan][2] (re, im) of d m                       */
     const double *df_tan;    /* [n_tan][N] d f along the channel                 */
     /* weight-only channels: seeds that move only the number density */
     int n_wgt;
     const double *df_wgt;    /* [n_wgt][N]                                       */
     int save_control;        /* 1: return per-size series and matrix elements;
                                 2: re-weight those of an earlier call instead of
                                 solving (src/lmie_schedule.c:1087-1124): only
                                 f and the weight-only channels are used, the Mie
                                 tangent channels return zero sums            */
     const double *saved1;    /* [4][N]          if save_control == 2            */
     const double *saved2;    /* [4][N][n_qang]  if save_control == 2            */
     int precision;           /* enum lmie_precision                              */
     int device;
     int keep;                /* 1: keep the device state for lmie_gpu_adjoint    */
} lmie_gpu_job;

typedef struct {
     /* raw f-weighted sums over sizes: ext, sca, bak, asym (src/lmie_core.c:537-596) */
     double sums[4];
     double *sums_ch;         /* [n_tan + n_wgt][4] their tangents               (caller) */
     /* expansion of the raw 4-element matrix, channel 0 = value, then tangents,
        each [6][n_coef] in the output order beta, alpha, zeta, delta, gamma, epsilon,
        before the (2 / C_sca) normalisation */
     double *gc_raw;          /* [1 + n_tan + n_wgt][6][n_coef] or NULL          (caller) */
     double *lc_raw;          /* same for the Legendre expansion, or NULL        (caller) */
     double *save1;           /* [4][N]          if save_control == 1            (caller) */
     double *save2;           /* [4][N][n_qang]  if save_control == 1            (caller) */
     void *state;             /* job->keep: device state for lmie_gpu_adjoint     */
     int precision;           /* arithmetic the angular contraction ran in: the
                                 job's, or LMIE_PRECISION_FP64 where the mixed
                                 kernels do not cover the job (forward tangent
                                 channels, more than 3 weight-only channels,
                                 save_control 1)                                   */
} lmie_gpu_result;

int lmie_gpu_forward(const lmie_gpu_job *job, lmie_gpu_result *res);

/* Reverse pass on the state kept by lmie_gpu_forward (job->keep = 1, value
   channel only): given the cotangents of the raw sums (4) and of the raw
   expansions (gc_raw / lc_raw of channel 0, [6][n_coef], NULL = zero), return
   per size the cotangents of f_i and (x_bar non-NULL) of the size parameter
   x_i, and the cotangent of m as (d/d mr, d/d mi) summed over sizes.  The
   state stays valid until lmie_gpu_release. */
typedef struct {
     const double *sums_bar;  /* [4]                                              */
     const double *gc_bar;    /* [6][n_coef] or NULL                              */
     const double *lc_bar;    /* [6][n_coef] or NULL                              */
} lmie_gpu_adj_in;

typedef struct {
     double *f_bar;           /* [N]                                      (caller) */
     double m_bar[2];         /* dL/d mr, dL/d mi                                  */
     double *x_bar;           /* [N] dL/d x_i, or NULL (not formed)       (caller) */
} lmie_gpu_adj_out;

int lmie_gpu_adjoint(void *state, const lmie_gpu_adj_in *in, lmie_gpu_adj_out *out);
void lmie_gpu_release(void *state);

/* Batches of independent cases, one kernel launch for all cases.  The host
   thread that runs case `slot` of a batch calls lmie_gpu_batch_enter, then
   the same lmie_gpu_forward / lmie_gpu_adjoint calls as for a single case, then
   lmie_gpu_batch_leave (also after a failure).  Those calls block until every
   case of the batch has made the same call or left; one of the threads then
   runs the pass for all of them and each returns its own case's result.
   Fused into single launches: mixed precision forwards with angles, no
   tangent channels, save_control 0 and keep = 1, and their reverse passes with
   a gc or lc cotangent; other cases run one after another inside the pass.
   No counterpart in the reference (its cases are separate calls).
     lmie_gpu_batch_fused   cases fused in pass 0 (forward) or 1 (reverse) */
void *lmie_gpu_batch_open(int n_case);
int lmie_gpu_batch_enter(void *batch, int slot);
void lmie_gpu_batch_leave(void);
int lmie_gpu_batch_fused(void *batch, int pass);
void lmie_gpu_batch_close(void *batch);

/* Scattering matrix at n_theta user angles from expansion coefficients
   (misc/rtutil_scat_int.c:510-797): coefs [n_set][6][ld], n_coef used,
   theta [n_theta] in degrees, pf [n_set][6][n_theta]; gsf = 1 for
   generalized-spherical-function coefficients, 0 for Legendre. */
int lmie_gpu_phase_func(int n_set, int n_coef, int ld, const double *coefs, int n_theta,
                        const double *theta, double *pf, int gsf, int device);

/* lmie_gpu_phase_func without any device allocation: no counterpart in the
   reference; for callers that cannot allocate device memory (code under
   stream capture, serving loops, harnesses that forbid hipMalloc).  Same
   arguments, kernel and results (bit for bit) as lmie_gpu_phase_func; the
   staging of coefficients, angles and phase function lives in a static
   device array of 2^21 doubles (16 MiB) per GPU, held until the process
   ends.  Calls whose staging, 6 n_set ld + 6 n_set n_theta + n_theta doubles
   (each part rounded up to 32), exceeds it fail with -1 before any device
   work.  The array keeps the last call's contents; calls on one GPU from
   threads of one process take turns on a lock per device (overlapping calls
   would race on the array). */
int lmie_gpu_phase_func_static(int n_set, int n_coef, int ld, const double *coefs, int n_theta,
                               const double *theta, double *pf, int gsf, int device);

/* Error text of the last failing call on this thread ("" if none). */
const char *lmie_gpu_last_error(void);

/* Host wall-clock time of entry k >= LMIE_T_HOST_GRID of enum lmie_timer
   (include/lmie_hip.h) for this thread's current call, kept when timing is on;
   lmie_gpu_host_times_clear zeroes those entries at the start of a call. */
void lmie_gpu_host_time(int k, double ms);
void lmie_gpu_host_times_clear(void);

/* Diagnostics (tests and benchmarks; no counterpart in the reference):
     lmie_gpu_gauss_rule        the Gauss-Legendre rule of the angle kernels,
                                n nodes ascending and weights (misc/rtutil_math.c:73-129)
     lmie_gpu_selftest_mfma16   one 16x16x32 16-bit matrix instruction on row-major
                                float inputs A (16x32), B (32x16) -> D (16x16)
                                (bf16 = 1: bf16 parts, else fp16)
     lmie_gpu_selftest_mfma32   one 32x32x16 fp16 matrix instruction, A (32x16), B (16x32)
                                -> D (32x32), row-major float inputs
     lmie_gpu_selftest_mfma_f64 one 16x16x4 FP64 matrix instruction, A (16x4), B (4x16)
     lmie_gpu_selftest_copy     a plain device copy of n_bytes, reps times: the
                                calibration point of the HBM traffic counters;
                                *ms = average kernel time
     lmie_gpu_gsf_table         the expansion tables of the angle kernels at n_mu
                                angles mu (any values in [-1, 1]), l = 0..n_l-1:
                                out[5][n_l][n_mu] = P^00, P^02, P^22, P^2-2, P_l
                                (misc/rtutil_math.c:867-1017, 506-513)
     lmie_gpu_pitau_table       the FP64 angle tables at n_mu angles mu, orders
                                1..n: bp[n][n_mu] = pi + tau, bm = pi - tau
                                (src/lmie_core.c:669-691) */
int lmie_gpu_gauss_rule(int n, double *x, double *wt, int device);
int lmie_gpu_gsf_table(int n_mu, const double *mu, int n_l, double *out, int device);
int lmie_gpu_pitau_table(int n_mu, const double *mu, int n, double *bp, double *bm, int device);
int lmie_gpu_selftest_mfma16(int bf16, const float *A, const float *B, float *D, int device);
int lmie_gpu_selftest_mfma32(const float *A, const float *B, float *D, int device);
int lmie_gpu_selftest_mfma_f64(const double *A, const double *B, double *D, int device);
int lmie_gpu_selftest_copy(size_t n_bytes, int reps, int device, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* LMIE_GPU_H */
