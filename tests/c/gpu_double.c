/*
 * gpu_double.c -- a test double of the device layer (include/lmie_gpu.h) for
 * host-only builds of lmie_amd/csrc/lmie_host.c: the host C layer linked
 * against it runs without a GPU, under AddressSanitizer and
 * UndefinedBehaviorSanitizer, and device failures can be forced from the host
 * (gpu_double_fail) to check the error paths.
 *
 * Results are not Mie solutions: every call fills its outputs with finite,
 * deterministic values (positive sums, so the host normalisation divides by a
 * nonzero scattering cross section) and checks the sizes it is handed by
 * writing every element of every output array.  The device state of
 * job->keep is a heap block that lmie_gpu_release frees, so a leaked state
 * shows up in the leak checker.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lmie_gpu.h"

static __thread char err[256];
static const char *fail_at;        /* name of the call to fail, or NULL */

/* Forces the next calls of the named device entry point ("forward",
   "adjoint", "phase_func", "batch_open", or NULL for none) to fail. */
void gpu_double_fail(const char *what) { fail_at = what; }

static int forced(const char *what) {
     if (fail_at && strcmp(fail_at, what) == 0) {
          snprintf(err, sizeof err, "injected failure of lmie_gpu_%s", what);
          return 1;
     }
     return 0;
}

const char *lmie_gpu_last_error(void) { return err; }

typedef struct {
     int N;
     double *block;                 /* stands for device memory kept for the reverse pass */
} state_t;

int lmie_gpu_forward(const lmie_gpu_job *J, lmie_gpu_result *R) {
     if (forced("forward")) return -1;
     const int N = J->n_size, nch = 1 + J->n_tan + J->n_wgt, C = J->n_coef, Q = J->n_qang;
     double fsum = 0.;
     for (int i = 0; i < N; ++i) fsum += J->f[i] + J->x[i] * 0. + J->n1[i] * 0. + J->n2[i] * 0.;
     for (int q = 0; q < 4; ++q) R->sums[q] = 1. + q + 1e-3 * fsum;
     for (int c = 1; c < nch; ++c)
          for (int q = 0; q < 4; ++q) R->sums_ch[4 * (c - 1) + q] = 1e-3 * (c + q);
     for (int g = 0; g < 2; ++g) {
          double *dst = g ? R->lc_raw : R->gc_raw;
          if (!dst || !(g ? J->calc_lc : J->calc_gc)) continue;
          for (int c = 0; c < nch; ++c)
               for (int k = 0; k < 6; ++k)
                    for (int l = 0; l < C; ++l) dst[((size_t) c * 6 + k) * C + l] = (k + 1.) / (l + 1. + c);
     }
     if (J->save_control == 1) {
          for (size_t e = 0; e < (size_t) 4 * N; ++e) R->save1[e] = 1e-3 * (double) (e % 7);
          for (size_t e = 0; e < (size_t) 4 * N * Q; ++e) R->save2[e] = 1e-3 * (double) (e % 11);
     }
     if (J->save_control == 2) {
          volatile double t = 0.;
          for (size_t e = 0; e < (size_t) 4 * N; ++e) t += J->saved1[e];
          for (size_t e = 0; e < (size_t) 4 * N * Q; ++e) t += J->saved2[e];
     }
     R->state = NULL;
     if (J->keep) {
          state_t *s = malloc(sizeof *s);
          if (!s) { snprintf(err, sizeof err, "out of memory"); return -1; }
          s->N = N;
          s->block = calloc((size_t) N, sizeof(double));
          R->state = s;
     }
     return 0;
}

int lmie_gpu_adjoint(void *state, const lmie_gpu_adj_in *in, lmie_gpu_adj_out *out) {
     if (forced("adjoint")) return -1;
     state_t *s = state;
     if (!s) { snprintf(err, sizeof err, "no device state"); return -1; }
     for (int i = 0; i < s->N; ++i) {
          out->f_bar[i] = 1e-6 * in->sums_bar[0] + s->block[i];
          if (out->x_bar) out->x_bar[i] = 1e-6 * i;
     }
     out->m_bar[0] = 1.;
     out->m_bar[1] = -1.;
     return 0;
}

void lmie_gpu_release(void *state) {
     state_t *s = state;
     if (!s) return;
     free(s->block);
     free(s);
}

/* batches: every case runs on its own (nothing is fused) */
void *lmie_gpu_batch_open(int n_case) {
     if (forced("batch_open")) return NULL;
     return calloc(1, sizeof(int) * (n_case > 0 ? n_case : 1));
}
int lmie_gpu_batch_enter(void *batch, int slot) { return batch && slot >= 0 ? 0 : -1; }
void lmie_gpu_batch_leave(void) {}
int lmie_gpu_batch_fused(void *batch, int pass) { (void) batch; (void) pass; return 0; }
void lmie_gpu_batch_close(void *batch) { free(batch); }

int lmie_gpu_phase_func(int n_set, int n_coef, int ld, const double *coefs, int n_theta,
                        const double *theta, double *pf, int gsf, int device) {
     (void) gsf; (void) device;
     if (forced("phase_func")) return -1;
     for (int s = 0; s < n_set; ++s)
          for (int k = 0; k < 6; ++k)
               for (int i = 0; i < n_theta; ++i)
                    pf[((size_t) s * 6 + k) * n_theta + i] =
                         coefs[((size_t) s * 6 + k) * ld + (n_coef - 1)] * 0. + theta[i] * 1e-3;
     return 0;
}

void lmie_gpu_host_time(int k, double ms) { (void) k; (void) ms; }
void lmie_gpu_host_times_clear(void) {}

/* timing entry points of the device layer (include/lmie_hip.h) */
int lmie_set_timing(int on) { (void) on; return 0; }
int lmie_get_timing(double *ms, int n) {
     for (int k = 0; k < n; ++k) ms[k] = 0.;
     return 0;
}
