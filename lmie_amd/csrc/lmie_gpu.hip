// lmie_gpu.hip -- gfx950 kernels of the size-integrated Mie path and the thin C
// ABI of include/lmie_gpu.h.
//
// Forward call (main stream; angle tables on two side streams):
//   gauss_nodes      Gauss-Legendre angles on [-1, 1]          (misc/rtutil_math.c:231-316)
//   pitau_table16    split 16-bit pi_n +- tau_n at the positive angles (mixed path)
//   pitau_table      FP64 pi_n +- tau_n (FP64 path)            (src/lmie_core.c:669-691)
//   gsf_table        P^{00}, P^{02}, P^{22}, P^{2-2}, P_l       (misc/rtutil_math.c:867-1017)
//   mie_chains       per size: D_n downward; ratio, psi_n, chi_n (src/lmie_core.c:298-405)
//   mie_terms        per (size, 8 orders): a_n, b_n, contraction operand, series partials
//                                                              (src/lmie_core.c:453-652)
//   mie_sums         series per size                          (src/lmie_core.c:537-596)
//   mie_coefs        per size with forward tangents (tangent passes)
//   angular_mix      S sums on the 16-bit matrix cores (split operands), Stokes in
//                    FP64, weighted sums per angle            (src/lmie_core.c:658-791)
//   angular_f64      the same on the FP64 matrix cores
//   reduce_parts     fixed-order sum of the per-tile partials (src/lmie_schedule.c:109-147)
//   expand_*         generalized-spherical / Legendre expansion (misc/rtutil_scat_int.c:20-480)
//   saved_sums       save_control 2: re-weighting of saved per-size arrays
//                                                              (src/lmie_schedule.c:1087-1124)
// Reverse call: expand_adjoint_*, adj_weights, adj_tables, adj_gbuild,
// adj_gsplit, adj_ugemm32 (the kernel matrix G and U = G A), rev_orders and
// rev_carry (per-size reverse sweep); angular_adj_f64 on the FP64 path.
//
// Batches (lmie_gpu_batch_*): the same kernels in batched form (name_b, and
// the batch flag of adj_ugemm32), one launch for all cases of a batch, with
// copy_jobs moving the staged uploads and results.
//
// HBM layouts: orders are rows and sizes (padded to N_p) columns, so the 64
// lanes of a wave (64 consecutive sizes) move 512 contiguous bytes per row;
// split 16-bit operands are blocked by 8 along the contraction index
// ([plane][K / 8][column][8]), one 16-byte fragment per lane.
#include <hip/hip_runtime.h>

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/lmie_gpu.h"
#include "../../include/lmie_hip.h"

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int KC = 16;     // orders per LDS stage of the contraction
constexpr int TILE = 64;   // size-point tile of the padded layouts

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Mixed path scaling: the split operands hold (2n+1)(a_n +- b_n) * kSA and
// (pi_n +- tau_n) / (n(n+1)) * kSB, so that their product is
// fv4_n (a_n +- b_n)(pi_n +- tau_n) * kSA * kSB with both factors inside the
// normal range of fp16 (|A| <= 4 * 7622, |B| <= 2^14); powers of two, so the
// scaling is exact.
constexpr double kSA = 4.0, kSB = 16384.0;
constexpr int KM = 32;     // orders per LDS stage of the mixed contraction (4 groups of 8)
// planes of the split angle table: (pi+tau) high, middle, low, (pi-tau) high, middle, low
constexpr int B16_PLANES = 6;

// 16-bit element types of the split operands
struct F16 {
     typedef _Float16 T;
     typedef v8h V;
     static __device__ __forceinline__ unsigned short bits(double v, double *rounded) {
          _Float16 h = (_Float16) (float) v;
          *rounded = (double) (float) h;
          return __builtin_bit_cast(unsigned short, h);
     }
     static __device__ __forceinline__ double value(unsigned short b) {
          return (double) (float) __builtin_bit_cast(_Float16, b);
     }
     static __device__ __forceinline__ v4f mfma(v4u a, v4u b, v4f c) {
          return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c, 0, 0, 0);
     }
     static __device__ __forceinline__ v16f mfma32(v4u a, v4u b, v16f c) {
          return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c, 0, 0, 0);
     }
};
struct BF16 {
     typedef __bf16 T;
     typedef v8b V;
     static __device__ __forceinline__ unsigned short bits(double v, double *rounded) {
          __bf16 h = (__bf16) (float) v;
          *rounded = (double) (float) h;
          return __builtin_bit_cast(unsigned short, h);
     }
     static __device__ __forceinline__ double value(unsigned short b) {
          return (double) __builtin_bit_cast(float, (unsigned) b << 16);
     }
     static __device__ __forceinline__ v4f mfma(v4u a, v4u b, v4f c) {
          return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8b, a), __builtin_bit_cast(v8b, b), c, 0, 0, 0);
     }
     static __device__ __forceinline__ v16f mfma32(v4u a, v4u b, v16f c) {
          return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8b, a), __builtin_bit_cast(v8b, b), c, 0, 0, 0);
     }
};

// high and low 16-bit parts of v: hi = round(v), lo = round(v - hi)
template <class E>
__device__ __forceinline__ void split16(double v, unsigned short *hi, unsigned short *lo) {
     double r, r2;
     *hi = E::bits(v, &r);
     *lo = E::bits(v - r, &r2);
}
// three parts: hi = round(v), mid = round(v - hi), lo = round(v - hi - mid)
// (hi and mid are split16's two parts)
template <class E>
__device__ __forceinline__ void split16x3(double v, unsigned short *hi, unsigned short *mid, unsigned short *lo) {
     double r, r2, r3;
     *hi = E::bits(v, &r);
     *mid = E::bits(v - r, &r2);
     *lo = E::bits((v - r) - r2, &r3);
}

// Row of the 16x16 f32 result held in lane l, element v, for the 16-bit
// matrix instructions (checked by lmie_gpu_selftest_mfma16).
__device__ __forceinline__ int mfma16_row(int l, int v) { return 4 * (l >> 4) + v; }
// Row of the 32x32 f32 result held in lane l, value v, for the 32x32x16
// 16-bit instructions (column l % 32; checked by lmie_gpu_selftest_mfma32).
__device__ __forceinline__ int mfma32_row(int l, int v) { return 8 * (v >> 2) + 4 * (l >> 5) + (v & 3); }

thread_local char g_err[512];
thread_local double g_times[LMIE_N_TIMERS];
int g_timing = 0;

int set_err(const char *fmt, ...) {
     va_list ap;
     va_start(ap, fmt);
     vsnprintf(g_err, sizeof g_err, fmt, ap);
     va_end(ap);
     return -1;
}

#define HIPCHECK(expr)                                                                     \
     do {                                                                                  \
          hipError_t e_ = (expr);                                                          \
          if (e_ != hipSuccess)                                                            \
               return set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),        \
                              __FILE__, __LINE__);                                         \
     } while (0)

inline int rup(int a, int b) { return (a + b - 1) / b * b; }

// ---------------------------------------------------------------------------
// complex helpers (double2 = re, im)
struct cd {
     double r, i;
};
__device__ __forceinline__ cd cmk(double r, double i) { return cd{r, i}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return cd{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cd cmul(cd a, cd b) { return cd{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return cd{a.r * s, a.i * s}; }
__device__ __forceinline__ cd cconj(cd a) { return cd{a.r, -a.i}; }
__device__ __forceinline__ double cabs2(cd a) { return a.r * a.r + a.i * a.i; }
// Smith's division: scaled by the larger component of the divisor, so no
// intermediate overflows where |b|^2 would.
__device__ __forceinline__ cd cdiv(cd a, cd b) {
     if (fabs(b.r) >= fabs(b.i)) {
          double t = b.i / b.r, d = b.r + b.i * t;
          return cd{(a.r + a.i * t) / d, (a.i - a.r * t) / d};
     }
     double t = b.r / b.i, d = b.r * t + b.i;
     return cd{(a.r * t + a.i) / d, (a.i * t - a.r) / d};
}
__device__ __forceinline__ cd crecip(cd b) { return cdiv(cd{1., 0.}, b); }
// 1/b = conj(b) / |b|^2 with one division, for divisors well inside the
// range (|b|^2 neither overflows nor underflows)
__device__ __forceinline__ cd crecip_fast(cd b) {
     const double q = 1. / (b.r * b.r + b.i * b.i);
     return cd{b.r * q, -b.i * q};
}
// 1/q from the hardware reciprocal and two Newton steps (about 1 ulp; no
// special-value handling), for the dependence chains of the recurrences
__device__ __forceinline__ double rcp_nr(double q) {
     double r = __builtin_amdgcn_rcp(q);
     double e = __builtin_fma(-q, r, 1.);
     r = __builtin_fma(r, e, r);
     e = __builtin_fma(-q, r, 1.);
     return __builtin_fma(r, e, r);
}
__device__ __forceinline__ cd crecip_nr(cd b) {
     const double q = rcp_nr(b.r * b.r + b.i * b.i);
     return cd{b.r * q, -b.i * q};
}
// Smith's division with the reciprocals of rcp_nr (about 1 ulp).  Both of
// Smith's cases in one instruction stream: with (p, q) the divisor's larger
// and smaller component and (u, v) the dividend's matching pair, the quotient
// is ((u + v t) / d, s (v - u t) / d), t = q / p, d = p + q t, s = +1 when
// the real part is the larger (else -1); the same operations as the two
// branches, so the same bits, without divergent lanes.
__device__ __forceinline__ cd cdiv_nr(cd a, cd b) {
     const bool rr = fabs(b.r) >= fabs(b.i);
     const double p = rr ? b.r : b.i, q = rr ? b.i : b.r;
     const double u = rr ? a.r : a.i, v = rr ? a.i : a.r;
     const double t = q * rcp_nr(p), rd = rcp_nr(p + q * t);
     const double im = (v - u * t) * rd;
     return cd{(u + v * t) * rd, rr ? im : -im};
}

// ---------------------------------------------------------------------------
// Gauss-Legendre angles: one thread per root.  Newton steps with the
// second-order correction, stopped at a step <= DBL_EPSILON
// (misc/rtutil_math.c:73-129); nodes ascending, weights
// (1 - y^2) 2 / (n P_{n-1}(y'))^2 with y' the last-but-one iterate
// (misc/rtutil_math.c:279-310).  Near mu = +-1 that weight formula is
// sensitive (about 1e-8 relative per ulp of y') to the iteration path, so the
// initial guesses come from the host (the C library's cos and tan, as the
// reference) and every operation here is a separately rounded IEEE operation
// (no contraction): the rule is the reference's rule bit for bit.
// The factors j / (j + 1) of the Legendre recurrence come from LDS when the
// degree fits (the same correctly rounded quotients, computed once per
// workgroup), else they are divided in the loop.
constexpr int GAUSS_LDS = 8192;

template <bool TAB>
__device__ __forceinline__ void gauss_newton(int n, const double *cj, double &y, double &pa, double &pb) {
#pragma clang fp contract(off)
     const double dn = (double) n, nn1 = (double) (n * (n + 1));
     double y0;
     int it = 0;
     do {
          double pm;
          pa = 1.;
          pb = 0.;
#pragma unroll 4
          for (int j = 0; j < n; ++j) {
               pm = pb;
               pb = pa;
               const double u = __dmul_rn(y, pb);
               const double c = TAB ? cj[j] : __ddiv_rn((double) j, __dadd_rn((double) j, 1.));
               pa = __dadd_rn(__dmul_rn(c, __dsub_rn(u, pm)), u);
          }
          const double s = __dsub_rn(1., __dmul_rn(y, y));
          const double d1 = __ddiv_rn(__dmul_rn(dn, __dsub_rn(pb, __dmul_rn(y, pa))), s);
          const double d2 = __ddiv_rn(__dsub_rn(__dmul_rn(__dmul_rn(2., y), d1), __dmul_rn(nn1, pa)), s);
          const double h = __ddiv_rn(pa, d1);
          y0 = y;
          y = __dsub_rn(y0, __dmul_rn(h, __dadd_rn(1., __ddiv_rn(__dmul_rn(h, d2), __dmul_rn(2., d1)))));
     } while (fabs(__dsub_rn(y, y0)) > 2.220446049250313e-16 && ++it < 100);
}

__device__ __forceinline__ void gauss_nodes_body(int n, const double *guess, double *mu, double *w) {
#pragma clang fp contract(off)
     __shared__ double cj[GAUSS_LDS];
     const bool tab = n <= GAUSS_LDS;
     if (tab) {
          for (int j = threadIdx.x; j < n; j += blockDim.x) cj[j] = __ddiv_rn((double) j, __dadd_rn((double) j, 1.));
          __syncthreads();
     }
     int i = blockIdx.x * blockDim.x + threadIdx.x;
     if (i >= (n + 1) / 2) return;
     double y = guess[i], pa, pb;
     if (tab) gauss_newton<true>(n, cj, y, pa, pb);
     else gauss_newton<false>(n, nullptr, y, pa, pb);
     const double dn = (double) n;
     double c = __dmul_rn(dn, pb);
     c = __ddiv_rn(2., __dmul_rn(c, c));
     const double off = __dmul_rn(1., y);
     mu[i] = __dsub_rn(0., off);
     mu[n - 1 - i] = __dadd_rn(0., off);
     w[i] = w[n - 1 - i] = __dmul_rn(__dsub_rn(1., __dmul_rn(off, y)), c);
}
__global__ void gauss_nodes(int n, const double *guess, double *mu, double *w) { gauss_nodes_body(n, guess, mu, w); }
// Batched form of each kernel used by fused batches (name_b): one launch for
// all cases, blockIdx.z picks the case, whose own grid (gx, gy blocks) may be
// smaller than the launch grid; its surplus blocks exit.
struct gauss_nodes_B {
     int n;
     const double *guess;
     double *mu;
     double *w;
     unsigned gx, gy;
};
__global__ void gauss_nodes_b(const gauss_nodes_B *cases) {
     const gauss_nodes_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     gauss_nodes_body(a.n, a.guess, a.mu, a.w);
}


// pi_n +- tau_n at the positive angles mu_{H+j}: one thread per angle, upward
// recurrences pi_{n+1} = mu pi_n + (n+1)/n (mu pi_n - pi_{n-1}),
// tau_n = n (mu pi_n - pi_{n-1}) - pi_{n-1} (src/lmie_core.c:669-691).
__global__ void pitau_table(int H, int Hp, int n1max, int Kp, const double *mu, double *Bp, double *Bm) {
     // separately rounded operations in the reference's order: the table is
     // the reference's pi_n, tau_n bit for bit
#pragma clang fp contract(off)
     int j = blockIdx.x * blockDim.x + threadIdx.x;
     if (j >= Hp) return;
     if (j >= H) {
          for (int n = 1; n <= Kp; ++n) Bp[(size_t) (n - 1) * Hp + j] = Bm[(size_t) (n - 1) * Hp + j] = 0.;
          return;
     }
     double m = mu[H + j], p0 = 0., p1 = 1.;
     for (int n = 1; n <= n1max; ++n) {
          double a1 = m * p1, a2 = a1 - p0;
          double tau = n * a2 - p0;
          Bp[(size_t) (n - 1) * Hp + j] = p1 + tau;
          Bm[(size_t) (n - 1) * Hp + j] = p1 - tau;
          p0 = p1;
          p1 = a1 + (n + 1.) / n * a2;
     }
     for (int n = n1max + 1; n <= Kp; ++n) Bp[(size_t) (n - 1) * Hp + j] = Bm[(size_t) (n - 1) * Hp + j] = 0.;
}

// Split 16-bit tables for the mixed path, blocked by 8 orders:
// [6 planes: (pi+tau) hi, mid, lo, (pi-tau) hi, mid, lo][Kb][Hp][8], value
// (pi_n +- tau_n) / (n (n+1)) * kSB.  A lane (angle) writes 16 bytes per plane
// per group of 8 orders; consecutive lanes are consecutive 16-byte chunks.
// The per-order factors (n+1)/n and kSB/(n(n+1)) come from LDS when n1max
// fits, else they are divided in the loop.
constexpr int PITAU_LDS = 4096;

template <class E, bool TAB>
__device__ __forceinline__ void pitau16_rows(int H, int Hp, int n1max, int Kb, const double *mu, const double *fq,
                                             const double *fs, unsigned short *B16, int j) {
     const size_t plane = (size_t) Kb * Hp * 8;
     const double m = j < H ? mu[H + j] : 0.;
     double p0 = 0., p1 = 1.;
     for (int g = 0; g < Kb; ++g) {
          unsigned short v[B16_PLANES][8];
          for (int e = 0; e < 8; ++e) {
               const int n = 8 * g + e + 1;
               double bp = 0., bm = 0.;
               if (j < H && n <= n1max) {
                    const double a1 = m * p1, a2 = a1 - p0;
                    const double tau = n * a2 - p0;
                    const double sc = TAB ? fs[n] : kSB / (n * (n + 1.));
                    bp = (p1 + tau) * sc;
                    bm = (p1 - tau) * sc;
                    p0 = p1;
                    p1 = a1 + (TAB ? fq[n] : (n + 1.) / n) * a2;
               }
               split16x3<E>(bp, &v[0][e], &v[1][e], &v[2][e]);
               split16x3<E>(bm, &v[3][e], &v[4][e], &v[5][e]);
          }
          for (int q = 0; q < B16_PLANES; ++q) {
               v4u w;
               for (int k = 0; k < 4; ++k) w[k] = (unsigned) v[q][2 * k] | ((unsigned) v[q][2 * k + 1] << 16);
               *(v4u *) (B16 + q * plane + ((size_t) g * Hp + j) * 8) = w;
          }
     }
}

template <class E>
__device__ __forceinline__ void pitau_table16_body(int H, int Hp, int n1max, int Kb, const double *mu, unsigned short *B16) {
     __shared__ double fq[PITAU_LDS + 1], fs[PITAU_LDS + 1];
     const bool tab = n1max <= PITAU_LDS;
     if (tab) {
          for (int n = threadIdx.x + 1; n <= n1max; n += blockDim.x) {
               fq[n] = (n + 1.) / n;
               fs[n] = kSB / (n * (n + 1.));
          }
          __syncthreads();
     }
     const int j = blockIdx.x * blockDim.x + threadIdx.x;
     if (j >= Hp) return;
     if (tab) pitau16_rows<E, true>(H, Hp, n1max, Kb, mu, fq, fs, B16, j);
     else pitau16_rows<E, false>(H, Hp, n1max, Kb, mu, nullptr, nullptr, B16, j);
}
template <class E>
__global__ void pitau_table16(int H, int Hp, int n1max, int Kb, const double *mu, unsigned short *B16) { pitau_table16_body<E>(H, Hp, n1max, Kb, mu, B16); }
struct pitau_table16_B {
     int H;
     int Hp;
     int n1max;
     int Kb;
     const double *mu;
     unsigned short *B16;
     unsigned gx, gy;
};
template <class E>
__global__ void pitau_table16_b(const pitau_table16_B *cases) {
     const pitau_table16_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     pitau_table16_body<E>(a.H, a.Hp, a.n1max, a.Kb, a.mu, a.B16);
}


// Value of lane `lane` (uniform) of a double: a scalar operand for the whole
// wave.
__device__ __forceinline__ double readlane_d(double v, int lane) {
     const unsigned long long u = (unsigned long long) __double_as_longlong(v);
     const unsigned lo = __builtin_amdgcn_readlane((unsigned) u, lane);
     const unsigned hi = __builtin_amdgcn_readlane((unsigned) (u >> 32), lane);
     return __longlong_as_double((long long) (((unsigned long long) hi << 32) | lo));
}

// Generalized spherical functions at the positive angles, l = 0 .. C-1, by
// their three-term recurrences in l (misc/rtutil_math.c:956-1016).  Layout
// [4][C][Hp] (+ plane 4: P_l for the Legendre expansion,
// misc/rtutil_math.c:506-513): lanes (angles) store contiguously for every l.
// The five recurrences are independent: blockIdx.y picks one (0: P00, 1: P02,
// 2: P22, 3: P2-2, 4: P_l), so five times as many threads run one fifth of the
// steps.  Columns H .. Hp-1 are written as zeros.
// Each step divides the reference's numerator by its divisor d_l (L1, sqrt(L1^2
// - 4) or L (L1^2 - 4)) in the reference's order of operations; the quotient
// is q = RN(x y) corrected once, q + y RN(x - q d), with y = RN(1 / d_l):
// Markstein's correction, which gives the correctly rounded x / d (the IEEE
// division's bits) when y is the correctly rounded reciprocal and q is
// within an ulp.  The divisors' reciprocals, the square roots and the
// Legendre factors (l-1)/l of 64 consecutive l are formed one per lane (in
// parallel) and each step reads its l's from that lane as scalars, so the
// dependence chain of a step is the numerator and three operations instead
// of a full division.
// FN is the recurrence (a template argument, so a step has no branch on it).
template <int FN>
__device__ __forceinline__ void gsf_plane(int H, int Hp, int C, const double *mu, double *P, int j, int lane) {
     // separately rounded operations in the reference's order (bitwise the
     // reference's functions)
#pragma clang fp contract(off)
     // every lane stays in the loops (its l's are read by the others); the
     // columns past H run with mu = 0 and store zeros
     const bool live = j < H;
     const double m = live ? mu[H + j] : 0., m2 = m * m;
     double a0 = 0., b0 = 0.;
     for (int l0 = 0; l0 < C; l0 += 64) {
          // this lane's order l0 + lane: divisor, its reciprocal, square root
          // (FN 1), Legendre factor (FN 4)
          double dv = 1., yv = 1., sav = 0.;
          {
               const int lc = l0 + lane;
               const double L = lc - 1, L1 = lc;
               if constexpr (FN == 4) {
                    yv = lc >= 2 ? (lc - 1) / (double) lc : 0.;
               } else if (lc >= 3) {
                    if constexpr (FN == 0) dv = L1;
                    else if constexpr (FN == 1) {
                         sav = sqrt(L * L - 4.);
                         dv = sqrt(L1 * L1 - 4.);
                    } else dv = L * (L1 * L1 - 4);
                    yv = 1. / dv;
               }
          }
          const int ne = min(64, C - l0);
          double *row = P + (size_t) l0 * Hp + j;
          auto put = [&](int e, double v) {
               row[(size_t) e * Hp] = live ? v : 0.;
               a0 = b0;
               b0 = v;
          };
          // l = 0, 1, 2 by their closed forms, then the recurrence steps (no
          // per-step test of l)
          int e0 = 0;
          if (l0 == 0) {
               for (int l = 0; l < 3 && l < C; ++l) {
                    double v;
                    if constexpr (FN == 4) v = l == 0 ? 1. : l == 1 ? m : readlane_d(yv, l) * (m * b0 - a0) + m * b0;
                    else if (l == 0) v = FN == 0 ? 1. : 0.;
                    else if (l == 1) v = FN == 0 ? m : 0.;
                    else if constexpr (FN == 0) v = (3. * m2 - 1.) / 2.;
                    else if constexpr (FN == 1) v = .25 * sqrt(6.) * (m2 - 1.);
                    else if constexpr (FN == 2) v = .25 * (1. + m) * (1. + m);
                    else v = .25 * (1. - m) * (1. - m);
                    put(l, v);
               }
               e0 = 3;
          }
          if constexpr (FN == 4) {
               // P_l = a (m P_{l-1} - P_{l-2}) + m P_{l-1}, a = (l-1)/l
               for (int e = e0; e < ne; ++e) {
                    const double a = readlane_d(yv, e), bb = m * b0;
                    put(e, a * (bb - a0) + bb);
               }
          } else {
               for (int e = e0; e < ne; ++e) {
                    const int l = l0 + e;
                    const double L = l - 1, L1 = l, t = 2. * L + 1.;
                    const double d = readlane_d(dv, e), y = readlane_d(yv, e);
                    double x;
                    if constexpr (FN == 0) {
                         const double da = t * m;
                         x = da * b0 - L * a0;
                    } else if constexpr (FN == 1) {
                         const double da = t * m;
                         x = da * b0 - readlane_d(sav, e) * a0;
                    } else {
                         const double k4 = t * 4, kb = L1 * (L * L - 4);
                         const double dc = t * (L * L1) * m;
                         x = FN == 2 ? (dc - k4) * b0 - kb * a0 : (dc + k4) * b0 - kb * a0;
                    }
                    const double q = x * y;
                    put(e, __builtin_fma(__builtin_fma(-q, d, x), y, q));
               }
          }
     }
}
__device__ __forceinline__ void gsf_table_body(int H, int Hp, int C, const double *mu, double *T) {
     const int j = blockIdx.x * blockDim.x + threadIdx.x, fn = blockIdx.y, lane = threadIdx.x & 63;
     if (j >= Hp) return;
     double *P = T + (size_t) fn * C * Hp;
     switch (fn) {
     case 0: gsf_plane<0>(H, Hp, C, mu, P, j, lane); break;
     case 1: gsf_plane<1>(H, Hp, C, mu, P, j, lane); break;
     case 2: gsf_plane<2>(H, Hp, C, mu, P, j, lane); break;
     case 3: gsf_plane<3>(H, Hp, C, mu, P, j, lane); break;
     default: gsf_plane<4>(H, Hp, C, mu, P, j, lane); break;
     }
}
__global__ void gsf_table(int H, int Hp, int C, const double *mu, double *T) { gsf_table_body(H, Hp, C, mu, T); }
struct gsf_table_B {
     int H;
     int Hp;
     int C;
     const double *mu;
     double *T;
     unsigned gx, gy;
};
__global__ void gsf_table_b(const gsf_table_B *cases) {
     const gsf_table_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     gsf_table_body(a.H, a.Hp, a.C, a.mu, a.T);
}


// ---------------------------------------------------------------------------
// Per-size Mie coefficients.  One thread per size point (NT tangent channels
// carried alongside).  Downward D_n(z) from n2 (stored for n <= n1) and
// downward ratio r_n = psi_n / psi_{n-1} from r_{n1} = 0 (stored), then one
// upward sweep for psi_n, chi_n, a_n, b_n, the series and the A operand
// (src/lmie_core.c:298-652).  Scratch rows are [n][Np] so every access of a
// wave is one contiguous 512-byte (or 1 KiB) line.
struct CoefArgs {
     int N, Np, n1max;
     const double *x;           // [Np]
     const double *sx, *cx;     // [Np] sin x, cos x (from the host's C library, as the reference)
     const int *n1, *n2;        // [Np]
     const int *kfill;          // [Np / TILE] rows to zero-fill up to
     double mr, mi;
     int nt;                    // tangent channels in this pass
     const double *dx;          // [nt][Np]
     const double *dm;          // [nt][2]
     double *scrD;              // [(1 + nt)][2][n1max + 1][Np]
     double *scrR;              // [(1 + nt)][n1max + 1][Np]
     double *A;                 // base planes [4][Kp][Np]
     double *At[4];             // tangent planes of this pass, each [4][Kp][Np]
     int Kp;
     double *S;                 // series [4][Np]: ext, sca, |c_bak|^2, asym
     double *St[4];             // tangent series of this pass, each [4][Np]
     int write_base;            // write A and S of the base channel
     double *CB;                // reverse-pass checkpoint: c_bak series [2][Np] or NULL
};

// Used for passes that carry tangent channels (NT >= 1); the value channel
// alone goes through mie_chains / mie_terms / mie_sums.
template <int NT>
__global__ __launch_bounds__(64) void mie_coefs(CoefArgs p) {
     // separately rounded operations in the reference's order and association
     // (src/lmie_core.c:298-611): the real recurrences reproduce the
     // reference's bits; complex divisions use Smith's scaling (ulp-level
     // differences that the downward recurrences damp)
#pragma clang fp contract(off)
     int i = blockIdx.x * blockDim.x + threadIdx.x;
     if (i >= p.Np) return;
     const size_t Np = p.Np, plane = (size_t) p.Kp * Np, srow = (size_t) (p.n1max + 1) * Np;
     int kf = p.kfill[i / TILE];
     if (i >= p.N) {
          // padded columns: zero operand rows and series
          for (int r = 0; r < kf; ++r)
               for (int c = 0; c < 4; ++c) {
                    if (p.write_base) p.A[c * plane + r * Np + i] = 0.;
                    for (int t = 0; t < NT; ++t) p.At[t][c * plane + r * Np + i] = 0.;
               }
          for (int q = 0; q < 4; ++q) {
               if (p.write_base) p.S[q * Np + i] = 0.;
               for (int t = 0; t < NT; ++t) p.St[t][q * Np + i] = 0.;
          }
          return;
     }
     const double x = p.x[i];
     const int n1 = p.n1[i], n2 = p.n2[i];
     const cd m = cmk(p.mr, -p.mi);
     const cd z = cscale(m, x);
     const cd m2 = cmul(m, m);
     cd dz[NT > 0 ? NT : 1], dmv[NT > 0 ? NT : 1];
     double dxv[NT > 0 ? NT : 1];
#pragma unroll
     for (int t = 0; t < NT; ++t) {
          dxv[t] = p.dx[t * Np + i];
          dmv[t] = cmk(p.dm[2 * t], p.dm[2 * t + 1]);
          dz[t] = cadd(cscale(dmv[t], x), cscale(m, dxv[t]));
     }
     double *Dre = p.scrD, *Dim = p.scrD + srow, *R = p.scrR;

     // downward D_j(z) = (j+1)/z - 1 / (D_{j+1} + (j+1)/z), D_{n2} = 0
     cd D = cmk(0., 0.), dD[NT > 0 ? NT : 1];
#pragma unroll
     for (int t = 0; t < NT; ++t) dD[t] = cmk(0., 0.);
     for (int j = n2 - 1; j >= 1; --j) {
          const cd c1 = cdiv(cmk(j + 1., 0.), z);
          const cd c2 = cadd(D, c1);
          D = csub(c1, crecip(c2));
#pragma unroll
          for (int t = 0; t < NT; ++t) {
               const cd c5 = cmul(cdiv(cscale(c1, -1.), z), dz[t]);
               dD[t] = cadd(c5, cdiv(cadd(dD[t], c5), cmul(c2, c2)));
          }
          if (j <= n1) {
               Dre[j * Np + i] = D.r;
               Dim[j * Np + i] = D.i;
#pragma unroll
               for (int t = 0; t < NT; ++t) {
                    p.scrD[(size_t) (1 + t) * 2 * srow + j * Np + i] = dD[t].r;
                    p.scrD[(size_t) (1 + t) * 2 * srow + srow + j * Np + i] = dD[t].i;
               }
          }
     }
     // downward ratio r_j = 1 / ((2j+1)/x - r_{j+1}), r_{n1} = 0
     {
          double r = 0., dr[NT > 0 ? NT : 1];
#pragma unroll
          for (int t = 0; t < NT; ++t) dr[t] = 0.;
          R[n1 * Np + i] = 0.;
#pragma unroll
          for (int t = 0; t < NT; ++t) p.scrR[(size_t) (1 + t) * srow + n1 * Np + i] = 0.;
          for (int j = n1 - 1; j >= 1; --j) {
               const double c = (2. * j + 1.) / x;
               const double rn = 1. / (c - r);
               const double a2 = -c / x, a3 = rn * rn;
#pragma unroll
               for (int t = 0; t < NT; ++t) {
                    dr[t] = -(a2 * dxv[t] - dr[t]) * a3;
                    p.scrR[(size_t) (1 + t) * srow + j * Np + i] = dr[t];
               }
               r = rn;
               R[j * Np + i] = r;
          }
     }
     // upward sweep
     const double sx = p.sx[i], cx = p.cx[i];
     const double ax = 1. / x;
     double psi0 = sx, chi0 = cx, chi1 = ax * cx + sx;
     double dpsi0[NT > 0 ? NT : 1], dchi0[NT > 0 ? NT : 1], dchi1[NT > 0 ? NT : 1];
#pragma unroll
     for (int t = 0; t < NT; ++t) {
          const double a2l = -ax / x * dxv[t];
          dpsi0[t] = dxv[t] * cx;
          dchi0[t] = dxv[t] * -sx;
          dchi1[t] = a2l * cx + ax * dchi0[t] + dxv[t] * cx;
     }
     double se = 0., ss = 0., sg = 0.;
     cd cb = cmk(0., 0.), ap = cmk(0., 0.), bp = cmk(0., 0.);
     double dse[NT > 0 ? NT : 1], dss[NT > 0 ? NT : 1], dsg[NT > 0 ? NT : 1];
     cd dcb[NT > 0 ? NT : 1], dap[NT > 0 ? NT : 1], dbp[NT > 0 ? NT : 1];
#pragma unroll
     for (int t = 0; t < NT; ++t) {
          dse[t] = dss[t] = dsg[t] = 0.;
          dcb[t] = dap[t] = dbp[t] = cmk(0., 0.);
     }
     double sgn = -1.;
     for (int j = 1; j <= n1; ++j) {
          const double r = R[j * Np + i];
          const double psi = r * psi0;
          const cd Dj = cmk(Dre[j * Np + i], Dim[j * Np + i]);
          const double jx = j / x;
          const cd zj = cmk(psi, chi1), zjm = cmk(psi0, chi0);
          // a_j with u = D_j / m + j / x; b_j with v = m D_j + j / x
          const cd u = cadd(cdiv(Dj, m), cmk(jx, 0.));
          const cd v = cadd(cmul(m, Dj), cmk(jx, 0.));
          const cd na = csub(cscale(u, psi), cmk(psi0, 0.)), da = csub(cmul(u, zj), zjm);
          const cd nb = csub(cscale(v, psi), cmk(psi0, 0.)), db = csub(cmul(v, zj), zjm);
          const cd a = cdiv(na, da), b = cdiv(nb, db);
          const double fv1 = 2. * j + 1., fv2 = (j + 1.) / j, fv4 = fv1 / (j * (j + 1.));
          const double fv3m = j > 1 ? (j + 1.) / ((double) j / (j - 1.)) : 0.;
          const double fv4m = j > 1 ? (2. * j - 1.) / ((j - 1.) * j) : 0.;
          (void) fv2;
          se += fv1 * (a.r + b.r);
          ss += fv1 * (cabs2(a) + cabs2(b));
          cb = cadd(cb, cscale(csub(a, b), fv1 * sgn));
          if (j > 1)
               sg += fv3m * ((ap.r * a.r + ap.i * a.i) + (bp.r * b.r + bp.i * b.i)) + fv4m * (ap.r * bp.r + ap.i * bp.i);
          if (p.write_base) {
               const size_t o = (size_t) (j - 1) * Np + i;
               p.A[o] = fv4 * (a.r + b.r);
               p.A[plane + o] = fv4 * (a.i + b.i);
               p.A[2 * plane + o] = fv4 * (a.r - b.r);
               p.A[3 * plane + o] = fv4 * (a.i - b.i);
          }
#pragma unroll
          for (int t = 0; t < NT; ++t) {
               const double drt = p.scrR[(size_t) (1 + t) * srow + j * Np + i];
               const double dpsi = drt * psi0 + r * dpsi0[t];
               const cd dD = cmk(p.scrD[(size_t) (1 + t) * 2 * srow + j * Np + i],
                                 p.scrD[(size_t) (1 + t) * 2 * srow + srow + j * Np + i]);
               const double djx = -jx * dxv[t] / x;
               const cd dzj = cmk(dpsi, dchi1[t]), dzjm = cmk(dpsi0[t], dchi0[t]);
               // du = dD/m - D dm / m^2 + d(j/x); dv = dm D + m dD + d(j/x)
               const cd du = cadd(csub(cdiv(dD, m), cmul(cdiv(Dj, m2), dmv[t])), cmk(djx, 0.));
               const cd dv = cadd(cadd(cmul(dmv[t], Dj), cmul(m, dD)), cmk(djx, 0.));
               const cd dna = csub(cadd(cscale(du, psi), cscale(u, dpsi)), cmk(dpsi0[t], 0.));
               const cd dda = csub(cadd(cmul(du, zj), cmul(u, dzj)), dzjm);
               const cd dnb = csub(cadd(cscale(dv, psi), cscale(v, dpsi)), cmk(dpsi0[t], 0.));
               const cd ddb = csub(cadd(cmul(dv, zj), cmul(v, dzj)), dzjm);
               const cd dat = csub(cdiv(dna, da), cdiv(cmul(na, dda), cmul(da, da)));
               const cd dbt = csub(cdiv(dnb, db), cdiv(cmul(nb, ddb), cmul(db, db)));
               dse[t] += fv1 * (dat.r + dbt.r);
               dss[t] += fv1 * 2. * ((a.r * dat.r + a.i * dat.i) + (b.r * dbt.r + b.i * dbt.i));
               dcb[t] = cadd(dcb[t], cscale(csub(dat, dbt), fv1 * sgn));
               if (j > 1)
                    dsg[t] += fv3m * (dap[t].r * a.r + dap[t].i * a.i + ap.r * dat.r + ap.i * dat.i +
                                      dbp[t].r * b.r + dbp[t].i * b.i + bp.r * dbt.r + bp.i * dbt.i) +
                              fv4m * (dap[t].r * bp.r + dap[t].i * bp.i + ap.r * dbp[t].r + ap.i * dbp[t].i);
               const size_t o = (size_t) (j - 1) * Np + i;
               p.At[t][o] = fv4 * (dat.r + dbt.r);
               p.At[t][plane + o] = fv4 * (dat.i + dbt.i);
               p.At[t][2 * plane + o] = fv4 * (dat.r - dbt.r);
               p.At[t][3 * plane + o] = fv4 * (dat.i - dbt.i);
               dap[t] = dat;
               dbp[t] = dbt;
               // advance tangent recurrences
               const double c = fv1 / x;
               const double dchi2 = -c * dxv[t] / x * chi1 + c * dchi1[t] - dchi0[t];
               dchi0[t] = dchi1[t];
               dchi1[t] = dchi2;
               dpsi0[t] = dpsi;
          }
          ap = a;
          bp = b;
          sgn = -sgn;
          const double chi2 = fv1 / x * chi1 - chi0;
          chi0 = chi1;
          chi1 = chi2;
          psi0 = psi;
     }
     // zero rows n1 .. kf-1 so the contraction can run whole chunks
     for (int r = n1; r < kf; ++r) {
          const size_t o = (size_t) r * Np + i;
          for (int c = 0; c < 4; ++c) {
               if (p.write_base) p.A[c * plane + o] = 0.;
#pragma unroll
               for (int t = 0; t < NT; ++t) p.At[t][c * plane + o] = 0.;
          }
     }
     if (p.CB) {
          p.CB[i] = cb.r;
          p.CB[Np + i] = cb.i;
     }
     if (p.write_base) {
          p.S[i] = se;
          p.S[Np + i] = ss;
          p.S[2 * Np + i] = cabs2(cb);
          p.S[3 * Np + i] = sg;
     }
#pragma unroll
     for (int t = 0; t < NT; ++t) {
          p.St[t][i] = dse[t];
          p.St[t][Np + i] = dss[t];
          p.St[t][2 * Np + i] = 2. * (cb.r * dcb[t].r + cb.i * dcb[t].i);
          p.St[t][3 * Np + i] = dsg[t];
     }
}

// Element i of row `row` of a [rows][Np] array: the row pointer is uniform
// over the wave (row and the base are), the lane's byte offset a 32-bit
// register, so each load is one scalar-based address and no 64-bit lane
// arithmetic.
// (double2: one 16-byte load, read as a built-in vector type.)
template <class T>
struct LdRowType {
     typedef T type;
};
template <>
struct LdRowType<double2> {
     typedef double type __attribute__((ext_vector_type(2)));
};
template <class T>
__device__ __forceinline__ T ld_row(const T *base, int row, size_t Np, unsigned ioff) {
     // the row pointer through readfirstlane: scalar registers, and the loop
     // optimiser cannot fold the lane offset into a 64-bit pointer per lane
     typedef typename LdRowType<T>::type L;
     const unsigned long long a = (unsigned long long) (base + (size_t) row * Np);
     const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned) a);
     const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned) (a >> 32));
     typedef const char __attribute__((address_space(1))) *gchar;
     typedef const L __attribute__((address_space(1))) *gL;
     const gchar rp = (gchar) (((unsigned long long) hi << 32) | lo);
     const L v = *(gL) (rp + ioff);
     if constexpr (std::is_same<T, double2>::value) return double2{v.x, v.y};
     else return v;
}

// 1/q from the hardware reciprocal r (about 2^-24 relative) and the series
// r (1 + e + e^2), e = 1 - q r: three dependent operations after the
// reciprocal instead of the four of two Newton steps, error e^3 (< 2^-72)
// plus one rounding.  For the D chain, whose steps are latency bound.
__device__ __forceinline__ double rcp_s3(double q) {
     const double r = __builtin_amdgcn_rcp(q);
     const double e = __builtin_fma(-q, r, 1.);
     return __builtin_fma(r, __builtin_fma(e, e, e), r);
}

// ---------------------------------------------------------------------------
// Value channel in three kernels: mie_chains runs the three recurrences that
// are serial in n, two threads per size in different waves (the D chain; the
// ratio chain and then the psi, chi chain); mie_terms forms a_n, b_n, the
// contraction operand and the partial series sums in parallel, one thread
// per (size, group of 8 orders); mie_sums adds the partial sums per size.  The
// serial part is a few operations per order, the parallel part fills the GPU.
// D_j is kept as a complex pair, one 16-byte element per (order, size), so
// each step of the D chain stores one element.
struct ChainArgs {
     int N, Np, n1max, n2rows;
     const double *x, *sx, *cx;
     const int *n1, *n2;
     double mr, mi;
     int d_all;                 // store D_j up to n2 - 1 (reverse pass) instead of n1
     double *D;                 // [n2rows][Np][2]: D_j (re, im) at row j
     double *PC;                // [2][n1max + 1][Np]: U_j (psi_j = s U_j), chi_j at row j
     double *PS;                // [Np]: s
};

__device__ __forceinline__ void mie_chains_body(ChainArgs p) {
#pragma clang fp contract(off)
     const int i = blockIdx.x * blockDim.x + threadIdx.x;
     if (i >= p.N) return;
     const size_t Np = p.Np;
     const double x = p.x[i];
     const int n1 = p.n1[i], n2 = p.n2[i], dtop = p.d_all ? n2 - 1 : n1;
     double2 *Dv = (double2 *) p.D;
     if (blockIdx.y == 0) {
          // downward D_j(z) = (j+1)/z - 1 / w_j, w_j = D_{j+1} + (j+1)/z,
          // D_{n2} = 0.  The chain carries w: w_{j-1} = (2j+1)/z - 1 / w_j,
          // w_{n2-1} = n2 / z, with 1 / w = conj(w) / |w|^2 (|w| stays between
          // about 1/|z| and n2/|z|) and each update one fused multiply-add per
          // component; D_j = (j+1)/z - 1 / w_j is formed beside the chain.  The
          // recurrence damps its last-bit differences.
          const cd z = cscale(cmk(p.mr, -p.mi), x), iz = crecip(z);
          cd w = cscale(iz, (double) n2);
          for (int j = n2 - 1; j >= 1; --j) {
               const cd c1 = cscale(iz, j + 1.), e = cscale(iz, 2. * j + 1.);
               const double q = rcp_s3(__builtin_fma(w.r, w.r, w.i * w.i));
               if (j <= dtop)
                    Dv[j * Np + i] = double2{__builtin_fma(-w.r, q, c1.r), __builtin_fma(w.i, q, c1.i)};
               w = cmk(__builtin_fma(-w.r, q, e.r), __builtin_fma(w.i, q, e.i));
          }
          return;
     }
     // the other chains in another wave (blockIdx.y = 1), one loop: psi by
     // Miller's downward recurrence, U_{n1} = 0, U_{n1-1} = 1,
     // U_{j-1} = (2j+1)/x U_j - U_{j+1}, psi_j = s U_j with s = sin x / U_0 (the
     // ratios r_j = U_j / U_{j-1} of the reference's ratio chain, whose
     // product from psi_0 = sin x is the same psi_j), beside the upward
     // chi_{j+1} = (2j+1)/x chi_j - chi_{j-1}; two independent dependence
     // chains of two operations per order and no round trip through memory.
     // U grows downward by at most the decay of psi_j between n1 - 1 and 0.
     const size_t ppl = (size_t) (p.n1max + 1) * Np;
     double *U = p.PC, *Chi = p.PC + ppl;
     const double ax = 1. / x, sx = p.sx[i], cx = p.cx[i];
     double ua = 0., ub = 1.;               // U_{j+1}, U_j, starting at j = n1 - 1
     U[n1 * Np + i] = 0.;
     U[(n1 - 1) * Np + i] = 1.;
     double chi0 = cx, chi1 = ax * cx + sx;
     Chi[i] = chi0;
     for (int t = 1; t <= n1; ++t) {
          Chi[t * Np + i] = chi1;
          const double fv1 = 2. * t + 1.;
          const double chi2 = fv1 * ax * chi1 - chi0;
          chi0 = chi1;
          chi1 = chi2;
          const int j = n1 - t;
          if (j >= 1) {
               const double un = (2. * j + 1.) * ax * ub - ua;
               U[(j - 1) * Np + i] = un;
               ua = ub;
               ub = un;
          }
     }
     p.PS[i] = sx / ub;
}
__global__ __launch_bounds__(64) void mie_chains(ChainArgs p) { mie_chains_body(p); }
struct mie_chains_B {
     ChainArgs p;
     unsigned gx, gy;
};
__global__ __launch_bounds__(64) void mie_chains_b(const mie_chains_B *cases) {
     const mie_chains_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     mie_chains_body(a.p);
}


// a_j, b_j (src/lmie_core.c:557-583), operations as mie_coefs
__device__ __forceinline__ void mie_ab_jx(double jx, cd m, cd im, cd Dj, double psi, double psi0, double chi1,
                                          double chi0, cd *a, cd *b) {
#pragma clang fp contract(off)
     const cd zj = cmk(psi, chi1), zjm = cmk(psi0, chi0);
     const cd u = cadd(cmul(Dj, im), cmk(jx, 0.));   // D/m with im = 1/m
     const cd v = cadd(cmul(m, Dj), cmk(jx, 0.));
     const cd na = csub(cscale(u, psi), cmk(psi0, 0.)), da = csub(cmul(u, zj), zjm);
     const cd nb = csub(cscale(v, psi), cmk(psi0, 0.)), db = csub(cmul(v, zj), zjm);
     *a = cdiv_nr(na, da);
     *b = cdiv_nr(nb, db);
}
// a_j, b_j of size parameter x (jx = j / x)
__device__ __forceinline__ void mie_ab_at(int j, double x, cd m, cd im, cd Dj, double psi, double psi0, double chi1,
                                          double chi0, cd *a, cd *b) {
     mie_ab_jx(j / x, m, im, Dj, psi, psi0, chi1, chi0, a, b);
}

// Size-independent factors of order j (the reference's expressions, so the
// same bits), one row each over j = 0 .. n: fv4 = (2j+1) / (j (j+1)),
// fv3m = (j+1) / (j / (j-1)) and fv4m = (2j-1) / ((j-1) j) for j > 1 (else 0),
// fv3 = (j+2) / ((j+1) / j).  The per-order kernels read them as uniform loads
// instead of dividing in every lane.
constexpr int FT_FV4 = 0, FT_FV3M = 1, FT_FV4M = 2, FT_FV3 = 3, FT_ROWS = 4;

__device__ __forceinline__ void order_factors_body(int n, double *ft) {
#pragma clang fp contract(off)
     const int j = blockIdx.x * blockDim.x + threadIdx.x;
     if (j > n) return;
     const size_t ld = (size_t) n + 1;
     const double fv1 = 2. * j + 1.;
     ft[FT_FV4 * ld + j] = j > 0 ? fv1 / (j * (j + 1.)) : 0.;
     ft[FT_FV3M * ld + j] = j > 1 ? (j + 1.) / ((double) j / (j - 1.)) : 0.;
     ft[FT_FV4M * ld + j] = j > 1 ? (2. * j - 1.) / ((j - 1.) * j) : 0.;
     ft[FT_FV3 * ld + j] = j > 0 ? (j + 2.) / ((j + 1.) / j) : 0.;
}
__global__ void order_factors(int n, double *ft) { order_factors_body(n, ft); }
struct order_factors_B {
     int n;
     double *ft;
     unsigned gx, gy;
};
__global__ void order_factors_b(const order_factors_B *cases) {
     const order_factors_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     order_factors_body(a.n, a.ft);
}


struct TermArgs {
     int N, Np, n1max, n2rows, Kp, Kb, ng;
     const double *ft;          // order_factors rows over 0 .. n1max + 1
     const double *x;
     const int *n1;
     const int *kfill;          // [Np / TILE]
     double mr, mi;
     const double *D, *PC, *PS; // from mie_chains: [n2rows][Np][2] D_j, [2][n1max + 1][Np] U_j and
                                // chi_j, [Np] s (psi_j = s U_j)
     double *A;                 // FP64 planes [4][Kp][Np] (E = void)
     unsigned short *A16;       // split planes [8][Kb][Np][8] (E = F16 / BF16)
     double *part;              // [5][ng][Np]: se, ss, cb re, cb im, sg of each group
};

template <class E>
__device__ __forceinline__ void mie_terms_body(TermArgs p) {
#pragma clang fp contract(off)
     constexpr bool MIX = !std::is_void<E>::value;
     const int i = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
     if (i >= p.Np || g >= p.ng) return;
     const size_t Np = p.Np;
     const size_t plane = (size_t) p.Kp * Np, p16 = (size_t) p.Kb * Np * 8, gp = (size_t) p.ng * Np;
     const size_t ld = (size_t) p.n1max + 2;
     const int kf = p.kfill[i / TILE];
     const int n1 = i < p.N ? p.n1[i] : 0;
     const int j0 = 8 * g + 1, j1 = min(8 * g + 8, n1);
     double se = 0., ss = 0., sg = 0.;
     cd cb = cmk(0., 0.);
     // split operand of the 8 orders, [plane][order pair]; orders past n1 stay 0
     unsigned buf[8][4];
#pragma unroll
     for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int k = 0; k < 4; ++k) buf[q][k] = 0u;
     if (j0 <= j1) {
          const double x = p.x[i];
          const cd m = cmk(p.mr, -p.mi), im = crecip(m);
          // the group's rows are the same in every lane: the D rows are all
          // loaded at once (rows clamped to the array; those past a lane's
          // n1 are read, not used), so the loads overlap
          const double2 *Dv = (const double2 *) p.D;
          const double *U = p.PC, *Chi = p.PC + (size_t) (p.n1max + 1) * Np;
          const unsigned o8 = (unsigned) i * 8u, o16 = (unsigned) i * 16u;
          const int pmax = p.n1max, dmax = p.n2rows - 1;
          const double sc = p.PS[i], ax = 1. / x;
          double ps[10], ch[10], dr[9], di[9];   // rows j0 - 2 .. j0 + 7 (psi, chi), j0 - 1 .. j0 + 7 (D)
          // U and chi of the group's rows from two seed rows each, by the
          // chains' own steps (the same bits as the stored rows): U downward
          // from rows j0 + 8, j0 + 7 (U_{n1} = 0, U_{n1-1} = 1 where those rows
          // fall inside, as the chain starts), chi upward from rows j0 - 2,
          // j0 - 1 (rows 0 and 1 for the first group).  Rows past a lane's n1
          // are computed from unused values and never used.
          {
               double u[11];                  // rows j0 - 2 .. j0 + 8
               u[10] = ld_row(U, min(j0 + 8, pmax), Np, o8);
               u[9] = ld_row(U, min(j0 + 7, pmax), Np, o8);
#pragma unroll
               for (int e = 8; e >= 0; --e) {
                    const int r = j0 - 2 + e, jj = r + 1;
                    const double un = (2. * jj + 1.) * ax * u[e + 1] - u[e + 2];
                    u[e] = r == n1 ? 0. : r == n1 - 1 ? 1. : un;
               }
#pragma unroll
               for (int e = 0; e < 10; ++e) ps[e] = sc * u[e];
               const bool first = j0 == 1;
               ch[0] = first ? 0. : ld_row(Chi, max(j0 - 2, 0), Np, o8);
               ch[1] = ld_row(Chi, j0 - 1, Np, o8);
               ch[2] = first ? ld_row(Chi, 1, Np, o8) : 0.;
#pragma unroll
               for (int e = 2; e < 10; ++e) {
                    const int t = j0 - 2 + e - 1;
                    const double fv1 = 2. * t + 1.;
                    const double c2 = fv1 * ax * ch[e - 1] - ch[e - 2];
                    if (e > 2 || !first) ch[e] = c2;
               }
          }
#pragma unroll
          for (int e = 0; e < 9; ++e) {
               const double2 v = ld_row(Dv, min(j0 - 1 + e, dmax), Np, o16);
               dr[e] = v.x;
               di[e] = v.y;
          }
          cd ap = cmk(0., 0.), bp = cmk(0., 0.);
          double psi0 = ps[1], chi0 = ch[1];
          if (j0 > 1) mie_ab_at(j0 - 1, x, m, im, cmk(dr[0], di[0]), psi0, ps[0], chi0, ch[0], &ap, &bp);
          double sgn = (j0 & 1) ? -1. : 1.;
          // the 8 orders unrolled: the split halves land in registers
#pragma unroll
          for (int e = 0; e < 8; ++e) {
               const int j = j0 + e;
               if (j <= j1) {
                    const double psi = ps[e + 2], chi = ch[e + 2];
                    cd a, b;
                    mie_ab_at(j, x, m, im, cmk(dr[e + 1], di[e + 1]), psi, psi0, chi, chi0, &a, &b);
                    const double fv1 = 2. * j + 1.;
                    se += fv1 * (a.r + b.r);
                    ss += fv1 * (cabs2(a) + cabs2(b));
                    cb = cadd(cb, cscale(csub(a, b), fv1 * sgn));
                    if (j > 1)
                         sg += p.ft[FT_FV3M * ld + j] * ((ap.r * a.r + ap.i * a.i) + (bp.r * b.r + bp.i * b.i)) +
                               p.ft[FT_FV4M * ld + j] * (ap.r * bp.r + ap.i * bp.i);
                    if constexpr (MIX) {
                         const double sc = fv1 * kSA;
                         unsigned short h[8];
                         split16<E>(sc * (a.r + b.r), &h[0], &h[1]);
                         split16<E>(sc * (a.i + b.i), &h[2], &h[3]);
                         split16<E>(sc * (a.r - b.r), &h[4], &h[5]);
                         split16<E>(sc * (a.i - b.i), &h[6], &h[7]);
#pragma unroll
                         for (int q = 0; q < 8; ++q) buf[q][e / 2] |= (unsigned) h[q] << (16 * (e & 1));
                    } else {
                         const size_t o = (size_t) (j - 1) * Np + i;
                         const double fv4 = p.ft[FT_FV4 * ld + j];
                         p.A[o] = fv4 * (a.r + b.r);
                         p.A[plane + o] = fv4 * (a.i + b.i);
                         p.A[2 * plane + o] = fv4 * (a.r - b.r);
                         p.A[3 * plane + o] = fv4 * (a.i - b.i);
                    }
                    ap = a;
                    bp = b;
                    psi0 = psi;
                    chi0 = chi;
                    sgn = -sgn;
               }
          }
     }
     // operand rows of orders beyond n1 (up to the tile's depth) are zero
     if constexpr (MIX) {
          if (8 * g < kf) {
#pragma unroll
               for (int q = 0; q < 8; ++q) {
                    const v4u w = {buf[q][0], buf[q][1], buf[q][2], buf[q][3]};
                    *(v4u *) (p.A16 + q * p16 + ((size_t) g * Np + i) * 8) = w;
               }
          }
     } else {
          for (int r = max(8 * g, n1); r < min(8 * g + 8, kf); ++r)
               for (int c = 0; c < 4; ++c) p.A[c * plane + (size_t) r * Np + i] = 0.;
     }
     // mie_sums reads the groups below (n1 + 7) / 8 only
     if (j0 > n1) return;
     double *pt = p.part + (size_t) g * Np + i;
     pt[0] = se;
     pt[gp] = ss;
     pt[2 * gp] = cb.r;
     pt[3 * gp] = cb.i;
     pt[4 * gp] = sg;
}
template <class E>
__global__ __launch_bounds__(64) void mie_terms(TermArgs p) { mie_terms_body<E>(p); }
struct mie_terms_B {
     TermArgs p;
     unsigned gx, gy;
};
template <class E>
__global__ __launch_bounds__(64) void mie_terms_b(const mie_terms_B *cases) {
     const mie_terms_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     mie_terms_body<E>(a.p);
}


// Series per size from the group partials (src/lmie_core.c:586-596): a block
// of 32 sizes x 8 lanes, lane l adds the groups g = l mod 8, then the eight
// lane sums are added in a fixed order (deterministic).
__device__ __forceinline__ void mie_sums_body(const double *part, int N, int Np, int ng, const int *n1, double *S, double *CB) {
     __shared__ double red[5][8][32];
     const int il = threadIdx.x & 31, l = threadIdx.x >> 5, i = blockIdx.x * 32 + il;
     const size_t gp = (size_t) ng * Np;
     double v[5] = {0., 0., 0., 0., 0.};
     if (i < Np) {
          const int gn = i < N ? (n1[i] + 7) / 8 : 0;
          for (int g = l; g < gn; g += 8) {
               const double *pt = part + (size_t) g * Np + i;
#pragma unroll
               for (int q = 0; q < 5; ++q) v[q] += pt[q * gp];
          }
     }
#pragma unroll
     for (int q = 0; q < 5; ++q) red[q][l][il] = v[q];
     __syncthreads();
     if (l || i >= Np) return;
     double se = 0., ss = 0., sg = 0.;
     cd cb = cmk(0., 0.);
     for (int k = 0; k < 8; ++k) {
          se += red[0][k][il];
          ss += red[1][k][il];
          cb = cadd(cb, cmk(red[2][k][il], red[3][k][il]));
          sg += red[4][k][il];
     }
     S[i] = se;
     S[Np + i] = ss;
     S[2 * Np + i] = cabs2(cb);
     S[3 * Np + i] = sg;
     if (CB) {
          CB[i] = cb.r;
          CB[Np + i] = cb.i;
     }
}
__global__ __launch_bounds__(256) void mie_sums(const double *part, int N, int Np, int ng, const int *n1, double *S, double *CB) { mie_sums_body(part, N, Np, ng, n1, S, CB); }
struct mie_sums_B {
     const double *part;
     int N;
     int Np;
     int ng;
     const int *n1;
     double *S;
     double *CB;
     unsigned gx, gy;
};
__global__ __launch_bounds__(256) void mie_sums_b(const mie_sums_B *cases) {
     const mie_sums_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     mie_sums_body(a.part, a.N, a.Np, a.ng, a.n1, a.S, a.CB);
}


// ---------------------------------------------------------------------------
// Weighted sums over sizes: out[k] = sum_i u_k[i] v_k[i] (+ sum_i u2_k[i] v2_k[i]),
// one workgroup per output, fixed-order tree reduction (deterministic).
struct DotJob {
     const double *u, *v, *u2, *v2;
};

__device__ __forceinline__ void weighted_dots_body(const DotJob *jobs, int n, double *out) {
     __shared__ double red[256];
     const DotJob jb = jobs[blockIdx.x];
     double s = 0.;
     int i = threadIdx.x;
     // eight elements in flight, added in the order of the plain loop
     for (; i + 7 * 256 < n; i += 8 * 256) {
          double a[8], b[8], a2[8], b2[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
               a[e] = jb.u ? jb.u[i + 256 * e] : 1.;
               b[e] = jb.v[i + 256 * e];
               a2[e] = jb.u2 ? jb.u2[i + 256 * e] : 0.;
               b2[e] = jb.u2 ? jb.v2[i + 256 * e] : 0.;
          }
#pragma unroll
          for (int e = 0; e < 8; ++e) {
               s += jb.u ? a[e] * b[e] : b[e];
               if (jb.u2) s += a2[e] * b2[e];
          }
     }
     for (; i < n; i += 256) {
          s += jb.u ? jb.u[i] * jb.v[i] : jb.v[i];
          if (jb.u2) s += jb.u2[i] * jb.v2[i];
     }
     red[threadIdx.x] = s;
     __syncthreads();
     for (int w = 128; w > 0; w >>= 1) {
          if ((int) threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
          __syncthreads();
     }
     if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void weighted_dots(const DotJob *jobs, int n, double *out) { weighted_dots_body(jobs, n, out); }
struct weighted_dots_B {
     const DotJob *jobs;
     int n;
     double *out;
     unsigned gx, gy;
};
__global__ __launch_bounds__(256) void weighted_dots_b(const weighted_dots_B *cases) {
     const weighted_dots_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     weighted_dots_body(a.jobs, a.n, a.out);
}


// ---------------------------------------------------------------------------
// Angular contraction, FP64 matrix cores.
//
// For a tile of sizes i and positive angles j the four complex sums
//   sp+ = sum_n S_n (pi+tau)_n   sm+ = sum_n T_n (pi-tau)_n
//   sp- = sum_n s_n S_n (pi-tau)_n   sm- = sum_n s_n T_n (pi+tau)_n,
// S_n = fv4 (a+b), T_n = fv4 (a-b), s_n = (-1)^(n-1) (the -mu mirror,
// src/lmie_core.c:617-652,682-686), are 8 real GEMM accumulators of
// v_mfma_f64_16x16x4 (operand lane layout: A[i][k] in lane i + 16 k, B[k][j] in
// lane j + 16 k; result D[l/16 + 4 v][l % 16] in lane l, element v -- checked by
// lmie_gpu_selftest_mfma_f64).  The
// epilogue forms S1, S2 and the four matrix elements (src/lmie_core.c:726-749),
// weights them by f_i and reduces over the sizes of the workgroup; partials per
// block of sizes go to HBM and are summed by reduce_parts.
struct AngArgs {
     int Np, Kp, Hp, H, Q;
     const double *A[3];        // base planes, tangent planes (channel 0, 1 of the pass)
     const double *Bp, *Bm;     // [Kp][Hp]
     const int *ktile;          // rows to contract per size tile of TM
     const double *W[4];        // weight channels [Np]: f, then weight-only tangents
     const double *WT[2];       // tangent df [Np]
     const double *f;           // base weight (for f dP of tangents)
     int tiles_per_block;       // size tiles per workgroup
     double *part;              // [gridDim.y][nch][4][Q]
     double *save2;             // [4][N][Q] per-size matrix (save_control 1) or NULL
     int N;
     double *sst;               // reverse-pass checkpoint [Np][Hp][8]: sp+, sm+, sp-, sm- (re, im) or NULL
};

__device__ __forceinline__ void stokes(double spr, double spi, double smr, double smi, double P[4]) {
     const double s1r = .5 * (spr + smr), s1i = .5 * (spi + smi);
     const double s2r = .5 * (spr - smr), s2i = .5 * (spi - smi);
     const double q1 = s1r * s1r + s1i * s1i, q2 = s2r * s2r + s2i * s2i;
     P[0] = (q1 + q2) * .5;
     P[1] = s1r * s2r + s1i * s2i;            // Re S1 S2*
     P[2] = -(q1 - q2) * .5;
     P[3] = s1i * s2r - s1r * s2i;            // Im S1 S2*
}

__device__ __forceinline__ void stokes_tan(double spr, double spi, double smr, double smi, double tpr,
                                           double tpi, double tmr, double tmi, double dP[4]) {
     const double s1r = .5 * (spr + smr), s1i = .5 * (spi + smi);
     const double s2r = .5 * (spr - smr), s2i = .5 * (spi - smi);
     const double t1r = .5 * (tpr + tmr), t1i = .5 * (tpi + tmi);
     const double t2r = .5 * (tpr - tmr), t2i = .5 * (tpi - tmi);
     const double dq1 = 2. * (t1r * s1r + t1i * s1i), dq2 = 2. * (t2r * s2r + t2i * s2i);
     // d(S1 S2*) = dS1 S2* + S1 dS2*
     const double dxr = t1r * s2r + t1i * s2i + s1r * t2r + s1i * t2i;
     const double dxi = t1i * s2r - t1r * s2i + s1i * t2r - s1r * t2i;
     dP[0] = (dq1 + dq2) * .5;
     dP[1] = dxr;
     dP[2] = -(dq1 - dq2) * .5;
     dP[3] = dxi;
}

template <int WM, int WN, int NT, int NW>
__global__ __launch_bounds__(256) void angular_f64(AngArgs p) {
     constexpr int TM = 32 * WM, TN = 32 * WN;
     constexpr int SA = TM + 16, SB = TN + 16;       // LDS row strides (bank spread)
     constexpr int NC = 1 + NT;
     __shared__ double sA[NC * 4 * KC * SA];
     __shared__ double sB[2 * KC * SB];
     constexpr int NE = (NW + NT) > 0 ? (NW + NT) : 1;
     __shared__ double sR[NE * 8 * 32 * WN];

     const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
     const int wm = wave & 1, wn = wave >> 1;
     const int l16 = lane & 15, lg = lane >> 4;
     const int jt0 = blockIdx.x * TN;
     const int tile0 = blockIdx.y * p.tiles_per_block;
     const size_t Np = p.Np, Hp = p.Hp, plane = (size_t) p.Kp * Np;

     double ew[NE][WN][8];
#pragma unroll
     for (int c = 0; c < NW + NT; ++c)
#pragma unroll
          for (int b = 0; b < WN; ++b)
#pragma unroll
               for (int q = 0; q < 8; ++q) ew[c][b][q] = 0.;

     for (int tt = 0; tt < p.tiles_per_block; ++tt) {
          const int tile = tile0 + tt;
          const int it0 = tile * TM;
          if (it0 >= p.Np) break;
          const int kmax = p.ktile[tile];
          v4d acc[NC][WM][WN][8];
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
               for (int a = 0; a < WM; ++a)
#pragma unroll
                    for (int b = 0; b < WN; ++b)
#pragma unroll
                         for (int q = 0; q < 8; ++q) acc[c][a][b][q] = v4d{0., 0., 0., 0.};

          for (int k0 = 0; k0 < kmax; k0 += KC) {
               __syncthreads();
               // stage A: NC*4 planes x KC rows x TM sizes
               for (int e = threadIdx.x; e < NC * 4 * KC * (TM / 2); e += 256) {
                    const int col2 = e % (TM / 2), rest = e / (TM / 2);
                    const int row = rest % KC, pl = rest / KC;
                    const int c = pl / 4, comp = pl % 4;
                    const double2 v = *(const double2 *) (p.A[c] + comp * plane + (size_t) (k0 + row) * Np + it0 + 2 * col2);
                    *(double2 *) (sA + (pl * KC + row) * SA + 2 * col2) = v;
               }
               for (int e = threadIdx.x; e < 2 * KC * (TN / 2); e += 256) {
                    const int col2 = e % (TN / 2), rest = e / (TN / 2);
                    const int row = rest % KC, pm = rest / KC;
                    const double *src = pm ? p.Bm : p.Bp;
                    const double2 v = *(const double2 *) (src + (size_t) (k0 + row) * Hp + jt0 + 2 * col2);
                    *(double2 *) (sB + (pm * KC + row) * SB + 2 * col2) = v;
               }
               __syncthreads();
#pragma unroll
               for (int ks = 0; ks < KC / 4; ++ks) {
                    const int r = ks * 4 + lg;                   // this lane's k row
                    const double sg = ((k0 + r) & 1) ? -1. : 1.;   // s_n, n = k0 + r + 1
                    double bp[WN], bm[WN];
#pragma unroll
                    for (int b = 0; b < WN; ++b) {
                         const int jc = wn * 16 * WN + 16 * b + l16;
                         bp[b] = sB[r * SB + jc];
                         bm[b] = sB[(KC + r) * SB + jc];
                    }
#pragma unroll
                    for (int c = 0; c < NC; ++c)
#pragma unroll
                         for (int a = 0; a < WM; ++a) {
                              const int ic = wm * 16 * WM + 16 * a + l16;
                              const double Sr = sA[((c * 4 + 0) * KC + r) * SA + ic];
                              const double Si = sA[((c * 4 + 1) * KC + r) * SA + ic];
                              const double Tr = sA[((c * 4 + 2) * KC + r) * SA + ic];
                              const double Ti = sA[((c * 4 + 3) * KC + r) * SA + ic];
#pragma unroll
                              for (int b = 0; b < WN; ++b) {
                                   v4d *q = acc[c][a][b];
                                   q[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(Sr, bp[b], q[0], 0, 0, 0);
                                   q[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(Si, bp[b], q[1], 0, 0, 0);
                                   q[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(Tr, bm[b], q[2], 0, 0, 0);
                                   q[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ti, bm[b], q[3], 0, 0, 0);
                                   q[4] = __builtin_amdgcn_mfma_f64_16x16x4f64(sg * Sr, bm[b], q[4], 0, 0, 0);
                                   q[5] = __builtin_amdgcn_mfma_f64_16x16x4f64(sg * Si, bm[b], q[5], 0, 0, 0);
                                   q[6] = __builtin_amdgcn_mfma_f64_16x16x4f64(sg * Tr, bp[b], q[6], 0, 0, 0);
                                   q[7] = __builtin_amdgcn_mfma_f64_16x16x4f64(sg * Ti, bp[b], q[7], 0, 0, 0);
                              }
                         }
               }
          }
          // epilogue: Stokes elements, weights, reduction over this lane's sizes
#pragma unroll
          for (int a = 0; a < WM; ++a)
#pragma unroll
               for (int v = 0; v < 4; ++v) {
                    const int i = it0 + wm * 16 * WM + 16 * a + lg + 4 * v;
                    double wv[NW > 0 ? NW : 1], wt[NT > 0 ? NT : 1];
#pragma unroll
                    for (int c = 0; c < NW; ++c) wv[c] = p.W[c][i];
#pragma unroll
                    for (int t = 0; t < NT; ++t) wt[t] = p.WT[t][i];
                    const double f0 = p.f[i];
#pragma unroll
                    for (int b = 0; b < WN; ++b) {
                         const v4d *q = acc[0][a][b];
                         double Pp[4], Pm[4];
                         stokes(q[0][v], q[1][v], q[2][v], q[3][v], Pp);
                         stokes(q[4][v], q[5][v], q[6][v], q[7][v], Pm);
#pragma unroll
                         for (int c = 0; c < NW; ++c)
#pragma unroll
                              for (int k = 0; k < 4; ++k) {
                                   ew[c][b][k] += wv[c] * Pp[k];
                                   ew[c][b][4 + k] += wv[c] * Pm[k];
                              }
#pragma unroll
                         for (int t = 0; t < NT; ++t) {
                              const v4d *d = acc[1 + t][a][b];
                              double dPp[4], dPm[4];
                              stokes_tan(q[0][v], q[1][v], q[2][v], q[3][v], d[0][v], d[1][v], d[2][v], d[3][v], dPp);
                              stokes_tan(q[4][v], q[5][v], q[6][v], q[7][v], d[4][v], d[5][v], d[6][v], d[7][v], dPm);
#pragma unroll
                              for (int k = 0; k < 4; ++k) {
                                   ew[NW + t][b][k] += wt[t] * Pp[k] + f0 * dPp[k];
                                   ew[NW + t][b][4 + k] += wt[t] * Pm[k] + f0 * dPm[k];
                              }
                         }
                         if (p.sst) {
                              const int j = jt0 + wn * 16 * WN + 16 * b + l16;
                              double *o = p.sst + ((size_t) i * p.Hp + j) * 8;
#pragma unroll
                              for (int c = 0; c < 8; ++c) o[c] = q[c][v];
                         }
                         if (p.save2 && i < p.N) {
                              const int j = jt0 + wn * 16 * WN + 16 * b + l16;
                              if (j < p.H) {
#pragma unroll
                                   for (int k = 0; k < 4; ++k) {
                                        p.save2[((size_t) k * p.N + i) * p.Q + p.H + j] = Pp[k];
                                        p.save2[((size_t) k * p.N + i) * p.Q + p.H - 1 - j] = Pm[k];
                                   }
                              }
                         }
                    }
               }
     }
     // reduce over the 4 lane groups (same angle column), then over the two
     // waves sharing the columns, then write this block's partial
#pragma unroll
     for (int c = 0; c < NW + NT; ++c)
#pragma unroll
          for (int b = 0; b < WN; ++b)
#pragma unroll
               for (int q = 0; q < 8; ++q) {
                    double s = ew[c][b][q];
                    s += __shfl_xor(s, 16);
                    s += __shfl_xor(s, 32);
                    ew[c][b][q] = s;
               }
     __syncthreads();
     if (wm == 1 && lg == 0) {
#pragma unroll
          for (int c = 0; c < NW + NT; ++c)
#pragma unroll
               for (int b = 0; b < WN; ++b)
#pragma unroll
                    for (int q = 0; q < 8; ++q) sR[((c * WN + b) * 8 + q) * 32 + wn * 16 + l16] = ew[c][b][q];
     }
     __syncthreads();
     if (wm == 0 && lg == 0) {
          const int nch = NW + NT;
          double *out = p.part + (size_t) blockIdx.y * nch * 4 * p.Q;
#pragma unroll
          for (int b = 0; b < WN; ++b) {
               const int j = jt0 + wn * 16 * WN + 16 * b + l16;
               if (j < p.H) {
#pragma unroll
                    for (int c = 0; c < NW + NT; ++c)
#pragma unroll
                         for (int k = 0; k < 4; ++k) {
                              const double vp = ew[c][b][k] + sR[((c * WN + b) * 8 + k) * 32 + wn * 16 + l16];
                              const double vm = ew[c][b][4 + k] + sR[((c * WN + b) * 8 + 4 + k) * 32 + wn * 16 + l16];
                              out[((size_t) c * 4 + k) * p.Q + p.H + j] = vp;
                              out[((size_t) c * 4 + k) * p.Q + p.H - 1 - j] = vm;
                         }
               }
          }
     }
}


// Staging of one LDS stage of a blocked split operand: plane-major rows of
// `cols` 16-byte columns, [plane][group of 8][column], `rows` = planes x G rows
// per stage.  Thread tid of NT copies elements tid + NT t; with R = NT / cols
// rows per step, element t of thread tid sits at the thread's base address
// plus a compile-time offset in (plane, group), and each stage adds one
// uniform stride, so no address arithmetic runs inside the stage loop.
template <int NT, int COLS, int G>
struct Stager {
     static constexpr int R = NT / COLS;
     static_assert(NT % COLS == 0 && (R % G == 0 || G % R == 0), "staging shape");
     const unsigned short *base;    // thread's element 0 at stage 0
     size_t plane, row;             // strides in unsigned shorts: plane, group of 8
     __device__ Stager(const unsigned short *src, size_t plane_, size_t row_, int col0, int tid)
         : plane(plane_), row(row_) {
          const int r = tid / COLS, col = tid % COLS;
          base = src + (size_t) (r / G) * plane + (size_t) (r % G) * row + (size_t) (col0 + col) * 8;
     }
     __device__ __forceinline__ const v4u *at(int t, int stage) const {
          return (const v4u *) (base + (size_t) ((R * t) / G) * plane + (size_t) ((R * t) % G) * row +
                                (size_t) stage * G * row);
     }
};

// The split 16-bit contraction on the 16x16x32 instruction: the four complex
// sums of the FP64 contraction with the size operand in two 16-bit parts
// (A = Ah + Al, 22 bits) and the angle operand in three (B = Bh + Bm + Bl,
// 33 bits), on the matrix cores with fp32 accumulation:
//   NP = 5:  A B ~ Ah Bh + [Ah Bm + Al Bh + Ah Bl + Al Bm]   (dropped: Al Bl ~ 2^-34)
//   NP = 3:  A B ~ Ah Bh + [Ah Bm + Al Bh]                   (dropped: ~ 2^-24 of A B)
// The leading product and the bracketed corrections accumulate separately and
// meet in FP64 in the epilogue: an fp32 accumulator rounds at every product
// it takes, relative to its own running value, so the corrections (2^-11 of
// the sums) would otherwise add four roundings of the sums' size per stage.
// The angle operand is the same for every size point, so its rounding errors
// add up coherently over the size integral (two parts of B leave expansion
// coefficients with relative errors of a few 1e-5); the size operand's
// rounding is independent from size to size.
// One workgroup of four waves (one per SIMD) per (size tile of 64, angle tile
// of 64) task, two LDS stages of 32 orders filled straight from memory
// (global_load_lds_dwordx4: a wave moves one 1 KiB row, 64 lanes x 16 bytes,
// into LDS without registers) while the matrix cores work on the other stage,
// one barrier per stage.  Waves in a 2 x 2 grid, each 32 sizes x 32 angles
// (2 x 2 blocks of 16 x 16), 8 sums per element: at +mu and at -mu, whose
// products flip the sign of the odd orders of the size operand
// (s_n = (-1)^(n-1)).  The (size tile, angle tile) of each workgroup comes
// from a host-built list dealt over the XCDs (xcd_tasks).  Partials per size
// tile: part[tile][NW][4][Q].
struct Mix2Args {
     int Np, Kb, Hb, H, Q;
     const unsigned short *A16;     // [8][Kb][Np][8]: S re h/l, S im h/l, T re h/l, T im h/l
     const unsigned short *B16;     // [6][Kb][Hb][8]: (pi+tau) h/m/l, (pi-tau) h/m/l
     const int *ktile;              // orders per size tile of 64 (multiple of KM)
     const int2 *task;              // (size tile, angle tile) per workgroup, (-1, -1) = none
     const double *W[4];            // weight channels [Np]
     double *part;                  // [nT][NW][4][Q]
};

// Shapes of the forward contraction (environment LMIE_MIX_SHAPE, read at every
// call).  A workgroup owns TM sizes x TN angles; its NT / 64 waves form a
// WM x WN grid, each wave 16 MA sizes x 16 NB angles; SEP: the corrections
// have accumulators of their own (else every product goes to one
// accumulator per sum); ST LDS stages, loads ST - 1 stages ahead; OCC
// workgroups per CU.
//   0: 64 x 128, eight waves (two per SIMD), one accumulator, 2 x 80 KiB
//   1: 64 x 64, four waves (one per SIMD), separate corrections, 2 x 56 KiB
//   2: 32 x 64, four waves, separate corrections, 2 x 40 KiB, two workgroups per CU
//   3: 64 x 64, eight waves (two per SIMD), separate corrections, 2 x 56 KiB
// DT > 0: every stage also touches each 128-byte line of the stage DT + 1
// ahead (one 4-byte direct load per line into a scratch slot of LDS that is
// never read), so that line is in L2 when its stage is loaded; the stage loads
// then wait for an L2 hit instead of a trip to memory.
constexpr int MIX_G = KM / 8;
constexpr int MIX_SHAPE_DEFAULT = 3;
struct MixS0 { static constexpr int NT = 512, TM = 64, TN = 128, WM = 2, WN = 4, MA = 2, NB = 2, ST = 2, OCC = 1, DT = 0; static constexpr bool SEP = false; };
struct MixS1 { static constexpr int NT = 256, TM = 64, TN = 64, WM = 2, WN = 2, MA = 2, NB = 2, ST = 2, OCC = 1, DT = 0; static constexpr bool SEP = true; };
struct MixS2 { static constexpr int NT = 256, TM = 32, TN = 64, WM = 2, WN = 2, MA = 1, NB = 2, ST = 2, OCC = 2, DT = 0; static constexpr bool SEP = true; };
struct MixS3 { static constexpr int NT = 512, TM = 64, TN = 64, WM = 4, WN = 2, MA = 1, NB = 2, ST = 2, OCC = 1, DT = 0; static constexpr bool SEP = true; };
template <class S> constexpr int mix_sa() { return 8 * MIX_G * S::TM; }              // v4u of A per stage
template <class S> constexpr int mix_sb() { return B16_PLANES * MIX_G * S::TN; }     // v4u of B per stage
// direct loads one wave issues per stage (its vmcnt events)
template <class S> constexpr int mix_loads() {
     return (8 * MIX_G * ((S::TM + 63) / 64) + B16_PLANES * MIX_G * ((S::TN + 63) / 64)) / (S::NT / 64);
}

typedef __attribute__((address_space(3))) void *lds_ptr;

// One wave's share of a stage: the A rows (plane, group) of the size tile and
// the B rows of the angle tile, cut into segments of at most 64 columns of 16
// bytes; wave w takes segments w, w + NT / 64, ...  Each segment is one direct
// load (lane l: column l of the segment).
// lines a stage touches, and direct loads per wave for them
template <class S> constexpr int mix_touch_lines() { return 8 * MIX_G * S::TM / 8 + B16_PLANES * MIX_G * S::TN / 8; }
template <class S> constexpr int mix_touch_loads() { return (mix_touch_lines<S>() + S::NT - 1) / S::NT; }

// Touch of stage c: thread tid loads one dword of lines tid, tid + NT, ...
// (the last load of a thread repeats a line when the count is not a multiple)
template <class S>
__device__ __forceinline__ void mix_stage_touch(const Mix2Args &p, unsigned *sink, int c, int it0, int jt0, int tid) {
     constexpr int LA = 8 * MIX_G * S::TM / 8, L = mix_touch_lines<S>();
     const size_t Np = p.Np, Hb = p.Hb;
     const size_t pA = (size_t) p.Kb * Np * 8, pB = (size_t) p.Kb * Hb * 8;
#pragma unroll
     for (int t = 0; t < mix_touch_loads<S>(); ++t) {
          int li = tid + S::NT * t;
          li = li < L ? li : L - 1;
          const unsigned short *src;
          if (li < LA) {
               const int row = li / (S::TM / 8), seg = li % (S::TM / 8), q = row / MIX_G, g = row % MIX_G;
               src = p.A16 + q * pA + ((size_t) (c * MIX_G + g) * Np + it0 + 8 * seg) * 8;
          } else {
               const int lb = li - LA, row = lb / (S::TN / 8), seg = lb % (S::TN / 8), q = row / MIX_G, g = row % MIX_G;
               src = p.B16 + q * pB + ((size_t) (c * MIX_G + g) * Hb + jt0 + 8 * seg) * 8;
          }
          __builtin_amdgcn_global_load_lds((const void *) src, (lds_ptr) sink, 4, 0, 0);
     }
}

template <class S>
__device__ __forceinline__ void mix_stage_load(const Mix2Args &p, v4u *sA, v4u *sB, int c, int it0, int jt0,
                                               int wave, int lane) {
     constexpr int NWAVE = S::NT / 64, SEGA = (S::TM + 63) / 64, SEGB = (S::TN + 63) / 64;
     constexpr int LA = S::TM < 64 ? S::TM : 64, LB = S::TN < 64 ? S::TN : 64;
     const size_t Np = p.Np, Hb = p.Hb;
     const size_t pA = (size_t) p.Kb * Np * 8, pB = (size_t) p.Kb * Hb * 8;
     if (lane < LA) {
#pragma unroll
          for (int t = 0; t < 8 * MIX_G * SEGA / NWAVE; ++t) {
               const int r = wave + NWAVE * t, row = r / SEGA, seg = r % SEGA, q = row / MIX_G, g = row % MIX_G;
               const unsigned short *src = p.A16 + q * pA + ((size_t) (c * MIX_G + g) * Np + it0 + 64 * seg + lane) * 8;
               __builtin_amdgcn_global_load_lds((const void *) src, (lds_ptr) (sA + row * S::TM + 64 * seg), 16, 0, 0);
          }
     }
     if (lane < LB) {
#pragma unroll
          for (int t = 0; t < B16_PLANES * MIX_G * SEGB / NWAVE; ++t) {
               const int r = wave + NWAVE * t, row = r / SEGB, seg = r % SEGB, q = row / MIX_G, g = row % MIX_G;
               const unsigned short *src = p.B16 + q * pB + ((size_t) (c * MIX_G + g) * Hb + jt0 + 64 * seg + lane) * 8;
               __builtin_amdgcn_global_load_lds((const void *) src, (lds_ptr) (sB + row * S::TN + 64 * seg), 16, 0, 0);
          }
     }
}

// The products of four sums (x_k times b_k, k = 0..3; x = (h, l) parts,
// b = (h, m, l) parts), interleaved so that consecutive instructions are
// independent: the leading product into m, the corrections into c.
template <class E, int NP>
__device__ __forceinline__ void mix_products(const v4u *x0, const v4u *b0, const v4u *x1, const v4u *b1,
                                             const v4u *x2, const v4u *b2, const v4u *x3, const v4u *b3,
                                             v4f *m, v4f *c) {
     m[0] = E::mfma(x0[0], b0[0], m[0]); m[1] = E::mfma(x1[0], b1[0], m[1]);
     m[2] = E::mfma(x2[0], b2[0], m[2]); m[3] = E::mfma(x3[0], b3[0], m[3]);
     c[0] = E::mfma(x0[0], b0[1], c[0]); c[1] = E::mfma(x1[0], b1[1], c[1]);
     c[2] = E::mfma(x2[0], b2[1], c[2]); c[3] = E::mfma(x3[0], b3[1], c[3]);
     c[0] = E::mfma(x0[1], b0[0], c[0]); c[1] = E::mfma(x1[1], b1[0], c[1]);
     c[2] = E::mfma(x2[1], b2[0], c[2]); c[3] = E::mfma(x3[1], b3[0], c[3]);
     if constexpr (NP >= 4) {
          c[0] = E::mfma(x0[0], b0[2], c[0]); c[1] = E::mfma(x1[0], b1[2], c[1]);
          c[2] = E::mfma(x2[0], b2[2], c[2]); c[3] = E::mfma(x3[0], b3[2], c[3]);
     }
     if constexpr (NP == 5) {
          c[0] = E::mfma(x0[1], b0[1], c[0]); c[1] = E::mfma(x1[1], b1[1], c[1]);
          c[2] = E::mfma(x2[1], b2[1], c[2]); c[3] = E::mfma(x3[1], b3[1], c[3]);
     }
}

template <class E, int NP, int NW, class S>
__device__ __forceinline__ void angular_mix_body(const Mix2Args &p, int2 tk) {
     constexpr int NT = S::NT, TM = S::TM, TN = S::TN, G = MIX_G, MA = S::MA, NB = S::NB, ST = S::ST;
     constexpr int SA = mix_sa<S>(), SB = mix_sb<S>();
     __shared__ v4u lds[ST * (SA + SB) + (S::DT > 0 ? 16 : 0)];
     static_assert(sizeof(lds) >= sizeof(double) * S::WM * 4 * 8 * TN, "epilogue buffer");
     unsigned *sink = (unsigned *) (lds + ST * (SA + SB));   // touch loads land here (never read)
     static_assert(S::WM * 16 * MA == TM && S::WN * 16 * NB == TN && S::WM * S::WN == NT / 64, "wave grid");

     if (tk.x < 0) return;
     const int st = tk.x, jt = tk.y;
     const int it0 = st * TM, jt0 = jt * TN, nch = p.ktile[st] / KM;
     const int lane = threadIdx.x & 63;
     const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
     const int wm = wave % S::WM, wn = wave / S::WM;
     const int l16 = lane & 15, lg = lane >> 4;
     // the first weight channel of this lane's sizes, loaded now so the
     // epilogue does not wait for it
     double w0[MA][4];
#pragma unroll
     for (int a = 0; a < MA; ++a)
#pragma unroll
          for (int v = 0; v < 4; ++v) w0[a][v] = p.W[0][it0 + wm * 16 * MA + 16 * a + mfma16_row(lane, v)];
     const v4u flip = {0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u};
     // stage buffer b: A rows at lds + b (SA + SB), B rows after them
     auto sA = [&](int b_) { return lds + b_ * (SA + SB); };
     auto sB = [&](int b_) { return lds + b_ * (SA + SB) + SA; };

     // leading product and corrections (the same accumulators without SEP)
     v4f am[MA][NB][8], ac_[S::SEP ? MA : 1][S::SEP ? NB : 1][8];
#pragma unroll
     for (int a = 0; a < MA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b)
#pragma unroll
               for (int k = 0; k < 8; ++k) am[a][b][k] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
     for (int a = 0; a < (S::SEP ? MA : 1); ++a)
#pragma unroll
          for (int b = 0; b < (S::SEP ? NB : 1); ++b)
#pragma unroll
               for (int k = 0; k < 8; ++k) ac_[a][b][k] = v4f{0.f, 0.f, 0.f, 0.f};
     auto acc_c = [&](int a, int b) -> v4f * { if constexpr (S::SEP) return ac_[a][b]; else return am[a][b]; };

     // stages 0 .. ST-2 in flight before the loop; stage c + ST - 1 is issued
     // at stage c into the buffer of stage c - 1
#pragma unroll
     for (int c = 0; c < ST - 1; ++c)
          if (c < nch) mix_stage_load<S>(p, sA(c), sB(c), c, it0, jt0, wave, lane);
     if constexpr (S::DT > 0)
          if (nch > 0) mix_stage_touch<S>(p, sink, min(ST - 1 + S::DT, nch - 1), it0, jt0, threadIdx.x);
     for (int c = 0; c < nch; ++c) {
          const int buf = c % ST;
          // this wave's rows of stage c have landed (the younger stages may
          // still be in flight); after the barrier every wave's have, and
          // every wave is done with stage c - 1
          // (the touch loads issued after them may stay in flight)
          if constexpr (S::DT > 0) __builtin_amdgcn_s_waitcnt(0x0F70 | mix_touch_loads<S>());
          else if (ST == 3 && c + 1 < nch) __builtin_amdgcn_s_waitcnt(0x0F70 | mix_loads<S>());
          else __builtin_amdgcn_s_waitcnt(0);
          // the plain barrier: each wave has waited for its own rows above, and
          // a workgroup fence here would also wait for the younger stages
          __builtin_amdgcn_s_barrier();
          __asm__ volatile("" ::: "memory");
          if (c + ST - 1 < nch) {
               const int nb = (c + ST - 1) % ST;
               mix_stage_load<S>(p, sA(nb), sB(nb), c + ST - 1, it0, jt0, wave, lane);
          }
          if constexpr (S::DT > 0)
               mix_stage_touch<S>(p, sink, min(c + ST + S::DT, nch - 1), it0, jt0, threadIdx.x);
          // B fragments of the angle blocks: (pi+tau) h/m/l, (pi-tau) h/m/l
          v4u bp[NB][3], bm[NB][3];
#pragma unroll
          for (int b = 0; b < NB; ++b) {
               const int jc = wn * 16 * NB + 16 * b + l16;
#pragma unroll
               for (int q = 0; q < 3; ++q) {
                    bp[b][q] = sB(buf)[(q * G + lg) * TN + jc];
                    bm[b][q] = sB(buf)[((3 + q) * G + lg) * TN + jc];
               }
          }
#pragma unroll
          for (int a = 0; a < MA; ++a) {
               const int ic = wm * 16 * MA + 16 * a + l16;
               v4u x[8];
#pragma unroll
               for (int k = 0; k < 8; ++k) x[k] = sA(buf)[(k * G + lg) * TM + ic];
               // +mu: sp = S (pi+tau), sm = T (pi-tau)
#pragma unroll
               for (int b = 0; b < NB; ++b)
                    mix_products<E, NP>(x + 0, bp[b], x + 2, bp[b], x + 4, bm[b], x + 6, bm[b], am[a][b], acc_c(a, b));
               // -mu: the odd orders change sign, sp = s S (pi-tau), sm = s T (pi+tau)
#pragma unroll
               for (int k = 0; k < 8; ++k) x[k] ^= flip;
#pragma unroll
               for (int b = 0; b < NB; ++b)
                    mix_products<E, NP>(x + 0, bm[b], x + 2, bm[b], x + 4, bp[b], x + 6, bp[b], am[a][b] + 4,
                                        acc_c(a, b) + 4);
          }
     }

     // epilogue: the sums in double (leading product plus corrections), the
     // Stokes elements, weighted sums over this lane's sizes per angle; the
     // partial sums of every (wave row block, lane group) go to LDS and one
     // thread per output adds them (no cross-lane shuffles), one weight
     // channel at a time (the Stokes elements are recomputed per channel), so
     // few doubles are live beside the accumulators
     const double unscale = 1. / (kSA * kSB);
     double *sR = (double *) lds;            // [WM][4 lane groups][8][TN]
     __syncthreads();                        // every wave is done with the stages
     double *out = p.part + (size_t) st * NW * 4 * p.Q;
#pragma unroll
     for (int c = 0; c < NW; ++c) {
          if (c > 0) __syncthreads();        // the previous channel's sums are read
#pragma unroll
          for (int b = 0; b < NB; ++b) {
               const int jl = wn * 16 * NB + 16 * b + l16;
               double ew[8];
#pragma unroll
               for (int k = 0; k < 8; ++k) ew[k] = 0.;
#pragma unroll
               for (int a = 0; a < MA; ++a)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                         const v4f *qm = am[a][b], *qc = acc_c(a, b);
                         double q[8];
#pragma unroll
                         for (int k = 0; k < 8; ++k)
                              q[k] = (S::SEP ? (double) qm[k][v] + (double) qc[k][v] : (double) qm[k][v]) * unscale;
                         const int i = it0 + wm * 16 * MA + 16 * a + mfma16_row(lane, v);
                         double Pp[4], Pm[4];
                         stokes(q[0], q[1], q[2], q[3], Pp);
                         stokes(q[4], q[5], q[6], q[7], Pm);
                         const double wv = c == 0 ? w0[a][v] : p.W[c][i];
#pragma unroll
                         for (int k = 0; k < 4; ++k) {
                              ew[k] += wv * Pp[k];
                              ew[4 + k] += wv * Pm[k];
                         }
                    }
#pragma unroll
               for (int k = 0; k < 8; ++k) sR[((wm * 4 + lg) * 8 + k) * TN + jl] = ew[k];
          }
          __syncthreads();
          for (int e = threadIdx.x; e < 8 * TN; e += NT) {
               const int jl = e % TN, k = e / TN, j = jt0 + jl;
               if (j >= p.H) continue;
               double v = 0.;
#pragma unroll
               for (int m = 0; m < S::WM * 4; ++m) v += sR[(m * 8 + k) * TN + jl];
               // k < 4: direct angle H + j; k >= 4: mirrored angle H - 1 - j
               out[((size_t) c * 4 + (k & 3)) * p.Q + (k < 4 ? p.H + j : p.H - 1 - j)] = v;
          }
     }
}
template <class E, int NP, int NW, class S>
__global__ __launch_bounds__(S::NT) __attribute__((amdgpu_waves_per_eu(S::NT / 256 * S::OCC)))
void angular_mix(Mix2Args p) {
     angular_mix_body<E, NP, NW, S>(p, p.task[blockIdx.x]);
}
// angular_mix over the cases of a batch: one task list for all cases, task
// (size tile, angle tile | case << 16), padding tasks (-1, case << 16).
template <class E, int NP, int NW, class S>
__global__ __launch_bounds__(S::NT) __attribute__((amdgpu_waves_per_eu(S::NT / 256 * S::OCC)))
void angular_mix_b(const Mix2Args *cases, const int2 *task) {
     const int2 t = task[blockIdx.x];
     const Mix2Args p = cases[t.y >> 16];
     angular_mix_body<E, NP, NW, S>(p, make_int2(t.x, t.y & 0xffff));
}


// ---------------------------------------------------------------------------
// Reverse pass.
//
// Transpose of the expansion (misc/rtutil_scat_int.c:63-230 read backwards):
// from the cotangents of the raw coefficients to the cotangents of the raw
// matrix elements F11, F33, F12, F34 at the Gauss angles.  Two kernels: the
// sums over l in chunks of EXA_L (one thread per positive node j and chunk,
// partials [chunk][24][Hp]), then the fixed-order sum over chunks.
constexpr int EXA_L = 32;

__device__ __forceinline__ void expand_adjoint_part_body(const double *T, const double *gbar, const double *lbar, int H, int Hp, int C, double *part) {
     const int j = blockIdx.x * blockDim.x + threadIdx.x, ck = blockIdx.y;
     if (j >= Hp) return;
     const size_t plane = (size_t) C * Hp;
     double Em[6] = {0, 0, 0, 0, 0, 0}, Ep[6] = {0, 0, 0, 0, 0, 0};   // GSF rows
     double Lm[6] = {0, 0, 0, 0, 0, 0}, Lp[6] = {0, 0, 0, 0, 0, 0};   // Legendre rows
     const int l1 = min(C, (ck + 1) * EXA_L);
     for (int l = ck * EXA_L; l < l1 && j < H; ++l) {
          const double sl = (l & 1) ? -1. : 1., fl = (2 * l + 1) / 2.;
          const size_t o = (size_t) l * Hp + j;
          if (gbar) {
               double cb[6];
               for (int k = 0; k < 6; ++k) cb[k] = gbar[(size_t) k * C + l] * fl;
               const double e1 = (cb[1] + cb[2]) / 2., e2 = (cb[1] - cb[2]) / 2.;
               const double p00 = T[o], p02 = T[plane + o], p22 = T[2 * plane + o], p2m = T[3 * plane + o];
               Em[0] += sl * cb[0] * p00;  Ep[0] += cb[0] * p00;
               Em[1] += sl * e1 * p2m;     Ep[1] += e1 * p22;
               Em[2] += sl * e2 * p22;     Ep[2] += e2 * p2m;
               Em[3] += sl * cb[3] * p00;  Ep[3] += cb[3] * p00;
               Em[4] += sl * cb[4] * p02;  Ep[4] += cb[4] * p02;
               Em[5] += sl * cb[5] * p02;  Ep[5] += cb[5] * p02;
          }
          if (lbar) {
               const double pl = T[4 * plane + o];
               for (int k = 0; k < 6; ++k) {
                    const double c = lbar[(size_t) k * C + l] * fl * (k == 5 ? -1. : 1.);
                    Lm[k] += sl * c * pl;
                    Lp[k] += c * pl;
               }
          }
     }
     double *out = part + (size_t) ck * 24 * Hp + j;
     for (int k = 0; k < 6; ++k) {
          out[(size_t) k * Hp] = Em[k];
          out[(size_t) (6 + k) * Hp] = Ep[k];
          out[(size_t) (12 + k) * Hp] = Lm[k];
          out[(size_t) (18 + k) * Hp] = Lp[k];
     }
}
__global__ __launch_bounds__(64) void expand_adjoint_part(const double *T, const double *gbar, const double *lbar, int H, int Hp, int C, double *part) { expand_adjoint_part_body(T, gbar, lbar, H, Hp, C, part); }
struct expand_adjoint_part_B {
     const double *T;
     const double *gbar;
     const double *lbar;
     int H;
     int Hp;
     int C;
     double *part;
     unsigned gx, gy;
};
__global__ __launch_bounds__(64) void expand_adjoint_part_b(const expand_adjoint_part_B *cases) {
     const expand_adjoint_part_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     expand_adjoint_part_body(a.T, a.gbar, a.lbar, a.H, a.Hp, a.C, a.part);
}


// Sums of the chunk partials: a workgroup of 32 nodes x 24 rows, one thread
// per (node, row) adding the chunks in order, then the row combinations per
// node through LDS.
constexpr int EXF_J = 32;
__device__ __forceinline__ void expand_adjoint_fin_body(const double *part, int nck, const double *w, int H, int Hp, int Q, double *pfbar) {
     __shared__ double sv[24][EXF_J];
     const int jj = threadIdx.x % EXF_J, k = threadIdx.x / EXF_J, j = blockIdx.x * EXF_J + jj;
     double v = 0.;
     if (j < H) {
          const double *pk = part + (size_t) k * Hp + j;
          const size_t step = (size_t) 24 * Hp;
          int ck = 0;
          for (; ck + 8 <= nck; ck += 8) {   // eight loads in flight, added in order
               double t[8];
#pragma unroll
               for (int e = 0; e < 8; ++e) t[e] = pk[(size_t) (ck + e) * step];
#pragma unroll
               for (int e = 0; e < 8; ++e) v += t[e];
          }
          for (; ck < nck; ++ck) v += pk[(size_t) ck * step];
     }
     sv[k][jj] = v;
     __syncthreads();
     if (k || j >= H) return;
     double vv[24];
     for (int r = 0; r < 24; ++r) vv[r] = sv[r][jj];
     const int ip = H + j, im = H - 1 - j;
     const double wq = w[ip];
     for (int s = 0; s < 2; ++s) {
          const double *E = vv + (s ? 6 : 0), *L = vv + (s ? 18 : 12);
          const int a = s ? ip : im;
          pfbar[a] = wq * (E[0] + E[1] + E[2] + L[0] + L[1]);
          pfbar[Q + a] = wq * (E[1] - E[2] + E[3] + L[2] + L[3]);
          pfbar[2 * Q + a] = wq * (E[4] + L[4]);
          pfbar[3 * Q + a] = wq * (E[5] + L[5]);
     }
}
__global__ __launch_bounds__(24 * EXF_J) void expand_adjoint_fin(const double *part, int nck, const double *w, int H, int Hp, int Q, double *pfbar) { expand_adjoint_fin_body(part, nck, w, H, Hp, Q, pfbar); }
struct expand_adjoint_fin_B {
     const double *part;
     int nck;
     const double *w;
     int H;
     int Hp;
     int Q;
     double *pfbar;
     unsigned gx, gy;
};
__global__ __launch_bounds__(24 * EXF_J) void expand_adjoint_fin_b(const expand_adjoint_fin_B *cases) {
     const expand_adjoint_fin_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     expand_adjoint_fin_body(a.part, a.nck, a.w, a.H, a.Hp, a.Q, a.pfbar);
}


// Cotangents of the four complex sums at one (size, angle) from those of the
// four matrix elements (gradient convention d/dRe + i d/dIm):
//   sp^ = (p0 + p1) sp/2 - (p2 + i p3) sm/2,  sm^ = (p0 - p1) sm/2 - (p2 - i p3) sp/2,
// the adjoint of P11 = (|sp|^2 + |sm|^2)/4, P33 = (|sp|^2 - |sm|^2)/4,
// P12 = -Re(sp sm*)/2, P34 = -Im(sp sm*)/2 (src/lmie_core.c:726-744).
__device__ __forceinline__ void stokes_adj(double spr, double spi, double smr, double smi, const double pb[4],
                                           double out[4]) {
     const double a = .5 * (pb[0] + pb[1]), c = .5 * (pb[0] - pb[1]), p2 = .5 * pb[2], p3 = .5 * pb[3];
     out[0] = a * spr - (p2 * smr - p3 * smi);
     out[1] = a * spi - (p2 * smi + p3 * smr);
     out[2] = c * smr - (p2 * spr + p3 * spi);
     out[3] = c * smi - (p2 * spi - p3 * spr);
}

// Angular part of the size-weight cotangent: fbar_i = sum_{j, +-, k} pfbar_k P_k(i, j),
// one wave per size point.
__global__ __launch_bounds__(256) void fbar_angular(const double *sst, const double *pfbar, int N, int H, int Hp,
                                                    int Q, double *fang) {
     const int lane = threadIdx.x & 63;
     const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
     if (i >= N) return;
     double acc = 0.;
     for (int j = lane; j < H; j += 64) {
          const double *q = sst + ((size_t) i * Hp + j) * 8;
          double Pp[4], Pm[4];
          stokes(q[0], q[1], q[2], q[3], Pp);
          stokes(q[4], q[5], q[6], q[7], Pm);
          for (int k = 0; k < 4; ++k)
               acc += pfbar[k * Q + H + j] * Pp[k] + pfbar[k * Q + H - 1 - j] * Pm[k];
     }
     for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
     if (lane == 0) fang[i] = acc;
}

// Transposed angular contraction on FP64 matrix cores: for every size i and
// order n,
//   Sbar_n = f_i sum_j [sp^+ (pi+tau)_n + s_n sp^- (pi-tau)_n],
//   Tbar_n = f_i sum_j [sm^+ (pi-tau)_n + s_n sm^- (pi+tau)_n],
// K = the positive angles j (16 per LDS stage), the cotangents of the sums
// formed while staging from the checkpointed sums.  Tile 32 sizes x 32 orders.
struct AdjArgs {
     int Np, Kp, Hp, H, Q;
     const double *sst;         // [Np][Hp][8]
     const double *pfbar;       // [4][Q]
     const double *Bp, *Bm;     // [Kp][Hp]
     const double *f;           // [Np]
     const int *ktile;          // orders per size tile of 32
     double *bar;               // [4][Kp][Np]: Sbar re, im, Tbar re, im
};

__global__ __launch_bounds__(256) void angular_adj_f64(AdjArgs p) {
     constexpr int JC = 16, TM = 32, TN = 32, SA = TM + 16, SB = TN + 16;
     __shared__ double sA[8 * JC * SA];
     __shared__ double sB[2 * JC * SB];
     const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
     const int wm = wave & 1, wn = wave >> 1;
     const int l16 = lane & 15, lg = lane >> 4;
     const int n0 = blockIdx.x * TN, it0 = blockIdx.y * TM;
     if (n0 >= p.ktile[blockIdx.y]) return;          // whole tile beyond these sizes' orders
     const size_t Np = p.Np, Hp = p.Hp, plane = (size_t) p.Kp * Np;
     v4d acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
     const int ncol = wn * 16 + l16;                  // this lane's output order (B column)
     const double sg = ((n0 + ncol) & 1) ? -1. : 1.;   // s_n, n = n0 + ncol + 1
     for (int j0 = 0; j0 < p.H; j0 += JC) {
          __syncthreads();
          // stage the cotangents of the sums: 32 sizes x 16 angles
          for (int e = threadIdx.x; e < TM * JC; e += 256) {
               const int ii = e % TM, jj = e / TM;
               const int i = it0 + ii, j = j0 + jj;
               double out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
               if (j < p.H) {
                    const double *q = p.sst + ((size_t) i * Hp + j) * 8;
                    double pb[4], pm[4];
                    for (int k = 0; k < 4; ++k) {
                         pb[k] = p.pfbar[k * p.Q + p.H + j];
                         pm[k] = p.pfbar[k * p.Q + p.H - 1 - j];
                    }
                    stokes_adj(q[0], q[1], q[2], q[3], pb, out);      // sp^+, sm^+
                    stokes_adj(q[4], q[5], q[6], q[7], pm, out + 4);  // sp^-, sm^-
               }
               for (int c = 0; c < 8; ++c) sA[(c * JC + jj) * SA + ii] = out[c];
          }
          // stage pi +- tau: rows = angles, columns = orders n0 .. n0 + 31
          for (int e = threadIdx.x; e < TN * JC; e += 256) {
               const int nn = e % TN, jj = e / TN;
               const int j = j0 + jj;
               const size_t o = (size_t) (n0 + nn) * Hp + j;
               sB[jj * SB + nn] = j < p.H ? p.Bp[o] : 0.;
               sB[(JC + jj) * SB + nn] = j < p.H ? p.Bm[o] : 0.;
          }
          __syncthreads();
#pragma unroll
          for (int ks = 0; ks < JC / 4; ++ks) {
               const int r = ks * 4 + lg;
               const int ic = wm * 16 + l16;
               const double bp = sB[r * SB + ncol], bm = sB[(JC + r) * SB + ncol];
               const double bps = sg * bp, bms = sg * bm;
               const double spr = sA[(0 * JC + r) * SA + ic], spi = sA[(1 * JC + r) * SA + ic];
               const double smr = sA[(2 * JC + r) * SA + ic], smi = sA[(3 * JC + r) * SA + ic];
               const double qpr = sA[(4 * JC + r) * SA + ic], qpi = sA[(5 * JC + r) * SA + ic];
               const double qmr = sA[(6 * JC + r) * SA + ic], qmi = sA[(7 * JC + r) * SA + ic];
               acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(spr, bp, acc[0], 0, 0, 0);
               acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(qpr, bms, acc[0], 0, 0, 0);
               acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(spi, bp, acc[1], 0, 0, 0);
               acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(qpi, bms, acc[1], 0, 0, 0);
               acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(smr, bm, acc[2], 0, 0, 0);
               acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(qmr, bps, acc[2], 0, 0, 0);
               acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(smi, bm, acc[3], 0, 0, 0);
               acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(qmi, bps, acc[3], 0, 0, 0);
          }
     }
     // result rows (sizes) l/16 + 4 v, column (order) l % 16
     for (int v = 0; v < 4; ++v) {
          const int i = it0 + wm * 16 + lg + 4 * v;
          const int n = n0 + ncol;            // row n of the planes holds order n + 1
          const double fi = p.f[i];
          for (int c = 0; c < 4; ++c) p.bar[c * plane + (size_t) n * Np + i] = fi * acc[c][v];
     }
}

// ---------------------------------------------------------------------------
// Reverse pass of the mixed path through the order-by-order kernel matrix.
//
// The cotangents of the four complex sums at an angle are linear in the sums
// (stokes_adj): at +mu_j sp^ = a_j sp - beta_j sm, sm^ = c_j sm - conj(beta_j) sp
// with a = (p0 + p1)/2, c = (p0 - p1)/2, beta = (p2 + i p3)/2 of that angle's
// matrix-element cotangents, and likewise (a', c', beta') at -mu_j.  Putting
// the forward sums sp = sum_m S_m (pi+tau)_m, ... into the transposed
// contraction gives, for every size,
//   Sbar_n = f (sum_m Gss(n,m) S_m - Gst(n,m) T_m)
//   Tbar_n = f (sum_m Gtt(n,m) T_m - conj(Gst(m,n)) S_m)
// with the size-independent matrices (P+- = pi +- tau at the positive nodes,
// s_n s_m = (-1)^(n+m) from the mirrored nodes)
//   Gss = sum_j a_j P+_n P+_m + s_n s_m a'_j P-_n P-_m,   Gtt = sum_j c_j P-_n P-_m + s_n s_m c'_j P+_n P+_m,
//   Gst = sum_j beta_j P+_n P-_m + s_n s_m beta'_j P-_n P+_m,
// and the angular part of the size-weight cotangent is the quadratic form
// sum_j pfbar . P(i, j) = 1/2 sum_n Re(conj S_n U_S,n + conj T_n U_T,n), U = Sbar / f.
// So the reverse contraction needs no per-(size, angle) checkpoint: five
// kernels build the weights, the split tables, G (split 16-bit, fp32
// accumulation), its split operand, and U = G A on the matrix cores.
struct AdjScal {
     double wmax;               // largest |weight| (weights are stored divided by it)
     unsigned gmax;             // bits of the largest |G| (non-negative float)
};

// Weights of the eight right-hand tables per positive node j (row q of wts):
// a, a', c, c', beta_r, beta_i, beta'_r, beta'_i, divided by their largest
// magnitude.  One workgroup.
__device__ __forceinline__ void adj_weights_body(const double *pfbar, int H, int Hq, int Q, double *wts, AdjScal *sc) {
     __shared__ double red[1024];
     double mx = 0.;
     for (int j = threadIdx.x; j < H; j += blockDim.x)
          for (int k = 0; k < 4; ++k)
               mx = fmax(mx, fmax(fabs(pfbar[k * Q + H + j]), fabs(pfbar[k * Q + H - 1 - j])));
     red[threadIdx.x] = mx;
     __syncthreads();
     for (int o = blockDim.x / 2; o > 0; o >>= 1) {
          if ((int) threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
          __syncthreads();
     }
     // |a|, |c| <= max(|p0|, |p1|) and |beta parts| <= |p| / 2: the bound of the pfbar
     // magnitudes bounds every weight
     const double wmax = red[0] > 0. ? red[0] : 1., iw = 1. / wmax;
     for (int j = threadIdx.x; j < Hq; j += blockDim.x) {
          double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          if (j < H) {
               const double *pp = pfbar + H + j, *pm = pfbar + H - 1 - j;
               v[0] = .5 * (pp[0] + pp[Q]);
               v[1] = .5 * (pm[0] + pm[Q]);
               v[2] = .5 * (pp[0] - pp[Q]);
               v[3] = .5 * (pm[0] - pm[Q]);
               v[4] = .5 * pp[2 * Q];
               v[5] = .5 * pp[3 * Q];
               v[6] = .5 * pm[2 * Q];
               v[7] = .5 * pm[3 * Q];
          }
          for (int q = 0; q < 8; ++q) wts[(size_t) q * Hq + j] = v[q] * iw;
     }
     if (threadIdx.x == 0) {
          sc->wmax = wmax;
          sc->gmax = 0u;
     }
}
__global__ __launch_bounds__(1024) void adj_weights(const double *pfbar, int H, int Hq, int Q, double *wts, AdjScal *sc) { adj_weights_body(pfbar, H, Hq, Q, wts, sc); }
struct adj_weights_B {
     const double *pfbar;
     int H;
     int Hq;
     int Q;
     double *wts;
     AdjScal *sc;
     unsigned gx, gy;
};
__global__ __launch_bounds__(1024) void adj_weights_b(const adj_weights_B *cases) {
     const adj_weights_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     adj_weights_body(a.pfbar, a.H, a.Hq, a.Q, a.wts, a.sc);
}


// Split tables with angles along the contraction: [plane][Hq / 8][Kq][8],
// 8 consecutive positive nodes per 16 bytes, from the forward's split tables
// B16 ([6][Kb][Hb][8], orders along the contraction): left (4 planes) P+ and
// P- of order r + 1 over r (r + 1) (r + 2) times kSB, high and middle parts
// (two of the B16 parts, transposed); right (16 planes) weight q times P+
// (q = 0, 3, 6, 7) or P- (q = 1, 2, 4, 5), from P = high + middle + low,
// high and low parts.  One thread per (order, group of 8 nodes).
template <class E>
__device__ __forceinline__ void adj_tables_body(const unsigned short *B16, int Kb, int Hb, const double *wts, int H, int Hq, int Kq, int n1max, unsigned short *L16, unsigned short *R16) {
     const int r = blockIdx.x * 64 + threadIdx.x, jg = blockIdx.y;
     if (r >= Kq) return;
     const size_t plane = (size_t) (Hq / 8) * Kq * 8, o = ((size_t) jg * Kq + r) * 8, bpl = (size_t) Kb * Hb * 8;
     unsigned short lv[4][8], rv[16][8];
     for (int e = 0; e < 8; ++e) {
          const int j = 8 * jg + e;
          double pp = 0., pm = 0.;
          if (j < H && r < n1max) {
               const size_t bo = ((size_t) (r / 8) * Hb + j) * 8 + (r % 8);
               lv[0][e] = B16[0 * bpl + bo];
               lv[1][e] = B16[1 * bpl + bo];
               lv[2][e] = B16[3 * bpl + bo];
               lv[3][e] = B16[4 * bpl + bo];
               pp = E::value(lv[0][e]) + E::value(lv[1][e]) + E::value(B16[2 * bpl + bo]);
               pm = E::value(lv[2][e]) + E::value(lv[3][e]) + E::value(B16[5 * bpl + bo]);
          } else {
#pragma unroll
               for (int q = 0; q < 4; ++q) lv[q][e] = 0;
          }
          for (int q = 0; q < 8; ++q) {
               const bool plus = q == 0 || q == 3 || q == 6 || q == 7;
               // the planes of the sigma-sigma products (q = 1, 3, 6, 7) carry
               // sigma_m = (-1)^r of their order (an exact sign change)
               const bool odd = (q == 1 || q == 3 || q == 6 || q == 7) && (r & 1);
               const double v = wts[(size_t) q * Hq + j] * (plus ? pp : pm);
               split16<E>(odd ? -v : v, &rv[2 * q][e], &rv[2 * q + 1][e]);
          }
     }
     for (int q = 0; q < 4; ++q) {
          v4u w;
          for (int k = 0; k < 4; ++k) w[k] = (unsigned) lv[q][2 * k] | ((unsigned) lv[q][2 * k + 1] << 16);
          *(v4u *) (L16 + q * plane + o) = w;
     }
     for (int q = 0; q < 16; ++q) {
          v4u w;
          for (int k = 0; k < 4; ++k) w[k] = (unsigned) rv[q][2 * k] | ((unsigned) rv[q][2 * k + 1] << 16);
          *(v4u *) (R16 + q * plane + o) = w;
     }
}
template <class E>
__global__ __launch_bounds__(64) void adj_tables(const unsigned short *B16, int Kb, int Hb, const double *wts, int H, int Hq, int Kq, int n1max, unsigned short *L16, unsigned short *R16) { adj_tables_body<E>(B16, Kb, Hb, wts, H, Hq, Kq, n1max, L16, R16); }
struct adj_tables_B {
     const unsigned short *B16;
     int Kb;
     int Hb;
     const double *wts;
     int H;
     int Hq;
     int Kq;
     int n1max;
     unsigned short *L16;
     unsigned short *R16;
     unsigned gx, gy;
};
template <class E>
__global__ __launch_bounds__(64) void adj_tables_b(const adj_tables_B *cases) {
     const adj_tables_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     adj_tables_body<E>(a.B16, a.Kb, a.Hb, a.wts, a.H, a.Hq, a.Kq, a.n1max, a.L16, a.R16);
}


// G on the 32x32x16 matrix instruction: one workgroup per 64 x 64 tile of
// (n, m), four waves of 32 x 32, contraction over the positive nodes in
// stages of 16 (two LDS stages, 80 KiB: two workgroups per CU).  Each G
// element is a sum of two products, the second with the factor
// sigma_n sigma_m = (-1)^(n+m): the right planes of those products carry
// sigma_m (adj_tables) and the left fragment is negated on the lanes of odd
// rows, so both products accumulate into one accumulator:
// 0 P+ a P+ + P- a' P-, 1 P- c P- + P+ c' P+, 2 P+ br P- + P- b'r P+,
// 3 P+ bi P- + P- b'i P+.  Output in fp32, Gf[4][Kq][Kq] = Gss, Gtt, Re Gst,
// Im Gst (row n, column m, scaled by kSB^2 / (wmax n (n+1) m (m+1))), and the
// largest magnitude.
template <class E>
__device__ __forceinline__ void adj_gbuild_body(const unsigned short *L16, const unsigned short *R16, int Hq, int Kq, float *Gf, AdjScal *sc,
                                                int tn, int tm) {
     constexpr int T = 64, G = 2, KS = 8 * G;
     constexpr int NL = 4 * G * T / 256, NR = 16 * G * T / 256;
     __shared__ v4u sL[2][4 * G][T];
     __shared__ v4u sR[2][16 * G][T];
     const int n0 = tn * T, m0 = tm * T, nch = Hq / KS;
     const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
     const int wa = wave & 1, wb = wave >> 1, l32 = lane & 31, lh = lane >> 5;
     const size_t plane = (size_t) (Hq / 8) * Kq * 8;
     // odd rows n (lane parity) negate the left fragment of the sigma-sigma products
     const unsigned sg = (lane & 1) ? 0x80008000u : 0u;
     const v4u neg = {sg, sg, sg, sg};
     v4u rl[NL], rr[NR];
     const Stager<256, T, G> stL(L16, plane, (size_t) Kq * 8, n0, threadIdx.x);
     const Stager<256, T, G> stR(R16, plane, (size_t) Kq * 8, m0, threadIdx.x);
     auto load = [&](int c) {
#pragma unroll
          for (int t = 0; t < NL; ++t) rl[t] = *stL.at(t, c);
#pragma unroll
          for (int t = 0; t < NR; ++t) rr[t] = *stR.at(t, c);
     };
     auto store = [&](int s) {
#pragma unroll
          for (int t = 0; t < NL; ++t) {
               const int e = threadIdx.x + 256 * t;
               sL[s][e / T][e % T] = rl[t];
          }
#pragma unroll
          for (int t = 0; t < NR; ++t) {
               const int e = threadIdx.x + 256 * t;
               sR[s][e / T][e % T] = rr[t];
          }
     };
     v16f acc[4];
#pragma unroll
     for (int k = 0; k < 4; ++k)
#pragma unroll
          for (int v = 0; v < 16; ++v) acc[k][v] = 0.f;
     load(0);
     store(0);
     if (nch > 1) load(1);
     for (int c = 0; c < nch; ++c) {
          const int s = c & 1;
          __syncthreads();
          // lane l: row (n or m) l % 32, nodes 8 (l / 32) + e of the stage
          v4u y[16], x[8];
#pragma unroll
          for (int q = 0; q < 16; ++q) y[q] = sR[s][q * G + lh][wb * 32 + l32];
          // left planes: 0 P+ hi, 1 P+ lo, 2 P- hi, 3 P- lo; 4..7 the same
          // negated on odd rows
#pragma unroll
          for (int q = 0; q < 4; ++q) x[q] = sL[s][q * G + lh][wa * 32 + l32];
#pragma unroll
          for (int q = 0; q < 4; ++q) x[4 + q] = x[q] ^ neg;
          if (c + 1 < nch) store(s ^ 1);
          if (c + 2 < nch) load(c + 2);
#define GB(k, L, R) acc[k] = E::mfma32(x[L], y[2 * (R)], acc[k]); acc[k] = E::mfma32(x[L], y[2 * (R) + 1], acc[k]); \
          acc[k] = E::mfma32(x[(L) + 1], y[2 * (R)], acc[k]);
          GB(0, 0, 0) GB(1, 2, 2) GB(2, 0, 4) GB(3, 0, 5)
          GB(0, 6, 1) GB(1, 4, 3) GB(2, 6, 6) GB(3, 6, 7)
#undef GB
     }
     float mx = 0.f;
#pragma unroll
     for (int v = 0; v < 16; ++v) {
          const int r = n0 + wa * 32 + mfma32_row(lane, v);
          const int cm = m0 + wb * 32 + l32;
          const float g0 = acc[0][v], g1 = acc[1][v], g2 = acc[2][v], g3 = acc[3][v];
          const size_t o = (size_t) r * Kq + cm, pl = (size_t) Kq * Kq;
          Gf[o] = g0;
          Gf[pl + o] = g1;
          Gf[2 * pl + o] = g2;
          Gf[3 * pl + o] = g3;
          mx = fmaxf(mx, fmaxf(fmaxf(fabsf(g0), fabsf(g1)), fmaxf(fabsf(g2), fabsf(g3))));
     }
     for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
     if (lane == 0) atomicMax(&sc->gmax, __float_as_uint(mx));
}
// Tile (row tile, column tile) of workgroup id L of a g x g tile grid.  The
// hardware deals workgroup ids round robin over the 8 XCDs, so XCD x runs ids
// x, x + 8, ...; those take the consecutive tiles start_x, start_x + 1, ... of
// the column-major order, so the workgroups in flight on one XCD share one or
// two right-hand column tiles (4 MB each at Kq 1920) and stream the left-hand
// rows together, from that XCD's L2.
__device__ __forceinline__ int2 xcd_tile(int L, int g) {
     const int nt = g * g, x = L & 7, k = L >> 3, q = nt >> 3, r = nt & 7;
     const int t = x * q + min(x, r) + k;
     return make_int2(t % g, t / g);
}
template <class E>
__global__ __launch_bounds__(256) void adj_gbuild(const unsigned short *L16, const unsigned short *R16, int Hq, int Kq, float *Gf, AdjScal *sc) {
     const int2 t = xcd_tile(blockIdx.x, Kq / 64);
     adj_gbuild_body<E>(L16, R16, Hq, Kq, Gf, sc, t.x, t.y);
}
struct adj_gbuild_B {
     const unsigned short *L16;
     const unsigned short *R16;
     int Hq;
     int Kq;
     float *Gf;
     AdjScal *sc;
     unsigned gx, gy;
};
template <class E>
__global__ __launch_bounds__(256) void adj_gbuild_b(const adj_gbuild_B *cases) {
     const adj_gbuild_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     adj_gbuild_body<E>(a.L16, a.R16, a.Hq, a.Kq, a.Gf, a.sc, blockIdx.x, blockIdx.y);
}


// Exponent k with |G| * 2^(14 - k) < 2^14 for every entry (0 for G = 0).
__device__ __forceinline__ int adj_gexp(const AdjScal *sc) {
     const float g = __uint_as_float(sc->gmax);
     if (!(g > 0.f)) return 0;
     int k;
     (void) frexpf(g, &k);
     return k;
}

// Split operand of U = G A: [12][Kq / 8][Kq][8], contraction index m in
// groups of 8, column n: Gss, -Re Gst, Im Gst, Gtt, -Re Gts, Im Gts (high,
// low) with Gts(n, m) = conj Gst(m, n), all times 2^(14 - k).  One thread per
// (group of 8 m, n).
template <class E>
__device__ __forceinline__ void adj_gsplit_body(const float *Gf, const AdjScal *sc, int Kq, unsigned short *G16) {
     const int n = blockIdx.x * 64 + threadIdx.x, mg = blockIdx.y;
     if (n >= Kq) return;
     const double s = ldexp(1., 14 - adj_gexp(sc));
     const size_t pl = (size_t) Kq * Kq, plane = (size_t) (Kq / 8) * Kq * 8, o = ((size_t) mg * Kq + n) * 8;
     unsigned short v[12][8];
     for (int e = 0; e < 8; ++e) {
          const int m = 8 * mg + e;
          const size_t nm = (size_t) n * Kq + m, mn = (size_t) m * Kq + n;
          const double val[6] = {Gf[nm], -(double) Gf[2 * pl + nm], Gf[3 * pl + nm], Gf[pl + nm],
                                 -(double) Gf[2 * pl + mn], -(double) Gf[3 * pl + mn]};
          for (int q = 0; q < 6; ++q) split16<E>(val[q] * s, &v[2 * q][e], &v[2 * q + 1][e]);
     }
     for (int q = 0; q < 12; ++q) {
          v4u w;
          for (int k = 0; k < 4; ++k) w[k] = (unsigned) v[q][2 * k] | ((unsigned) v[q][2 * k + 1] << 16);
          *(v4u *) (G16 + q * plane + o) = w;
     }
}
template <class E>
__global__ __launch_bounds__(64) void adj_gsplit(const float *Gf, const AdjScal *sc, int Kq, unsigned short *G16) { adj_gsplit_body<E>(Gf, sc, Kq, G16); }
struct adj_gsplit_B {
     const float *Gf;
     const AdjScal *sc;
     int Kq;
     unsigned short *G16;
     unsigned gx, gy;
};
template <class E>
__global__ __launch_bounds__(64) void adj_gsplit_b(const adj_gsplit_B *cases) {
     const adj_gsplit_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     adj_gsplit_body<E>(a.Gf, a.sc, a.Kq, a.G16);
}


// Operands of U = G A (adj_ugemm32).  Accumulators: U_S re, im, U_T re, im;
// output Uf[4][Kq][Np] in fp32, unscaled to U = Sbar / f.
struct UArgs {
     int Np, Kb, Kq;
     const unsigned short *A16;     // [8][Kb][Np][8]
     const unsigned short *G16;     // [12][Kq / 8][Kq][8]
     const int *ktile;              // per size tile of 64
     const int2 *task;              // (size tile, order tile), heaviest first
     const AdjScal *sc;
     float *Uf;                     // [4][Kq][Np]
};

// U = G A with the 32x32x16 matrix instruction: one workgroup per (size tile
// of TS, order tile of 64) task, stages of 16 orders in two LDS buffers
// (112 KiB for TS = 128, 80 KiB for TS = 64).  Waves in a grid of TS / 32 / WA
// size blocks x 2 order blocks, each wave 32 orders (matrix rows) x 32 WA
// sizes, 4 accumulators per element (U_S, U_T re, im).  TS = 128, WA = 1:
// eight waves, two per SIMD, which hide each other's LDS and barrier waits;
// WA = 2: four waves, each operand fragment read for twice the products;
// TS = 64: four waves, two workgroups per CU.
template <class E, int WA, int TS, bool B>
__global__ __launch_bounds__(TS * 4 / WA) __attribute__((amdgpu_waves_per_eu(TS == 64 ? 2 : 1)))
void adj_ugemm32(UArgs p0, const UArgs *cases, const int2 *btask) {
     constexpr int TO = 64, G = 2;                           // orders, groups of 8 per stage
     constexpr int NT = TS * 4 / WA, WS = TS / 32 / WA;      // threads, size-block waves
     constexpr int NA = 8 * G * TS / NT, NG = 12 * G * TO / NT;
     static_assert(NA * NT == 8 * G * TS && NG * NT == 12 * G * TO, "staging split");
     __shared__ v4u sA[2][8 * G][TS];
     __shared__ v4u sG[2][12 * G][TO];
     // B: one task list for the cases of a batch, task (size tile, order tile | case << 16)
     int2 tk;
     UArgs p;
     if constexpr (B) {
          const int2 t = btask[blockIdx.x];
          p = cases[t.y >> 16];
          tk = make_int2(t.x, t.y & 0xffff);
     } else {
          p = p0;
          tk = p.task[blockIdx.x];
     }
     if (tk.x < 0) return;
     const int it0 = tk.x * TS, n0 = tk.y * TO, nch = p.ktile[tk.x] / (8 * G);
     const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
     const int wm = wave % WS, wn = wave / WS, l32 = lane & 31, lh = lane >> 5;
     const size_t Np = p.Np, Kq = p.Kq;
     const size_t pA = (size_t) p.Kb * Np * 8, pG = (Kq / 8) * Kq * 8;
     const v4u neg = {0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
     v4u ra[NA], rg[NG];
     const Stager<NT, TS, G> stA(p.A16, pA, Np * 8, it0, threadIdx.x);
     const Stager<NT, TO, G> stG(p.G16, pG, Kq * 8, n0, threadIdx.x);
     auto load = [&](int c) {
#pragma unroll
          for (int t = 0; t < NA; ++t) ra[t] = *stA.at(t, c);
#pragma unroll
          for (int t = 0; t < NG; ++t) rg[t] = *stG.at(t, c);
     };
     auto store = [&](int s) {
#pragma unroll
          for (int t = 0; t < NA; ++t) {
               const int e = threadIdx.x + NT * t;
               sA[s][e / TS][e % TS] = ra[t];
          }
#pragma unroll
          for (int t = 0; t < NG; ++t) {
               const int e = threadIdx.x + NT * t;
               sG[s][e / TO][e % TO] = rg[t];
          }
     };
     v16f acc[WA][4];
#pragma unroll
     for (int a = 0; a < WA; ++a)
#pragma unroll
          for (int k = 0; k < 4; ++k)
#pragma unroll
               for (int v = 0; v < 16; ++v) acc[a][k][v] = 0.f;
     if (nch > 0) {
          load(0);
          store(0);
          load(min(1, nch - 1));
     }
     // terms per accumulator: (G plane pair, A plane hi, A plane lo); A planes
     // 0..7 as A16, 8 = -A+r hi, 9 = -A+r lo, 10 = -A-r hi, 11 = -A-r lo
     constexpr int term[4][3][3] = {{{0, 0, 1}, {1, 4, 5}, {2, 6, 7}},     // U_S re
                                    {{0, 2, 3}, {1, 6, 7}, {2, 10, 11}},   // U_S im
                                    {{3, 4, 5}, {4, 0, 1}, {5, 2, 3}},     // U_T re
                                    {{3, 6, 7}, {4, 2, 3}, {5, 8, 9}}};    // U_T im
     for (int c = 0; c < nch; ++c) {
          const int s = c & 1;
          __syncthreads();
          v4u g[12];
#pragma unroll
          for (int q = 0; q < 12; ++q) g[q] = sG[s][q * G + lh][wn * 32 + l32];
          v4u x[WA][12];
#pragma unroll
          for (int a = 0; a < WA; ++a) {
#pragma unroll
               for (int q = 0; q < 8; ++q) x[a][q] = sA[s][q * G + lh][(wm * WA + a) * 32 + l32];
               x[a][8] = x[a][0] ^ neg;
               x[a][9] = x[a][1] ^ neg;
               x[a][10] = x[a][4] ^ neg;
               x[a][11] = x[a][5] ^ neg;
          }
          // hi x hi, lo(G) x hi(A), hi(G) x lo(A) for every (term, size block,
          // accumulator): 4 WA independent accumulators between dependent products
#pragma unroll
          for (int t = 0; t < 3; ++t) {
#pragma unroll
               for (int pr = 0; pr < 3; ++pr) {
#pragma unroll
                    for (int a = 0; a < WA; ++a)
#pragma unroll
                         for (int k = 0; k < 4; ++k) {
                              const int gp = term[k][t][0], xh = term[k][t][1], xl = term[k][t][2];
                              const v4u gv = g[2 * gp + (pr == 1 ? 1 : 0)];
                              const v4u xv = x[a][pr == 2 ? xl : xh];
                              acc[a][k] = E::mfma32(gv, xv, acc[a][k]);
                         }
               }
               // the stage body is one basic block: the last stages store and
               // load a repeat of the last stage (the other LDS buffer is not
               // read again, the repeated rows are in bounds)
               if (t == 0) {
                    store(s ^ 1);
                    load(min(c + 2, nch - 1));
               }
          }
     }
     // U = acc n (n+1) wmax / (2^(14 - k) kSB^2 kSA); rows are orders
     const double base = p.sc->wmax * ldexp(1., adj_gexp(p.sc) - 14) / (kSB * kSB * kSA);
#pragma unroll
     for (int a = 0; a < WA; ++a)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
               const int r = n0 + wn * 32 + mfma32_row(lane, v);
               const int i = it0 + (wm * WA + a) * 32 + l32;
               const double f = base * (r + 1.) * (r + 2.);
#pragma unroll
               for (int k = 0; k < 4; ++k) p.Uf[((size_t) k * Kq + r) * Np + i] = (float) (acc[a][k][v] * f);
          }
}


// Shape of the reverse contraction (environment LMIE_UGEMM_WAVES, read at
// every call): 8 (default) eight waves on 128 sizes, 4 four waves on 128
// sizes, 2 four waves on 64 sizes with two workgroups per CU.
int ugemm_waves() {
     const char *e = getenv("LMIE_UGEMM_WAVES");
     const int r = e ? atoi(e) : 8;
     return (r == 4 || r == 2) ? r : 8;
}

// Per-size reverse sweep (adjoint of src/lmie_core.c:298-611) over the
// forward's kept recurrences (D_j up to n2 - 1, psi_j, chi_j, the reciprocals
// 1 / (D_{j+1} + (j+1)/z)), in two kernels:
//   rev_orders one thread per (size, chunk of RCH orders j in [1, n2)):
//              for j <= n1 the cotangents of a_j, b_j from the angular part
//              (Sbar, Tbar) and the series, through a = (u psi - psi0) /
//              (u zeta - zeta0), u = D/m + j/x and b (v = m D + j/x), to the
//              local cotangent of D_j and the direct part of the cotangent of
//              m; and the adjoint of the downward recurrence
//              D_j = (j+1)/z - 1/(D_{j+1} + (j+1)/z), a first-order linear
//              recurrence upward in j carrying the D cotangent to z = m x,
//              which the chunk returns as an affine function of the carry
//              entering it;
//   rev_carry  one thread per size: chains the chunks, then the per-size totals.
// Outputs per size: the cotangent of m (re, im as d/dRe + i d/dIm), of f_i and
// (when asked for) of x_i.
constexpr int RCH = 64;

struct RevArgs {
     int N, Np, Kp, n1max, n2rows, nch;
     const double *ft;          // order_factors rows over 0 .. n1max + 1
     const double *x, *f;
     const int *n1, *n2;
     double mr, mi;
     double im1r, im1i, im2r, im2i;   // 1/m, 1/m^2 (uniform, formed on the host)
     const double *D, *PC, *PS; // forward recurrences (mie_chains, D up to n2 - 1)
     const double *CB;          // [2][Np]
     const double *S;           // series [4][Np]
     const double *bar;         // [4][Kp][Np] f-weighted cotangents (FP64 path), or
     const float *Uf;           // [4][Kq][Np] U = Sbar / f (mixed path; fang from U)
     int Kq;
     const double *fang;        // [Np] (FP64 path)
     double sb0, sb1, sb2, sb3; // cotangents of the raw sums
     double *scan;              // [nch][REV_ROWS][Np]: P, C0, Q, Z0 (re, im), m cotangent (re, im),
                                // f quadratic form, local x cotangent
     double *mbar;              // [2][Np]
     double *fbar;              // [Np]
     double *xbar;              // [Np] cotangent of x_i, or null (not formed)
};
constexpr int REV_ROWS = 12;

// 1/m and 1/m^2 of m = mr - i mi for the reverse sweep, by Smith's division
static void host_crecip(double br, double bi, double *rr, double *ri) {
     if (fabs(br) >= fabs(bi)) {
          const double t = bi / br, d = br + bi * t;
          *rr = 1. / d;
          *ri = -t / d;
     } else {
          const double t = br / bi, d = br * t + bi;
          *rr = t / d;
          *ri = -1. / d;
     }
}
static void rev_recips(RevArgs &r) {
     const double mr = r.mr, mi = -r.mi;
     host_crecip(mr, mi, &r.im1r, &r.im1i);
     host_crecip(mr * mr - mi * mi, 2. * mr * mi, &r.im2r, &r.im2i);
}

// a_j, b_j for the reverse sweep, with the reciprocals the adjoint reuses:
// ra = 1/(u zeta_j - zeta_{j-1}), rb = 1/(v zeta_j - zeta_{j-1}) of the
// denominators of a and b, u = D/m + j/x, v = m D + j/x; a = (u psi_j -
// psi_{j-1}) ra.  Not the forward's bits (conj(d) / |d|^2 and a product with
// the reciprocal instead of Smith's division; a few ulp): the reverse pass
// needs the values, and d a / d u = (psi_j - a zeta_j) ra reuses ra.  |d|^2
// stays far inside the double range: |d| >= |a numerator| and zeta_j of the
// orders j <= n1(x) grows at most like (2j-1)!! / x^j.
struct AbRev {
     cd a, b, ra, rb;
};
__device__ __forceinline__ AbRev mie_ab_rev(double jx, cd m, cd im, cd Dj, double psi, double psi0, double chi1,
                                            double chi0) {
     AbRev o;
     const cd zj = cmk(psi, chi1), zjm = cmk(psi0, chi0);
     const cd u = cadd(cmul(Dj, im), cmk(jx, 0.)), v = cadd(cmul(m, Dj), cmk(jx, 0.));
     o.ra = crecip_nr(csub(cmul(u, zj), zjm));
     o.rb = crecip_nr(csub(cmul(v, zj), zjm));
     o.a = cmul(csub(cscale(u, psi), cmk(psi0, 0.)), o.ra);
     o.b = cmul(csub(cscale(v, psi), cmk(psi0, 0.)), o.rb);
     return o;
}

// Per chunk, with the carry entering it as c, every carry and D cotangent
// inside is affine in c: carry out = P c + C0, z cotangent = Q c + Z0.
// The order j of the loop is the same in every lane (the chunk is the
// workgroup's), so the rows read are uniform: each window load takes a row
// clamped to the array's extent (in bounds for every lane), unconditionally,
// so no wait follows it before the value's use an order later.  Rows of psi,
// chi and U past a lane's n1 are read but never used (every use is under
// j <= n1); D past n2 - 1 is replaced by 0 (D_{n2} = 0 starts the
// recurrence).  UP: the angular cotangents come from U (mixed path), else
// from the FP64 cotangent planes.
template <bool UP>
__device__ __forceinline__ void rev_orders_body(RevArgs p) {
     const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
     if (i >= p.Np) return;
     const size_t Np = p.Np, plane = (size_t) p.Kp * Np, ld = (size_t) p.n1max + 2;
     const unsigned o4 = (unsigned) i * 4u, o8 = (unsigned) i * 8u, o16 = (unsigned) i * 16u;
     cd P = cmk(1., 0.), C0 = cmk(0., 0.), Q = cmk(0., 0.), Z0 = cmk(0., 0.), mb = cmk(0., 0.);
     double fq = 0., xq = 0.;
     const int n1 = i < p.N ? p.n1[i] : 0, n2 = i < p.N ? p.n2[i] : 0;
     const int lo = 1 + k * RCH, hi = min(n2, lo + RCH);
     if (lo < hi) {
          const double x = p.x[i], f = p.f[i];
          // m, 1/m, 1/m^2 from the arguments: scalar operands, no registers per lane
          const cd m = cmk(p.mr, -p.mi), im1 = cmk(p.im1r, p.im1i), im2 = cmk(p.im2r, p.im2i);
          const cd iz = cscale(im1, 1. / x);   // 1/z, z = m x
          const double2 *Dv = (const double2 *) p.D;
          const double *U = p.PC, *Chi = p.PC + (size_t) (p.n1max + 1) * Np;
          const double sc = p.PS[i];
          const int pmax = p.n1max, dmax = p.n2rows - 1;
          // a_n, b_n recomputed from the kept recurrences, one order ahead for
          // the asymmetry series.  The recurrences are read through a window
          // (psi, chi at j-1, j, j+1; D at j, j+1; the U row of j) whose next
          // entries are loaded one order ahead of their use.
          auto rpc = [&](int n) {   // (psi_n, chi_n)
               const int r = min(n, pmax);
               return double2{sc * ld_row(U, r, Np, o8), ld_row(Chi, r, Np, o8)};
          };
          auto rD = [&](int n) {
               const double2 v = ld_row(Dv, min(n, dmax), Np, o16);
               return n < n2 ? cmk(v.x, v.y) : cmk(0., 0.);
          };
          const size_t uq = (size_t) p.Kq * Np;
          auto rdu = [&](int q, int n) { return UP ? ld_row(p.Uf + q * uq, min(n, p.Kq) - 1, Np, o4) : 0.f; };
          const double2 pcm = rpc(lo - 1), pc0 = rpc(lo), pc1 = rpc(lo + 1);
          double psm = pcm.x, ps0 = pc0.x, ps1 = pc1.x;
          double chm = pcm.y, ch0 = pc0.y, ch1 = pc1.y;
          cd D0 = rD(lo), D1 = rD(lo + 1);
          float u0 = rdu(0, lo), u1 = rdu(1, lo), u2 = rdu(2, lo), u3 = rdu(3, lo);
          cd am = cmk(0., 0.), bm = cmk(0., 0.);
          AbRev cur, nxt;
          cur.a = cur.b = cur.ra = cur.rb = cmk(0., 0.);
          nxt = cur;
          if (lo <= n1) {
               if (lo > 1) {
                    const double2 pc2 = rpc(lo - 2);
                    const AbRev pr = mie_ab_rev((lo - 1) / x, m, im1, rD(lo - 1), psm, pc2.x, chm, pc2.y);
                    am = pr.a;
                    bm = pr.b;
               }
               cur = mie_ab_rev(lo / x, m, im1, D0, ps0, psm, ch0, chm);
          }
          const cd cbv = cmk(p.CB[i], p.CB[Np + i]);
          // U path: every cotangent below is linear in f (U = Sbar / f and the
          // series parts carry f), so the sweep runs with f = 1 and the chunk's
          // f-linear results are scaled once at the end
          const double fw = UP ? 1. : f;
          const double sb0 = p.sb0 * fw, sb1 = p.sb1 * fw, sb2 = p.sb2 * fw, sb3 = p.sb3 * fw;
          double sgn = (lo & 1) ? -1. : 1.;
          const double ix = 1. / x;
          for (int j = lo; j < hi; ++j) {
               // the window's next entries
               const double2 pc2 = rpc(j + 2);
               const double ps2 = pc2.x, ch2 = pc2.y;
               const cd D2 = rD(j + 2);
               const float v0 = rdu(0, j + 1), v1 = rdu(1, j + 1), v2 = rdu(2, j + 1), v3 = rdu(3, j + 1);
               cd lc = cmk(0., 0.);
               if (j <= n1) {
                    if (j < n1) nxt = mie_ab_rev((j + 1) * ix, m, im1, D1, ps1, ps0, ch1, ch0);
                    const cd a = cur.a, b = cur.b, an = nxt.a, bn = nxt.b;
                    const double psi = ps0;
                    const cd zj = cmk(ps0, ch0);
                    const cd Dj = D0;
                    const double fv1 = 2. * j + 1., fv4 = p.ft[FT_FV4 * ld + j];
                    cd Sb, Tb;
                    if constexpr (UP) {
                         const cd us = cmk(u0, u1), ut = cmk(u2, u3);
                         // 1/2 Re(conj S U_S + conj T U_T), S = fv4 (a + b), T = fv4 (a - b)
                         fq += .5 * fv4 * ((a.r + b.r) * us.r + (a.i + b.i) * us.i + (a.r - b.r) * ut.r + (a.i - b.i) * ut.i);
                         Sb = us;
                         Tb = ut;
                    } else {
                         Sb = cmk(ld_row(p.bar, j - 1, Np, o8), ld_row(p.bar + plane, j - 1, Np, o8));
                         Tb = cmk(ld_row(p.bar + 2 * plane, j - 1, Np, o8), ld_row(p.bar + 3 * plane, j - 1, Np, o8));
                    }
                    // angular part
                    cd ab = cscale(cadd(Sb, Tb), fv4), bb = cscale(csub(Sb, Tb), fv4);
                    // series parts: ext, sca, bak, asym
                    ab.r += sb0 * fv1;
                    bb.r += sb0 * fv1;
                    ab = cadd(ab, cscale(a, 2. * sb1 * fv1));
                    bb = cadd(bb, cscale(b, 2. * sb1 * fv1));
                    ab = cadd(ab, cscale(cbv, 2. * sb2 * fv1 * sgn));
                    bb = csub(bb, cscale(cbv, 2. * sb2 * fv1 * sgn));
                    if (j < n1) {
                         const double fv3 = p.ft[FT_FV3 * ld + j];
                         ab = cadd(ab, cadd(cscale(an, sb3 * fv3), cscale(b, sb3 * fv4)));
                         bb = cadd(bb, cadd(cscale(bn, sb3 * fv3), cscale(a, sb3 * fv4)));
                    }
                    if (j > 1) {
                         const double fv3m = p.ft[FT_FV3M * ld + j];
                         ab = cadd(ab, cscale(am, sb3 * fv3m));
                         bb = cadd(bb, cscale(bm, sb3 * fv3m));
                    }
                    // a = (u psi - psi0) / (u zeta - zeta0), da/du = (psi - a zeta) / den
                    const cd dua = cmul(csub(cmk(psi, 0.), cmul(a, zj)), cur.ra);
                    const cd dvb = cmul(csub(cmk(psi, 0.), cmul(b, zj)), cur.rb);
                    const cd ub = cmul(cconj(dua), ab), vb = cmul(cconj(dvb), bb);
                    if (p.xbar) {
                         // d a / d x at fixed D_j (psi_j, psi_{j-1}, chi_j, chi_{j-1}
                         // and j / x move): with P = psi_j - a zeta_j the definition
                         // of a gives psi_{j-1} - a zeta_{j-1} = u P, and with
                         // psi_j' = psi_{j-1} - j psi_j / x, psi_{j-1}' = j psi_{j-1} / x - psi_j
                         // (the same for chi) it reduces to
                         // (d a / d u) (1 + (D / m)^2 - j (j + 1) / x^2); b likewise
                         // with m D.  The D_j part goes through the z cotangent.
                         const double c0 = 1. - j * (j + 1.) * (ix * ix);
                         const cd dm = cmul(Dj, im1), md = cmul(m, Dj);
                         const cd wu = cmul(dm, dm), wv = cmul(md, md);
                         xq += (c0 + wu.r) * ub.r + wu.i * ub.i + (c0 + wv.r) * vb.r + wv.i * vb.i;
                    }
                    // u = D/m + j/x, v = m D + j/x
                    lc = cadd(cmul(cconj(im1), ub), cmul(cconj(m), vb));
                    mb = cadd(mb, cadd(cmul(cconj(cscale(cmul(Dj, im2), -1.)), ub), cmul(cconj(Dj), vb)));
                    sgn = -sgn;
                    am = a;
                    bm = b;
                    cur = nxt;
               }
               // D_j = c1 - 1/c2, c2 = D_{j+1} + c1, c1 = (j+1)/z, D_{n2} = 0; 1/c2
               // by the forward's own operations (the same bits).  D cotangent
               // = P c + (C0 + lc); carry = g D cotangent
               const cd c1 = cscale(iz, j + 1.);
               const cd rc = crecip_nr(cadd(D1, c1));
               const cd ic22 = cmul(rc, rc);
               const cd g = cconj(ic22);
               const cd h = cconj(cmul(cadd(cmk(1., 0.), ic22), cscale(cmul(c1, iz), -1.)));
               const cd t = cadd(C0, lc);
               Q = cadd(Q, cmul(h, P));
               Z0 = cadd(Z0, cmul(h, t));
               P = cmul(g, P);
               C0 = cmul(g, t);
               psm = ps0; ps0 = ps1; ps1 = ps2;
               chm = ch0; ch0 = ch1; ch1 = ch2;
               D0 = D1; D1 = D2;
               u0 = v0; u1 = v1; u2 = v2; u3 = v3;
          }
     }
     // rev_carry reads the chunks below (n2 - 1 + RCH - 1) / RCH only
     if (lo >= n2) return;
     if constexpr (UP) {
          const double f = p.f[i];
          C0 = cscale(C0, f);
          Z0 = cscale(Z0, f);
          mb = cscale(mb, f);
          xq *= f;
     }
     double *o = p.scan + (size_t) k * REV_ROWS * Np + i;
     o[0] = P.r; o[Np] = P.i; o[2 * Np] = C0.r; o[3 * Np] = C0.i;
     o[4 * Np] = Q.r; o[5 * Np] = Q.i; o[6 * Np] = Z0.r; o[7 * Np] = Z0.i;
     o[8 * Np] = mb.r; o[9 * Np] = mb.i; o[10 * Np] = fq; o[11 * Np] = xq;
}
template <bool UP>
__global__ __launch_bounds__(64) void rev_orders(RevArgs p) { rev_orders_body<UP>(p); }
struct rev_orders_B {
     RevArgs p;
     unsigned gx, gy;
};
__global__ __launch_bounds__(64) void rev_orders_b(const rev_orders_B *cases) {
     const rev_orders_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     rev_orders_body<true>(a.p);
}


__device__ __forceinline__ void rev_carry_body(RevArgs p) {
     const int i = blockIdx.x * blockDim.x + threadIdx.x;
     if (i >= p.Np) return;
     const size_t Np = p.Np;
     if (i >= p.N) {
          p.mbar[i] = p.mbar[Np + i] = 0.;
          p.fbar[i] = 0.;
          if (p.xbar) p.xbar[i] = 0.;
          return;
     }
     const double x = p.x[i];
     const int nk = (p.n2[i] - 1 + RCH - 1) / RCH;
     // chain the chunks upward from j = 1 with carry 0
     cd carry = cmk(0., 0.), zb = cmk(0., 0.), mb = cmk(0., 0.);
     double fq = 0., xq = 0.;
     for (int k = 0; k < nk; ++k) {
          const double *q = p.scan + (size_t) k * REV_ROWS * Np + i;
          const cd P = cmk(q[0], q[Np]), C0 = cmk(q[2 * Np], q[3 * Np]);
          const cd Q = cmk(q[4 * Np], q[5 * Np]), Z0 = cmk(q[6 * Np], q[7 * Np]);
          zb = cadd(zb, cadd(cmul(Q, carry), Z0));
          carry = cadd(cmul(P, carry), C0);
          mb = cadd(mb, cmk(q[8 * Np], q[9 * Np]));
          fq += q[10 * Np];
          xq += q[11 * Np];
     }
     mb = cadd(mb, cscale(zb, x));
     // z = m x with m = mr - i mi: the x cotangent of z is Re(conj(m) zbar)
     if (p.xbar) p.xbar[i] = xq + p.mr * zb.r - p.mi * zb.i;
     p.mbar[i] = mb.r;
     p.mbar[Np + i] = mb.i;
     // size-weight cotangent: angular part plus the series part
     const double fser = p.sb0 * p.S[i] + p.sb1 * p.S[Np + i] + p.sb2 * p.S[2 * Np + i] + p.sb3 * p.S[3 * Np + i];
     p.fbar[i] = (p.Uf ? fq : p.fang[i]) + fser;
}
__global__ __launch_bounds__(64) void rev_carry(RevArgs p) { rev_carry_body(p); }
struct rev_carry_B {
     RevArgs p;
     unsigned gx, gy;
};
__global__ __launch_bounds__(64) void rev_carry_b(const rev_carry_B *cases) {
     const rev_carry_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     rev_carry_body(a.p);
}


// Sum of the per-block partials per (pass channel, element, angle); channel c
// of the pass goes to channel map.g[c] of the output.
struct ChMap {
     int g[6];
};

// Eight lanes per output element, each summing every eighth block, then the
// eight partials in a fixed order: deterministic, and eight times the
// threads of a plain per-element loop.
__device__ __forceinline__ void reduce_parts_body(const double *part, int nblk, int nchp, int Q, ChMap map, double *out) {
     const int per = nchp * 4 * Q;
     const int e = blockIdx.x * 32 + (threadIdx.x >> 3), l = threadIdx.x & 7;
     double s = 0.;
     if (e < per) {
          int b = l;
          for (; b + 56 < nblk; b += 64) {   // eight partials in flight, added in order
               double t[8];
#pragma unroll
               for (int k = 0; k < 8; ++k) t[k] = part[(size_t) (b + 8 * k) * per + e];
#pragma unroll
               for (int k = 0; k < 8; ++k) s += t[k];
          }
          for (; b < nblk; b += 8) s += part[(size_t) b * per + e];
     }
     s += __shfl_xor(s, 1);
     s += __shfl_xor(s, 2);
     s += __shfl_xor(s, 4);
     if (e >= per || l) return;
     const int c = e / (4 * Q), rem = e % (4 * Q);
     out[(size_t) map.g[c] * 4 * Q + rem] = s;
}
__global__ __launch_bounds__(256) void reduce_parts(const double *part, int nblk, int nchp, int Q, ChMap map, double *out) { reduce_parts_body(part, nblk, nchp, Q, map, out); }
struct reduce_parts_B {
     const double *part;
     int nblk;
     int nchp;
     int Q;
     ChMap map;
     double *out;
     unsigned gx, gy;
};
__global__ __launch_bounds__(256) void reduce_parts_b(const reduce_parts_B *cases) {
     const reduce_parts_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     reduce_parts_body(a.part, a.nblk, a.nchp, a.Q, a.map, a.out);
}


// ---------------------------------------------------------------------------
// Expansion.  For channel c the quadrature-weighted element rows at the
// mirrored (-) and direct (+) angles of every positive node j:
//   GSF rows  0..5: F11, F11+F33, F11-F33, F33, F12, F34 at -mu_j
//            6..11: the same at +mu_j            (misc/rtutil_scat_int.c:173-183)
//   Legendre 0..5: F11, F11, F33, F33, F12, F34 at -mu_j, 6..11 at +mu_j.
__device__ __forceinline__ void expand_rows_body(const double *pf, int nch, int Q, int H, int Hp, const double *w, int gsf, double *E) {
     const int j = blockIdx.x * blockDim.x + threadIdx.x;
     const int c = blockIdx.y;
     if (j >= Hp || c >= nch) return;
     double *Ec = E + (size_t) c * 12 * Hp;
     if (j >= H) {
          for (int r = 0; r < 12; ++r) Ec[r * Hp + j] = 0.;
          return;
     }
     const double *P = pf + (size_t) c * 4 * Q;
     const int ip = H + j, im = H - 1 - j;
     const double wq = w[ip];
     for (int s = 0; s < 2; ++s) {
          const int a = s ? ip : im;
          const double f11 = P[a] * wq, f33 = P[Q + a] * wq, f12 = P[2 * Q + a] * wq, f34 = P[3 * Q + a] * wq;
          double *row = Ec + (size_t) 6 * s * Hp;
          if (gsf) {
               row[0 * Hp + j] = f11;
               row[1 * Hp + j] = f11 + f33;
               row[2 * Hp + j] = f11 - f33;
               row[3 * Hp + j] = f33;
          } else {
               row[0 * Hp + j] = f11;
               row[1 * Hp + j] = f11;
               row[2 * Hp + j] = f33;
               row[3 * Hp + j] = f33;
          }
          row[4 * Hp + j] = f12;
          row[5 * Hp + j] = f34;
     }
}
__global__ void expand_rows(const double *pf, int nch, int Q, int H, int Hp, const double *w, int gsf, double *E) { expand_rows_body(pf, nch, Q, H, Hp, w, gsf, E); }
struct expand_rows_B {
     const double *pf;
     int nch;
     int Q;
     int H;
     int Hp;
     const double *w;
     int gsf;
     double *E;
     unsigned gx, gy;
};
__global__ void expand_rows_b(const expand_rows_B *cases) {
     const expand_rows_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     expand_rows_body(a.pf, a.nch, a.Q, a.H, a.Hp, a.w, a.gsf, a.E);
}


// One wave per (l, channel): twelve dot products over the positive nodes,
// folded with (-1)^l for the mirrored half, then the alpha/zeta recombination
// and the (2l+1)/2 factor (misc/rtutil_scat_int.c:185-213, :436-469).
__device__ __forceinline__ void expand_coefs_body(const double *T, const double *E, int H, int Hp, int C, int gsf, double *coefs) {
     const int lane = threadIdx.x & 63;
     const int l = blockIdx.x * 4 + (threadIdx.x >> 6);
     const int c = blockIdx.y;
     if (l >= C) return;
     const size_t plane = (size_t) C * Hp;
     const double *P00 = T + (gsf ? 0 : 4 * plane) + (size_t) l * Hp, *P02 = T + plane + (size_t) l * Hp;
     const double *P22 = T + 2 * plane + (size_t) l * Hp, *P2m = T + 3 * plane + (size_t) l * Hp;
     const double *Ec = E + (size_t) c * 12 * Hp;
     double d[12];
#pragma unroll
     for (int r = 0; r < 12; ++r) d[r] = 0.;
     for (int j = lane; j < H; j += 64) {
          const double p00 = P00[j];
          if (gsf) {
               const double p02 = P02[j], p22 = P22[j], p2m = P2m[j];
               d[0] += p00 * Ec[0 * Hp + j];  d[6] += p00 * Ec[6 * Hp + j];
               d[1] += p2m * Ec[1 * Hp + j];  d[7] += p22 * Ec[7 * Hp + j];
               d[2] += p22 * Ec[2 * Hp + j];  d[8] += p2m * Ec[8 * Hp + j];
               d[3] += p00 * Ec[3 * Hp + j];  d[9] += p00 * Ec[9 * Hp + j];
               d[4] += p02 * Ec[4 * Hp + j];  d[10] += p02 * Ec[10 * Hp + j];
               d[5] += p02 * Ec[5 * Hp + j];  d[11] += p02 * Ec[11 * Hp + j];
          } else {
#pragma unroll
               for (int r = 0; r < 12; ++r) d[r] += p00 * Ec[r * Hp + j];
          }
     }
#pragma unroll
     for (int r = 0; r < 12; ++r) {
          double s = d[r];
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
          d[r] = s;
     }
     if (lane == 0) {
          const double sl = (l & 1) ? -1. : 1., fl = (2 * l + 1) / 2.;
          double e[6];
#pragma unroll
          for (int k = 0; k < 6; ++k) e[k] = d[k] * sl + d[6 + k];
          double *out = coefs + (size_t) c * 6 * C;
          if (gsf) {
               const double t = e[1];
               e[1] = (t + e[2]) / 2.;
               e[2] = (t - e[2]) / 2.;
          } else {
               e[5] = -e[5];
          }
#pragma unroll
          for (int k = 0; k < 6; ++k) out[(size_t) k * C + l] = e[k] * fl;
     }
}
__global__ __launch_bounds__(256) void expand_coefs(const double *T, const double *E, int H, int Hp, int C, int gsf, double *coefs) { expand_coefs_body(T, E, H, Hp, C, gsf, coefs); }
struct expand_coefs_B {
     const double *T;
     const double *E;
     int H;
     int Hp;
     int C;
     int gsf;
     double *coefs;
     unsigned gx, gy;
};
__global__ __launch_bounds__(256) void expand_coefs_b(const expand_coefs_B *cases) {
     const expand_coefs_B a = cases[blockIdx.z];
     if (blockIdx.x >= a.gx || blockIdx.y >= a.gy) return;
     expand_coefs_body(a.T, a.E, a.H, a.Hp, a.C, a.gsf, a.coefs);
}


// Scattering matrix at user angles from coefficients: one thread per
// (set, angle) (misc/rtutil_scat_int.c:510-589, :662-710).
__global__ void phase_at_angles(int nset, int ncoef, int ld, const double *X, int nth, const double *theta,
                                double *pf, int gsf) {
#pragma clang fp contract(off)
     const int e = blockIdx.x * blockDim.x + threadIdx.x;
     if (e >= nset * nth) return;
     const int s = e / nth, i = e % nth;
     const double *x = X + (size_t) s * 6 * ld;
     const double m = cos(theta[i] * (kPi / 180.)), m2 = m * m;
     double acc[6] = {0., 0., 0., 0., 0., 0.};
     if (gsf) {
          double a00 = 0, b00 = 0, a02 = 0, b02 = 0, a22 = 0, b22 = 0, a2m = 0, b2m = 0;
          for (int l = 0; l < ncoef; ++l) {
               double v00, v02, v22, v2m;
               if (l == 0) { v00 = 1.; v02 = v22 = v2m = 0.; }
               else if (l == 1) { v00 = m; v02 = v22 = v2m = 0.; }
               else if (l == 2) {
                    v00 = (3. * m2 - 1.) / 2.;
                    v02 = .25 * sqrt(6.) * (m2 - 1.);
                    v22 = .25 * (1. + m) * (1. + m);
                    v2m = .25 * (1. - m) * (1. - m);
               } else {
                    const double L = l - 1, L1 = l, t = 2. * L + 1.;
                    const double sa = sqrt(L * L - 4.), sb = sqrt(L1 * L1 - 4.);
                    const double k4 = t * 4, kb = L1 * (L * L - 4), kd = L * (L1 * L1 - 4);
                    const double da = t * m, dc = t * (L * L1) * m;
                    v00 = (da * b00 - L * a00) / L1;
                    v02 = (da * b02 - sa * a02) / sb;
                    v22 = ((dc - k4) * b22 - kb * a22) / kd;
                    v2m = ((dc + k4) * b2m - kb * a2m) / kd;
               }
               acc[0] += x[l] * v00;
               acc[3] += x[3 * ld + l] * v00;
               if (l >= 2) {
                    acc[4] += x[4 * ld + l] * v02;
                    acc[5] += x[5 * ld + l] * v02;
                    acc[1] += (x[ld + l] + x[2 * ld + l]) * v22;
                    acc[2] += (x[ld + l] - x[2 * ld + l]) * v2m;
               }
               a00 = b00; b00 = v00; a02 = b02; b02 = v02; a22 = b22; b22 = v22; a2m = b2m; b2m = v2m;
          }
          double *o = pf + (size_t) s * 6 * nth;
          o[i] = acc[0];
          o[3 * nth + i] = acc[3];
          o[4 * nth + i] = acc[4];
          o[5 * nth + i] = acc[5];
          o[nth + i] = (acc[1] + acc[2]) / 2.;
          o[2 * nth + i] = (acc[1] - acc[2]) / 2.;
     } else {
          double pm = 1., pc = m;
          for (int l = 0; l < ncoef; ++l) {
               double v;
               if (l == 0) v = 1.;
               else if (l == 1) v = m;
               else {
                    const double a = (l - 1) / (double) l, b = m * pc;
                    v = a * (b - pm) + b;
                    pm = pc;
                    pc = v;
               }
               for (int k = 0; k < 6; ++k) acc[k] += x[k * ld + l] * v;
          }
          double *o = pf + (size_t) s * 6 * nth;
          for (int k = 0; k < 6; ++k) o[k * nth + i] = k == 5 ? -acc[k] : acc[k];
     }
}

// ---------------------------------------------------------------------------
// Workspace: device buffers and a stream, reused across calls, one per
// concurrent caller (a mutex-guarded pool), so the API stays reentrant.
struct Buf {
     void *p = nullptr;
     size_t n = 0;
     int get(size_t bytes) {
          if (bytes <= n) return 0;
          if (p) (void) hipFree(p);
          p = nullptr;
          n = 0;
          bytes = bytes < 256 ? 256 : bytes;
          if (hipMalloc(&p, bytes) != hipSuccess) return set_err("hipMalloc(%zu bytes) failed", bytes);
          n = bytes;
          return 0;
     }
     template <class T> T *as() const { return (T *) p; }
};

struct PinBuf {      // pinned host memory, grown on demand
     void *p = nullptr;
     size_t n = 0;
     int get(size_t bytes) {
          if (bytes <= n) return 0;
          if (p) (void) hipHostFree(p);
          p = nullptr;
          n = 0;
          if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return set_err("hipHostMalloc(%zu bytes) failed", bytes);
          n = bytes;
          return 0;
     }
};

// Many small copies as one: each job moves n doubles (src NULL: writes zeros).
struct CopyJob {
     const double *src;
     double *dst;
     size_t n;
};
__global__ __launch_bounds__(256) void copy_jobs(const CopyJob *jobs) {
     const CopyJob j = jobs[blockIdx.y];
     for (size_t e = (size_t) blockIdx.x * 256 + threadIdx.x; e < j.n; e += (size_t) gridDim.x * 256)
          j.dst[e] = j.src ? j.src[e] : 0.;
}

// Uploads of a batch gathered into one host-to-device copy and one copy_jobs
// launch.  Sizes are whole doubles.
struct Upload {
     struct Item { size_t off; void *dst; size_t n8; bool zero; };
     std::vector<unsigned char> h;
     std::vector<Item> items;
     int add(void *dst, const void *src, size_t bytes) {
          if (bytes % 8) return set_err("internal: upload of %zu bytes", bytes);
          const size_t off = h.size();
          h.resize(off + bytes);
          memcpy(h.data() + off, src, bytes);
          items.push_back(Item{off, dst, bytes / 8, false});
          return 0;
     }
     int zero(void *dst, size_t bytes) {
          if (bytes % 8) return set_err("internal: fill of %zu bytes", bytes);
          items.push_back(Item{0, dst, bytes / 8, true});
          return 0;
     }
     // drops what was queued after the given sizes (a case that failed)
     void rewind(size_t n_items, size_t n_bytes) {
          items.resize(n_items);
          h.resize(n_bytes);
     }
     int flush(Buf &dev, PinBuf &pin, hipStream_t s) { return flush_all({this}, dev, pin, s); }
     // the staged uploads of several callers as one copy and one copy_jobs
     // launch, each one's bytes going straight into the pinned buffer (which
     // must not be refilled before the copy has run)
     static int flush_all(const std::vector<const Upload *> &ups, Buf &dev, PinBuf &pin, hipStream_t s) {
          size_t bytes = 0, n = 0;
          for (const Upload *u : ups) {
               bytes += u->h.size();
               n += u->items.size();
          }
          if (n == 0) return 0;
          if (n > 65535) return set_err("batch of %zu uploads (at most 65535)", n);
          const size_t data = (bytes + 255) & ~(size_t) 255, total = data + sizeof(CopyJob) * n;
          if (dev.get(total) || pin.get(total)) return -1;
          unsigned char *base = dev.as<unsigned char>(), *hp = (unsigned char *) pin.p;
          CopyJob *jobs = (CopyJob *) (hp + data);
          size_t off = 0, k = 0;
          for (const Upload *u : ups) {
               memcpy(hp + off, u->h.data(), u->h.size());
               for (const Item &it : u->items)
                    jobs[k++] = CopyJob{it.zero ? nullptr : (const double *) (base + off + it.off), (double *) it.dst,
                                        it.n8};
               off += u->h.size();
          }
          HIPCHECK(hipMemcpyAsync(base, hp, total, hipMemcpyHostToDevice, s));
          hipLaunchKernelGGL(copy_jobs, dim3(32, (unsigned) n), dim3(256), 0, s, (const CopyJob *) (base + data));
          HIPCHECK(hipGetLastError());
          return 0;
     }
};

// Results of a batch gathered by one copy_jobs launch into one device-to-host
// copy; flush waits for the stream and scatters into the callers' arrays.
struct Download {
     struct Item { const void *src; void *dst; size_t off, n8; };
     std::vector<Item> items;
     size_t total = 0;
     int add(void *dst, const void *src, size_t bytes) {
          if (bytes % 8) return set_err("internal: download of %zu bytes", bytes);
          items.push_back(Item{src, dst, total, bytes / 8});
          total += bytes;
          return 0;
     }
     int flush(Buf &dev, PinBuf &pin, hipStream_t s) {
          if (items.empty()) return 0;
          if (items.size() > 65535) return set_err("batch of %zu results (at most 65535)", items.size());
          const size_t data = (total + 255) & ~(size_t) 255, jb = sizeof(CopyJob) * items.size();
          if (dev.get(data + jb) || pin.get(data + jb)) return -1;
          unsigned char *base = dev.as<unsigned char>(), *hp = (unsigned char *) pin.p;
          CopyJob *jobs = (CopyJob *) (hp + data);
          for (size_t k = 0; k < items.size(); ++k)
               jobs[k] = CopyJob{(const double *) items[k].src, (double *) (base + items[k].off), items[k].n8};
          HIPCHECK(hipMemcpyAsync(base + data, hp + data, jb, hipMemcpyHostToDevice, s));
          hipLaunchKernelGGL(copy_jobs, dim3(32, (unsigned) items.size()), dim3(256), 0, s,
                             (const CopyJob *) (base + data));
          HIPCHECK(hipGetLastError());
          HIPCHECK(hipMemcpyAsync(hp, base, total, hipMemcpyDeviceToHost, s));
          HIPCHECK(hipStreamSynchronize(s));
          for (const Item &it : items) memcpy(it.dst, hp + it.off, 8 * it.n8);
          return 0;
     }
};

struct Workspace {
     int device = -1;
     hipStream_t stream = nullptr;
     hipStream_t side = nullptr;     // angle tables, concurrent with the per-size recurrences
     hipStream_t side2 = nullptr;    // expansion tables, from the Gauss rule on
     hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr, sums = nullptr, nodes = nullptr, join3 = nullptr;
     hipEvent_t uploaded = nullptr;  // a batch case's inputs are on the device (its own stream)
     std::vector<double> hguess;     // host staging of the Gauss root guesses
     std::vector<int2> htask;        // host staging of the contraction task list
     hipEvent_t ev[2 * LMIE_N_TIMERS + 2];
     Buf x, sx, cx, f, n1, n2, kfill, ktile32, ktile64, dx, dm, dft, dfw;
     Buf guess, mu, w, Bp, Bm, B16, A16, T, scrD, scrR, A, S, part, pf, E, coefs, dots, dotjobs, save2, order;
     std::vector<double> hsc;       // host staging of sin x, cos x
     // reverse-pass checkpoint and buffers
     Buf CB, sst, pfbar, bar, fang, mbar, fbar, gbar, expart;
     Buf wts, scal, L16, R16, Gf, G16, Uf, tasks;   // mixed reverse pass
     Buf chD, chPC, chPS, chPart, revScan;   // per-size recurrences of the value channel
     Buf ftab;                      // order_factors of the value channel
     Buf sw, s1, s2;                // save_control 2: channel weights, saved per-size arrays
     Buf bargs, bup, bdown;         // per-case arguments and staging of a batch launched from here
     PinBuf pup, pdown;
     PinBuf pres;                   // pinned landing area of a call's results (one DMA each)
     Upload fin;                    // a single call's inputs, one copy (forward_impl)
     Buf finDev;
     PinBuf finPin;
     int home = -1, busy = 0;       // batch slot of this workspace (-1: general pool)
     int n2rows = 0, ng = 0;
     int ushape = 8, ntask = 0;     // reverse contraction: shape and task count of the pending pass
     int valid = 0;
     int N = 0, Np = 0, Kp = 0, Hp = 0, H = 0, Q = 0, C = 0, n1max = 0, n2max = 0, calc_gc = 0, calc_lc = 0;
     int mixed = 0, bf16 = 0, Kb = 0, Hb = 0;
     std::vector<int> kt64, kt128;  // contraction depth per size tile of 64, 128
     Buf ktile128;
     double mr = 0., mi = 0.;
};

pthread_mutex_t g_pool_mu = PTHREAD_MUTEX_INITIALIZER;
std::vector<Workspace *> g_pool;
std::vector<Workspace *> g_home;     // workspaces of batch slots (Workspace::home)

// Makes device d current on this thread (hipSetDevice only when it is not).
int use_device(int d) {
     int cur = -1;
     if (hipGetDevice(&cur) != hipSuccess || cur != d) HIPCHECK(hipSetDevice(d));
     return 0;
}

int ws_create(int device, Workspace **out) {
     Workspace *w = new Workspace();
     w->device = device;
     HIPCHECK(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
     HIPCHECK(hipStreamCreateWithFlags(&w->side, hipStreamNonBlocking));
     HIPCHECK(hipStreamCreateWithFlags(&w->side2, hipStreamNonBlocking));
     HIPCHECK(hipEventCreateWithFlags(&w->fork, hipEventDisableTiming));
     HIPCHECK(hipEventCreateWithFlags(&w->join, hipEventDisableTiming));
     HIPCHECK(hipEventCreateWithFlags(&w->join2, hipEventDisableTiming));
     HIPCHECK(hipEventCreateWithFlags(&w->sums, hipEventDisableTiming));
     HIPCHECK(hipEventCreateWithFlags(&w->nodes, hipEventDisableTiming));
     HIPCHECK(hipEventCreateWithFlags(&w->join3, hipEventDisableTiming));
     HIPCHECK(hipEventCreateWithFlags(&w->uploaded, hipEventDisableTiming));
     for (auto &e : w->ev) HIPCHECK(hipEventCreate(&e));
     *out = w;
     return 0;
}


int ws_acquire(int device, Workspace **out) {
     pthread_mutex_lock(&g_pool_mu);
     Workspace *w = nullptr;
     for (size_t k = 0; k < g_pool.size(); ++k)
          if (g_pool[k]->device == device) {
               w = g_pool[k];
               g_pool.erase(g_pool.begin() + k);
               break;
          }
     pthread_mutex_unlock(&g_pool_mu);
     if (use_device(device)) return -1;
     if (!w && ws_create(device, &w)) return -1;
     *out = w;
     return 0;
}

// The workspace of slot `slot` of a batch on `device` (the current device):
// the same one at every batch call while it is free, so a retrieval loop that
// passes its cases in the same order finds buffers of the right sizes.
int ws_acquire_home(int device, int slot, Workspace **out) {
     pthread_mutex_lock(&g_pool_mu);
     Workspace *w = nullptr;
     for (Workspace *h : g_home)
          if (h->device == device && h->home == slot && !h->busy) {
               w = h;
               break;
          }
     if (w) w->busy = 1;
     pthread_mutex_unlock(&g_pool_mu);
     if (!w) {
          if (ws_create(device, &w)) return -1;
          w->home = slot;
          w->busy = 1;
          pthread_mutex_lock(&g_pool_mu);
          g_home.push_back(w);
          pthread_mutex_unlock(&g_pool_mu);
     }
     *out = w;
     return 0;
}

void ws_release(Workspace *w) {
     pthread_mutex_lock(&g_pool_mu);
     if (w->home >= 0) w->busy = 0;
     else g_pool.push_back(w);
     pthread_mutex_unlock(&g_pool_mu);
}

struct Timer {
     Workspace *w;
     int on;
     void mark(int id, int end, hipStream_t st = nullptr) {
          if (on) (void) hipEventRecord(w->ev[2 * id + end], st ? st : w->stream);
     }
     void collect(const int *ids, int n) {
          for (int k = 0; k < LMIE_T_HOST_GRID; ++k) g_times[k] = 0.;
          if (!on) return;
          for (int k = 0; k < n; ++k) {
               float ms = 0.f;
               if (hipEventElapsedTime(&ms, w->ev[2 * ids[k]], w->ev[2 * ids[k] + 1]) == hipSuccess)
                    g_times[ids[k]] = ms;
          }
     }
};

// The FP64 contraction of a pass without tangent channels, with its operand
// stages moved by direct loads (global_load_lds_dwordx4, 16 bytes a lane
// straight into LDS, no registers) into ST LDS buffers: stages up to ST - 1
// ahead load while the matrix cores work on this one, one barrier per stage.
// One size tile per workgroup, the largest size tiles first.  On 32 x 32
// tiles (WM = WN = 1) each wave holds 8 accumulators of 16 x 16 in ordinary
// registers and two workgroups share a CU; with 16 accumulators (64 x 32
// tiles) the compiler keeps them in ordinary registers across the stage loop
// and copies them to the matrix-core registers and back at every stage,
// which costs as much as the loads saved.  The products and their order are
// those of angular_f64; the partials are per size tile (angular_f64 sums four
// tiles per partial), so the sums group differently (1e-15 relative).
template <int WM, int WN, int NW, int ST, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC))) void angular_f64_dma(AngArgs p) {
     constexpr int TM = 32 * WM, TN = 32 * WN;
     constexpr int SA = TM + 16, SB = TN + 16;       // LDS row strides (bank spread)
     constexpr int AST = 4 * KC * SA, BST = 2 * KC * SB;
     // direct loads one wave issues per stage (its vmcnt events)
     constexpr int NLOAD = KC + KC / 2;
     static_assert(TM * 8 <= 1024 && TN * 8 <= 1024 && KC % 4 == 0, "a row is one direct load of at most 64 lanes");
     static_assert(ST == 2 || ST == 3, "stages");
     static_assert(NLOAD < 64, "vmcnt field");
     __shared__ double sA[ST * AST];
     __shared__ double sB[ST * BST];
     double *sR = sA;                    // [NW][WN][8][32], after the stages
     static_assert(NW * 8 * 32 * WN <= ST * AST, "epilogue buffer");

     const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
     const int wm = wave & 1, wn = wave >> 1;
     const int l16 = lane & 15, lg = lane >> 4;
     const int jt0 = blockIdx.x * TN;
     const int tile = gridDim.y - 1 - blockIdx.y;
     const int it0 = tile * TM;
     const size_t Np = p.Np, Hp = p.Hp, plane = (size_t) p.Kp * Np;
     const int nch = p.ktile[tile] / KC;

     // stage of orders k0 .. k0 + KC - 1 into buffer b_: A rows (plane, order)
     // of TM sizes and B rows ((pi+tau) / (pi-tau), order) of TN angles, one
     // direct load per row, rows dealt over the four waves
     auto stage_load = [&](int b_, int k0) {
#pragma unroll
          for (int u = 0; u < KC; ++u) {
               const int row = wave + 4 * u, pl = row / KC, r = row % KC;
               if (lane < TM / 2)
                    __builtin_amdgcn_global_load_lds(
                         (const void *) (p.A[0] + pl * plane + (size_t) (k0 + r) * Np + it0 + 2 * lane),
                         (lds_ptr) (sA + b_ * AST + row * SA), 16, 0, 0);
          }
#pragma unroll
          for (int u = 0; u < KC / 2; ++u) {
               const int row = wave + 4 * u, pm = row / KC, r = row % KC;
               const double *src = pm ? p.Bm : p.Bp;
               if (lane < TN / 2)
                    __builtin_amdgcn_global_load_lds((const void *) (src + (size_t) (k0 + r) * Hp + jt0 + 2 * lane),
                                                     (lds_ptr) (sB + b_ * BST + row * SB), 16, 0, 0);
          }
     };

     v4d acc[WM][WN][8];
#pragma unroll
     for (int a = 0; a < WM; ++a)
#pragma unroll
          for (int b = 0; b < WN; ++b)
#pragma unroll
               for (int q = 0; q < 8; ++q) acc[a][b][q] = v4d{0., 0., 0., 0.};

     // stages 0 .. ST - 2 in flight before the first is computed
#pragma unroll
     for (int k = 0; k < ST - 1; ++k)
          if (k < nch) stage_load(k, k * KC);
     for (int kc = 0; kc < nch; ++kc) {
          const int buf = kc % ST, k0 = kc * KC;
          // this wave's rows of the stage have landed (the younger stage may
          // stay in flight: the loads complete in order), then everyone's;
          // every wave is also done with the buffer of stage kc - 1, which
          // takes stage kc + ST - 1
          // (vmcnt in bits 3:0 and 15:14, lgkmcnt and expcnt left open)
          if (ST == 3 && kc + 1 < nch) __builtin_amdgcn_s_waitcnt(0x0F70 | (NLOAD & 15) | ((NLOAD >> 4) << 14));
          else __builtin_amdgcn_s_waitcnt(0);
          __builtin_amdgcn_s_barrier();
          __asm__ volatile("" ::: "memory");
          if (kc + ST - 1 < nch) stage_load((kc + ST - 1) % ST, (kc + ST - 1) * KC);
          const double *a0 = sA + buf * AST, *b0 = sB + buf * BST;
#pragma unroll
          for (int ks = 0; ks < KC / 4; ++ks) {
               const int r = ks * 4 + lg;                   // this lane's k row
               const double sg = ((k0 + r) & 1) ? -1. : 1.;   // s_n, n = k0 + r + 1
               // the -mu products take s_n on the angle operand: the same
               // products as s_n on the size operand (a sign), fewer multiplies
               double bp[WN], bm[WN], sbp[WN], sbm[WN];
#pragma unroll
               for (int b = 0; b < WN; ++b) {
                    const int jc = wn * 16 * WN + 16 * b + l16;
                    bp[b] = b0[r * SB + jc];
                    bm[b] = b0[(KC + r) * SB + jc];
                    sbp[b] = sg * bp[b];
                    sbm[b] = sg * bm[b];
               }
#pragma unroll
               for (int a = 0; a < WM; ++a) {
                    const int ic = wm * 16 * WM + 16 * a + l16;
                    const double Sr = a0[(0 * KC + r) * SA + ic];
                    const double Si = a0[(1 * KC + r) * SA + ic];
                    const double Tr = a0[(2 * KC + r) * SA + ic];
                    const double Ti = a0[(3 * KC + r) * SA + ic];
#pragma unroll
                    for (int b = 0; b < WN; ++b) {
                         v4d *q = acc[a][b];
                         q[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(Sr, bp[b], q[0], 0, 0, 0);
                         q[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(Si, bp[b], q[1], 0, 0, 0);
                         q[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(Tr, bm[b], q[2], 0, 0, 0);
                         q[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ti, bm[b], q[3], 0, 0, 0);
                         q[4] = __builtin_amdgcn_mfma_f64_16x16x4f64(Sr, sbm[b], q[4], 0, 0, 0);
                         q[5] = __builtin_amdgcn_mfma_f64_16x16x4f64(Si, sbm[b], q[5], 0, 0, 0);
                         q[6] = __builtin_amdgcn_mfma_f64_16x16x4f64(Tr, sbp[b], q[6], 0, 0, 0);
                         q[7] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ti, sbp[b], q[7], 0, 0, 0);
                    }
               }
          }
     }

     // epilogue: Stokes elements, weights, reduction over this lane's sizes
     double ew[NW][WN][8];
#pragma unroll
     for (int c = 0; c < NW; ++c)
#pragma unroll
          for (int b = 0; b < WN; ++b)
#pragma unroll
               for (int q = 0; q < 8; ++q) ew[c][b][q] = 0.;
#pragma unroll
     for (int a = 0; a < WM; ++a)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
               const int i = it0 + wm * 16 * WM + 16 * a + lg + 4 * v;
               double wv[NW];
#pragma unroll
               for (int c = 0; c < NW; ++c) wv[c] = p.W[c][i];
#pragma unroll
               for (int b = 0; b < WN; ++b) {
                    const v4d *q = acc[a][b];
                    double Pp[4], Pm[4];
                    stokes(q[0][v], q[1][v], q[2][v], q[3][v], Pp);
                    stokes(q[4][v], q[5][v], q[6][v], q[7][v], Pm);
#pragma unroll
                    for (int c = 0; c < NW; ++c)
#pragma unroll
                         for (int k = 0; k < 4; ++k) {
                              ew[c][b][k] += wv[c] * Pp[k];
                              ew[c][b][4 + k] += wv[c] * Pm[k];
                         }
                    if (p.sst) {
                         const int j = jt0 + wn * 16 * WN + 16 * b + l16;
                         double *o = p.sst + ((size_t) i * p.Hp + j) * 8;
#pragma unroll
                         for (int c = 0; c < 8; ++c) o[c] = q[c][v];
                    }
                    if (p.save2 && i < p.N) {
                         const int j = jt0 + wn * 16 * WN + 16 * b + l16;
                         if (j < p.H) {
#pragma unroll
                              for (int k = 0; k < 4; ++k) {
                                   p.save2[((size_t) k * p.N + i) * p.Q + p.H + j] = Pp[k];
                                   p.save2[((size_t) k * p.N + i) * p.Q + p.H - 1 - j] = Pm[k];
                              }
                         }
                    }
               }
          }
     // reduce over the 4 lane groups (same angle column), then over the two
     // waves sharing the columns, then write this block's partial
#pragma unroll
     for (int c = 0; c < NW; ++c)
#pragma unroll
          for (int b = 0; b < WN; ++b)
#pragma unroll
               for (int q = 0; q < 8; ++q) {
                    double v = ew[c][b][q];
                    v += __shfl_xor(v, 16);
                    v += __shfl_xor(v, 32);
                    ew[c][b][q] = v;
               }
     __syncthreads();                    // every wave is done with the stages (sR aliases them)
     if (wm == 1 && lg == 0) {
#pragma unroll
          for (int c = 0; c < NW; ++c)
#pragma unroll
               for (int b = 0; b < WN; ++b)
#pragma unroll
                    for (int q = 0; q < 8; ++q) sR[((c * WN + b) * 8 + q) * 32 + wn * 16 + l16] = ew[c][b][q];
     }
     __syncthreads();
     if (wm == 0 && lg == 0) {
          double *out = p.part + (size_t) tile * NW * 4 * p.Q;
#pragma unroll
          for (int b = 0; b < WN; ++b) {
               const int j = jt0 + wn * 16 * WN + 16 * b + l16;
               if (j < p.H) {
#pragma unroll
                    for (int c = 0; c < NW; ++c)
#pragma unroll
                         for (int k = 0; k < 4; ++k) {
                              const double vp = ew[c][b][k] + sR[((c * WN + b) * 8 + k) * 32 + wn * 16 + l16];
                              const double vm = ew[c][b][4 + k] + sR[((c * WN + b) * 8 + 4 + k) * 32 + wn * 16 + l16];
                              out[((size_t) c * 4 + k) * p.Q + p.H + j] = vp;
                              out[((size_t) c * 4 + k) * p.Q + p.H - 1 - j] = vm;
                         }
               }
          }
     }
}

// Operand staging of the FP64 contraction without tangent channels
// (environment LMIE_F64_STAGING, read at every call): 1 (default) direct
// loads into two LDS buffers on 32 x 32 tiles (angular_f64_dma, 16 x 16 per
// wave, two workgroups per CU); 0 the loads through registers into one
// buffer on 64 x 32 tiles (angular_f64).  Measured on the two-mode workload
// (32768 sizes, 3810 angles): 20.9 against 43.7 ms; the direct loads on
// 64 x 32 tiles took 26.1 ms with three buffers and 27.0 with two.
int f64_staging() {
     const char *e = getenv("LMIE_F64_STAGING");
     return e && atoi(e) == 0 ? 0 : 1;
}

// Size tile of the FP64 contraction of a pass: 32 with tangent channels or
// the direct loads, else 64
int f64_tile(int n_tan, int staging) { return n_tan > 0 || staging != 0 ? 32 : 64; }

// Size tiles per workgroup of the FP64 contraction: one for the direct-load
// kernel, four for angular_f64 (each workgroup's partial takes one slot)
int f64_tiles_per_block(int n_tan, int staging) { return n_tan == 0 && staging != 0 ? 1 : 4; }

// Launch sequence of the contraction for one pass (up to 2 tangent channels
// and up to 4 weight channels).
template <int WM, int WN, int NT, int NW>
int launch_angular(const AngArgs &a, int n_tiles, int staging, hipStream_t s, int *nblk_out) {
     constexpr int TM = 32 * WM, TN = 32 * WN;
     (void) TM;
     const int tpb = f64_tiles_per_block(NT, staging);
     const int nblk = (n_tiles + tpb - 1) / tpb;
     AngArgs b = a;
     b.tiles_per_block = tpb;
     dim3 grid(a.Hp / TN, nblk);
     if constexpr (NT == 0) {
          if constexpr (WM == 1 && WN == 1) {
               if (staging != 0) hipLaunchKernelGGL((angular_f64_dma<WM, WN, NW, 2, 2>), grid, dim3(256), 0, s, b);
               else hipLaunchKernelGGL((angular_f64<WM, WN, NT, NW>), grid, dim3(256), 0, s, b);
          } else {
               if (staging == 0) hipLaunchKernelGGL((angular_f64<WM, WN, NT, NW>), grid, dim3(256), 0, s, b);
               else return set_err("internal: FP64 staging %d with %d x %d tiles", staging, TM, TN);
          }
     } else {
          hipLaunchKernelGGL((angular_f64<WM, WN, NT, NW>), grid, dim3(256), 0, s, b);
     }
     HIPCHECK(hipGetLastError());
     *nblk_out = nblk;
     return 0;
}

template <int WM, int WN, int NT>
int launch_angular_nw(int nw, const AngArgs &a, int n_tiles, int staging, hipStream_t s, int *nblk) {
     switch (nw) {
     case 0:
          if (NT == 0) return set_err("internal: contraction pass without channels");
          return launch_angular<WM, WN, NT, (NT > 0 ? 0 : 1)>(a, n_tiles, staging, s, nblk);
     case 1: return launch_angular<WM, WN, NT, 1>(a, n_tiles, staging, s, nblk);
     case 2: return launch_angular<WM, WN, NT, 2>(a, n_tiles, staging, s, nblk);
     case 3: return launch_angular<WM, WN, NT, 3>(a, n_tiles, staging, s, nblk);
     default: return launch_angular<WM, WN, NT, 4>(a, n_tiles, staging, s, nblk);
     }
}

int launch_coefs(int nt, const CoefArgs &c, hipStream_t s) {
     dim3 grid(c.Np / 64);
     switch (nt) {
     case 1: hipLaunchKernelGGL(mie_coefs<1>, grid, dim3(64), 0, s, c); break;
     case 2: hipLaunchKernelGGL(mie_coefs<2>, grid, dim3(64), 0, s, c); break;
     case 3: hipLaunchKernelGGL(mie_coefs<3>, grid, dim3(64), 0, s, c); break;
     default: hipLaunchKernelGGL(mie_coefs<4>, grid, dim3(64), 0, s, c); break;
     }
     HIPCHECK(hipGetLastError());
     return 0;
}

// Task list of a tiled contraction over (size tile, column tile) pairs for
// the XCD layout of the chip: the hardware deals workgroup ids round robin
// over the 8 XCDs, so id 8k + x is the k-th workgroup of XCD x.  Size tiles go
// heaviest first (ktile descending) to the XCDs in turn; inside an XCD they
// are taken in groups of `group`, and a group runs column tile by column tile.
// The workgroups in flight on one XCD then share a column tile's operand
// (angle or order tables) across the group and a size tile's operand across
// the column tiles, both from that XCD's L2.  ncol(s) = column tiles of size
// tile s.  Padding entries are (-1, -1).
template <class F>
void xcd_tasks(const std::vector<int> &kt, F ncol, int group, std::vector<int2> &out) {
     const int nT = (int) kt.size();
     std::vector<int> ord(nT);
     for (int t = 0; t < nT; ++t) ord[t] = t;
     std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return kt[a] > kt[b]; });
     std::vector<std::vector<int2>> per(8);
     for (int x = 0; x < 8; ++x) {
          std::vector<int> mine;
          for (int k = x; k < nT; k += 8) mine.push_back(ord[k]);
          for (size_t g0 = 0; g0 < mine.size(); g0 += group) {
               const size_t g1 = std::min(mine.size(), g0 + group);
               int cmax = 0;
               for (size_t k = g0; k < g1; ++k) cmax = std::max(cmax, ncol(mine[k]));
               for (int c = 0; c < cmax; ++c)
                    for (size_t k = g0; k < g1; ++k)
                         if (c < ncol(mine[k])) per[x].push_back(make_int2(mine[k], c));
          }
     }
     size_t len = 0;
     for (auto &v : per) len = std::max(len, v.size());
     out.assign(8 * len, make_int2(-1, -1));
     for (int x = 0; x < 8; ++x)
          for (size_t k = 0; k < per[x].size(); ++k) out[8 * k + x] = per[x][k];
}

// Split products of the forward mixed contraction (environment
// LMIE_MIX_PRODUCTS, read at every call): 5 (default), 4 (without A_lo B_mid;
// shape 3 only, else 5) or 3.
int mix_products() {
     const char *e = getenv("LMIE_MIX_PRODUCTS");
     const int r = e ? atoi(e) : 5;
     return (r == 3 || r == 4) ? r : 5;
}

// Shape of the forward mixed contraction (environment LMIE_MIX_SHAPE, read at
// every call; the shapes are listed at MixS0): size tile, angle tile.
struct MixShapeInfo { int id, tm, tn, nt; };
MixShapeInfo mix_shape() {
     const char *e = getenv("LMIE_MIX_SHAPE");
     const int id = e ? atoi(e) : MIX_SHAPE_DEFAULT;
     if (id == 0) return {0, MixS0::TM, MixS0::TN, MixS0::NT};
     if (id == 2) return {2, MixS2::TM, MixS2::TN, MixS2::NT};
     if (id == 3) return {3, MixS3::TM, MixS3::TN, MixS3::NT};
     return {1, MixS1::TM, MixS1::TN, MixS1::NT};
}

// Size tiles per group of xcd_tasks in the forward mixed contraction
// (environment LMIE_MIX_GROUP, read at every call, default 4): the 32
// workgroups in flight on one XCD are then 4 size tiles x 8 angle tiles, so
// each angle tile's operand is read into that XCD's L2 for 4 size tiles and
// each size tile's (the larger operand per tile) for 8 angle tiles.
int mix_group() {
     const char *e = getenv("LMIE_MIX_GROUP");
     const int g = e ? atoi(e) : 4;
     return (g < 1 || g > 64) ? 4 : g;
}

// Size tiles per group of xcd_tasks (environment LMIE_TILE_GROUP, read at
// every call, default 1: the workgroups of one XCD take the column tiles of
// one size tile together).
int tile_group() {
     const char *e = getenv("LMIE_TILE_GROUP");
     const int g = e ? atoi(e) : 1;
     return (g < 1 || g > 64) ? 1 : g;
}

// Initial guesses of the Gauss-Legendre roots, cos(t + (n-1)/(8 n^3) / tan t)
// with t = (4 i + 3) pi / (4 n + 2) (misc/rtutil_math.c:93-97), on the host.
void gauss_guesses(int n, double *g) {
     const int ia = 4 * n + 2;
     const double da = (n - 1) / (8. * pow((double) n, 3));
     for (int i = 0; i < (n + 1) / 2; ++i) {
          const double db = (4. * i + 3.) / ia * kPi;
          g[i] = cos(db + da / tan(db));
     }
}


int forward_impl(const lmie_gpu_job *J, lmie_gpu_result *R, Workspace *w) {
     const auto t_entry = std::chrono::steady_clock::now();
     const int N = J->n_size, Np = rup(N, 2 * TILE);   // whole tiles of 128 sizes
     const int n1max = J->n1_max, Q = J->n_qang, C = J->n_coef;
     const int H = Q / 2, Hp = rup(H > 0 ? H : 1, TILE), Kp = rup(n1max, 32);
     const int nt = J->n_tan, nw = J->n_wgt, nch = 1 + nt + nw;
     if (N <= 0) return set_err("no size points");
     if (Q % 2) return set_err("odd n_qang");
     for (int i = 0; i < N; ++i)
          if (J->n1[i] > n1max || J->n1[i] < 1 || J->n2[i] < J->n1[i])
               return set_err("size point %d: order n1 = %d outside [1, %d] (n2 = %d)", i, J->n1[i], n1max, J->n2[i]);
     hipStream_t s = w->stream;
     Timer tm{w, g_timing};
     w->valid = 0;
     // mixed path: value and weight-only channels on the split 16-bit kernel;
     // forward tangents and per-size matrix output stay on the FP64 kernel
     const int mixed = (J->precision == LMIE_PRECISION_MIXED || J->precision == LMIE_PRECISION_MIXED_BF16) &&
                       nt == 0 && nw <= 3 && J->save_control != 1 && Q > 0;
     const int bf16 = J->precision == LMIE_PRECISION_MIXED_BF16;
     const int kgran = mixed ? KM : KC;
     const int Kb = rup(n1max, KM) / 8;
     // angle padding of the split tables: tiles of 128 angles (whole angle
     // tiles of every contraction shape)
     const MixShapeInfo ms = mix_shape();
     const int Hb = rup(H > 0 ? H : 1, 128);

     // host-side tables: per-tile contraction depths
     const int nt32 = Np / 32, nt64 = Np / 64;
     std::vector<int> n1h(Np, 0), n2h(Np, 0), kt32(nt32), kt64(nt64), kfill(nt64);
     for (int i = 0; i < N; ++i) { n1h[i] = J->n1[i]; n2h[i] = J->n2[i]; }
     for (int t = 0; t < nt32; ++t) {
          int m = 0;
          for (int i = t * 32; i < t * 32 + 32; ++i) m = n1h[i] > m ? n1h[i] : m;
          kt32[t] = rup(m, kgran);
     }
     for (int t = 0; t < nt64; ++t) {
          kt64[t] = kt32[2 * t] > kt32[2 * t + 1] ? kt32[2 * t] : kt32[2 * t + 1];
          kfill[t] = 0;
     }
     // operand rows are zero up to the depth of the 128-size tile (the
     // reverse contraction runs whole 128-size tiles)
     std::vector<int> kt128(nt64 / 2);
     for (int t = 0; t < nt64 / 2; ++t) kt128[t] = std::max(kt64[2 * t], kt64[2 * t + 1]);
     for (int t = 0; t < nt64; ++t) kfill[t] = kt128[t / 2];
     // uploads
     if (w->x.get(sizeof(double) * Np) || w->f.get(sizeof(double) * Np) || w->n1.get(sizeof(int) * Np) ||
         w->n2.get(sizeof(int) * Np) || w->kfill.get(sizeof(int) * nt64) || w->ktile32.get(sizeof(int) * nt32) ||
         w->ktile64.get(sizeof(int) * nt64) || w->dx.get(sizeof(double) * Np * (nt ? nt : 1)) ||
         w->dm.get(sizeof(double) * 2 * (nt ? nt : 1)) || w->dft.get(sizeof(double) * Np * (nt ? nt : 1)) ||
         w->dfw.get(sizeof(double) * Np * (nw ? nw : 1)) || w->sx.get(sizeof(double) * Np) ||
         w->cx.get(sizeof(double) * Np) || w->S.get(sizeof(double) * 4 * Np * (1 + nt)))
          return -1;
     // every input of the pipeline goes up in one pinned host-to-device copy
     // and one copy_jobs launch (a dozen small pageable copies cost more host
     // time than the data)
     Upload &up = w->fin;
     up.items.clear();
     up.h.clear();
     // x, f zero-padded to Np; sin x, cos x of every size from the host C
     // library, as the reference (padding 0, 1)
     w->hsc.assign(4 * (size_t) Np, 0.);
     {
          double *hx = w->hsc.data(), *hf = hx + Np, *hs = hf + Np, *hc = hs + Np;
          for (int i = 0; i < Np; ++i) {
               hx[i] = i < N ? J->x[i] : 0.;
               hf[i] = i < N ? J->f[i] : 0.;
               hs[i] = i < N ? (J->sx ? J->sx[i] : sin(J->x[i])) : 0.;
               hc[i] = i < N ? (J->cx ? J->cx[i] : cos(J->x[i])) : 1.;
          }
          if (up.add(w->x.p, hx, sizeof(double) * Np) || up.add(w->f.p, hf, sizeof(double) * Np) ||
              up.add(w->sx.p, hs, sizeof(double) * Np) || up.add(w->cx.p, hc, sizeof(double) * Np) ||
              up.add(w->n1.p, n1h.data(), sizeof(int) * Np) || up.add(w->n2.p, n2h.data(), sizeof(int) * Np) ||
              up.add(w->kfill.p, kfill.data(), sizeof(int) * nt64) ||
              up.add(w->ktile32.p, kt32.data(), sizeof(int) * nt32) ||
              up.add(w->ktile64.p, kt64.data(), sizeof(int) * nt64))
               return -1;
     }
     // seed columns zero-padded to Np
     auto add_padded = [&](Buf &dst, const double *src, int rows) -> int {
          std::vector<double> pad((size_t) Np * rows, 0.);
          for (int t = 0; t < rows; ++t) memcpy(pad.data() + (size_t) t * Np, src + (size_t) t * N, sizeof(double) * N);
          return up.add(dst.p, pad.data(), sizeof(double) * pad.size());
     };
     if (nt && (add_padded(w->dx, J->dx, nt) || add_padded(w->dft, J->df_tan, nt) ||
                up.add(w->dm.p, J->dm, sizeof(double) * 2 * nt)))
          return -1;
     if (nw && add_padded(w->dfw, J->df_wgt, nw)) return -1;
     if (Q > 0) {
          // initial guesses of the Gauss roots (host C library)
          if (w->guess.get(sizeof(double) * H)) return -1;
          w->hguess.resize(H > 0 ? H : 1);
          gauss_guesses(Q, w->hguess.data());
          if (up.add(w->guess.p, w->hguess.data(), sizeof(double) * H)) return -1;
     }
     // pinned landing area of the results: expansions [2][nch][6][C], then the sums
     if (w->pres.get(sizeof(double) * ((size_t) nch * 12 * C + 4 * nch))) return -1;
     if (mixed) {
          // the contraction's task list
          xcd_tasks(ms.tm == 64 ? kt64 : kt32, [&](int) { return (H + ms.tn - 1) / ms.tn; }, mix_group(), w->htask);
          if (w->order.get(sizeof(int2) * w->htask.size()) ||
              up.add(w->order.p, w->htask.data(), sizeof(int2) * w->htask.size()))
               return -1;
     }
     // series dot products: channel 0, then tangent, then weight-only channels
     std::vector<DotJob> dot_jobs;
     {
          const double *S0 = w->S.as<double>(), *f = w->f.as<double>();
          for (int q = 0; q < 4; ++q) dot_jobs.push_back(DotJob{f, S0 + (size_t) q * Np, nullptr, nullptr});
          for (int t = 0; t < nt; ++t)
               for (int q = 0; q < 4; ++q)
                    dot_jobs.push_back(DotJob{w->dft.as<double>() + (size_t) t * Np, S0 + (size_t) q * Np, f,
                                              w->S.as<double>() + (size_t) 4 * Np * (1 + t) + (size_t) q * Np});
          for (int t = 0; t < nw; ++t)
               for (int q = 0; q < 4; ++q)
                    dot_jobs.push_back(DotJob{w->dfw.as<double>() + (size_t) t * Np, S0 + (size_t) q * Np, nullptr, nullptr});
          if (w->dotjobs.get(sizeof(DotJob) * dot_jobs.size()) || w->dots.get(sizeof(double) * dot_jobs.size()) ||
              up.add(w->dotjobs.p, dot_jobs.data(), sizeof(DotJob) * dot_jobs.size()))
               return -1;
     }
     if (up.flush(w->finDev, w->finPin, s)) return -1;
     const double prep_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_entry).count();
     tm.mark(LMIE_T_TOTAL, 0);

     // angle tables, on the side streams while the per-size recurrences run
     // (side_tables, launched after the first per-size kernel)
     hipStream_t ts = w->side;
     HIPCHECK(hipEventRecord(w->fork, s));
     HIPCHECK(hipStreamWaitEvent(ts, w->fork, 0));
     tm.mark(LMIE_T_TABLES, 0, ts);
     auto side_tables = [&]() -> int {
     if (Q > 0) {
          hipStream_t s = ts;
          if (w->mu.get(sizeof(double) * Q) || w->w.get(sizeof(double) * Q) ||
              (!mixed && (w->Bp.get(sizeof(double) * (size_t) Kp * Hp) || w->Bm.get(sizeof(double) * (size_t) Kp * Hp))))
               return -1;
          hipLaunchKernelGGL(gauss_nodes, dim3((H + 63) / 64), dim3(64), 0, s, Q, (const double *) w->guess.as<double>(),
                             w->mu.as<double>(), w->w.as<double>());
          // the expansion tables on a third stream from here: they run beside
          // the per-size recurrences rather than beside the contraction
          if (J->calc_gc || J->calc_lc) {
               const int planes = J->calc_lc ? 5 : 4;
               if (w->T.get(sizeof(double) * planes * (size_t) C * Hp)) return -1;
               HIPCHECK(hipEventRecord(w->nodes, s));
               HIPCHECK(hipStreamWaitEvent(w->side2, w->nodes, 0));
               hipLaunchKernelGGL(gsf_table, dim3(Hp / 64, planes), dim3(64), 0, w->side2, H, Hp, C,
                                  (const double *) w->mu.as<double>(), w->T.as<double>());
          }
          HIPCHECK(hipEventRecord(w->join3, w->side2));
          if (mixed) {
               if (w->B16.get(sizeof(unsigned short) * B16_PLANES * (size_t) Kb * Hb * 8)) return -1;
               if (bf16)
                    hipLaunchKernelGGL(pitau_table16<BF16>, dim3(Hb / 64), dim3(64), 0, s, H, Hb, n1max, Kb,
                                       (const double *) w->mu.as<double>(), w->B16.as<unsigned short>());
               else
                    hipLaunchKernelGGL(pitau_table16<F16>, dim3(Hb / 64), dim3(64), 0, s, H, Hb, n1max, Kb,
                                       (const double *) w->mu.as<double>(), w->B16.as<unsigned short>());
               // the mixed contraction needs only the split tables; the FP64
               // tables below serve the reverse pass, which runs after join2
               HIPCHECK(hipEventRecord(w->join, s));
          }
          if (!mixed) {
               hipLaunchKernelGGL(pitau_table, dim3(Hp / 64), dim3(64), 0, s, H, Hp, n1max, Kp, w->mu.as<double>(),
                                  w->Bp.as<double>(), w->Bm.as<double>());
          }
          // the FP64 contraction needs the tables above; the expansion tables
          // below only the expansion
          if (!mixed) HIPCHECK(hipEventRecord(w->join, s));
          HIPCHECK(hipGetLastError());
     }
     tm.mark(LMIE_T_TABLES, 1, ts);
     if (Q <= 0) HIPCHECK(hipEventRecord(w->join, ts));
     return 0;
     };

     // per-size coefficients (passes of up to 4 tangent channels)
     const size_t plane = (size_t) Kp * Np, srow = (size_t) (n1max + 1) * Np;
     const int nt_pass = nt < 4 ? nt : 4;
     if ((!mixed && w->A.get(sizeof(double) * 4 * plane * (1 + nt))) || w->S.get(sizeof(double) * 4 * Np * (1 + nt)) ||
         (mixed && w->A16.get(sizeof(unsigned short) * 8 * (size_t) Kb * Np * 8)) ||
         (nt && (w->scrD.get(sizeof(double) * 2 * srow * (1 + nt_pass)) ||
                 w->scrR.get(sizeof(double) * srow * (1 + nt_pass)))))
          return -1;
     const int keep = J->keep;
     if (keep) {
          if (nt || J->save_control == 1) return set_err("the reverse-pass checkpoint covers the value channel only");
          if (w->CB.get(sizeof(double) * 2 * Np) ||
              (Q > 0 && !mixed && w->sst.get(sizeof(double) * 8 * (size_t) Np * Hp)))
               return -1;
     }
     tm.mark(LMIE_T_COEFS, 0);
     if (nt == 0) {
          // value channel: serial recurrences, parallel terms, per-size sums
          int n2max = 0, kfmax = 0;
          for (int i = 0; i < N; ++i) n2max = J->n2[i] > n2max ? J->n2[i] : n2max;
          for (int t = 0; t < nt64; ++t) kfmax = kfill[t] > kfmax ? kfill[t] : kfmax;
          const int n2rows = n2max + 1, ng = kfmax / 8;
          if (w->chD.get(sizeof(double) * 2 * (size_t) n2rows * Np) ||
              w->chPC.get(sizeof(double) * 2 * (size_t) (n1max + 1) * Np) || w->chPS.get(sizeof(double) * Np) ||
              w->chPart.get(sizeof(double) * 5 * (size_t) (ng > 0 ? ng : 1) * Np))
               return -1;
          ChainArgs ca;
          ca.N = N; ca.Np = Np; ca.n1max = n1max; ca.n2rows = n2rows;
          ca.x = w->x.as<double>(); ca.sx = w->sx.as<double>(); ca.cx = w->cx.as<double>();
          ca.n1 = w->n1.as<int>(); ca.n2 = w->n2.as<int>(); ca.mr = J->mr; ca.mi = J->mi;
          ca.d_all = keep; ca.D = w->chD.as<double>(); ca.PC = w->chPC.as<double>(); ca.PS = w->chPS.as<double>();
          hipLaunchKernelGGL(mie_chains, dim3(Np / 64, 2), dim3(64), 0, s, ca);
          if (side_tables()) return -1;
          if (w->ftab.get(sizeof(double) * FT_ROWS * (size_t) (n1max + 2))) return -1;
          hipLaunchKernelGGL(order_factors, dim3((n1max + 2 + 255) / 256), dim3(256), 0, s, n1max + 1,
                             w->ftab.as<double>());
          TermArgs ta;
          ta.N = N; ta.Np = Np; ta.n1max = n1max; ta.n2rows = n2rows; ta.Kp = Kp; ta.Kb = Kb; ta.ng = ng;
          ta.ft = w->ftab.as<double>();
          ta.x = w->x.as<double>(); ta.n1 = w->n1.as<int>(); ta.kfill = w->kfill.as<int>();
          ta.mr = J->mr; ta.mi = J->mi; ta.D = w->chD.as<double>(); ta.PC = w->chPC.as<double>();
          ta.PS = w->chPS.as<double>();
          ta.A = mixed ? nullptr : w->A.as<double>();
          ta.A16 = mixed ? w->A16.as<unsigned short>() : nullptr;
          ta.part = w->chPart.as<double>();
          if (ng > 0) {
               const dim3 grid(Np / 64, ng);
               if (!mixed) hipLaunchKernelGGL(mie_terms<void>, grid, dim3(64), 0, s, ta);
               else if (bf16) hipLaunchKernelGGL(mie_terms<BF16>, grid, dim3(64), 0, s, ta);
               else hipLaunchKernelGGL(mie_terms<F16>, grid, dim3(64), 0, s, ta);
          }
          hipLaunchKernelGGL(mie_sums, dim3((Np + 31) / 32), dim3(256), 0, s, (const double *) w->chPart.as<double>(), N, Np,
                             ng, (const int *) w->n1.as<int>(), w->S.as<double>(), keep ? w->CB.as<double>() : nullptr);
          HIPCHECK(hipGetLastError());
          w->n2rows = n2rows;
          w->ng = ng;
     }
     if (nt > 0 && side_tables()) return -1;
     for (int t0 = 0; nt > 0 && t0 < nt; t0 += 4) {
          CoefArgs c;
          memset(&c, 0, sizeof c);
          c.N = N; c.Np = Np; c.n1max = n1max;
          c.x = w->x.as<double>(); c.sx = w->sx.as<double>(); c.cx = w->cx.as<double>(); c.n1 = w->n1.as<int>(); c.n2 = w->n2.as<int>(); c.kfill = w->kfill.as<int>();
          c.mr = J->mr; c.mi = J->mi;
          int k = nt - t0 < 4 ? nt - t0 : 4;
          if (k < 0) k = 0;
          c.nt = k;
          c.dx = w->dx.as<double>() + (size_t) t0 * Np;
          c.dm = w->dm.as<double>() + 2 * t0;
          c.scrD = w->scrD.as<double>();
          c.scrR = w->scrR.as<double>();
          c.A = w->A.as<double>();
          c.Kp = Kp;
          c.S = w->S.as<double>();
          for (int q = 0; q < k; ++q) {
               c.At[q] = w->A.as<double>() + 4 * plane * (1 + t0 + q);
               c.St[q] = w->S.as<double>() + (size_t) 4 * Np * (1 + t0 + q);
          }
          c.write_base = t0 == 0;
          if (launch_coefs(k, c, s)) return -1;
     }
     tm.mark(LMIE_T_COEFS, 1);

     // series sums (jobs uploaded with the inputs)
     {
          // on the side stream after the angle tables (beside the contraction);
          // the main stream joins it (join2) before the expansion
          HIPCHECK(hipEventRecord(w->sums, s));
          HIPCHECK(hipStreamWaitEvent(ts, w->sums, 0));
          tm.mark(LMIE_T_REDUCE, 0);
          hipLaunchKernelGGL(weighted_dots, dim3(dot_jobs.size()), dim3(256), 0, ts, w->dotjobs.as<DotJob>(), Np,
                             w->dots.as<double>());
          HIPCHECK(hipGetLastError());
          HIPCHECK(hipEventRecord(w->join2, ts));
     }

     // angular contraction and reduction (after the angle tables)
     HIPCHECK(hipStreamWaitEvent(s, w->join, 0));
     if (Q > 0) {
          if (w->pf.get(sizeof(double) * (size_t) nch * 4 * Q)) return -1;
          if (J->save_control == 1 && w->save2.get(sizeof(double) * (size_t) 4 * N * Q)) return -1;
          // pass plan: pass 0 carries channel 0, up to 2 tangent and 3 weight
          // channels; further passes carry the rest (recomputing the base sums)
          struct Pass { int ntp, nwp, t0, w0, nblk; size_t off; };
          std::vector<Pass> plan;
          size_t tot = 0;
          const int f64st = f64_staging();
          for (int td = 0, wd = 0, ps = 0; ps == 0 || td < nt || wd < nw; ++ps) {
               Pass P;
               P.ntp = nt - td < 2 ? nt - td : 2;
               const int base = ps == 0 ? 1 : 0;
               const int ww = nw - wd < 4 - base ? nw - wd : 4 - base;
               P.nwp = base + ww;
               P.t0 = td;
               P.w0 = wd;
               const int tpb = f64_tiles_per_block(P.ntp, f64st);
               P.nblk = ((Np / f64_tile(P.ntp, f64st)) + tpb - 1) / tpb;
               P.off = tot;
               tot += (size_t) P.nblk * (P.nwp + P.ntp) * 4 * Q;
               plan.push_back(P);
               td += P.ntp;
               wd += ww;
          }
          if (mixed) {
               plan.resize(1);
               plan[0].ntp = 0;
               plan[0].nwp = 1 + nw;
               plan[0].t0 = plan[0].w0 = 0;
               plan[0].nblk = Np / ms.tm;
               plan[0].off = 0;
               tot = (size_t) plan[0].nblk * plan[0].nwp * 4 * Q;
          }
          if (w->part.get(sizeof(double) * tot)) return -1;
          tm.mark(LMIE_T_ANGULAR, 0);
          if (mixed) {
               Mix2Args m;
               memset(&m, 0, sizeof m);
               m.Np = Np; m.Kb = Kb; m.Hb = Hb; m.H = H; m.Q = Q;
               m.A16 = w->A16.as<unsigned short>();
               m.B16 = w->B16.as<unsigned short>();
               m.ktile = ms.tm == 64 ? w->ktile64.as<int>() : w->ktile32.as<int>();
               m.task = w->order.as<int2>();
               m.W[0] = w->f.as<double>();
               for (int q = 0; q < nw; ++q) m.W[1 + q] = w->dfw.as<double>() + (size_t) q * Np;
               m.part = w->part.as<double>();
               const dim3 grid((unsigned) w->htask.size());
               // five split products (three with LMIE_MIX_PRODUCTS=3), in the
               // shape of LMIE_MIX_SHAPE
#define LMIE_MIX(E, NP, NWC, S) hipLaunchKernelGGL((angular_mix<E, NP, NWC, S>), grid, dim3(S::NT), 0, s, m)
#define LMIE_MIX_NW(E, NP, S)                                                                            \
               do {                                                                                      \
                    if (nw == 0) LMIE_MIX(E, NP, 1, S); else if (nw == 1) LMIE_MIX(E, NP, 2, S);        \
                    else if (nw == 2) LMIE_MIX(E, NP, 3, S); else LMIE_MIX(E, NP, 4, S);                \
               } while (0)
#define LMIE_MIX_NP(E, S)                                                                                \
               do {                                                                                      \
                    if (mix_products() == 3) LMIE_MIX_NW(E, 3, S);                                       \
                    else if (mix_products() == 4 && std::is_same<S, MixS3>::value) LMIE_MIX_NW(E, 4, MixS3); \
                    else LMIE_MIX_NW(E, 5, S);                                                          \
               } while (0)
#define LMIE_MIX_S(S) do { if (bf16) LMIE_MIX_NP(BF16, S); else LMIE_MIX_NP(F16, S); } while (0)
               if (ms.id == 0) LMIE_MIX_S(MixS0);
               else if (ms.id == 2) LMIE_MIX_S(MixS2);
               else if (ms.id == 3) LMIE_MIX_S(MixS3);
               else LMIE_MIX_S(MixS1);
#undef LMIE_MIX_S
#undef LMIE_MIX_NP
#undef LMIE_MIX_NW
#undef LMIE_MIX
               HIPCHECK(hipGetLastError());
          }
          for (size_t ps = 0; ps < plan.size() && !mixed; ++ps) {
               Pass &P = plan[ps];
               AngArgs a;
               memset(&a, 0, sizeof a);
               a.Np = Np; a.Kp = Kp; a.Hp = Hp; a.H = H; a.Q = Q; a.N = N;
               a.A[0] = w->A.as<double>();
               for (int q = 0; q < P.ntp; ++q) a.A[1 + q] = w->A.as<double>() + 4 * plane * (1 + P.t0 + q);
               a.Bp = w->Bp.as<double>(); a.Bm = w->Bm.as<double>();
               a.f = w->f.as<double>();
               int wi = 0;
               if (ps == 0) a.W[wi++] = w->f.as<double>();
               for (int q = 0; wi < P.nwp; ++q) a.W[wi++] = w->dfw.as<double>() + (size_t) (P.w0 + q) * Np;
               for (int q = 0; q < P.ntp; ++q) a.WT[q] = w->dft.as<double>() + (size_t) (P.t0 + q) * Np;
               a.part = w->part.as<double>() + P.off;
               a.save2 = (ps == 0 && J->save_control == 1) ? w->save2.as<double>() : nullptr;
               a.sst = (ps == 0 && keep) ? w->sst.as<double>() : nullptr;
               int nblk = 0, rc;
               if (P.ntp == 0 && f64_tile(0, f64st) == 32) {
                    a.ktile = w->ktile32.as<int>();
                    rc = launch_angular_nw<1, 1, 0>(P.nwp, a, Np / 32, f64st, s, &nblk);
               } else if (P.ntp == 0) {
                    a.ktile = w->ktile64.as<int>();
                    rc = launch_angular_nw<2, 1, 0>(P.nwp, a, Np / 64, f64st, s, &nblk);
               } else {
                    a.ktile = w->ktile32.as<int>();
                    rc = P.ntp == 1 ? launch_angular_nw<1, 1, 1>(P.nwp, a, Np / 32, f64st, s, &nblk)
                                    : launch_angular_nw<1, 1, 2>(P.nwp, a, Np / 32, f64st, s, &nblk);
               }
               if (rc) return -1;
               if (nblk != P.nblk) return set_err("internal: partial block count");
          }
          tm.mark(LMIE_T_ANGULAR, 1);
          tm.mark(LMIE_T_REDUCE, 0);
          for (size_t ps = 0; ps < plan.size(); ++ps) {
               const Pass &P = plan[ps];
               ChMap map;
               int c = 0;
               if (ps == 0) map.g[c++] = 0;
               for (int q = 0; c < P.nwp; ++q) map.g[c++] = 1 + nt + P.w0 + q;
               for (int q = 0; q < P.ntp; ++q) map.g[c++] = 1 + P.t0 + q;
               const int per = (P.nwp + P.ntp) * 4 * Q;
               hipLaunchKernelGGL(reduce_parts, dim3((per + 31) / 32), dim3(256), 0, s,
                                  (const double *) (w->part.as<double>() + P.off), P.nblk, P.nwp + P.ntp, Q, map,
                                  w->pf.as<double>());
          }
          HIPCHECK(hipGetLastError());
          tm.mark(LMIE_T_REDUCE, 1);

          // expansion (after the expansion tables)
          HIPCHECK(hipStreamWaitEvent(s, w->join2, 0));
          HIPCHECK(hipStreamWaitEvent(s, w->join3, 0));
          tm.mark(LMIE_T_EXPAND, 0);
          if (w->E.get(sizeof(double) * (size_t) nch * 12 * Hp) || w->coefs.get(sizeof(double) * (size_t) nch * 6 * C))
               return -1;
          for (int gsf = 1; gsf >= 0; --gsf) {
               double *dst = gsf ? R->gc_raw : R->lc_raw;
               if (!(gsf ? J->calc_gc : J->calc_lc) || !dst) continue;
               double *land = (double *) w->pres.p + (gsf ? 0 : (size_t) nch * 6 * C);
               hipLaunchKernelGGL(expand_rows, dim3(Hp / 64, nch), dim3(64), 0, s, (const double *) w->pf.as<double>(),
                                  nch, Q, H, Hp, (const double *) w->w.as<double>(), gsf, w->E.as<double>());
               hipLaunchKernelGGL(expand_coefs, dim3((C + 3) / 4, nch), dim3(256), 0, s, (const double *) w->T.as<double>(),
                                  (const double *) w->E.as<double>(), H, Hp, C, gsf, w->coefs.as<double>());
               HIPCHECK(hipGetLastError());
               HIPCHECK(hipMemcpyAsync(land, w->coefs.p, sizeof(double) * (size_t) nch * 6 * C, hipMemcpyDeviceToHost, s));
          }
          tm.mark(LMIE_T_EXPAND, 1);
          if (J->save_control == 1 && R->save2)
               HIPCHECK(hipMemcpyAsync(R->save2, w->save2.p, sizeof(double) * (size_t) 4 * N * Q, hipMemcpyDeviceToHost, s));
     } else {
          tm.mark(LMIE_T_REDUCE, 1);
     }
     HIPCHECK(hipStreamWaitEvent(s, w->join2, 0));   // the series dots
     HIPCHECK(hipStreamWaitEvent(s, w->join3, 0));   // the expansion tables
     tm.mark(LMIE_T_TOTAL, 1);
     // results: through the pinned landing area (the expansions above, the sums here)
     const double *dots = (const double *) w->pres.p + (size_t) nch * 12 * C;
     HIPCHECK(hipMemcpyAsync((void *) dots, w->dots.p, sizeof(double) * 4 * nch, hipMemcpyDeviceToHost, s));
     if (J->save_control == 1 && R->save1)
          for (int q = 0; q < 4; ++q)
               HIPCHECK(hipMemcpyAsync(R->save1 + (size_t) q * N, w->S.as<double>() + (size_t) q * Np,
                                       sizeof(double) * N, hipMemcpyDeviceToHost, s));
     HIPCHECK(hipStreamSynchronize(s));
     if (Q > 0)
          for (int gsf = 1; gsf >= 0; --gsf) {
               double *dst = gsf ? R->gc_raw : R->lc_raw;
               if ((gsf ? J->calc_gc : J->calc_lc) && dst)
                    memcpy(dst, (const double *) w->pres.p + (gsf ? 0 : (size_t) nch * 6 * C),
                           sizeof(double) * (size_t) nch * 6 * C);
          }
     for (int q = 0; q < 4; ++q) R->sums[q] = dots[q];
     for (int c = 1; c < nch; ++c)
          for (int q = 0; q < 4; ++q) R->sums_ch[4 * (c - 1) + q] = dots[4 * c + q];
     const int ids[] = {LMIE_T_TABLES, LMIE_T_COEFS, LMIE_T_ANGULAR, LMIE_T_REDUCE, LMIE_T_EXPAND, LMIE_T_TOTAL};
     tm.collect(ids, Q > 0 ? 6 : 2);
     lmie_gpu_host_time(LMIE_T_HOST_PREP, prep_ms);
     w->valid = 1;
     w->N = N; w->Np = Np; w->Kp = Kp; w->Hp = Hp; w->H = H; w->Q = Q; w->C = C; w->n1max = n1max;
     w->n2max = 0;
     for (int i = 0; i < N; ++i) w->n2max = J->n2[i] > w->n2max ? J->n2[i] : w->n2max;
     w->calc_gc = J->calc_gc; w->calc_lc = J->calc_lc; w->mr = J->mr; w->mi = J->mi;
     w->mixed = mixed; w->bf16 = bf16; w->Kb = Kb; w->Hb = Hb; w->kt64 = kt64; w->kt128 = kt128;
     R->precision = mixed ? J->precision : LMIE_PRECISION_FP64;
     return 0;
}

// Re-weighting of the per-size series and matrix elements of an earlier call
// (save_control = 2, src/lmie_schedule.c:1087-1124): per channel c with
// weights W[c][i], pf[c][q][j] = sum_i W[c][i] s2[q][i][j] and
// dots[c][q] = sum_i W[c][i] s1[q][i].  Eight lanes per output, each adding
// every eighth size, then the eight partials in a fixed order.
__global__ __launch_bounds__(256) void saved_sums(const double *W, const double *s1, const double *s2, int N, int Q,
                                                  double *pf, double *dots) {
     const int per = 4 * (Q + 1), c = blockIdx.y;
     const int e = blockIdx.x * 32 + (threadIdx.x >> 3), l = threadIdx.x & 7;
     const double *w = W + (size_t) c * N;
     double s = 0.;
     if (e < 4 * Q) {
          const int q = e / Q, j = e % Q;
          for (int i = l; i < N; i += 8) s += w[i] * s2[((size_t) q * N + i) * Q + j];
     } else if (e < per) {
          const int q = e - 4 * Q;
          for (int i = l; i < N; i += 8) s += w[i] * s1[(size_t) q * N + i];
     }
     s += __shfl_xor(s, 1);
     s += __shfl_xor(s, 2);
     s += __shfl_xor(s, 4);
     if (l || e >= per) return;
     if (e < 4 * Q) pf[(size_t) c * 4 * Q + e] = s;
     else dots[4 * c + e - 4 * Q] = s;
}

// save_control = 2: the raw sums and expansions from the saved per-size
// arrays and new weights (FP64), on the main stream.  Channels as the
// forward: value, Mie tangents (zero: the saved arrays carry none), weights.
int saved_impl(const lmie_gpu_job *J, lmie_gpu_result *R, Workspace *w) {
     const int N = J->n_size, Q = J->n_qang, C = J->n_coef;
     const int nt = J->n_tan, nw = J->n_wgt, nch = 1 + nt + nw;
     const int H = Q / 2, Hp = rup(H > 0 ? H : 1, TILE);
     if (!J->saved1 || (Q > 0 && !J->saved2)) return set_err("save_control = 2 needs the saved per-size arrays");
     if (N < 1) return set_err("no size points");
     hipStream_t s = w->stream;
     Timer tm{w, g_timing};
     if (w->sw.get(sizeof(double) * (size_t) nch * N) || w->s1.get(sizeof(double) * 4 * (size_t) N) ||
         (Q > 0 && w->s2.get(sizeof(double) * 4 * (size_t) N * Q)) ||
         w->pf.get(sizeof(double) * (size_t) nch * 4 * (Q > 0 ? Q : 1)) || w->dots.get(sizeof(double) * 4 * nch))
          return -1;
     double *W = w->sw.as<double>();
     HIPCHECK(hipMemcpyAsync(W, J->f, sizeof(double) * N, hipMemcpyHostToDevice, s));
     if (nt) HIPCHECK(hipMemsetAsync(W + (size_t) N, 0, sizeof(double) * (size_t) nt * N, s));
     if (nw)
          HIPCHECK(hipMemcpyAsync(W + (size_t) (1 + nt) * N, J->df_wgt, sizeof(double) * (size_t) nw * N,
                                  hipMemcpyHostToDevice, s));
     HIPCHECK(hipMemcpyAsync(w->s1.p, J->saved1, sizeof(double) * 4 * (size_t) N, hipMemcpyHostToDevice, s));
     if (Q > 0)
          HIPCHECK(hipMemcpyAsync(w->s2.p, J->saved2, sizeof(double) * 4 * (size_t) N * Q, hipMemcpyHostToDevice, s));
     tm.mark(LMIE_T_TOTAL, 0);
     if (Q > 0) {
          // Gauss rule and expansion tables (as the forward)
          if (w->mu.get(sizeof(double) * Q) || w->w.get(sizeof(double) * Q) || w->guess.get(sizeof(double) * H))
               return -1;
          w->hguess.resize(H > 0 ? H : 1);
          gauss_guesses(Q, w->hguess.data());
          HIPCHECK(hipMemcpyAsync(w->guess.p, w->hguess.data(), sizeof(double) * H, hipMemcpyHostToDevice, s));
          hipLaunchKernelGGL(gauss_nodes, dim3((H + 63) / 64), dim3(64), 0, s, Q, (const double *) w->guess.as<double>(),
                             w->mu.as<double>(), w->w.as<double>());
          if (J->calc_gc || J->calc_lc) {
               const int planes = J->calc_lc ? 5 : 4;
               if (w->T.get(sizeof(double) * planes * (size_t) C * Hp)) return -1;
               hipLaunchKernelGGL(gsf_table, dim3(Hp / 64, planes), dim3(64), 0, s, H, Hp, C,
                                  (const double *) w->mu.as<double>(), w->T.as<double>());
          }
     }
     const int per = 4 * (Q + 1);
     hipLaunchKernelGGL(saved_sums, dim3((per + 31) / 32, nch), dim3(256), 0, s, (const double *) W,
                        (const double *) w->s1.as<double>(), (const double *) w->s2.as<double>(), N, Q,
                        w->pf.as<double>(), w->dots.as<double>());
     HIPCHECK(hipGetLastError());
     if (Q > 0) {
          if (w->E.get(sizeof(double) * (size_t) nch * 12 * Hp) || w->coefs.get(sizeof(double) * (size_t) nch * 6 * C))
               return -1;
          for (int gsf = 1; gsf >= 0; --gsf) {
               double *dst = gsf ? R->gc_raw : R->lc_raw;
               if (!(gsf ? J->calc_gc : J->calc_lc) || !dst) continue;
               hipLaunchKernelGGL(expand_rows, dim3(Hp / 64, nch), dim3(64), 0, s, (const double *) w->pf.as<double>(),
                                  nch, Q, H, Hp, (const double *) w->w.as<double>(), gsf, w->E.as<double>());
               hipLaunchKernelGGL(expand_coefs, dim3((C + 3) / 4, nch), dim3(256), 0, s, (const double *) w->T.as<double>(),
                                  (const double *) w->E.as<double>(), H, Hp, C, gsf, w->coefs.as<double>());
               HIPCHECK(hipGetLastError());
               HIPCHECK(hipMemcpyAsync(dst, w->coefs.p, sizeof(double) * (size_t) nch * 6 * C, hipMemcpyDeviceToHost, s));
          }
     }
     tm.mark(LMIE_T_TOTAL, 1);
     std::vector<double> dots((size_t) 4 * nch);
     HIPCHECK(hipMemcpyAsync(dots.data(), w->dots.p, sizeof(double) * 4 * nch, hipMemcpyDeviceToHost, s));
     HIPCHECK(hipStreamSynchronize(s));
     for (int q = 0; q < 4; ++q) R->sums[q] = dots[q];
     for (int c = 1; c < nch; ++c)
          for (int q = 0; q < 4; ++q) R->sums_ch[4 * (c - 1) + q] = dots[4 * c + q];
     const int ids[] = {LMIE_T_TOTAL};
     tm.collect(ids, 1);
     w->valid = 0;
     return 0;
}

// Mixed reverse contraction: weights, split tables, G, its split operand and
// U = G A (Uf, [4][Kq][Np]).
// Buffers, task list and depth uploads of the mixed reverse contraction, before
// the timed pipeline (host work and transfers the kernels do not wait on).
int adjoint_mixed_prepare(Workspace *w, hipStream_t s) {
     const int Np = w->Np, H = w->H, n1max = w->n1max;
     const int Kq = rup(n1max, 64), Hq = rup(H, KM);
     if (rup(n1max, KM) > w->Kb * 8) return set_err("internal: split operand depth");
     if (w->wts.get(sizeof(double) * 8 * (size_t) Hq) || w->scal.get(sizeof(AdjScal)) ||
         w->L16.get(sizeof(unsigned short) * 4 * (size_t) Hq * Kq) ||
         w->R16.get(sizeof(unsigned short) * 16 * (size_t) Hq * Kq) ||
         w->Gf.get(sizeof(float) * 4 * (size_t) Kq * Kq) || w->G16.get(sizeof(unsigned short) * 12 * (size_t) Kq * Kq) ||
         w->Uf.get(sizeof(float) * 4 * (size_t) Kq * Np))
          return -1;
     // (size tile of 128 or 64, order tile of 64) tasks over the XCDs
     w->ushape = ugemm_waves();
     const std::vector<int> &kts = w->ushape == 2 ? w->kt64 : w->kt128;
     std::vector<int2> task;
     xcd_tasks(kts, [&](int t) { return (kts[t] + 63) / 64; }, tile_group(), task);
     w->ntask = (int) task.size();
     if (w->ktile128.get(sizeof(int) * kts.size())) return -1;
     HIPCHECK(hipMemcpyAsync(w->ktile128.p, kts.data(), sizeof(int) * kts.size(), hipMemcpyHostToDevice, s));
     if (!task.empty()) {
          if (w->tasks.get(sizeof(int2) * task.size())) return -1;
          HIPCHECK(hipMemcpyAsync(w->tasks.p, task.data(), sizeof(int2) * task.size(), hipMemcpyHostToDevice, s));
     }
     return 0;
}

int adjoint_mixed_angular(Workspace *w, hipStream_t s, Timer &tm) {
     const int Np = w->Np, H = w->H, Q = w->Q, n1max = w->n1max;
     const int Kq = rup(n1max, 64), Hq = rup(H, KM);
     const int shape = w->ushape;
     const unsigned ntask = (unsigned) w->ntask;
     if (ntask == 0) {
          HIPCHECK(hipMemsetAsync(w->Uf.p, 0, sizeof(float) * 4 * (size_t) Kq * Np, s));
          return 0;
     }
     AdjScal *sc = w->scal.as<AdjScal>();
     hipLaunchKernelGGL(adj_weights, dim3(1), dim3(1024), 0, s, (const double *) w->pfbar.as<double>(), H, Hq, Q,
                        w->wts.as<double>(), sc);
     UArgs u;
     u.Np = Np; u.Kb = w->Kb; u.Kq = Kq;
     u.A16 = w->A16.as<unsigned short>(); u.G16 = w->G16.as<unsigned short>();
     u.ktile = w->ktile128.as<int>(); u.task = w->tasks.as<int2>(); u.sc = sc; u.Uf = w->Uf.as<float>();
#define LMIE_ADJ(E)                                                                                            \
     do {                                                                                                     \
          hipLaunchKernelGGL(adj_tables<E>, dim3(Kq / 64, Hq / 8), dim3(64), 0, s,                              \
                             (const unsigned short *) w->B16.as<unsigned short>(), w->Kb, w->Hb,                      \
                             (const double *) w->wts.as<double>(), H, Hq, Kq, n1max, w->L16.as<unsigned short>(),    \
                             w->R16.as<unsigned short>());                                                            \
          hipLaunchKernelGGL(adj_gbuild<E>, dim3((Kq / 64) * (Kq / 64)), dim3(256), 0, s,                       \
                             (const unsigned short *) w->L16.as<unsigned short>(),                             \
                             (const unsigned short *) w->R16.as<unsigned short>(), Hq, Kq, w->Gf.as<float>(), sc); \
          hipLaunchKernelGGL(adj_gsplit<E>, dim3(Kq / 64, Kq / 8), dim3(64), 0, s, (const float *) w->Gf.as<float>(), \
                             (const AdjScal *) sc, Kq, w->G16.as<unsigned short>());                           \
          tm.mark(LMIE_T_ADJ_GEMM, 0);                                                                          \
          if (shape == 4)                                                                                       \
               hipLaunchKernelGGL((adj_ugemm32<E, 2, 128, false>), dim3(ntask), dim3(256), 0, s, u, nullptr, nullptr);  \
          else if (shape == 2)                                                                                  \
               hipLaunchKernelGGL((adj_ugemm32<E, 1, 64, false>), dim3(ntask), dim3(256), 0, s, u, nullptr, nullptr);   \
          else                                                                                                  \
               hipLaunchKernelGGL((adj_ugemm32<E, 1, 128, false>), dim3(ntask), dim3(512), 0, s, u, nullptr, nullptr);  \
          tm.mark(LMIE_T_ADJ_GEMM, 1);                                                                          \
     } while (0)
     if (w->bf16) LMIE_ADJ(BF16);
     else LMIE_ADJ(F16);
#undef LMIE_ADJ
     HIPCHECK(hipGetLastError());
     return 0;
}

int adjoint_impl(Workspace *w, const lmie_gpu_adj_in *in, lmie_gpu_adj_out *out) {
     if (!w->valid) return set_err("no forward state");
     const int N = w->N, Np = w->Np, Kp = w->Kp, Hp = w->Hp, H = w->H, Q = w->Q, C = w->C;
     hipStream_t s = w->stream;
     Timer tm{w, g_timing};
     if (w->pfbar.get(sizeof(double) * 4 * (size_t) (Q > 0 ? Q : 1)) || w->fang.get(sizeof(double) * Np) ||
         w->mbar.get(sizeof(double) * 2 * Np) || w->fbar.get(sizeof(double) * 2 * Np) ||
         w->gbar.get(sizeof(double) * 12 * (size_t) (C > 0 ? C : 1)) ||
         w->bar.get(sizeof(double) * 4 * (size_t) Kp * Np))
          return -1;
     const int use_gc = in->gc_bar && w->calc_gc, use_lc = in->lc_bar && w->calc_lc;
     const int use_u = w->mixed && Q > 0 && (use_gc || use_lc);
     // the cotangent inputs and the contraction's task list go to the device
     // before the timed pipeline
     if (use_u && adjoint_mixed_prepare(w, s)) return -1;
     {
          // the m-cotangent sums at the end of the pass
          DotJob jobs[2] = {{nullptr, w->mbar.as<double>(), nullptr, nullptr},
                            {nullptr, w->mbar.as<double>() + Np, nullptr, nullptr}};
          if (w->dotjobs.get(sizeof(jobs)) || w->dots.get(sizeof(double) * 2) ||
              w->pres.get(sizeof(double) * (2 + 2 * (size_t) N)))
               return -1;
          HIPCHECK(hipMemcpyAsync(w->dotjobs.p, jobs, sizeof(jobs), hipMemcpyHostToDevice, s));
     }
     double *g = w->gbar.as<double>();
     if (Q > 0 && use_gc) HIPCHECK(hipMemcpyAsync(g, in->gc_bar, sizeof(double) * 6 * C, hipMemcpyHostToDevice, s));
     if (Q > 0 && use_lc)
          HIPCHECK(hipMemcpyAsync(g + 6 * C, in->lc_bar, sizeof(double) * 6 * C, hipMemcpyHostToDevice, s));
     tm.mark(LMIE_T_TOTAL, 0);
     HIPCHECK(hipMemsetAsync(w->fang.p, 0, sizeof(double) * Np, s));
     if (Q > 0 && (use_gc || use_lc)) {
          tm.mark(LMIE_T_ADJ_ANGULAR, 0);
          const int nck = (C + EXA_L - 1) / EXA_L;
          if (w->expart.get(sizeof(double) * 24 * (size_t) nck * Hp)) return -1;
          hipLaunchKernelGGL(expand_adjoint_part, dim3(Hp / 64, nck), dim3(64), 0, s, (const double *) w->T.as<double>(),
                             use_gc ? (const double *) g : nullptr, use_lc ? (const double *) (g + 6 * C) : nullptr,
                             H, Hp, C, w->expart.as<double>());
          hipLaunchKernelGGL(expand_adjoint_fin, dim3((H + EXF_J - 1) / EXF_J), dim3(24 * EXF_J), 0, s, (const double *) w->expart.as<double>(),
                             nck, (const double *) w->w.as<double>(), H, Hp, Q, w->pfbar.as<double>());
     }
     if (w->mixed && Q > 0 && (use_gc || use_lc)) {
          if (adjoint_mixed_angular(w, s, tm)) return -1;
          tm.mark(LMIE_T_ADJ_ANGULAR, 1);
     } else if (Q > 0 && (use_gc || use_lc)) {
          hipLaunchKernelGGL(fbar_angular, dim3((N + 3) / 4), dim3(256), 0, s, (const double *) w->sst.as<double>(),
                             (const double *) w->pfbar.as<double>(), N, H, Hp, Q, w->fang.as<double>());
          AdjArgs aa;
          aa.Np = Np; aa.Kp = Kp; aa.Hp = Hp; aa.H = H; aa.Q = Q;
          aa.sst = w->sst.as<double>(); aa.pfbar = w->pfbar.as<double>();
          aa.Bp = w->Bp.as<double>(); aa.Bm = w->Bm.as<double>();
          aa.f = w->f.as<double>(); aa.ktile = w->ktile32.as<int>(); aa.bar = w->bar.as<double>();
          hipLaunchKernelGGL(angular_adj_f64, dim3(Kp / 32, Np / 32), dim3(256), 0, s, aa);
          HIPCHECK(hipGetLastError());
          tm.mark(LMIE_T_ADJ_ANGULAR, 1);
     } else {
          tm.mark(LMIE_T_ADJ_ANGULAR, 0);
          HIPCHECK(hipMemsetAsync(w->bar.p, 0, sizeof(double) * 4 * (size_t) Kp * Np, s));
          tm.mark(LMIE_T_ADJ_ANGULAR, 1);
     }
     tm.mark(LMIE_T_ADJ_COEFS, 0);
     const int nch = (w->n2rows - 1 + RCH - 1) / RCH;
     if (w->revScan.get(sizeof(double) * REV_ROWS * (size_t) (nch > 0 ? nch : 1) * Np)) return -1;
     RevArgs r;
     r.N = N; r.Np = Np; r.Kp = Kp; r.n1max = w->n1max; r.n2rows = w->n2rows; r.nch = nch;
     r.ft = w->ftab.as<double>();
     r.x = w->x.as<double>(); r.f = w->f.as<double>();
     r.n1 = w->n1.as<int>(); r.n2 = w->n2.as<int>(); r.mr = w->mr; r.mi = w->mi; rev_recips(r);
     r.D = w->chD.as<double>(); r.PC = w->chPC.as<double>(); r.PS = w->chPS.as<double>();
     r.CB = w->CB.as<double>(); r.S = w->S.as<double>(); r.bar = w->bar.as<double>();
     r.Uf = use_u ? w->Uf.as<float>() : nullptr; r.Kq = rup(w->n1max, 64);
     r.fang = w->fang.as<double>();
     r.sb0 = in->sums_bar[0]; r.sb1 = in->sums_bar[1]; r.sb2 = in->sums_bar[2]; r.sb3 = in->sums_bar[3];
     r.scan = w->revScan.as<double>();
     r.mbar = w->mbar.as<double>(); r.fbar = w->fbar.as<double>();
     r.xbar = out->x_bar ? w->fbar.as<double>() + Np : nullptr;
     if (nch > 0) {
          if (use_u) hipLaunchKernelGGL(rev_orders<true>, dim3(Np / 64, nch), dim3(64), 0, s, r);
          else hipLaunchKernelGGL(rev_orders<false>, dim3(Np / 64, nch), dim3(64), 0, s, r);
     }
     hipLaunchKernelGGL(rev_carry, dim3(Np / 64), dim3(64), 0, s, r);
     HIPCHECK(hipGetLastError());
     tm.mark(LMIE_T_ADJ_COEFS, 1);
     // sum of the per-size m cotangents; results through the pinned landing area
     {
          hipLaunchKernelGGL(weighted_dots, dim3(2), dim3(256), 0, s, (const DotJob *) w->dotjobs.as<DotJob>(), Np,
                             w->dots.as<double>());
          HIPCHECK(hipGetLastError());
          double *land = (double *) w->pres.p;
          HIPCHECK(hipMemcpyAsync(land, w->dots.p, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
          HIPCHECK(hipMemcpyAsync(land + 2, w->fbar.p, sizeof(double) * N, hipMemcpyDeviceToHost, s));
          if (out->x_bar)
               HIPCHECK(hipMemcpyAsync(land + 2 + N, w->fbar.as<double>() + Np, sizeof(double) * N,
                                       hipMemcpyDeviceToHost, s));
          tm.mark(LMIE_T_TOTAL, 1);
          HIPCHECK(hipStreamSynchronize(s));
          memcpy(out->f_bar, land + 2, sizeof(double) * N);
          if (out->x_bar) memcpy(out->x_bar, land + 2 + N, sizeof(double) * N);
          // m = mr - i mi: dL/dmr = Re(mbar), dL/dmi = -Im(mbar)
          out->m_bar[0] = land[0];
          out->m_bar[1] = -land[1];
     }
     const int ids[] = {LMIE_T_ADJ_ANGULAR, LMIE_T_ADJ_COEFS, LMIE_T_TOTAL, LMIE_T_ADJ_GEMM};
     tm.collect(ids, use_u ? 4 : 3);
     return 0;
}

// ---------------------------------------------------------------------------
// Batches of independent cases in one launch per kernel.  The host thread
// that runs case k of a batch (lmie_gpu_batch_enter) hands its forward and
// later its reverse pass to the batch instead of running them; when every
// case has arrived (or left), one thread runs the pass for all cases: each
// case keeps its own workspace (buffers, reverse-pass state), and every kernel
// is launched once, with blockIdx.z (or the task list) picking the case.
// Fused: mixed precision, value channel only, angles, save_control 0, kept
// state (the forward of a reverse pass); other cases run one by one.


struct BatchBlob {   // host staging of per-case argument arrays, one upload
     std::vector<unsigned char> h;
     template <class T> size_t add(const std::vector<T> &v) {
          const size_t off = (h.size() + 255) & ~(size_t) 255;
          h.resize(off + sizeof(T) * (v.empty() ? 1 : v.size()));
          if (!v.empty()) memcpy(h.data() + off, v.data(), sizeof(T) * v.size());
          return off;
     }
};

template <class T> unsigned max_g(const std::vector<T> &v, unsigned T::*g) {
     unsigned m = 0;
     for (const T &x : v) m = std::max(m, x.*g);
     return m;
}

struct FwdCase {
     int slot;
     const lmie_gpu_job *J;
     lmie_gpu_result *R;
     Workspace *w;
     int N, Np, n1max, Q, C, H, Hp, Kp, Kb, Hb, nt64, n2rows, ng, planes;
     int nmix;                      // size tiles of the forward contraction's shape
     std::vector<int2> task;
};

int fused_forward_ok(const lmie_gpu_job *J) {
     return (J->precision == LMIE_PRECISION_MIXED || J->precision == LMIE_PRECISION_MIXED_BF16) && J->n_tan == 0 &&
            J->n_wgt == 0 && J->save_control == 0 && J->keep && J->n_qang > 0 && J->n_qang % 2 == 0 &&
            (J->calc_gc || J->calc_lc);
}

// Buffers and uploads of one case (the mixed value-channel branch of
// forward_impl); the uploads are queued on up, no kernels.
int fused_prepare(FwdCase &c, Upload &up) {
     const lmie_gpu_job *J = c.J;
     Workspace *w = c.w;
     const int N = J->n_size, Np = rup(N, 2 * TILE), n1max = J->n1_max, Q = J->n_qang, C = J->n_coef;
     const int H = Q / 2, Hp = rup(H, TILE), Kp = rup(n1max, 32), Kb = rup(n1max, KM) / 8,
               Hb = rup(H, 128);
     if (N <= 0) return set_err("no size points");
     for (int i = 0; i < N; ++i)
          if (J->n1[i] > n1max || J->n1[i] < 1 || J->n2[i] < J->n1[i])
               return set_err("size point %d: order n1 = %d outside [1, %d] (n2 = %d)", i, J->n1[i], n1max, J->n2[i]);
     w->valid = 0;
     const int nt32 = Np / 32, nt64 = Np / 64;
     std::vector<int> n1h(Np, 0), n2h(Np, 0), kt32(nt32), kt64(nt64), kfill(nt64), kt128(nt64 / 2);
     for (int i = 0; i < N; ++i) { n1h[i] = J->n1[i]; n2h[i] = J->n2[i]; }
     for (int t = 0; t < nt32; ++t) {
          int m = 0;
          for (int i = t * 32; i < t * 32 + 32; ++i) m = std::max(m, n1h[i]);
          kt32[t] = rup(m, KM);
     }
     for (int t = 0; t < nt64; ++t) kt64[t] = std::max(kt32[2 * t], kt32[2 * t + 1]);
     for (int t = 0; t < nt64 / 2; ++t) kt128[t] = std::max(kt64[2 * t], kt64[2 * t + 1]);
     for (int t = 0; t < nt64; ++t) kfill[t] = kt128[t / 2];
     int n2max = 0, kfmax = 0;
     for (int i = 0; i < N; ++i) n2max = std::max(n2max, J->n2[i]);
     for (int t = 0; t < nt64; ++t) kfmax = std::max(kfmax, kfill[t]);
     const int n2rows = n2max + 1, ng = kfmax / 8, planes = J->calc_lc ? 5 : 4;
     if (w->x.get(sizeof(double) * Np) || w->f.get(sizeof(double) * Np) || w->n1.get(sizeof(int) * Np) ||
         w->n2.get(sizeof(int) * Np) || w->kfill.get(sizeof(int) * nt64) || w->ktile32.get(sizeof(int) * nt32) ||
         w->ktile64.get(sizeof(int) * nt64) || w->sx.get(sizeof(double) * Np) || w->cx.get(sizeof(double) * Np) ||
         w->guess.get(sizeof(double) * H) || w->mu.get(sizeof(double) * Q) || w->w.get(sizeof(double) * Q) ||
         w->T.get(sizeof(double) * planes * (size_t) C * Hp) ||
         w->B16.get(sizeof(unsigned short) * B16_PLANES * (size_t) Kb * Hb * 8) || w->S.get(sizeof(double) * 4 * Np) ||
         w->A16.get(sizeof(unsigned short) * 8 * (size_t) Kb * Np * 8) || w->CB.get(sizeof(double) * 2 * Np) ||
         w->chD.get(sizeof(double) * 2 * (size_t) n2rows * Np) ||
         w->chPC.get(sizeof(double) * 2 * (size_t) (n1max + 1) * Np) || w->chPS.get(sizeof(double) * Np) ||
         w->chPart.get(sizeof(double) * 5 * (size_t) (ng > 0 ? ng : 1) * Np) ||
         w->ftab.get(sizeof(double) * FT_ROWS * (size_t) (n1max + 2)) || w->dotjobs.get(sizeof(DotJob) * 4) ||
         w->dots.get(sizeof(double) * 4) || w->pf.get(sizeof(double) * 4 * (size_t) Q) ||
         w->part.get(sizeof(double) * (size_t) (Np / 32) * 4 * Q) || w->E.get(sizeof(double) * 12 * (size_t) Hp) ||
         w->coefs.get(sizeof(double) * 12 * (size_t) C))
          return -1;
     // x, f zero-padded to Np; sin x, cos x from the host C library (padding 0, 1)
     w->hsc.assign(4 * (size_t) Np, 0.);
     double *hx = w->hsc.data(), *hf = hx + Np, *hs = hf + Np, *hc = hs + Np;
     for (int i = 0; i < Np; ++i) {
          hx[i] = i < N ? J->x[i] : 0.;
          hf[i] = i < N ? J->f[i] : 0.;
          hs[i] = i < N ? (J->sx ? J->sx[i] : sin(J->x[i])) : 0.;
          hc[i] = i < N ? (J->cx ? J->cx[i] : cos(J->x[i])) : 1.;
     }
     w->hguess.resize(H);
     gauss_guesses(Q, w->hguess.data());
     DotJob jobs[4];
     for (int q = 0; q < 4; ++q) jobs[q] = DotJob{w->f.as<double>(), w->S.as<double>() + (size_t) q * Np, nullptr, nullptr};
     if (up.add(w->x.p, hx, sizeof(double) * Np) || up.add(w->f.p, hf, sizeof(double) * Np) ||
         up.add(w->sx.p, hs, sizeof(double) * Np) || up.add(w->cx.p, hc, sizeof(double) * Np) ||
         up.add(w->n1.p, n1h.data(), sizeof(int) * Np) || up.add(w->n2.p, n2h.data(), sizeof(int) * Np) ||
         up.add(w->kfill.p, kfill.data(), sizeof(int) * nt64) || up.add(w->ktile32.p, kt32.data(), sizeof(int) * nt32) ||
         up.add(w->ktile64.p, kt64.data(), sizeof(int) * nt64) ||
         up.add(w->guess.p, w->hguess.data(), sizeof(double) * H) || up.add(w->dotjobs.p, jobs, sizeof(jobs)))
          return -1;
     const MixShapeInfo ms = mix_shape();
     xcd_tasks(ms.tm == 64 ? kt64 : kt32, [&](int) { return (H + ms.tn - 1) / ms.tn; }, mix_group(), c.task);
     c.nmix = Np / ms.tm;
     c.N = N; c.Np = Np; c.n1max = n1max; c.Q = Q; c.C = C; c.H = H; c.Hp = Hp; c.Kp = Kp; c.Kb = Kb; c.Hb = Hb;
     c.nt64 = nt64; c.n2rows = n2rows; c.ng = ng; c.planes = planes;
     w->N = N; w->Np = Np; w->Kp = Kp; w->Hp = Hp; w->H = H; w->Q = Q; w->C = C; w->n1max = n1max;
     w->n2max = n2max; w->calc_gc = J->calc_gc; w->calc_lc = J->calc_lc; w->mr = J->mr; w->mi = J->mi;
     w->mixed = 1; w->bf16 = J->precision == LMIE_PRECISION_MIXED_BF16; w->Kb = Kb; w->Hb = Hb;
     w->kt64 = kt64; w->kt128 = kt128; w->n2rows = n2rows; w->ng = ng;
     return 0;
}

// One launch per kernel over the prepared cases, on the streams of workspace L.
template <class E>
int fused_launch(std::vector<FwdCase> &cs, Upload &up, Workspace *L) {
     hipStream_t s = L->stream, ts = L->side, t2 = L->side2;
     if (up.flush(L->bup, L->pup, s)) return -1;
     const unsigned n = (unsigned) cs.size();
     std::vector<mie_chains_B> chains(n);
     std::vector<gauss_nodes_B> gauss(n);
     std::vector<gsf_table_B> gsf(n);
     std::vector<pitau_table16_B> pitau(n);
     std::vector<order_factors_B> fac(n);
     std::vector<mie_terms_B> terms(n);
     std::vector<mie_sums_B> sums(n);
     std::vector<weighted_dots_B> dots(n);
     std::vector<Mix2Args> mix(n);
     std::vector<reduce_parts_B> red(n);
     std::vector<expand_rows_B> rows[2];
     std::vector<expand_coefs_B> coefs[2];
     std::vector<int2> task;
     for (unsigned k = 0; k < n; ++k) {
          const FwdCase &c = cs[k];
          Workspace *w = c.w;
          ChainArgs ca;
          ca.N = c.N; ca.Np = c.Np; ca.n1max = c.n1max; ca.n2rows = c.n2rows;
          ca.x = w->x.as<double>(); ca.sx = w->sx.as<double>(); ca.cx = w->cx.as<double>();
          ca.n1 = w->n1.as<int>(); ca.n2 = w->n2.as<int>(); ca.mr = c.J->mr; ca.mi = c.J->mi;
          ca.d_all = 1; ca.D = w->chD.as<double>(); ca.PC = w->chPC.as<double>(); ca.PS = w->chPS.as<double>();
          chains[k] = mie_chains_B{ca, (unsigned) c.Np / 64, 2};
          gauss[k] = gauss_nodes_B{c.Q, w->guess.as<double>(), w->mu.as<double>(), w->w.as<double>(),
                                   (unsigned) (c.H + 63) / 64, 1};
          gsf[k] = gsf_table_B{c.H, c.Hp, c.C, w->mu.as<double>(), w->T.as<double>(), (unsigned) c.Hp / 64,
                               (unsigned) c.planes};
          pitau[k] = pitau_table16_B{c.H, c.Hb, c.n1max, c.Kb, w->mu.as<double>(), w->B16.as<unsigned short>(),
                                     (unsigned) c.Hb / 64, 1};
          fac[k] = order_factors_B{c.n1max + 1, w->ftab.as<double>(), (unsigned) (c.n1max + 2 + 255) / 256, 1};
          TermArgs ta;
          ta.N = c.N; ta.Np = c.Np; ta.n1max = c.n1max; ta.n2rows = c.n2rows; ta.Kp = c.Kp; ta.Kb = c.Kb; ta.ng = c.ng;
          ta.ft = w->ftab.as<double>(); ta.x = w->x.as<double>(); ta.n1 = w->n1.as<int>();
          ta.kfill = w->kfill.as<int>(); ta.mr = c.J->mr; ta.mi = c.J->mi; ta.D = w->chD.as<double>();
          ta.PC = w->chPC.as<double>(); ta.PS = w->chPS.as<double>(); ta.A = nullptr; ta.A16 = w->A16.as<unsigned short>();
          ta.part = w->chPart.as<double>();
          terms[k] = mie_terms_B{ta, (unsigned) c.Np / 64, (unsigned) c.ng};
          sums[k] = mie_sums_B{w->chPart.as<double>(), c.N, c.Np, c.ng, w->n1.as<int>(), w->S.as<double>(),
                               w->CB.as<double>(), (unsigned) (c.Np + 31) / 32, 1};
          dots[k] = weighted_dots_B{w->dotjobs.as<DotJob>(), c.Np, w->dots.as<double>(), 4, 1};
          Mix2Args &m = mix[k];
          memset(&m, 0, sizeof m);
          m.Np = c.Np; m.Kb = c.Kb; m.Hb = c.Hb; m.H = c.H; m.Q = c.Q;
          m.A16 = w->A16.as<unsigned short>(); m.B16 = w->B16.as<unsigned short>();
          m.ktile = c.nmix == c.Np / 64 ? w->ktile64.as<int>() : w->ktile32.as<int>();
          m.task = nullptr; m.W[0] = w->f.as<double>();
          m.part = w->part.as<double>();
          for (const int2 &t : c.task)
               task.push_back(make_int2(t.x, ((t.x < 0 ? 0 : t.y) & 0xffff) | (int) (k << 16)));
          ChMap map;
          memset(&map, 0, sizeof map);
          map.g[0] = 0;
          red[k] = reduce_parts_B{w->part.as<double>(), c.nmix, 1, c.Q, map, w->pf.as<double>(),
                                  (unsigned) (4 * c.Q + 31) / 32, 1};
          for (int gs = 0; gs < 2; ++gs) {
               const unsigned on = (gs ? c.J->calc_gc : c.J->calc_lc) ? 1u : 0u;
               rows[gs].push_back(expand_rows_B{w->pf.as<double>(), 1, c.Q, c.H, c.Hp, w->w.as<double>(), gs,
                                                w->E.as<double>(), (unsigned) c.Hp / 64, on});
               // generalized spherical into coefs[0 .. 6C), Legendre into [6C .. 12C)
               coefs[gs].push_back(expand_coefs_B{w->T.as<double>(), w->E.as<double>(), c.H, c.Hp, c.C, gs,
                                                  w->coefs.as<double>() + (gs ? 0 : 6 * (size_t) c.C),
                                                  (unsigned) (c.C + 3) / 4, on});
          }
     }
     BatchBlob b;
     const size_t o_chains = b.add(chains), o_gauss = b.add(gauss), o_gsf = b.add(gsf), o_pitau = b.add(pitau),
                  o_fac = b.add(fac), o_terms = b.add(terms), o_sums = b.add(sums), o_dots = b.add(dots),
                  o_mix = b.add(mix), o_task = b.add(task), o_red = b.add(red), o_rows0 = b.add(rows[0]),
                  o_rows1 = b.add(rows[1]), o_coefs0 = b.add(coefs[0]), o_coefs1 = b.add(coefs[1]);
     if (L->bargs.get(b.h.size())) return -1;
     unsigned char *d = L->bargs.as<unsigned char>();
     HIPCHECK(hipMemcpyAsync(d, b.h.data(), b.h.size(), hipMemcpyHostToDevice, s));
#define LMIE_B(T, off) ((const T *) (d + (off)))
     // main stream: per-size recurrences; side streams: the angle tables
     HIPCHECK(hipEventRecord(L->fork, s));
     HIPCHECK(hipStreamWaitEvent(ts, L->fork, 0));
     hipLaunchKernelGGL(mie_chains_b, dim3(max_g(chains, &mie_chains_B::gx), 2, n), dim3(64), 0, s,
                        LMIE_B(mie_chains_B, o_chains));
     hipLaunchKernelGGL(gauss_nodes_b, dim3(max_g(gauss, &gauss_nodes_B::gx), 1, n), dim3(64), 0, ts,
                        LMIE_B(gauss_nodes_B, o_gauss));
     HIPCHECK(hipEventRecord(L->nodes, ts));
     HIPCHECK(hipStreamWaitEvent(t2, L->nodes, 0));
     hipLaunchKernelGGL(gsf_table_b, dim3(max_g(gsf, &gsf_table_B::gx), max_g(gsf, &gsf_table_B::gy), n), dim3(64), 0,
                        t2, LMIE_B(gsf_table_B, o_gsf));
     HIPCHECK(hipEventRecord(L->join3, t2));
     hipLaunchKernelGGL(pitau_table16_b<E>, dim3(max_g(pitau, &pitau_table16_B::gx), 1, n), dim3(64), 0, ts,
                        LMIE_B(pitau_table16_B, o_pitau));
     HIPCHECK(hipEventRecord(L->join, ts));
     hipLaunchKernelGGL(order_factors_b, dim3(max_g(fac, &order_factors_B::gx), 1, n), dim3(256), 0, s,
                        LMIE_B(order_factors_B, o_fac));
     const unsigned tgy = max_g(terms, &mie_terms_B::gy);
     if (tgy > 0)
          hipLaunchKernelGGL(mie_terms_b<E>, dim3(max_g(terms, &mie_terms_B::gx), tgy, n), dim3(64), 0, s,
                             LMIE_B(mie_terms_B, o_terms));
     hipLaunchKernelGGL(mie_sums_b, dim3(max_g(sums, &mie_sums_B::gx), 1, n), dim3(256), 0, s,
                        LMIE_B(mie_sums_B, o_sums));
     hipLaunchKernelGGL(weighted_dots_b, dim3(4, 1, n), dim3(256), 0, s, LMIE_B(weighted_dots_B, o_dots));
     // contraction after the split tables, one task list for all cases
     HIPCHECK(hipStreamWaitEvent(s, L->join, 0));
     if (!task.empty())
     {
          const dim3 g((unsigned) task.size());
          const int np = mix_products(), id = mix_shape().id;
#define LMIE_MIXB(NP, S) hipLaunchKernelGGL((angular_mix_b<E, NP, 1, S>), g, dim3(S::NT), 0, s, LMIE_B(Mix2Args, o_mix), LMIE_B(int2, o_task))
          if (id == 0) { if (np == 3) LMIE_MIXB(3, MixS0); else LMIE_MIXB(5, MixS0); }
          else if (id == 2) { if (np == 3) LMIE_MIXB(3, MixS2); else LMIE_MIXB(5, MixS2); }
          else if (id == 3) { if (np == 3) LMIE_MIXB(3, MixS3); else if (np == 4) LMIE_MIXB(4, MixS3); else LMIE_MIXB(5, MixS3); }
          else { if (np == 3) LMIE_MIXB(3, MixS1); else LMIE_MIXB(5, MixS1); }
#undef LMIE_MIXB
     }
     hipLaunchKernelGGL(reduce_parts_b, dim3(max_g(red, &reduce_parts_B::gx), 1, n), dim3(256), 0, s,
                        LMIE_B(reduce_parts_B, o_red));
     // expansion after the expansion tables: generalized spherical, then Legendre
     HIPCHECK(hipStreamWaitEvent(s, L->join3, 0));
     for (int gs = 1; gs >= 0; --gs) {
          if (max_g(rows[gs], &expand_rows_B::gy) == 0) continue;
          hipLaunchKernelGGL(expand_rows_b, dim3(max_g(rows[gs], &expand_rows_B::gx), 1, n), dim3(64), 0, s,
                             LMIE_B(expand_rows_B, gs ? o_rows1 : o_rows0));
          hipLaunchKernelGGL(expand_coefs_b, dim3(max_g(coefs[gs], &expand_coefs_B::gx), 1, n), dim3(256), 0, s,
                             LMIE_B(expand_coefs_B, gs ? o_coefs1 : o_coefs0));
          HIPCHECK(hipGetLastError());
     }
#undef LMIE_B
     HIPCHECK(hipGetLastError());
     std::vector<double> hd(4 * (size_t) n);
     Download dn;
     for (unsigned k = 0; k < n; ++k) {
          const FwdCase &c = cs[k];
          const size_t cb = sizeof(double) * 6 * c.C;
          if ((c.J->calc_gc && c.R->gc_raw && dn.add(c.R->gc_raw, c.w->coefs.p, cb)) ||
              (c.J->calc_lc && c.R->lc_raw && dn.add(c.R->lc_raw, c.w->coefs.as<double>() + 6 * (size_t) c.C, cb)) ||
              dn.add(hd.data() + 4 * k, c.w->dots.p, sizeof(double) * 4))
               return -1;
     }
     if (dn.flush(L->bdown, L->pdown, s)) return -1;
     for (unsigned k = 0; k < n; ++k) {
          for (int q = 0; q < 4; ++q) cs[k].R->sums[q] = hd[4 * k + q];
          cs[k].w->valid = 1;
     }
     return 0;
}

struct RevCase {
     int slot;
     Workspace *w;
     const lmie_gpu_adj_in *in;
     lmie_gpu_adj_out *out;
     int nck, Kq, Hq, nch, use_gc, use_lc;
     std::vector<int2> task;
};

int fused_reverse_ok(Workspace *w, const lmie_gpu_adj_in *in) {
     return w->valid && w->mixed && w->Q > 0 && ((in->gc_bar && w->calc_gc) || (in->lc_bar && w->calc_lc));
}

// Buffers and uploads of one reverse pass (the mixed branch of adjoint_impl),
// queued on up.
int fused_rev_prepare(RevCase &c, Upload &up) {
     Workspace *w = c.w;
     const int Np = w->Np, Hp = w->Hp, H = w->H, Q = w->Q, C = w->C, n1max = w->n1max;
     const int Kq = rup(n1max, 64), Hq = rup(H, KM);
     if (rup(n1max, KM) > w->Kb * 8) return set_err("internal: split operand depth");
     c.use_gc = c.in->gc_bar && w->calc_gc;
     c.use_lc = c.in->lc_bar && w->calc_lc;
     c.nck = (C + EXA_L - 1) / EXA_L;
     c.Kq = Kq; c.Hq = Hq;
     c.nch = (w->n2rows - 1 + RCH - 1) / RCH;
     const size_t nk = w->kt128.size(), nk2 = nk + (nk & 1);   // whole doubles
     if (w->pfbar.get(sizeof(double) * 4 * (size_t) Q) || w->fang.get(sizeof(double) * Np) ||
         w->mbar.get(sizeof(double) * 2 * Np) || w->fbar.get(sizeof(double) * 2 * Np) ||
         w->gbar.get(sizeof(double) * 12 * (size_t) C) ||
         w->expart.get(sizeof(double) * 24 * (size_t) c.nck * Hp) || w->wts.get(sizeof(double) * 8 * (size_t) Hq) ||
         w->scal.get(sizeof(AdjScal)) || w->L16.get(sizeof(unsigned short) * 4 * (size_t) Hq * Kq) ||
         w->R16.get(sizeof(unsigned short) * 16 * (size_t) Hq * Kq) ||
         w->Gf.get(sizeof(float) * 4 * (size_t) Kq * Kq) ||
         w->G16.get(sizeof(unsigned short) * 12 * (size_t) Kq * Kq) ||
         w->Uf.get(sizeof(float) * 4 * (size_t) Kq * Np) ||
         w->revScan.get(sizeof(double) * REV_ROWS * (size_t) (c.nch > 0 ? c.nch : 1) * Np) ||
         w->ktile128.get(sizeof(int) * nk2) || w->dotjobs.get(sizeof(DotJob) * 2) || w->dots.get(sizeof(double) * 2))
          return -1;
     double *g = w->gbar.as<double>();
     std::vector<int> kt(w->kt128);
     kt.resize(nk2, 0);
     const std::vector<int> &kts = w->kt128;
     xcd_tasks(kts, [&](int t) { return (kts[t] + 63) / 64; }, tile_group(), c.task);
     DotJob jobs[2] = {{nullptr, w->mbar.as<double>(), nullptr, nullptr},
                       {nullptr, w->mbar.as<double>() + Np, nullptr, nullptr}};
     if ((c.use_gc && up.add(g, c.in->gc_bar, sizeof(double) * 6 * C)) ||
         (c.use_lc && up.add(g + 6 * C, c.in->lc_bar, sizeof(double) * 6 * C)) ||
         up.zero(w->fang.p, sizeof(double) * Np) || up.add(w->ktile128.p, kt.data(), sizeof(int) * nk2) ||
         (c.task.empty() && up.zero(w->Uf.p, sizeof(float) * 4 * (size_t) Kq * Np)) ||
         up.add(w->dotjobs.p, jobs, sizeof(jobs)))
          return -1;
     return 0;
}

template <class E>
int fused_rev_launch(std::vector<RevCase> &cs, Upload &up, Workspace *L) {
     hipStream_t s = L->stream;
     if (up.flush(L->bup, L->pup, s)) return -1;
     const unsigned n = (unsigned) cs.size();
     std::vector<expand_adjoint_part_B> epart(n);
     std::vector<expand_adjoint_fin_B> efin(n);
     std::vector<adj_weights_B> wts(n);
     std::vector<adj_tables_B> tabs(n);
     std::vector<adj_gbuild_B> gb(n);
     std::vector<adj_gsplit_B> gsp(n);
     std::vector<UArgs> ua(n);
     std::vector<int2> task;
     std::vector<rev_orders_B> ro(n);
     std::vector<rev_carry_B> rc(n);
     std::vector<weighted_dots_B> dots(n);
     for (unsigned k = 0; k < n; ++k) {
          const RevCase &c = cs[k];
          Workspace *w = c.w;
          const int H = w->H, Hp = w->Hp, Q = w->Q, C = w->C, Np = w->Np;
          const double *g = w->gbar.as<double>();
          AdjScal *sc = w->scal.as<AdjScal>();
          epart[k] = expand_adjoint_part_B{w->T.as<double>(), c.use_gc ? g : nullptr, c.use_lc ? g + 6 * C : nullptr,
                                           H, Hp, C, w->expart.as<double>(), (unsigned) Hp / 64, (unsigned) c.nck};
          efin[k] = expand_adjoint_fin_B{w->expart.as<double>(), c.nck, w->w.as<double>(), H, Hp, Q,
                                         w->pfbar.as<double>(), (unsigned) (H + EXF_J - 1) / EXF_J, 1};
          wts[k] = adj_weights_B{w->pfbar.as<double>(), H, c.Hq, Q, w->wts.as<double>(), sc, 1, 1};
          tabs[k] = adj_tables_B{w->B16.as<unsigned short>(), w->Kb, w->Hb, w->wts.as<double>(), H, c.Hq, c.Kq,
                                 w->n1max, w->L16.as<unsigned short>(), w->R16.as<unsigned short>(),
                                 (unsigned) c.Kq / 64, (unsigned) c.Hq / 8};
          gb[k] = adj_gbuild_B{w->L16.as<unsigned short>(), w->R16.as<unsigned short>(), c.Hq, c.Kq,
                               w->Gf.as<float>(), sc, (unsigned) c.Kq / 64, (unsigned) c.Kq / 64};
          gsp[k] = adj_gsplit_B{w->Gf.as<float>(), sc, c.Kq, w->G16.as<unsigned short>(), (unsigned) c.Kq / 64,
                                (unsigned) c.Kq / 8};
          UArgs &u = ua[k];
          u.Np = Np; u.Kb = w->Kb; u.Kq = c.Kq;
          u.A16 = w->A16.as<unsigned short>(); u.G16 = w->G16.as<unsigned short>();
          u.ktile = w->ktile128.as<int>(); u.task = nullptr; u.sc = sc; u.Uf = w->Uf.as<float>();
          for (const int2 &t : c.task)
               task.push_back(make_int2(t.x, ((t.x < 0 ? 0 : t.y) & 0xffff) | (int) (k << 16)));
          RevArgs r;
          r.N = w->N; r.Np = Np; r.Kp = w->Kp; r.n1max = w->n1max; r.n2rows = w->n2rows; r.nch = c.nch;
          r.ft = w->ftab.as<double>(); r.x = w->x.as<double>(); r.f = w->f.as<double>();
          r.n1 = w->n1.as<int>(); r.n2 = w->n2.as<int>(); r.mr = w->mr; r.mi = w->mi; rev_recips(r);
          r.D = w->chD.as<double>(); r.PC = w->chPC.as<double>(); r.PS = w->chPS.as<double>();
          r.CB = w->CB.as<double>(); r.S = w->S.as<double>(); r.bar = nullptr;   // read only without Uf
          r.Uf = w->Uf.as<float>(); r.Kq = c.Kq; r.fang = w->fang.as<double>();
          r.sb0 = c.in->sums_bar[0]; r.sb1 = c.in->sums_bar[1]; r.sb2 = c.in->sums_bar[2]; r.sb3 = c.in->sums_bar[3];
          r.scan = w->revScan.as<double>(); r.mbar = w->mbar.as<double>(); r.fbar = w->fbar.as<double>();
          r.xbar = c.out->x_bar ? w->fbar.as<double>() + Np : nullptr;
          ro[k] = rev_orders_B{r, (unsigned) Np / 64, (unsigned) c.nch};
          rc[k] = rev_carry_B{r, (unsigned) Np / 64, 1};
          dots[k] = weighted_dots_B{w->dotjobs.as<DotJob>(), Np, w->dots.as<double>(), 2, 1};
     }
     BatchBlob b;
     const size_t o_ep = b.add(epart), o_ef = b.add(efin), o_w = b.add(wts), o_t = b.add(tabs), o_gb = b.add(gb),
                  o_gs = b.add(gsp), o_u = b.add(ua), o_task = b.add(task), o_ro = b.add(ro), o_rc = b.add(rc),
                  o_d = b.add(dots);
     if (L->bargs.get(b.h.size())) return -1;
     unsigned char *d = L->bargs.as<unsigned char>();
     HIPCHECK(hipMemcpyAsync(d, b.h.data(), b.h.size(), hipMemcpyHostToDevice, s));
#define LMIE_B(T, off) ((const T *) (d + (off)))
     hipLaunchKernelGGL(expand_adjoint_part_b, dim3(max_g(epart, &expand_adjoint_part_B::gx),
                                                    max_g(epart, &expand_adjoint_part_B::gy), n),
                        dim3(64), 0, s, LMIE_B(expand_adjoint_part_B, o_ep));
     hipLaunchKernelGGL(expand_adjoint_fin_b, dim3(max_g(efin, &expand_adjoint_fin_B::gx), 1, n), dim3(24 * EXF_J), 0, s,
                        LMIE_B(expand_adjoint_fin_B, o_ef));
     hipLaunchKernelGGL(adj_weights_b, dim3(1, 1, n), dim3(1024), 0, s, LMIE_B(adj_weights_B, o_w));
     hipLaunchKernelGGL(adj_tables_b<E>, dim3(max_g(tabs, &adj_tables_B::gx), max_g(tabs, &adj_tables_B::gy), n),
                        dim3(64), 0, s, LMIE_B(adj_tables_B, o_t));
     hipLaunchKernelGGL(adj_gbuild_b<E>, dim3(max_g(gb, &adj_gbuild_B::gx), max_g(gb, &adj_gbuild_B::gy), n),
                        dim3(256), 0, s, LMIE_B(adj_gbuild_B, o_gb));
     hipLaunchKernelGGL(adj_gsplit_b<E>, dim3(max_g(gsp, &adj_gsplit_B::gx), max_g(gsp, &adj_gsplit_B::gy), n),
                        dim3(64), 0, s, LMIE_B(adj_gsplit_B, o_gs));
     if (!task.empty())
          hipLaunchKernelGGL((adj_ugemm32<E, 1, 128, true>), dim3((unsigned) task.size()), dim3(512), 0, s,
                             UArgs{}, LMIE_B(UArgs, o_u), LMIE_B(int2, o_task));
     const unsigned rgy = max_g(ro, &rev_orders_B::gy);
     if (rgy > 0)
          hipLaunchKernelGGL(rev_orders_b, dim3(max_g(ro, &rev_orders_B::gx), rgy, n), dim3(64), 0, s,
                             LMIE_B(rev_orders_B, o_ro));
     hipLaunchKernelGGL(rev_carry_b, dim3(max_g(rc, &rev_carry_B::gx), 1, n), dim3(64), 0, s,
                        LMIE_B(rev_carry_B, o_rc));
     hipLaunchKernelGGL(weighted_dots_b, dim3(2, 1, n), dim3(256), 0, s, LMIE_B(weighted_dots_B, o_d));
#undef LMIE_B
     HIPCHECK(hipGetLastError());
     std::vector<double> mb(2 * (size_t) n);
     Download dn;
     for (unsigned k = 0; k < n; ++k)
          if (dn.add(mb.data() + 2 * k, cs[k].w->dots.p, sizeof(double) * 2) ||
              dn.add(cs[k].out->f_bar, cs[k].w->fbar.p, sizeof(double) * cs[k].w->N) ||
              (cs[k].out->x_bar &&
               dn.add(cs[k].out->x_bar, cs[k].w->fbar.as<double>() + cs[k].w->Np, sizeof(double) * cs[k].w->N)))
               return -1;
     if (dn.flush(L->bdown, L->pdown, s)) return -1;
     for (unsigned k = 0; k < n; ++k) {
          // m = mr - i mi: dL/dmr = Re(mbar), dL/dmi = -Im(mbar)
          cs[k].out->m_bar[0] = mb[2 * k];
          cs[k].out->m_bar[1] = -mb[2 * k + 1];
     }
     return 0;
}

// The collective: cases register their pass; the last to arrive (or to
// leave without arriving) runs it for all.
struct Batch {
     int n = 0, device = 0;
     pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
     pthread_cond_t cv = PTHREAD_COND_INITIALIZER;
     int arrived[2] = {0, 0}, left[2] = {0, 0}, state[2] = {0, 0};   // state: 0 open, 1 running, 2 done
     std::vector<char> in[2];
     std::vector<const lmie_gpu_job *> job;
     std::vector<lmie_gpu_result *> res;
     std::vector<Workspace *> ws;
     std::vector<const lmie_gpu_adj_in *> ain;
     std::vector<lmie_gpu_adj_out *> aout;
     std::vector<int> rc[2];
     std::vector<std::string> err[2];
     int fused_cases[2] = {0, 0};
     // forward cases prepared by their own threads before they arrive
     // (buffers, host tables and staged uploads), so the thread that runs
     // the pass only launches
     std::vector<char> prepped;
     std::vector<FwdCase> pre;
     std::vector<Upload> preup;
     std::vector<int> prerc;
     std::vector<std::string> preerr;
     // the same for the reverse pass
     std::vector<char> rprepped;
     std::vector<RevCase> rpre;
     std::vector<Upload> rpreup;
     std::vector<int> rprerc;
     std::vector<std::string> rpreerr;
};

thread_local Batch *t_batch = nullptr;
thread_local int t_slot = -1, t_phase = 0;

int forward_one(const lmie_gpu_job *job, lmie_gpu_result *res);
int adjoint_one(Workspace *w, const lmie_gpu_adj_in *in, lmie_gpu_adj_out *out);

// A forward case of a fused batch prepared on its own thread, before it
// arrives: its workspace, host tables and staged uploads (fused_prepare).
void batch_prepare(Batch *b, int k, const lmie_gpu_job *J, lmie_gpu_result *res) {
     g_err[0] = 0;
     FwdCase &c = b->pre[k];
     c.slot = k; c.J = J; c.R = res; c.w = nullptr;
     int rc = use_device(J->device);
     if (rc == 0) rc = ws_acquire_home(J->device, k, &c.w);
     if (rc == 0) {
          // the case's inputs go up from here, on its own stream, one pinned
          // copy; the launch waits for the event
          Upload &up = b->preup[k];
          up.items.clear();
          up.h.clear();
          rc = fused_prepare(c, up);
          if (rc == 0) rc = up.flush(c.w->finDev, c.w->finPin, c.w->stream);
          if (rc == 0 && hipEventRecord(c.w->uploaded, c.w->stream) != hipSuccess)
               rc = set_err("hipEventRecord failed");
          if (rc) {
               (void) hipStreamSynchronize(c.w->stream);
               ws_release(c.w);
               c.w = nullptr;
          }
     }
     b->prerc[k] = rc;
     b->preerr[k] = rc ? g_err : "";
     b->prepped[k] = 1;
}

// A reverse case of a fused batch prepared on its own thread, before it
// arrives (fused_rev_prepare).  Its uploads stay staged on the host: the pass
// sends those of all cases as one copy (a few tens of KB per case; a copy, a
// scatter launch and an event per case from every case's thread contend for
// the runtime).
void batch_rev_prepare(Batch *b, int k, Workspace *w, const lmie_gpu_adj_in *in, lmie_gpu_adj_out *out) {
     g_err[0] = 0;
     RevCase &c = b->rpre[k];
     c.slot = k; c.w = w; c.in = in; c.out = out;
     Upload &up = b->rpreup[k];
     up.items.clear();
     up.h.clear();
     int rc = use_device(w->device);
     if (rc == 0) rc = fused_rev_prepare(c, up);
     b->rprerc[k] = rc;
     b->rpreerr[k] = rc ? g_err : "";
     b->rprepped[k] = 1;
}

void batch_run_forward(Batch *b) {
     std::vector<FwdCase> f16, b16;
     Upload up16, upb;
     Workspace *L = nullptr;           // the workspace whose streams launch the batch
     for (int k = 0; k < b->n; ++k) {
          if (!b->in[0][k]) continue;
          const lmie_gpu_job *J = b->job[k];
          g_err[0] = 0;
          int rc;
          if (b->prepped[k] && (b->prerc[k] || J->device != b->device)) {
               // prepared on another device, or the preparation failed
               if (b->prerc[k] == 0) {
                    ws_release(b->pre[k].w);
                    rc = forward_one(J, b->res[k]);
               } else {
                    b->rc[0][k] = b->prerc[k];
                    b->err[0][k] = b->preerr[k];
                    continue;
               }
          } else if (b->prepped[k]) {
               // its inputs are on the way on its own stream (batch_prepare)
               FwdCase &c = b->pre[k];
               if (!L) L = c.w;
               if (hipStreamWaitEvent(L->stream, c.w->uploaded, 0) != hipSuccess) {
                    (void) hipStreamSynchronize(c.w->stream);
                    if (c.w != L) ws_release(c.w);
                    b->rc[0][k] = set_err("hipStreamWaitEvent failed");
                    b->err[0][k] = g_err;
                    continue;
               }
               const bool bf = J->precision == LMIE_PRECISION_MIXED_BF16;
               (bf ? b16 : f16).push_back(std::move(c));
               continue;
          } else if (!fused_forward_ok(J) || J->device != b->device) {
               rc = forward_one(J, b->res[k]);
          } else {
               FwdCase c;
               c.slot = k; c.J = J; c.R = b->res[k]; c.w = nullptr;
               rc = ws_acquire_home(J->device, k, &c.w);
               if (rc == 0) {
                    if (!L) L = c.w;
                    const bool bf = J->precision == LMIE_PRECISION_MIXED_BF16;
                    Upload &up = bf ? upb : up16;
                    const size_t ni = up.items.size(), nb = up.h.size();
                    rc = fused_prepare(c, up);
                    if (rc == 0) {
                         (bf ? b16 : f16).push_back(std::move(c));
                         continue;
                    }
                    up.rewind(ni, nb);
                    if (c.w != L) ws_release(c.w);
               }
          }
          b->rc[0][k] = rc;
          b->err[0][k] = rc ? g_err : "";
     }
     if (!L) return;
     bool L_kept = false;
     for (int t = 0; t < 2; ++t) {     // one arithmetic type per launch
          std::vector<FwdCase> &v = t ? b16 : f16;
          if (v.empty()) continue;
          g_err[0] = 0;
          const int rc = t ? fused_launch<BF16>(v, upb, L) : fused_launch<F16>(v, up16, L);
          if (rc) (void) hipStreamSynchronize(L->stream);
          for (FwdCase &c : v) {
               b->rc[0][c.slot] = rc;
               b->err[0][c.slot] = rc ? g_err : "";
               c.R->state = rc ? nullptr : c.w;
               c.R->precision = c.J->precision;
               if (rc == 0 && c.w == L) L_kept = true;
               if (rc && c.w != L) ws_release(c.w);
               b->fused_cases[0]++;
          }
     }
     if (!L_kept) ws_release(L);
}

void batch_run_reverse(Batch *b) {
     std::vector<RevCase> f16, b16;
     Upload up16, upb;
     std::vector<const Upload *> staged;
     Workspace *L = nullptr;
     for (int k = 0; k < b->n; ++k) {
          if (!b->in[1][k]) continue;
          Workspace *w = b->ws[k];
          g_err[0] = 0;
          int rc;
          if (b->rprepped[k] && b->rprerc[k]) {
               b->rc[1][k] = b->rprerc[k];
               b->err[1][k] = b->rpreerr[k];
               continue;
          } else if (b->rprepped[k] && w->device == b->device) {
               // its uploads are staged (batch_rev_prepare)
               RevCase &c = b->rpre[k];
               if (!L) L = w;
               staged.push_back(&b->rpreup[k]);
               (w->bf16 ? b16 : f16).push_back(std::move(c));
               continue;
          } else if (!fused_reverse_ok(w, b->ain[k]) || w->device != b->device) {
               rc = adjoint_one(w, b->ain[k], b->aout[k]);
          } else {
               RevCase c;
               c.slot = k; c.w = w; c.in = b->ain[k]; c.out = b->aout[k];
               if (!L) L = w;
               Upload &up = w->bf16 ? upb : up16;
               const size_t ni = up.items.size(), nb = up.h.size();
               rc = fused_rev_prepare(c, up);
               if (rc == 0) {
                    (w->bf16 ? b16 : f16).push_back(std::move(c));
                    continue;
               }
               up.rewind(ni, nb);
          }
          b->rc[1][k] = rc;
          b->err[1][k] = rc ? g_err : "";
     }
     // every case's uploads in one copy on the launching stream, before
     // either arithmetic's launches
     int up_rc = 0;
     if (L) {
          g_err[0] = 0;
          staged.push_back(&up16);
          staged.push_back(&upb);
          up_rc = Upload::flush_all(staged, L->bup, L->pup, L->stream);
          up16.items.clear(); up16.h.clear();
          upb.items.clear(); upb.h.clear();
     }
     const std::string up_err = up_rc ? g_err : "";
     for (int t = 0; t < 2; ++t) {
          std::vector<RevCase> &v = t ? b16 : f16;
          if (v.empty()) continue;
          g_err[0] = 0;
          if (up_rc) snprintf(g_err, sizeof g_err, "%s", up_err.c_str());
          const int rc = up_rc ? up_rc : t ? fused_rev_launch<BF16>(v, upb, L) : fused_rev_launch<F16>(v, up16, L);
          if (rc) (void) hipStreamSynchronize(L->stream);
          for (RevCase &c : v) {
               b->rc[1][c.slot] = rc;
               b->err[1][c.slot] = rc ? g_err : "";
               b->fused_cases[1]++;
          }
     }
}

void batch_fire(Batch *b, int p) {   // with b->mu held; runs phase p if complete
     if (b->state[p] != 0 || b->arrived[p] + b->left[p] != b->n) return;
     if (b->arrived[p] == 0) {
          b->state[p] = 2;
          return;
     }
     b->state[p] = 1;
     pthread_mutex_unlock(&b->mu);
     (void) use_device(b->device);
     if (p == 0) batch_run_forward(b);
     else batch_run_reverse(b);
     pthread_mutex_lock(&b->mu);
     b->state[p] = 2;
     pthread_cond_broadcast(&b->cv);
}

int batch_arrive(int p, const lmie_gpu_job *job, lmie_gpu_result *res, Workspace *w, const lmie_gpu_adj_in *ain,
                 lmie_gpu_adj_out *aout) {
     Batch *b = t_batch;
     const int k = t_slot;
     // a fused case does its own preparation first, beside the others
     if (p == 0 && fused_forward_ok(job)) batch_prepare(b, k, job, res);
     if (p == 1 && fused_reverse_ok(w, ain)) batch_rev_prepare(b, k, w, ain, aout);
     pthread_mutex_lock(&b->mu);
     if (p == 0) {
          b->job[k] = job;
          b->res[k] = res;
          if (b->arrived[0] == 0) b->device = job->device;
     } else {
          b->ws[k] = w;
          b->ain[k] = ain;
          b->aout[k] = aout;
          if (b->arrived[1] == 0) b->device = w->device;
     }
     b->in[p][k] = 1;
     b->arrived[p]++;
     t_phase = p + 1;
     batch_fire(b, p);
     while (b->state[p] != 2) pthread_cond_wait(&b->cv, &b->mu);
     const int rc = b->rc[p][k];
     const std::string e = b->err[p][k];
     pthread_mutex_unlock(&b->mu);
     if (rc) return set_err("%s", e.c_str());
     return 0;
}

// Plain 16-byte-per-lane copy: the calibration point for HBM traffic counters
// (n16 elements of 16 bytes read and written once).
__global__ __launch_bounds__(256) void copy16(const v4u *src, v4u *dst, size_t n16) {
     for (size_t e = (size_t) blockIdx.x * blockDim.x + threadIdx.x; e < n16; e += (size_t) gridDim.x * blockDim.x)
          dst[e] = src[e];
}

// Lane-layout check of the FP64 matrix instruction used by the contraction:
// D = A (16x4) B (4x16), all row-major, one wave.
__global__ void mfma_f64_probe(const double *A, const double *B, double *D) {
     const int l = threadIdx.x;
     const double a = A[(l & 15) * 4 + (l >> 4)];
     const double b = B[(l >> 4) * 16 + (l & 15)];
     v4d c = {0., 0., 0., 0.};
     c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
     for (int v = 0; v < 4; ++v) D[((l >> 4) + 4 * v) * 16 + (l & 15)] = c[v];
}

// Lane-layout check of the 32x32x16 fp16 matrix instruction: D = A (32x16)
// B (16x32), row-major float inputs, one wave.  Operand lane l: row (column)
// l % 32, elements 8 (l / 32) .. +7 of the contraction; result lane l, value v:
// column l % 32, row given by mfma32_row (checked by lmie_gpu_selftest_mfma32).


__global__ void mfma32_probe(const float *A, const float *B, float *D) {
     const int l = threadIdx.x;
     unsigned short av[8], bv[8];
     double r;
     for (int e = 0; e < 8; ++e) {
          av[e] = F16::bits(A[(l & 31) * 16 + 8 * (l >> 5) + e], &r);
          bv[e] = F16::bits(B[(8 * (l >> 5) + e) * 32 + (l & 31)], &r);
     }
     v4u a, b;
     for (int k = 0; k < 4; ++k) {
          a[k] = (unsigned) av[2 * k] | ((unsigned) av[2 * k + 1] << 16);
          b[k] = (unsigned) bv[2 * k] | ((unsigned) bv[2 * k + 1] << 16);
     }
     v16f c;
     for (int v = 0; v < 16; ++v) c[v] = 0.f;
     c = F16::mfma32(a, b, c);
     for (int v = 0; v < 16; ++v) D[mfma32_row(l, v) * 32 + (l & 31)] = c[v];
}

// Lane-layout check of the 16-bit matrix instructions used by the mixed
// contraction: D = A (16x32) B (32x16), row-major float inputs (exactly
// representable in the 16-bit type), one wave.
template <class E>
__global__ void mfma16_probe(const float *A, const float *B, float *D) {
     const int l = threadIdx.x;
     unsigned short av[8], bv[8];
     double r;
     for (int e = 0; e < 8; ++e) {
          av[e] = E::bits(A[(l & 15) * 32 + 8 * (l >> 4) + e], &r);
          bv[e] = E::bits(B[(8 * (l >> 4) + e) * 16 + (l & 15)], &r);
     }
     v4u a, b;
     for (int k = 0; k < 4; ++k) {
          a[k] = (unsigned) av[2 * k] | ((unsigned) av[2 * k + 1] << 16);
          b[k] = (unsigned) bv[2 * k] | ((unsigned) bv[2 * k + 1] << 16);
     }
     v4f c = {0.f, 0.f, 0.f, 0.f};
     c = E::mfma(a, b, c);
     for (int v = 0; v < 4; ++v) D[mfma16_row(l, v) * 16 + (l & 15)] = c[v];
}

}  // namespace

extern "C" {

const char *lmie_gpu_last_error(void) { return g_err; }

int lmie_set_timing(int on) {
     g_timing = on ? 1 : 0;
     return 0;
}

void lmie_gpu_host_time(int k, double ms) {
     if (g_timing && k >= LMIE_T_HOST_GRID && k < LMIE_N_TIMERS) g_times[k] = ms;
}

void lmie_gpu_host_times_clear(void) {
     for (int k = LMIE_T_HOST_GRID; k < LMIE_N_TIMERS; ++k) g_times[k] = 0.;
}

int lmie_get_timing(double *ms, int n) {
     for (int k = 0; k < n && k < LMIE_N_TIMERS; ++k) ms[k] = g_times[k];
     return 0;
}

int lmie_gpu_forward(const lmie_gpu_job *job, lmie_gpu_result *res) {
     g_err[0] = 0;
     if (!job || !res) return set_err("null argument");
     if (t_batch && t_phase == 0) {
          res->state = nullptr;
          return batch_arrive(0, job, res, nullptr, nullptr, nullptr);
     }
     return forward_one(job, res);
}

int lmie_gpu_adjoint(void *state, const lmie_gpu_adj_in *in, lmie_gpu_adj_out *out) {
     g_err[0] = 0;
     if (!state || !in || !out || !in->sums_bar || !out->f_bar) return set_err("null argument");
     if (t_batch && t_phase == 1) return batch_arrive(1, nullptr, nullptr, (Workspace *) state, in, out);
     return adjoint_one((Workspace *) state, in, out);
}

void *lmie_gpu_batch_open(int n_case) {
     g_err[0] = 0;
     if (n_case < 1 || n_case > 32767) {
          set_err("batch of %d cases (1 .. 32767)", n_case);
          return nullptr;
     }
     Batch *b = new Batch();
     b->n = n_case;
     for (int p = 0; p < 2; ++p) {
          b->in[p].assign(n_case, 0);
          b->rc[p].assign(n_case, 0);
          b->err[p].assign(n_case, std::string());
     }
     b->job.assign(n_case, nullptr);
     b->res.assign(n_case, nullptr);
     b->ws.assign(n_case, nullptr);
     b->ain.assign(n_case, nullptr);
     b->aout.assign(n_case, nullptr);
     b->prepped.assign(n_case, 0);
     b->pre.resize(n_case);
     b->preup.resize(n_case);
     b->prerc.assign(n_case, 0);
     b->preerr.assign(n_case, std::string());
     b->rprepped.assign(n_case, 0);
     b->rpre.resize(n_case);
     b->rpreup.resize(n_case);
     b->rprerc.assign(n_case, 0);
     b->rpreerr.assign(n_case, std::string());
     return b;
}

int lmie_gpu_batch_enter(void *batch, int slot) {
     g_err[0] = 0;
     Batch *b = (Batch *) batch;
     if (!b || slot < 0 || slot >= b->n) return set_err("batch slot %d", slot);
     if (t_batch) return set_err("this thread is already in a batch");
     t_batch = b;
     t_slot = slot;
     t_phase = 0;
     return 0;
}

void lmie_gpu_batch_leave(void) {
     Batch *b = t_batch;
     if (!b) return;
     pthread_mutex_lock(&b->mu);
     for (int p = t_phase; p < 2; ++p) b->left[p]++;
     for (int p = 0; p < 2; ++p) batch_fire(b, p);
     pthread_mutex_unlock(&b->mu);
     t_batch = nullptr;
     t_slot = -1;
     t_phase = 0;
}

int lmie_gpu_batch_fused(void *batch, int pass) {
     Batch *b = (Batch *) batch;
     return b && pass >= 0 && pass < 2 ? b->fused_cases[pass] : 0;
}

void lmie_gpu_batch_close(void *batch) { delete (Batch *) batch; }
}

namespace {
int forward_one(const lmie_gpu_job *job, lmie_gpu_result *res) {
     if (job->precision != LMIE_PRECISION_FP64 && job->precision != LMIE_PRECISION_MIXED &&
         job->precision != LMIE_PRECISION_MIXED_BF16)
          return set_err("unknown precision %d", job->precision);
     if (job->save_control == 2 && job->keep) return set_err("save_control = 2 keeps no reverse-pass state");
     Workspace *w = nullptr;
     if (ws_acquire(job->device, &w)) return -1;
     res->state = nullptr;
     res->precision = LMIE_PRECISION_FP64;   /* the re-weighting of save_control 2 is FP64 */
     int rc = job->save_control == 2 ? saved_impl(job, res, w) : forward_impl(job, res, w);
     if (rc) {
          (void) hipStreamSynchronize(w->side);
          (void) hipStreamSynchronize(w->side2);
          (void) hipStreamSynchronize(w->stream);
     }
     if (rc == 0 && job->keep) {
          res->state = w;          // owned by the caller until lmie_gpu_release
          return 0;
     }
     ws_release(w);
     return rc;
}

int adjoint_one(Workspace *w, const lmie_gpu_adj_in *in, lmie_gpu_adj_out *out) {
     if (use_device(w->device)) return -1;
     int rc = adjoint_impl(w, in, out);
     if (rc) (void) hipStreamSynchronize(w->stream);
     return rc;
}
}  // namespace

extern "C" {

void lmie_gpu_release(void *state) {
     if (state) ws_release((Workspace *) state);
}

int lmie_gpu_phase_func(int n_set, int n_coef, int ld, const double *coefs, int n_theta, const double *theta,
                        double *pf, int gsf, int device) {
     g_err[0] = 0;
     if (n_set <= 0 || n_theta <= 0) return 0;
     if (n_coef > ld || n_coef < 0) return set_err("n_coef %d outside [0, %d]", n_coef, ld);
     HIPCHECK(hipSetDevice(device));
     double *dX = nullptr, *dT = nullptr, *dP = nullptr;
     const size_t nx = (size_t) n_set * 6 * ld, np = (size_t) n_set * 6 * n_theta;
     HIPCHECK(hipMalloc(&dX, sizeof(double) * (nx ? nx : 1)));
     HIPCHECK(hipMalloc(&dT, sizeof(double) * n_theta));
     HIPCHECK(hipMalloc(&dP, sizeof(double) * np));
     HIPCHECK(hipMemcpy(dX, coefs, sizeof(double) * nx, hipMemcpyHostToDevice));
     HIPCHECK(hipMemcpy(dT, theta, sizeof(double) * n_theta, hipMemcpyHostToDevice));
     const int tot = n_set * n_theta;
     hipLaunchKernelGGL(phase_at_angles, dim3((tot + 127) / 128), dim3(128), 0, 0, n_set, n_coef, ld,
                        (const double *) dX, n_theta, (const double *) dT, dP, gsf);
     HIPCHECK(hipGetLastError());
     HIPCHECK(hipMemcpy(pf, dP, sizeof(double) * np, hipMemcpyDeviceToHost));
     HIPCHECK(hipFree(dX));
     HIPCHECK(hipFree(dT));
     HIPCHECK(hipFree(dP));
     return 0;
}

// Staging of lmie_gpu_phase_func_static: the coefficients, the angles and the
// phase function of one call, in one file-scope device array in place of the
// three device blocks lmie_gpu_phase_func allocates per call.  Every GPU the
// library runs on holds its own copy of the array (16 MiB) until the process
// ends.  In host code the name `scratch` is not a device address: the address
// of each device's copy is taken once (hipGetSymbolAddress, with that device
// current) and kept in pf_scratch[device]; only that address is passed on.
constexpr size_t PF_SCRATCH_DOUBLES = (size_t) 1 << 21;
constexpr int PF_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[PF_SCRATCH_DOUBLES];
static double *pf_scratch[PF_MAX_DEVICES];     // guarded by pf_lock[device]
static std::mutex pf_lock[PF_MAX_DEVICES];

static int phase_func_static_locked(int n_set, int n_coef, int ld, const double *coefs, int n_theta,
                                    const double *theta, double *pf, int gsf, int device, size_t ot, size_t op) {
     HIPCHECK(hipSetDevice(device));
     if (!pf_scratch[device]) {
          void *p = nullptr;
          HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
          if (!p) return set_err("hipGetSymbolAddress gave no address for the phase-function scratch");
          pf_scratch[device] = (double *) p;
     }
     double *dX = pf_scratch[device], *dT = dX + ot, *dP = dX + op;
     const size_t nx = (size_t) n_set * 6 * ld, np = (size_t) n_set * 6 * n_theta;
     if (nx) HIPCHECK(hipMemcpy(dX, coefs, sizeof(double) * nx, hipMemcpyHostToDevice));
     HIPCHECK(hipMemcpy(dT, theta, sizeof(double) * n_theta, hipMemcpyHostToDevice));
     const int tot = n_set * n_theta;
     hipLaunchKernelGGL(phase_at_angles, dim3((tot + 127) / 128), dim3(128), 0, 0, n_set, n_coef, ld,
                        (const double *) dX, n_theta, (const double *) dT, dP, gsf);
     HIPCHECK(hipGetLastError());
     HIPCHECK(hipMemcpy(pf, dP, sizeof(double) * np, hipMemcpyDeviceToHost));
     return 0;
}

int lmie_gpu_phase_func_static(int n_set, int n_coef, int ld, const double *coefs, int n_theta,
                               const double *theta, double *pf, int gsf, int device) {
     g_err[0] = 0;
     if (n_set <= 0 || n_theta <= 0) return 0;
     if (n_coef > ld || n_coef < 0) return set_err("n_coef %d outside [0, %d]", n_coef, ld);
     if (device < 0 || device >= PF_MAX_DEVICES)
          return set_err("device %d outside [0, %d)", device, PF_MAX_DEVICES);
     // layout in the array (doubles): coefficients at 0, angles at ot, phase
     // function at op, each start a multiple of 32 (256 bytes); every factor
     // is bounded first so that the products cannot wrap
     const size_t cap = PF_SCRATCH_DOUBLES;
     if ((size_t) n_set > cap || (size_t) ld > cap || (size_t) n_theta > cap)
          return set_err("lmie_gpu_phase_func_static: n_set %d, ld %d, n_theta %d exceed its static scratch of %zu doubles",
                         n_set, ld, n_theta, cap);
     const size_t nx = (size_t) n_set * 6 * ld, np = (size_t) n_set * 6 * n_theta;
     const size_t ot = (nx + 31) / 32 * 32, op = (ot + (size_t) n_theta + 31) / 32 * 32;
     if (op + np > cap)
          return set_err("lmie_gpu_phase_func_static: %zu doubles of staging (coefficients %zu, angles %d, "
                         "phase function %zu) exceed its static scratch of %zu doubles",
                         op + np, nx, n_theta, np, cap);
     std::lock_guard<std::mutex> hold(pf_lock[device]);
     return phase_func_static_locked(n_set, n_coef, ld, coefs, n_theta, theta, pf, gsf, device, ot, op);
}

int lmie_gpu_gauss_rule(int n, double *x, double *wt, int device) {
     if (n < 1) return set_err("n < 1");
     HIPCHECK(hipSetDevice(device));
     std::vector<double> g((n + 1) / 2);
     gauss_guesses(n, g.data());
     double *d = nullptr;
     HIPCHECK(hipMalloc(&d, sizeof(double) * (g.size() + 2 * (size_t) n)));
     HIPCHECK(hipMemcpy(d, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice));
     hipLaunchKernelGGL(gauss_nodes, dim3(((n + 1) / 2 + 63) / 64), dim3(64), 0, 0, n, (const double *) d,
                        d + g.size(), d + g.size() + n);
     HIPCHECK(hipGetLastError());
     HIPCHECK(hipMemcpy(x, d + g.size(), sizeof(double) * n, hipMemcpyDeviceToHost));
     HIPCHECK(hipMemcpy(wt, d + g.size() + n, sizeof(double) * n, hipMemcpyDeviceToHost));
     HIPCHECK(hipFree(d));
     return 0;
}

int lmie_gpu_gsf_table(int n_mu, const double *mu, int n_l, double *out, int device) {
     if (n_mu < 1 || n_l < 1) return set_err("n_mu < 1 or n_l < 1");
     if (!mu || !out) return set_err("missing argument");
     HIPCHECK(hipSetDevice(device));
     // gsf_table reads the angles at mu[H + j], j < H, and writes [5][n_l][Hp]
     const int H = n_mu, Hp = rup(H, 64);
     const size_t tab = (size_t) 5 * n_l * Hp;
     double *d = nullptr;
     HIPCHECK(hipMalloc(&d, sizeof(double) * (2 * (size_t) H + tab)));
     HIPCHECK(hipMemcpy(d + H, mu, sizeof(double) * H, hipMemcpyHostToDevice));
     hipLaunchKernelGGL(gsf_table, dim3(Hp / 64, 5), dim3(64), 0, 0, H, Hp, n_l, (const double *) d, d + 2 * H);
     HIPCHECK(hipGetLastError());
     std::vector<double> t(tab);
     HIPCHECK(hipMemcpy(t.data(), d + 2 * H, sizeof(double) * tab, hipMemcpyDeviceToHost));
     HIPCHECK(hipFree(d));
     for (size_t r = 0; r < (size_t) 5 * n_l; ++r) memcpy(out + r * H, t.data() + r * Hp, sizeof(double) * H);
     return 0;
}

int lmie_gpu_pitau_table(int n_mu, const double *mu, int n, double *bp, double *bm, int device) {
     if (n_mu < 1 || n < 1) return set_err("n_mu < 1 or n < 1");
     if (!mu || !bp || !bm) return set_err("missing argument");
     HIPCHECK(hipSetDevice(device));
     // pitau_table reads the angles at mu[H + j], j < H, and writes [n][Hp] twice
     const int H = n_mu, Hp = rup(H, 64);
     const size_t tab = (size_t) n * Hp;
     double *d = nullptr;
     HIPCHECK(hipMalloc(&d, sizeof(double) * (2 * (size_t) H + 2 * tab)));
     HIPCHECK(hipMemcpy(d + H, mu, sizeof(double) * H, hipMemcpyHostToDevice));
     hipLaunchKernelGGL(pitau_table, dim3(Hp / 64), dim3(64), 0, 0, H, Hp, n, n, (const double *) d, d + 2 * H,
                        d + 2 * H + tab);
     HIPCHECK(hipGetLastError());
     std::vector<double> t(2 * tab);
     HIPCHECK(hipMemcpy(t.data(), d + 2 * H, sizeof(double) * 2 * tab, hipMemcpyDeviceToHost));
     HIPCHECK(hipFree(d));
     for (size_t r = 0; r < (size_t) n; ++r) {
          memcpy(bp + r * H, t.data() + r * Hp, sizeof(double) * H);
          memcpy(bm + r * H, t.data() + tab + r * Hp, sizeof(double) * H);
     }
     return 0;
}

int lmie_gpu_selftest_mfma16(int bf16, const float *A, const float *B, float *D, int device) {
     HIPCHECK(hipSetDevice(device));
     float *d = nullptr;
     HIPCHECK(hipMalloc(&d, sizeof(float) * (512 + 512 + 256)));
     HIPCHECK(hipMemcpy(d, A, sizeof(float) * 512, hipMemcpyHostToDevice));
     HIPCHECK(hipMemcpy(d + 512, B, sizeof(float) * 512, hipMemcpyHostToDevice));
     if (bf16)
          hipLaunchKernelGGL(mfma16_probe<BF16>, dim3(1), dim3(64), 0, 0, (const float *) d, (const float *) (d + 512), d + 1024);
     else
          hipLaunchKernelGGL(mfma16_probe<F16>, dim3(1), dim3(64), 0, 0, (const float *) d, (const float *) (d + 512), d + 1024);
     HIPCHECK(hipGetLastError());
     HIPCHECK(hipMemcpy(D, d + 1024, sizeof(float) * 256, hipMemcpyDeviceToHost));
     HIPCHECK(hipFree(d));
     return 0;
}

// Copies n_bytes (a multiple of 16) device to device `reps` times with copy16
// and returns the average kernel time in ms through *ms.
int lmie_gpu_selftest_copy(size_t n_bytes, int reps, int device, double *ms) {
     if (n_bytes == 0 || n_bytes % 16 || reps < 1) return set_err("copy size must be a positive multiple of 16");
     HIPCHECK(hipSetDevice(device));
     void *a = nullptr, *b = nullptr;
     hipEvent_t e0, e1;
     HIPCHECK(hipMalloc(&a, n_bytes));
     HIPCHECK(hipMalloc(&b, n_bytes));
     HIPCHECK(hipMemset(a, 1, n_bytes));
     HIPCHECK(hipEventCreate(&e0));
     HIPCHECK(hipEventCreate(&e1));
     const size_t n16 = n_bytes / 16;
     HIPCHECK(hipEventRecord(e0, 0));
     for (int r = 0; r < reps; ++r)
          hipLaunchKernelGGL(copy16, dim3(2048), dim3(256), 0, 0, (const v4u *) a, (v4u *) b, n16);
     HIPCHECK(hipGetLastError());
     HIPCHECK(hipEventRecord(e1, 0));
     HIPCHECK(hipEventSynchronize(e1));
     float t = 0.f;
     HIPCHECK(hipEventElapsedTime(&t, e0, e1));
     if (ms) *ms = t / reps;
     (void) hipEventDestroy(e0);
     (void) hipEventDestroy(e1);
     HIPCHECK(hipFree(a));
     HIPCHECK(hipFree(b));
     return 0;
}

int lmie_gpu_selftest_mfma32(const float *A, const float *B, float *D, int device) {
     HIPCHECK(hipSetDevice(device));
     float *d = nullptr;
     HIPCHECK(hipMalloc(&d, sizeof(float) * (512 + 512 + 1024)));
     HIPCHECK(hipMemcpy(d, A, sizeof(float) * 512, hipMemcpyHostToDevice));
     HIPCHECK(hipMemcpy(d + 512, B, sizeof(float) * 512, hipMemcpyHostToDevice));
     hipLaunchKernelGGL(mfma32_probe, dim3(1), dim3(64), 0, 0, (const float *) d, (const float *) (d + 512), d + 1024);
     HIPCHECK(hipGetLastError());
     HIPCHECK(hipMemcpy(D, d + 1024, sizeof(float) * 1024, hipMemcpyDeviceToHost));
     HIPCHECK(hipFree(d));
     return 0;
}

int lmie_gpu_selftest_mfma_f64(const double *A, const double *B, double *D, int device) {
     HIPCHECK(hipSetDevice(device));
     double *d = nullptr;
     HIPCHECK(hipMalloc(&d, sizeof(double) * (64 + 64 + 256)));
     HIPCHECK(hipMemcpy(d, A, sizeof(double) * 64, hipMemcpyHostToDevice));
     HIPCHECK(hipMemcpy(d + 64, B, sizeof(double) * 64, hipMemcpyHostToDevice));
     hipLaunchKernelGGL(mfma_f64_probe, dim3(1), dim3(64), 0, 0, (const double *) d, (const double *) (d + 64), d + 128);
     HIPCHECK(hipGetLastError());
     HIPCHECK(hipMemcpy(D, d + 128, sizeof(double) * 256, hipMemcpyDeviceToHost));
     HIPCHECK(hipFree(d));
     return 0;
}

}  // extern "C"
