// Block-sparse kernels, BiCGStab with point-block Jacobi preconditioning, and
// the global-implicit Newton step on the device (include/pfx_gi_solver.h).
//
// Layout: a block row of the BSR matrix is a group of G lanes (G = 8, 16 or
// 32 >= N); lane r owns row r of the block row, so the N lanes of a group
// read the N consecutive doubles of one column of a column-major block at a
// time and the whole block stream is read once, in order.  Every kernel is
// HBM-bound: the matrix (7 N^2 doubles per block row) dominates the bytes of
// a Krylov iteration, two products per iteration.
//
// Reductions are deterministic: each workgroup writes partial sums, one small
// kernel adds them in a fixed order, and the host reads the 16-byte result
// (BiCGStab's scalars are needed on the host to steer the iteration anyway).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "pfx_device.h"
#include "../../include/pfx_gi_solver.h"

namespace {

#define KCHK(x)                                                               \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "pfx_krylov: %s failed: %s (%s:%d)\n", #x,              \
              hipGetErrorString(e_), __FILE__, __LINE__);                     \
      return PFX_E_DEVICE;                                                    \
    }                                                                         \
  } while (0)

constexpr int BLOCK = 256;
constexpr int MAX_GRID = 4096;  // workgroups of the grid-stride kernels

// sum of (a, b) and (c, d) over the workgroup into partial[2 * blockIdx.x]
__device__ __forceinline__ void block_reduce2(double a, double b, double *partial) {
  __shared__ double sh[2][BLOCK / 64];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = a;
    sh[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int i = 0; i < BLOCK / 64; i++) {
      s0 += sh[0][i];
      s1 += sh[1][i];
    }
    partial[2 * blockIdx.x] = s0;
    partial[2 * blockIdx.x + 1] = s1;
  }
}

// out[0..1] = sums of the nparts partial pairs, in a fixed order
__global__ __launch_bounds__(BLOCK) void k_reduce2(int nparts, const double *__restrict__ partial,
                                                   double *__restrict__ out) {
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nparts; i += BLOCK) {
    a += partial[2 * i];
    b += partial[2 * i + 1];
  }
  __shared__ double sh[2][BLOCK];
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sh[0][0];
    out[1] = sh[1][0];
  }
}

// min over the workgroup / over the partials (the linear-form update ratio)
__global__ __launch_bounds__(BLOCK) void k_min_ratio(int64_t n, const double *__restrict__ m,
                                                     const double *__restrict__ y, double *__restrict__ partial) {
  double mn = 1.0e20;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const double c = m[i], d = y[i];
    if (c <= d) mn = fmin(mn, fabs(c / d));
  }
  for (int o = 32; o > 0; o >>= 1) mn = fmin(mn, __shfl_xor(mn, o, 64));
  __shared__ double sh[BLOCK / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = sh[0];
    for (int i = 1; i < BLOCK / 64; i++) v = fmin(v, sh[i]);
    partial[2 * blockIdx.x] = v;
    partial[2 * blockIdx.x + 1] = 0.0;
  }
}

__global__ __launch_bounds__(BLOCK) void k_reduce_min(int nparts, const double *__restrict__ partial,
                                                      double *__restrict__ out) {
  double v = 1.0e20;
  for (int i = threadIdx.x; i < nparts; i += BLOCK) v = fmin(v, partial[2 * i]);
  __shared__ double sh[BLOCK];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sh[0];
    out[1] = 0.0;
  }
}

// max |v| over the workgroup / over the partials (the residual's infinity
// norm for the divergence bound)
__global__ __launch_bounds__(BLOCK) void k_max_abs(int64_t n, const double *__restrict__ v,
                                                   double *__restrict__ partial) {
  double mx = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    mx = fmax(mx, fabs(v[i]));
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  __shared__ double sh[BLOCK / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sh[0];
    for (int i = 1; i < BLOCK / 64; i++) r = fmax(r, sh[i]);
    partial[2 * blockIdx.x] = r;
    partial[2 * blockIdx.x + 1] = 0.0;
  }
}

__global__ __launch_bounds__(BLOCK) void k_reduce_max(int nparts, const double *__restrict__ partial,
                                                      double *__restrict__ out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += BLOCK) v = fmax(v, partial[2 * i]);
  __shared__ double sh[BLOCK];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sh[0];
    out[1] = 0.0;
  }
}

// (a, b) and (c, d) over n entries; c == nullptr skips the second
__global__ __launch_bounds__(BLOCK) void k_dot2(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                                const double *__restrict__ c, const double *__restrict__ d,
                                                double *__restrict__ partial) {
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    s0 += a[i] * b[i];
    if (c) s1 += c[i] * d[i];
  }
  block_reduce2(s0, s1, partial);
}

// y = A x, optionally with the partial sums of (y, a1) and (y, y).  NT > 0
// fixes N at compile time: a block's N column loads are then all issued
// before the first multiply (N loads in flight per lane).
template <int G, int NT>
__global__ __launch_bounds__(BLOCK) void k_bsr_spmv(int64_t nrows, int N, const int32_t *__restrict__ rp,
                                                    const int32_t *__restrict__ ci,
                                                    const double *__restrict__ vals, const double *__restrict__ x,
                                                    double *__restrict__ y, const double *__restrict__ a1,
                                                    int self2, double *__restrict__ partial) {
  if (NT > 0) N = NT;
  const int r = threadIdx.x % G;
  const int64_t NN = (int64_t)N * N;
  double d1 = 0.0, d2 = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * (BLOCK / G) + threadIdx.x / G; row < nrows;
       row += (int64_t)gridDim.x * (BLOCK / G)) {
    if (r < N) {
      double acc0 = 0.0, acc1 = 0.0;
      const int b0 = rp[row], b1 = rp[row + 1];
      for (int bk = b0; bk < b1; bk++) {
        const double *A = vals + (int64_t)bk * NN + r;
        const double *xc = x + (int64_t)ci[bk] * N;
        if constexpr (NT > 0) {
          double a[NT], xv[NT];
#pragma unroll
          for (int k = 0; k < NT; k++) {
            a[k] = __builtin_nontemporal_load(A + k * NT);
            xv[k] = xc[k];
          }
#pragma unroll
          for (int k = 0; k < NT; k++) {
            if (k & 1) acc1 += a[k] * xv[k];
            else acc0 += a[k] * xv[k];
          }
        } else {
          int k = 0;
          for (; k + 1 < N; k += 2) {
            acc0 += __builtin_nontemporal_load(A + (int64_t)k * N) * xc[k];
            acc1 += __builtin_nontemporal_load(A + (int64_t)(k + 1) * N) * xc[k + 1];
          }
          if (k < N) acc0 += __builtin_nontemporal_load(A + (int64_t)k * N) * xc[k];
        }
      }
      const double yv = acc0 + acc1;
      y[row * N + r] = yv;
      if (a1) d1 += a1[row * N + r] * yv;
      if (self2) d2 += yv * yv;
    }
  }
  if (partial) block_reduce2(d1, d2, partial);
}

// group argmax (ties to the larger lane) within G lanes
template <int G>
__device__ __forceinline__ void group_argmax(double &v, int &idx) {
  for (int o = G / 2; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, G);
    const int j = __shfl_xor(idx, o, G);
    if (w > v || (w == v && j > idx)) {
      v = w;
      idx = j;
    }
  }
}

// LDS written by lanes of this wavefront becomes visible to the wavefront
// (a group of G lanes never spans two wavefronts)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inverse of each diagonal block by Gauss-Jordan with partial pivoting,
// lane r holding row r of the block and of the growing inverse; the lane
// that pivots at step j ends with row j of the inverse.  The pivot row is
// broadcast through LDS (one store by its lane, same-address reads by the
// group).  A zero pivot counts in *fail.
template <int G>
__global__ __launch_bounds__(BLOCK) void k_diag_inverse(int64_t nrows, int N, const int32_t *__restrict__ rp,
                                                        const int32_t *__restrict__ ci,
                                                        const double *__restrict__ vals, double *__restrict__ dinv,
                                                        int *__restrict__ fail) {
  const int r = threadIdx.x % G;
  const bool live = r < N;
  const int64_t NN = (int64_t)N * N;
  __shared__ double prow[BLOCK / G][2 * G];
  double *pr = prow[threadIdx.x / G];
  for (int64_t row = (int64_t)blockIdx.x * (BLOCK / G) + threadIdx.x / G; row < nrows;
       row += (int64_t)gridDim.x * (BLOCK / G)) {
    int db = -1;
    for (int bk = rp[row]; bk < rp[row + 1]; bk++)
      if (ci[bk] == row) db = bk;
    double a[G], e[G];
#pragma unroll
    for (int k = 0; k < G; k++) {
      a[k] = (live && k < N && db >= 0) ? vals[(int64_t)db * NN + (int64_t)k * N + r] : (k == r ? 1.0 : 0.0);
      e[k] = (k == r) ? 1.0 : 0.0;
    }
    bool done = !live;
    int mystep = -1;
    bool singular = db < 0;
#pragma unroll
    for (int j = 0; j < G; j++) {
      if (j < N) {
        double cand = done ? -1.0 : fabs(a[j]);
        int p = r;
        group_argmax<G>(cand, p);
        // columns before j of the pivot row are already zero
        if (r == p) {
#pragma unroll
          for (int k = j; k < G; k++) pr[k] = a[k];
#pragma unroll
          for (int k = 0; k < G; k++) pr[G + k] = e[k];
        }
        wave_sync();
        double pa[G], pe[G];
#pragma unroll
        for (int k = 0; k < G; k++) {
          pa[k] = k >= j ? pr[k] : 0.0;
          pe[k] = pr[G + k];
        }
        wave_sync();
        double piv = pa[j];
        if (piv == 0.0) {
          singular = true;
          piv = 1.0;
        }
        const double rinv = 1.0 / piv;
        if (r == p) {
          done = true;
          mystep = j;
#pragma unroll
          for (int k = 0; k < G; k++) {
            a[k] = pa[k] * rinv;
            e[k] = pe[k] * rinv;
          }
        } else {
          const double f = a[j] * rinv;
#pragma unroll
          for (int k = 0; k < G; k++) {
            if (k > j) a[k] -= f * pa[k];
            e[k] -= f * pe[k];
          }
          a[j] = 0.0;
        }
      }
    }
    if (live && mystep >= 0) {
#pragma unroll
      for (int k = 0; k < G; k++)
        if (k < N) dinv[row * NN + (int64_t)k * N + mystep] = e[k];
    }
    if (singular && r == 0) atomicAdd(fail, 1);
  }
}

// out = D^-1 v for the group's block row, v_r held by lane r
template <int G>
__device__ __forceinline__ double apply_dinv(const double *__restrict__ D, int N, int r, double v) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < G; k++) {
    const double vk = __shfl(v, k, G);
    if (k < N && r < N) acc += D[(int64_t)k * N + r] * vk;
  }
  return acc;
}

// BiCGStab start from x = 0: r = rhat = b, p = v = 0; partial (b, b)
__global__ __launch_bounds__(BLOCK) void k_bcgs_init(int64_t n, const double *__restrict__ b, double *__restrict__ x,
                                                     double *__restrict__ r, double *__restrict__ rhat,
                                                     double *__restrict__ p, double *__restrict__ v,
                                                     double *__restrict__ partial) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    rhat[i] = bi;
    p[i] = 0.0;
    v[i] = 0.0;
    s += bi * bi;
  }
  block_reduce2(s, 0.0, partial);
}

// p = r + beta (p - omega v);  phat = D^-1 p
template <int G>
__global__ __launch_bounds__(BLOCK) void k_bcgs_p(int64_t nrows, int N, const double *__restrict__ r,
                                                  const double *__restrict__ v, double *__restrict__ p,
                                                  double beta, double omega, const double *__restrict__ dinv,
                                                  double *__restrict__ phat) {
  const int l = threadIdx.x % G;
  const int64_t NN = (int64_t)N * N;
  for (int64_t row = (int64_t)blockIdx.x * (BLOCK / G) + threadIdx.x / G; row < nrows;
       row += (int64_t)gridDim.x * (BLOCK / G)) {
    const int64_t i = row * N + l;
    double pv = 0.0;
    if (l < N) {
      pv = r[i] + beta * (p[i] - omega * v[i]);
      p[i] = pv;
    }
    const double z = apply_dinv<G>(dinv + row * NN, N, l, pv);
    if (l < N) phat[i] = z;
  }
}

// s = r - alpha v;  shat = D^-1 s;  partial (s, s)
template <int G>
__global__ __launch_bounds__(BLOCK) void k_bcgs_s(int64_t nrows, int N, const double *__restrict__ r,
                                                  const double *__restrict__ v, double alpha,
                                                  const double *__restrict__ dinv, double *__restrict__ s,
                                                  double *__restrict__ shat, double *__restrict__ partial) {
  const int l = threadIdx.x % G;
  const int64_t NN = (int64_t)N * N;
  double ss = 0.0;
  for (int64_t row = (int64_t)blockIdx.x * (BLOCK / G) + threadIdx.x / G; row < nrows;
       row += (int64_t)gridDim.x * (BLOCK / G)) {
    const int64_t i = row * N + l;
    double sv = 0.0;
    if (l < N) {
      sv = r[i] - alpha * v[i];
      s[i] = sv;
      ss += sv * sv;
    }
    const double z = apply_dinv<G>(dinv + row * NN, N, l, sv);
    if (l < N) shat[i] = z;
  }
  block_reduce2(ss, 0.0, partial);
}

// x += alpha phat + omega shat;  r = s - omega t;  partial (r, r), (rhat, r)
__global__ __launch_bounds__(BLOCK) void k_bcgs_xr(int64_t n, double *__restrict__ x, const double *__restrict__ phat,
                                                   const double *__restrict__ shat, const double *__restrict__ s,
                                                   const double *__restrict__ t, double alpha, double omega,
                                                   double *__restrict__ r, const double *__restrict__ rhat,
                                                   double *__restrict__ partial) {
  double rr = 0.0, hr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    x[i] += alpha * phat[i] + omega * shat[i];
    const double rv = s[i] - omega * t[i];
    r[i] = rv;
    rr += rv * rv;
    hr += rhat[i] * rv;
  }
  block_reduce2(rr, hr, partial);
}

__global__ void k_zero_int(int *__restrict__ v) { *v = 0; }

// x += alpha phat (BiCGStab's early exit on a small s)
__global__ __launch_bounds__(BLOCK) void k_axpy(int64_t n, double alpha, const double *__restrict__ xp,
                                                double *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    y[i] += alpha * xp[i];
}

// Newton update with PMRTCheckUpdatePre's change to the step (pm_rt.F90:
// 734-840): log form y <- sign(y) min(|y|, max_dlnC), m <- m exp(-y); linear
// form y <- y * scale, m <- m - y.  partial (y, y) and (x, x), x = ln m or m.
__global__ __launch_bounds__(BLOCK) void k_newton_update(int64_t n, double *__restrict__ m, double *__restrict__ y,
                                                         int use_log, double max_dlnC, double scale,
                                                         double *__restrict__ partial, double *__restrict__ rel) {
  double yy = 0.0, xx = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    double d = y[i];
    double mi = m[i];
    // |update / old solution entry| (PMRTCheckUpdatePost, pm_rt.F90:903-906)
    if (rel) {
      const double dd = use_log ? copysign(1.0, d) * fmin(fabs(d), max_dlnC) : d * scale;
      rel[i] = fabs(dd / (use_log ? log(mi) : mi));
    }
    if (use_log) {
      d = copysign(1.0, d) * fmin(fabs(d), max_dlnC);
      mi = mi * exp(-d);
      const double lm = log(mi);
      xx += lm * lm;
    } else {
      d = d * scale;
      mi = mi - d;
      xx += mi * mi;
    }
    y[i] = d;
    m[i] = mi;
    yy += d * d;
  }
  block_reduce2(yy, xx, partial);
}

// ---------------------------------------------------------------------------
// Several right-hand sides of one scalar (block size 1) matrix, BiCGStab for
// each with Jacobi preconditioning, every scalar of the iteration on the
// device.  Vectors are cell-major [n][C] (the transport vector's layout, C
// components per cell); a workgroup holds RB = BLOCK / C whole rows, thread t
// owns component t % C of row t / C, so each thread always sees the same
// component and the per-component partial sums of a workgroup are fixed-order
// sums.  One-workgroup kernels turn the partial sums into each component's
// scalars and convergence status; the element kernels skip components whose
// status is set.  No value goes to the host between iterations: a solve is
// stream-ordered and can be captured in a graph.
struct mrhs_sc {
  double rho, rho_new, alpha, omega, beta;
  double tol, rnorm0, rnorm, dtol;
  int status;    // 0 running, PFX_KSP_* once decided
  int conv;      // the convergence reason of this component (rtol or atol)
  int it;        // iterations taken
  int pending;   // converged on |s| this iteration: x += alpha phat owed
};

__device__ __forceinline__ int mrhs_comps(int C) { return (BLOCK / C) * C; }

// per-component sums over this workgroup's threads: thread c < C of the
// workgroup writes partial[(blockIdx.x * C + c) * 2 + {0,1}]
__device__ __forceinline__ void mrhs_block_sum(double a, double b, int C, double *partial) {
  __shared__ double sh[2][BLOCK];
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  const int used = mrhs_comps(C);
  if ((int)threadIdx.x < C) {
    double s0 = 0.0, s1 = 0.0;
    for (int t = threadIdx.x; t < used; t += C) {
      s0 += sh[0][t];
      s1 += sh[1][t];
    }
    partial[((int64_t)blockIdx.x * C + threadIdx.x) * 2] = s0;
    partial[((int64_t)blockIdx.x * C + threadIdx.x) * 2 + 1] = s1;
  }
}

// true in the workgroup that finishes last (its partial sums, and every
// other workgroup's, are then visible to it); that workgroup turns the
// partial sums into the scalars and re-arms the counter
__device__ __forceinline__ bool mrhs_last_block(int *counter) {
  __shared__ int last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(counter, 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (last) {
    __threadfence();
    if (threadIdx.x == 0) *counter = 0;
  }
  return last;
}

// the two sums of component c over the nb workgroups, in workgroup order
__device__ __forceinline__ void mrhs_sums(int nb, int C, int c, const double *partial, double &a, double &b) {
  a = 0.0;
  b = 0.0;
  for (int k = 0; k < nb; k++) {
    a += partial[((int64_t)k * C + c) * 2];
    b += partial[((int64_t)k * C + c) * 2 + 1];
  }
}

__device__ __forceinline__ void mrhs_set_done(int C, const mrhs_sc *sc, int *done) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1;
    for (int k = 0; k < C; k++) all = all && sc[k].status != 0;
    *done = all;
  }
}

// grid-stride loop over the rows of a workgroup's threads
#define MRHS_ROWS(n, C)                                                       \
  const int c_ = threadIdx.x % (C);                                           \
  const int rl_ = threadIdx.x / (C);                                          \
  const bool on_ = (int)threadIdx.x < mrhs_comps(C);                          \
  const int64_t rstep_ = (int64_t)gridDim.x * (BLOCK / (C));                  \
  for (int64_t row = (int64_t)blockIdx.x * (BLOCK / (C)) + rl_; on_ && row < (n); row += rstep_)

struct mrhs_args {
  int64_t n;
  int C, nb, maxit, reserved;
  const int32_t *rp, *ci;
  const double *vals, *dinv, *b;
  double *x, *r, *rh, *p, *v, *s, *t, *ph, *sh;
  double *partial;
  mrhs_sc *sc;
  int *counter, *done, *fail;
  double rtol, atol, dtol;
};

__global__ __launch_bounds__(BLOCK) void k_m_diag(int64_t n, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ ci, const double *__restrict__ vals,
                                                  double *__restrict__ dinv, int *__restrict__ fail) {
  const int64_t row = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (row >= n) return;
  double d = 0.0;
  for (int32_t k = rp[row]; k < rp[row + 1]; k++)
    if (ci[k] == row) d = vals[k];
  if (d == 0.0 || !(d == d)) atomicAdd(fail, 1);
  dinv[row] = d != 0.0 ? 1.0 / d : 0.0;
}

// x = 0, r = rhat = b, p = v = 0; then the first scalars and the start
// tests of every component
__global__ __launch_bounds__(BLOCK) void k_m_init(const mrhs_args A) {
  const int C = A.C;
  double bb = 0.0;
  MRHS_ROWS(A.n, C) {
    const int64_t i = row * C + c_;
    const double bi = A.b[i];
    A.x[i] = 0.0;
    A.r[i] = bi;
    A.rh[i] = bi;
    A.p[i] = 0.0;
    A.v[i] = 0.0;
    bb += bi * bi;
  }
  mrhs_block_sum(bb, 0.0, C, A.partial);
  if (!mrhs_last_block(A.counter)) return;
  const int c = threadIdx.x;
  if (c < C) {
    double sum, unused;
    mrhs_sums(A.nb, C, c, A.partial, sum, unused);
    mrhs_sc S;
    S.rnorm0 = sqrt(sum);
    S.rnorm = S.rnorm0;
    S.tol = fmax(A.rtol * S.rnorm0, A.atol);
    S.dtol = A.dtol * S.rnorm0;
    S.conv = A.rtol * S.rnorm0 >= A.atol ? PFX_KSP_CONVERGED_RTOL : PFX_KSP_CONVERGED_ATOL;
    S.rho = 1.0;
    S.alpha = 1.0;
    S.omega = 1.0;
    S.rho_new = sum;
    S.beta = 0.0;
    S.it = 0;
    S.pending = 0;
    S.status = 0;
    if (*A.fail > 0) S.status = PFX_KSP_DIVERGED_PC_FAILED;
    else if (!isfinite(S.rnorm0)) S.status = PFX_KSP_DIVERGED_NANORINF;
    else if (S.rnorm0 <= S.tol) S.status = S.conv;
    A.sc[c] = S;
  }
  mrhs_set_done(C, A.sc, A.done);
}

// p = r + beta (p - omega v);  phat = D^-1 p
__global__ __launch_bounds__(BLOCK) void k_m_p(const mrhs_args A) {
  if (*A.done) return;
  const int C = A.C;
  MRHS_ROWS(A.n, C) {
    if (A.sc[c_].status != 0) continue;
    const int64_t i = row * C + c_;
    const double pv = A.r[i] + A.sc[c_].beta * (A.p[i] - A.sc[c_].omega * A.v[i]);
    A.p[i] = pv;
    A.ph[i] = A.dinv[row] * pv;
  }
}

// v = A phat, partial (rhat, v); then alpha = rho_new / (rhat, v)
__global__ __launch_bounds__(BLOCK) void k_m_v(const mrhs_args A) {
  if (*A.done) return;
  const int C = A.C;
  double a = 0.0;
  MRHS_ROWS(A.n, C) {
    if (A.sc[c_].status != 0) continue;
    double acc = 0.0;
    for (int32_t k = A.rp[row]; k < A.rp[row + 1]; k++) acc += A.vals[k] * A.ph[(int64_t)A.ci[k] * C + c_];
    const int64_t i = row * C + c_;
    A.v[i] = acc;
    a += A.rh[i] * acc;
  }
  mrhs_block_sum(a, 0.0, C, A.partial);
  if (!mrhs_last_block(A.counter)) return;
  const int c = threadIdx.x;
  if (c < C && A.sc[c].status == 0) {
    double rv, unused;
    mrhs_sums(A.nb, C, c, A.partial, rv, unused);
    if (rv == 0.0 || !isfinite(rv)) {
      A.sc[c].status = isfinite(rv) ? PFX_KSP_DIVERGED_BREAKDOWN : PFX_KSP_DIVERGED_NANORINF;
      A.sc[c].it += 1;
    } else {
      A.sc[c].alpha = A.sc[c].rho_new / rv;
    }
  }
}

// s = r - alpha v;  shat = D^-1 s;  partial (s, s); then BiCGStab's early
// exit: |s| <= tol converges with x += alpha phat owed
__global__ __launch_bounds__(BLOCK) void k_m_s(const mrhs_args A) {
  if (*A.done) return;
  const int C = A.C;
  double ss = 0.0;
  MRHS_ROWS(A.n, C) {
    if (A.sc[c_].status != 0) continue;
    const int64_t i = row * C + c_;
    const double si = A.r[i] - A.sc[c_].alpha * A.v[i];
    A.s[i] = si;
    A.sh[i] = A.dinv[row] * si;
    ss += si * si;
  }
  mrhs_block_sum(ss, 0.0, C, A.partial);
  if (!mrhs_last_block(A.counter)) return;
  const int c = threadIdx.x;
  if (c < C && A.sc[c].status == 0) {
    double sum, unused;
    mrhs_sums(A.nb, C, c, A.partial, sum, unused);
    const double sn = sqrt(sum);
    if (!isfinite(sn)) {
      A.sc[c].status = PFX_KSP_DIVERGED_NANORINF;
      A.sc[c].it += 1;
    } else if (sn <= A.sc[c].tol) {
      A.sc[c].rnorm = sn;
      A.sc[c].status = A.sc[c].conv;
      A.sc[c].pending = 1;
      A.sc[c].it += 1;
    }
  }
}

// t = A shat, partial (t, s), (t, t); then omega = (t, s) / (t, t)
__global__ __launch_bounds__(BLOCK) void k_m_t(const mrhs_args A) {
  if (*A.done) return;
  const int C = A.C;
  double ts = 0.0, tt = 0.0;
  MRHS_ROWS(A.n, C) {
    if (A.sc[c_].status != 0) continue;
    double acc = 0.0;
    for (int32_t k = A.rp[row]; k < A.rp[row + 1]; k++) acc += A.vals[k] * A.sh[(int64_t)A.ci[k] * C + c_];
    const int64_t i = row * C + c_;
    A.t[i] = acc;
    ts += A.s[i] * acc;
    tt += acc * acc;
  }
  mrhs_block_sum(ts, tt, C, A.partial);
  if (!mrhs_last_block(A.counter)) return;
  const int c = threadIdx.x;
  if (c < C && A.sc[c].status == 0) {
    double sts, stt;
    mrhs_sums(A.nb, C, c, A.partial, sts, stt);
    if (stt == 0.0 || !isfinite(stt)) {
      A.sc[c].status = isfinite(stt) ? PFX_KSP_DIVERGED_BREAKDOWN : PFX_KSP_DIVERGED_NANORINF;
      A.sc[c].it += 1;
    } else {
      A.sc[c].omega = sts / stt;
    }
  }
}

// x += alpha phat + omega shat;  r = s - omega t;  partial (r, r),
// (rhat, r) (a component converged on |s| takes x += alpha phat only); then
// the residual tests, and the next iteration's start tests (iteration
// limit, breakdown) and beta
__global__ __launch_bounds__(BLOCK) void k_m_xr(const mrhs_args A) {
  if (*A.done) return;
  const int C = A.C;
  double rr = 0.0, hr = 0.0;
  MRHS_ROWS(A.n, C) {
    const int64_t i = row * C + c_;
    if (A.sc[c_].pending) {
      A.x[i] += A.sc[c_].alpha * A.ph[i];
    } else if (A.sc[c_].status == 0) {
      A.x[i] += A.sc[c_].alpha * A.ph[i] + A.sc[c_].omega * A.sh[i];
      const double rv = A.s[i] - A.sc[c_].omega * A.t[i];
      A.r[i] = rv;
      rr += rv * rv;
      hr += A.rh[i] * rv;
    }
  }
  mrhs_block_sum(rr, hr, C, A.partial);
  if (!mrhs_last_block(A.counter)) return;
  const int c = threadIdx.x;
  if (c < C) {
    mrhs_sc S = A.sc[c];
    if (S.pending) {
      S.pending = 0;
    } else if (S.status == 0) {
      double srr, shr;
      mrhs_sums(A.nb, C, c, A.partial, srr, shr);
      S.it += 1;
      S.rho = S.rho_new;
      S.rho_new = shr;
      S.rnorm = sqrt(srr);
      if (!isfinite(S.rnorm)) S.status = PFX_KSP_DIVERGED_NANORINF;
      else if (S.rnorm <= S.tol) S.status = S.conv;
      else if (S.rnorm >= S.dtol) S.status = PFX_KSP_DIVERGED_DTOL;
      else if (S.it >= A.maxit) S.status = PFX_KSP_DIVERGED_ITS;
      else if (S.rho_new == 0.0 || S.omega == 0.0) S.status = PFX_KSP_DIVERGED_BREAKDOWN;
      else S.beta = (S.rho_new / S.rho) * (S.alpha / S.omega);
    }
    A.sc[c] = S;
  }
  mrhs_set_done(C, A.sc, A.done);
}

int grid_for(int64_t work_items) {
  int64_t g = (work_items + BLOCK - 1) / BLOCK;
  if (g > MAX_GRID) g = MAX_GRID;
  if (g < 1) g = 1;
  return (int)g;
}

int group_size_for(int N) { return N <= 8 ? 8 : (N <= 16 ? 16 : 32); }

}  // namespace

struct pfx_bsr {
  int device = 0;
  int64_t nrows = 0, nnzb = 0;
  int32_t N = 0, G = 16;
  const int32_t *rp = nullptr, *ci = nullptr;
  double *dinv = nullptr;
  double *r = nullptr, *rhat = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *t = nullptr;
  double *phat = nullptr, *shat = nullptr;
  double *partial = nullptr, *red = nullptr, *host2 = nullptr;
  int *fail = nullptr, *host_fail = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // several right-hand sides at block size 1 (pfx_bsr_solve_multi): C
  // components, vectors [nrows][C], the per-component scalars, the
  // all-decided flag and its pinned host copy
  int mC = 0;
  double *mvec = nullptr, *mdinv = nullptr, *mpartial = nullptr;
  mrhs_sc *msc = nullptr, *mhost_sc = nullptr;
  int *mdone = nullptr, *mhost_done = nullptr;
  int *mcounter = nullptr;   // workgroups done in the current reduction (0 between launches)
  bool mpending = false;   // a stream-ordered solve whose statistics are unread
  hipStream_t mstream = nullptr;
};

// shape of a handle, for the other translation units of the library
namespace pfx_internal {
const int32_t *bsr_row_ptr(const pfx_bsr *h) { return h ? h->rp : nullptr; }
int64_t bsr_rows(const pfx_bsr *h) { return h ? h->nrows : -1; }
int32_t bsr_block(const pfx_bsr *h) { return h ? h->N : -1; }
int64_t bsr_nnzb(const pfx_bsr *h) { return h ? h->nnzb : -1; }
}  // namespace pfx_internal

namespace {

// launch helpers over the block-row groups
int rows_grid(const pfx_bsr *h) {
  const int64_t rows_per_block = BLOCK / h->G;
  int64_t g = (h->nrows + rows_per_block - 1) / rows_per_block;
  if (g > MAX_GRID) g = MAX_GRID;
  if (g < 1) g = 1;
  return (int)g;
}

int read2(pfx_bsr *h, int nparts, hipStream_t s, double *a, double *b) {
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(BLOCK), 0, s, nparts, h->partial, h->red);
  KCHK(hipGetLastError());
  KCHK(hipMemcpyAsync(h->host2, h->red, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  KCHK(hipStreamSynchronize(s));
  *a = h->host2[0];
  if (b) *b = h->host2[1];
  return 0;
}

// max |v_i| over n entries (nparts workgroups)
int max_abs(pfx_bsr *h, int nparts, hipStream_t s, int64_t n, const double *v, double *out) {
  hipLaunchKernelGGL(k_max_abs, dim3(nparts), dim3(BLOCK), 0, s, n, v, h->partial);
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(BLOCK), 0, s, nparts, h->partial, h->red);
  KCHK(hipGetLastError());
  KCHK(hipMemcpyAsync(h->host2, h->red, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  KCHK(hipStreamSynchronize(s));
  *out = h->host2[0];
  return 0;
}

// ConvergenceTest's divergence bound on the residual (solver.F90:137)
constexpr double kMaxNorm = 1.0e20;

int spmv_launch(pfx_bsr *h, const double *vals, const double *x, double *y, const double *a1, int self2,
                double *partial, hipStream_t s) {
  const int g = rows_grid(h);
  if (h->N == 15)
    hipLaunchKernelGGL((k_bsr_spmv<16, 15>), dim3(g), dim3(BLOCK), 0, s, h->nrows, h->N, h->rp, h->ci, vals, x, y,
                       a1, self2, partial);
  else if (h->G == 8)
    hipLaunchKernelGGL((k_bsr_spmv<8, 0>), dim3(g), dim3(BLOCK), 0, s, h->nrows, h->N, h->rp, h->ci, vals, x, y, a1,
                       self2, partial);
  else if (h->G == 16)
    hipLaunchKernelGGL((k_bsr_spmv<16, 0>), dim3(g), dim3(BLOCK), 0, s, h->nrows, h->N, h->rp, h->ci, vals, x, y,
                       a1, self2, partial);
  else
    hipLaunchKernelGGL((k_bsr_spmv<32, 0>), dim3(g), dim3(BLOCK), 0, s, h->nrows, h->N, h->rp, h->ci, vals, x, y,
                       a1, self2, partial);
  KCHK(hipGetLastError());
  return g;
}

template <int G>
void launch_rows(pfx_bsr *h, int which, hipStream_t s, int g, const double *vals, double beta, double omega,
                 double alpha) {
  if (which == 0)
    hipLaunchKernelGGL(k_diag_inverse<G>, dim3(g), dim3(BLOCK), 0, s, h->nrows, h->N, h->rp, h->ci, vals, h->dinv,
                       h->fail);
  else if (which == 1)
    hipLaunchKernelGGL(k_bcgs_p<G>, dim3(g), dim3(BLOCK), 0, s, h->nrows, h->N, h->r, h->v, h->p, beta, omega,
                       h->dinv, h->phat);
  else
    hipLaunchKernelGGL(k_bcgs_s<G>, dim3(g), dim3(BLOCK), 0, s, h->nrows, h->N, h->r, h->v, alpha, h->dinv, h->s,
                       h->shat, h->partial);
}

int rows_kernel(pfx_bsr *h, int which, hipStream_t s, const double *vals, double beta, double omega, double alpha) {
  const int g = rows_grid(h);
  if (h->G == 8) launch_rows<8>(h, which, s, g, vals, beta, omega, alpha);
  else if (h->G == 16) launch_rows<16>(h, which, s, g, vals, beta, omega, alpha);
  else launch_rows<32>(h, which, s, g, vals, beta, omega, alpha);
  KCHK(hipGetLastError());
  return g;
}

}  // namespace

extern "C" int pfx_bsr_create(int64_t nrows, int32_t N, int64_t nnzb, const int32_t *row_ptr,
                              const int32_t *col_idx, pfx_bsr **out) {
  if (!out || !row_ptr || !col_idx || nrows < 1 || N < 1 || N > 32 || nnzb < nrows) return PFX_E_ARG;
  if (nrows > (int64_t)0x7fffffff || nnzb > (int64_t)0x7fffffff) return PFX_E_ARG;
  *out = nullptr;
  pfx_bsr *h = new pfx_bsr;
  KCHK(hipGetDevice(&h->device));
  h->nrows = nrows;
  h->nnzb = nnzb;
  h->N = N;
  h->G = group_size_for(N);
  h->rp = row_ptr;
  h->ci = col_idx;
  const size_t nv = (size_t)nrows * N * sizeof(double);
  double **vecs[] = {&h->r, &h->rhat, &h->p, &h->v, &h->s, &h->t, &h->phat, &h->shat};
  bool ok = hipMalloc((void **)&h->dinv, (size_t)nrows * N * N * sizeof(double)) == hipSuccess;
  for (double **vp : vecs) ok = ok && hipMalloc((void **)vp, nv) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->partial, 2 * MAX_GRID * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->red, 2 * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->fail, sizeof(int)) == hipSuccess;
  ok = ok && hipHostMalloc((void **)&h->host2, 2 * sizeof(double), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipHostMalloc((void **)&h->host_fail, sizeof(int), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipEventCreate(&h->e0) == hipSuccess && hipEventCreate(&h->e1) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    pfx_bsr_destroy(h);
    return PFX_E_NOMEM;
  }
  *out = h;
  return 0;
}

extern "C" int pfx_bsr_destroy(pfx_bsr *h) {
  if (!h) return 0;
  double *vecs[] = {h->dinv, h->r, h->rhat, h->p, h->v, h->s, h->t, h->phat, h->shat, h->partial, h->red};
  for (double *v : vecs)
    if (v) (void)hipFree(v);
  if (h->fail) (void)hipFree(h->fail);
  if (h->host2) (void)hipHostFree(h->host2);
  if (h->host_fail) (void)hipHostFree(h->host_fail);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->mvec) (void)hipFree(h->mvec);
  if (h->mdinv) (void)hipFree(h->mdinv);
  if (h->mpartial) (void)hipFree(h->mpartial);
  if (h->msc) (void)hipFree(h->msc);
  if (h->mdone) (void)hipFree(h->mdone);
  if (h->mcounter) (void)hipFree(h->mcounter);
  if (h->mhost_sc) (void)hipHostFree(h->mhost_sc);
  if (h->mhost_done) (void)hipHostFree(h->mhost_done);
  delete h;
  return 0;
}

extern "C" int pfx_bsr_spmv(pfx_bsr *h, const double *values, const double *x, double *y, void *stream) {
  if (!h || !values || !x || !y) return PFX_E_ARG;
  KCHK(hipSetDevice(h->device));
  hipStream_t s = (hipStream_t)stream;
  const int g = spmv_launch(h, values, x, y, nullptr, 0, nullptr, s);
  if (g < 0) return g;
  KCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int pfx_bsr_solve(pfx_bsr *h, const double *values, const double *b, double *x, const pfx_krylov_opts *o,
                             void *stream, pfx_krylov_stats *st) {
  if (!h || !values || !b || !x) return PFX_E_ARG;
  const double rtol = (o && o->rtol > 0.0) ? o->rtol : 1.0e-5;
  const double atol = (o && o->atol > 0.0) ? o->atol : 1.0e-50;
  const double dtol = (o && o->dtol > 0.0) ? o->dtol : 1.0e5;
  const int maxit = (o && o->max_iterations > 0) ? o->max_iterations : 10000;
  KCHK(hipSetDevice(h->device));
  hipStream_t s = (hipStream_t)stream;
  pfx_krylov_stats S;
  memset(&S, 0, sizeof(S));
  const int64_t n = h->nrows * h->N;
  const int ge = grid_for(n);
  KCHK(hipEventRecord(h->e0, s));

  // point-block Jacobi: invert the diagonal blocks
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, h->fail);
  int rc = rows_kernel(h, 0, s, values, 0.0, 0.0, 0.0);
  if (rc < 0) return rc;
  KCHK(hipMemcpyAsync(h->host_fail, h->fail, sizeof(int), hipMemcpyDeviceToHost, s));
  hipLaunchKernelGGL(k_bcgs_init, dim3(ge), dim3(BLOCK), 0, s, n, b, x, h->r, h->rhat, h->p, h->v, h->partial);
  KCHK(hipGetLastError());
  double bb = 0.0;
  if ((rc = read2(h, ge, s, &bb, nullptr))) return rc;
  S.rnorm0 = sqrt(bb);
  S.rnorm = S.rnorm0;
  const double tol = fmax(rtol * S.rnorm0, atol);
  const int conv = rtol * S.rnorm0 >= atol ? PFX_KSP_CONVERGED_RTOL : PFX_KSP_CONVERGED_ATOL;
  if (*h->host_fail > 0) {
    S.reason = PFX_KSP_DIVERGED_PC_FAILED;
  } else if (!std::isfinite(S.rnorm0)) {
    S.reason = PFX_KSP_DIVERGED_NANORINF;
  } else if (S.rnorm0 <= tol) {
    S.reason = conv;
  }
  double rho = 1.0, alpha = 1.0, omega = 1.0, rho_new = bb;
  int it = 0;
  while (S.reason == 0) {
    if (it >= maxit) {
      S.reason = PFX_KSP_DIVERGED_ITS;
      break;
    }
    if (rho_new == 0.0 || omega == 0.0) {
      S.reason = PFX_KSP_DIVERGED_BREAKDOWN;
      break;
    }
    it++;
    const double beta = (rho_new / rho) * (alpha / omega);
    if ((rc = rows_kernel(h, 1, s, values, beta, omega, 0.0)) < 0) return rc;
    int g = spmv_launch(h, values, h->phat, h->v, h->rhat, 0, h->partial, s);
    if (g < 0) return g;
    double rv = 0.0;
    if ((rc = read2(h, g, s, &rv, nullptr))) return rc;
    if (rv == 0.0 || !std::isfinite(rv)) {
      S.reason = std::isfinite(rv) ? PFX_KSP_DIVERGED_BREAKDOWN : PFX_KSP_DIVERGED_NANORINF;
      break;
    }
    alpha = rho_new / rv;
    if ((g = rows_kernel(h, 2, s, values, 0.0, 0.0, alpha)) < 0) return g;
    double ss = 0.0;
    if ((rc = read2(h, g, s, &ss, nullptr))) return rc;
    const double snorm = sqrt(ss);
    if (!std::isfinite(snorm)) {
      S.reason = PFX_KSP_DIVERGED_NANORINF;
      break;
    }
    if (snorm <= tol) {
      hipLaunchKernelGGL(k_axpy, dim3(ge), dim3(BLOCK), 0, s, n, alpha, h->phat, x);
      KCHK(hipGetLastError());
      S.rnorm = snorm;
      S.reason = conv;
      break;
    }
    if ((g = spmv_launch(h, values, h->shat, h->t, h->s, 1, h->partial, s)) < 0) return g;
    double ts = 0.0, tt = 0.0;
    if ((rc = read2(h, g, s, &ts, &tt))) return rc;
    if (tt == 0.0 || !std::isfinite(tt)) {
      S.reason = std::isfinite(tt) ? PFX_KSP_DIVERGED_BREAKDOWN : PFX_KSP_DIVERGED_NANORINF;
      break;
    }
    omega = ts / tt;
    hipLaunchKernelGGL(k_bcgs_xr, dim3(ge), dim3(BLOCK), 0, s, n, x, h->phat, h->shat, h->s, h->t, alpha, omega,
                       h->r, h->rhat, h->partial);
    KCHK(hipGetLastError());
    double rr = 0.0, hr = 0.0;
    if ((rc = read2(h, ge, s, &rr, &hr))) return rc;
    rho = rho_new;
    rho_new = hr;
    S.rnorm = sqrt(rr);
    if (!std::isfinite(S.rnorm)) S.reason = PFX_KSP_DIVERGED_NANORINF;
    else if (S.rnorm <= tol) S.reason = conv;
    else if (S.rnorm >= dtol * S.rnorm0) S.reason = PFX_KSP_DIVERGED_DTOL;
  }
  S.iterations = it;
  KCHK(hipEventRecord(h->e1, s));
  KCHK(hipEventSynchronize(h->e1));
  float ms = 0.f;
  KCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
  S.ms = ms;
  if (st) *st = S;
  return 0;
}

namespace {

int ensure_multi(pfx_bsr *h, int C) {
  if (h->mC == C && h->mvec) return 0;
  if (h->mvec) KCHK(hipFree(h->mvec));
  if (h->mdinv) KCHK(hipFree(h->mdinv));
  if (h->mpartial) KCHK(hipFree(h->mpartial));
  h->mvec = h->mdinv = h->mpartial = nullptr;
  h->mC = 0;
  const size_t nv = (size_t)h->nrows * C;
  // r, rhat, p, v, s, t, phat, shat, and a right-hand side and solution
  // (pfx_os_transport_step's)
  if (hipMalloc((void **)&h->mvec, 10 * nv * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&h->mdinv, (size_t)h->nrows * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&h->mpartial, (size_t)MAX_GRID * C * 2 * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return PFX_E_NOMEM;
  }
  if (!h->msc) {
    if (hipMalloc((void **)&h->msc, PFX_MAX_PRIMARY * sizeof(mrhs_sc)) != hipSuccess ||
        hipMalloc((void **)&h->mdone, sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&h->mcounter, sizeof(int)) != hipSuccess ||
        hipMemset(h->mcounter, 0, sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&h->mhost_sc, PFX_MAX_PRIMARY * sizeof(mrhs_sc), hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc((void **)&h->mhost_done, sizeof(int), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return PFX_E_NOMEM;
    }
  }
  h->mC = C;
  return 0;
}

// the components' outcome as one pfx_krylov_stats: iterations summed, the
// worst reason (a divergence before an atol convergence before rtol), the
// largest norms
void multi_stats(const mrhs_sc *sc, int C, pfx_krylov_stats *st) {
  pfx_krylov_stats S;
  memset(&S, 0, sizeof(S));
  S.reason = PFX_KSP_CONVERGED_RTOL;
  for (int c = 0; c < C; c++) {
    S.iterations += sc[c].it;
    const int r = sc[c].status == 0 ? PFX_KSP_DIVERGED_ITS : sc[c].status;
    if (r < 0 || (S.reason > 0 && r == PFX_KSP_CONVERGED_ATOL)) S.reason = r;
    S.rnorm0 = fmax(S.rnorm0, sc[c].rnorm0);
    S.rnorm = fmax(S.rnorm, sc[c].rnorm);
  }
  *st = S;
}

}  // namespace

extern "C" int pfx_bsr_solve_multi(pfx_bsr *h, const double *values, const double *B, double *X, int32_t ncomp,
                                   const pfx_krylov_opts *o, void *stream, int32_t stream_ordered,
                                   pfx_krylov_stats *st) {
  if (!h || !values || !B || !X || h->N != 1 || ncomp < 1 || ncomp > PFX_MAX_PRIMARY) return PFX_E_ARG;
  const double rtol = (o && o->rtol > 0.0) ? o->rtol : 1.0e-5;
  const double atol = (o && o->atol > 0.0) ? o->atol : 1.0e-50;
  const double dtol = (o && o->dtol > 0.0) ? o->dtol : 1.0e5;
  const int maxit = (o && o->max_iterations > 0) ? o->max_iterations : 10000;
  if (stream_ordered && maxit > PFX_KSP_MAX_ENQUEUED) return PFX_E_ARG;
  KCHK(hipSetDevice(h->device));
  int rc = ensure_multi(h, ncomp);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int C = ncomp;
  const int64_t n = h->nrows;
  const size_t nv = (size_t)n * C;
  double *r = h->mvec, *rh = r + nv, *p = rh + nv, *v = p + nv, *sv = v + nv, *t = sv + nv, *ph = t + nv,
         *sh = ph + nv;
  const int64_t rb = BLOCK / C;
  int nb = (int)std::min<int64_t>((n + rb - 1) / rb, MAX_GRID);
  if (nb < 1) nb = 1;
  const int gd = (int)std::min<int64_t>((n + BLOCK - 1) / BLOCK, 1 << 20);
  if (!stream_ordered) KCHK(hipEventRecord(h->e0, s));
  mrhs_args A;
  memset(&A, 0, sizeof(A));
  A.n = n;
  A.C = C;
  A.nb = nb;
  A.maxit = maxit;
  A.rp = h->rp;
  A.ci = h->ci;
  A.vals = values;
  A.dinv = h->mdinv;
  A.b = B;
  A.x = X;
  A.r = r;
  A.rh = rh;
  A.p = p;
  A.v = v;
  A.s = sv;
  A.t = t;
  A.ph = ph;
  A.sh = sh;
  A.partial = h->mpartial;
  A.sc = h->msc;
  A.counter = h->mcounter;
  A.done = h->mdone;
  A.fail = h->fail;
  A.rtol = rtol;
  A.atol = atol;
  A.dtol = dtol;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, h->fail);
  hipLaunchKernelGGL(k_m_diag, dim3(gd), dim3(BLOCK), 0, s, n, h->rp, h->ci, values, h->mdinv, h->fail);
  hipLaunchKernelGGL(k_m_init, dim3(nb), dim3(BLOCK), 0, s, A);
  KCHK(hipGetLastError());
  // five launches per iteration; each reduction's last workgroup turns the
  // partial sums into the scalars
  auto iteration = [&]() {
    hipLaunchKernelGGL(k_m_p, dim3(nb), dim3(BLOCK), 0, s, A);
    hipLaunchKernelGGL(k_m_v, dim3(nb), dim3(BLOCK), 0, s, A);
    hipLaunchKernelGGL(k_m_s, dim3(nb), dim3(BLOCK), 0, s, A);
    hipLaunchKernelGGL(k_m_t, dim3(nb), dim3(BLOCK), 0, s, A);
    hipLaunchKernelGGL(k_m_xr, dim3(nb), dim3(BLOCK), 0, s, A);
  };
  if (stream_ordered) {
    // exactly maxit iterations enqueued; decided components skip their work
    for (int it = 0; it < maxit; it++) iteration();
    KCHK(hipGetLastError());
    h->mpending = true;
    h->mstream = s;
    return 0;
  }
  // steered from the host in rounds of a few iterations: one 4-byte read
  // per round, not four per iteration
  KCHK(hipMemcpyAsync(h->mhost_done, h->mdone, sizeof(int), hipMemcpyDeviceToHost, s));
  KCHK(hipStreamSynchronize(s));
  int it = 0;
  while (!*h->mhost_done && it < maxit) {
    const int round = std::min(4, maxit - it);
    for (int k = 0; k < round; k++) iteration();
    it += round;
    KCHK(hipGetLastError());
    KCHK(hipMemcpyAsync(h->mhost_done, h->mdone, sizeof(int), hipMemcpyDeviceToHost, s));
    KCHK(hipStreamSynchronize(s));
  }
  KCHK(hipMemcpyAsync(h->mhost_sc, h->msc, C * sizeof(mrhs_sc), hipMemcpyDeviceToHost, s));
  KCHK(hipEventRecord(h->e1, s));
  KCHK(hipEventSynchronize(h->e1));
  h->mpending = false;
  if (st) {
    multi_stats(h->mhost_sc, C, st);
    float ms = 0.f;
    KCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
    st->ms = ms;
  }
  return 0;
}

extern "C" int pfx_bsr_multi_stats(pfx_bsr *h, void *stream, pfx_krylov_stats *st) {
  if (!h || !st || !h->msc || h->mC < 1) return PFX_E_ARG;
  KCHK(hipSetDevice(h->device));
  hipStream_t s = (hipStream_t)stream;
  KCHK(hipMemcpyAsync(h->mhost_sc, h->msc, h->mC * sizeof(mrhs_sc), hipMemcpyDeviceToHost, s));
  KCHK(hipStreamSynchronize(s));
  h->mpending = false;
  multi_stats(h->mhost_sc, h->mC, st);
  return 0;
}

namespace pfx_internal {
int bsr_multi_io(pfx_bsr *h, int C, double **B, double **X) {
  int rc = ensure_multi(h, C);
  if (rc) return rc;
  const size_t nv = (size_t)h->nrows * C;
  *B = h->mvec + 8 * nv;
  *X = h->mvec + 9 * nv;
  return 0;
}
}  // namespace pfx_internal

// ---------------------------------------------------------------------------
// the global-implicit Newton step

struct pfx_gi_solver {
  pfx_tables *tables = nullptr;
  int device = 0;
  int32_t nx = 0, ny = 0, nz = 0, N = 0;
  int64_t n = 0, nnzb = 0;
  int32_t *rp = nullptr, *ci = nullptr;
  double *vals = nullptr, *res = nullptr, *y = nullptr;
  double *rel = nullptr;   // |dx / x_old| per unknown (ITOL_RELATIVE_UPDATE)
  // the old state a cut time step restores: free ions and exchange state
  double *old_m = nullptr, *old_ionx = nullptr;
  int32_t R = 0;
  pfx_bsr *bsr = nullptr;
};

extern "C" int pfx_gi_solver_destroy(pfx_gi_solver *s) {
  if (!s) return 0;
  if (s->bsr) pfx_bsr_destroy(s->bsr);
  if (s->rp) (void)hipFree(s->rp);
  if (s->ci) (void)hipFree(s->ci);
  if (s->vals) (void)hipFree(s->vals);
  if (s->res) (void)hipFree(s->res);
  if (s->y) (void)hipFree(s->y);
  if (s->old_m) (void)hipFree(s->old_m);
  if (s->rel) (void)hipFree(s->rel);
  if (s->old_ionx) (void)hipFree(s->old_ionx);
  delete s;
  return 0;
}

extern "C" int pfx_gi_solver_create(pfx_tables *tables, const pfx_grid *grid, pfx_gi_solver **out) {
  if (!tables || !grid || !out || grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return PFX_E_ARG;
  *out = nullptr;
  int N, M, K, R;
  pfx_internal::table_sizes(tables, &N, &M, &K, &R);
  int64_t nnzb = 0;
  int rc = pfx_gi_structure(grid, nullptr, nullptr, &nnzb);
  if (rc) return rc;
  pfx_gi_solver *s = new pfx_gi_solver;
  KCHK(hipGetDevice(&s->device));
  s->tables = tables;
  s->nx = grid->nx;
  s->ny = grid->ny;
  s->nz = grid->nz;
  s->N = N;
  s->R = R;
  s->n = (int64_t)grid->nx * grid->ny * grid->nz;
  s->nnzb = nnzb;
  bool ok = hipMalloc((void **)&s->rp, (s->n + 1) * sizeof(int32_t)) == hipSuccess;
  ok = ok && hipMalloc((void **)&s->ci, nnzb * sizeof(int32_t)) == hipSuccess;
  ok = ok && hipMalloc((void **)&s->vals, (size_t)nnzb * N * N * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void **)&s->res, (size_t)s->n * N * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void **)&s->y, (size_t)s->n * N * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void **)&s->old_m, (size_t)s->n * N * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void **)&s->rel, (size_t)s->n * N * sizeof(double)) == hipSuccess;
  if (R > 0) ok = ok && hipMalloc((void **)&s->old_ionx, (size_t)s->n * R * sizeof(double)) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    pfx_gi_solver_destroy(s);
    return PFX_E_NOMEM;
  }
  if ((rc = pfx_gi_structure(grid, s->rp, s->ci, &nnzb)) || (rc = pfx_bsr_create(s->n, N, nnzb, s->rp, s->ci, &s->bsr))) {
    pfx_gi_solver_destroy(s);
    return rc;
  }
  *out = s;
  return 0;
}

extern "C" int pfx_gi_solver_step(pfx_gi_solver *s, pfx_cells *cells, const pfx_grid *grid, const pfx_react_opts *o,
                                  const pfx_newton_opts *no, pfx_newton_stats *st) {
  if (!s || !cells || !grid || !o || !cells->pri_molal) return PFX_E_ARG;
  if (grid->nx != s->nx || grid->ny != s->ny || grid->nz != s->nz || cells->ncells != s->n) return PFX_E_ARG;
  const auto wall0 = std::chrono::steady_clock::now();
  const double atol = (no && no->atol > 0.0) ? no->atol : 1.0e-50;
  const double rtol = (no && no->rtol > 0.0) ? no->rtol : 1.0e-8;
  const double stol = (no && no->stol > 0.0) ? no->stol : 1.0e-8;
  const int maxit = (no && no->max_iterations > 0) ? no->max_iterations : 50;
  const double itol_res = no ? no->itol_res : 0.0;
  const double itol_upd = no ? no->itol_update : 0.0;
  const double itol_rel = no ? no->itol_rel_update : 0.0;
  pfx_krylov_opts ko;
  memset(&ko, 0, sizeof(ko));
  if (no) ko = no->krylov;
  KCHK(hipSetDevice(s->device));
  hipStream_t hs = (hipStream_t)o->stream;
  pfx_newton_stats S;
  memset(&S, 0, sizeof(S));
  pfx_bsr *h = s->bsr;
  const int64_t n = s->n * s->N;
  const int ge = grid_for(n);
  double tms[2];
  // the capped-Q/K derivative for the first PFX_JAC_CAPPED_ITERS Newton
  // iterations, the reference's after them
  pfx_react_opts oe = *o;

  auto assemble = [&](double *fnorm) -> int {
    oe.jacobian_flags = S.iterations < PFX_JAC_CAPPED_ITERS ? o->jacobian_flags
                                                            : (o->jacobian_flags & ~PFX_JAC_CAPPED_QK);
    int rc = pfx_gi_residual_jacobian(s->tables, cells, grid, &oe, s->rp, s->res, s->vals);
    if (rc) return rc;
    pfx_tables_timing(s->tables, tms);
    S.ms_assembly += tms[0] + tms[1];
    hipLaunchKernelGGL(k_dot2, dim3(ge), dim3(BLOCK), 0, hs, n, s->res, s->res, nullptr, nullptr, h->partial);
    KCHK(hipGetLastError());
    double ff = 0.0;
    if ((rc = read2(h, ge, hs, &ff, nullptr))) return rc;
    *fnorm = sqrt(ff);
    return 0;
  };

  int rc = assemble(&S.fnorm0);
  if (rc) return rc;
  S.fnorm = S.fnorm0;
  // the initial residual decides nothing but NaN: one iteration at least
  // (newton_min_iterations, convergence.F90:242-245)
  if (!std::isfinite(S.fnorm0)) S.reason = PFX_SNES_DIVERGED_FNORM_NAN;
  const int use_log = grid->log_formulation ? 1 : 0;
  // max_dlnC of the network (reaction%max_dlnC)
  double max_dlnC = 5.0;
  pfx_internal::table_max_dlnC(s->tables, &max_dlnC);
  while (S.reason == 0) {
    if (S.iterations >= maxit) {
      S.reason = PFX_SNES_DIVERGED_MAX_IT;
      break;
    }
    pfx_krylov_stats ks;
    if ((rc = pfx_bsr_solve(h, s->vals, s->res, s->y, &ko, (void *)hs, &ks))) return rc;
    S.linear_iterations += ks.iterations;
    S.ms_linear += ks.ms;
    if (ks.reason < 0) {
      S.reason = PFX_SNES_DIVERGED_LINEAR_SOLVE;
      break;
    }
    double scale = 1.0;
    if (!use_log) {
      hipLaunchKernelGGL(k_min_ratio, dim3(ge), dim3(BLOCK), 0, hs, n, cells->pri_molal, s->y, h->partial);
      hipLaunchKernelGGL(k_reduce_min, dim3(1), dim3(BLOCK), 0, hs, ge, h->partial, h->red);
      KCHK(hipGetLastError());
      KCHK(hipMemcpyAsync(h->host2, h->red, 2 * sizeof(double), hipMemcpyDeviceToHost, hs));
      KCHK(hipStreamSynchronize(hs));
      const double min_ratio = h->host2[0];
      if (min_ratio < 1.0) scale = min_ratio * 0.99;
    }
    hipLaunchKernelGGL(k_newton_update, dim3(ge), dim3(BLOCK), 0, hs, n, cells->pri_molal, s->y, use_log, max_dlnC,
                       scale, h->partial, itol_rel > 0.0 ? s->rel : nullptr);
    KCHK(hipGetLastError());
    double yy = 0.0, xx = 0.0;
    if ((rc = read2(h, ge, hs, &yy, &xx))) return rc;
    S.iterations++;
    if ((rc = assemble(&S.fnorm))) return rc;
    // SNESConvergedDefault's order
    if (!std::isfinite(S.fnorm)) S.reason = PFX_SNES_DIVERGED_FNORM_NAN;
    else if (S.fnorm < atol) S.reason = PFX_SNES_CONVERGED_FNORM_ABS;
    else if (S.fnorm <= rtol * S.fnorm0) S.reason = PFX_SNES_CONVERGED_FNORM_RELATIVE;
    else if (sqrt(yy) < stol * sqrt(xx)) S.reason = PFX_SNES_CONVERGED_SNORM_RELATIVE;
    // PMRTCheckUpdatePost's relative update, then ConvergenceTest's
    // infinity norms and max_norm (||F||_inf <= ||F||_2: without ITOL only a
    // 2-norm above the bound needs the infinity norm)
    if (itol_rel > 0.0) {
      double rmax = 0.0;
      if ((rc = max_abs(h, ge, hs, n, s->rel, &rmax))) return rc;
      if (rmax < itol_rel) S.reason = PFX_SNES_CONVERGED_RELATIVE_UPDATE;
    }
    if (itol_res > 0.0 || (std::isfinite(S.fnorm) && S.fnorm > kMaxNorm)) {
      double finf = 0.0;
      if ((rc = max_abs(h, ge, hs, n, s->res, &finf))) return rc;
      if (itol_res > 0.0 && finf < itol_res) S.reason = PFX_SNES_CONVERGED_ITOL_RES;
      if (itol_upd > 0.0) {
        double uinf = 0.0;
        if ((rc = max_abs(h, ge, hs, n, s->y, &uinf))) return rc;
        if (uinf < itol_upd) S.reason = PFX_SNES_CONVERGED_ITOL_UPDATE;
      }
      if (finf > kMaxNorm) S.reason = PFX_SNES_DIVERGED_MAX_NORM;
    } else if (itol_upd > 0.0) {
      double uinf = 0.0;
      if ((rc = max_abs(h, ge, hs, n, s->y, &uinf))) return rc;
      if (uinf < itol_upd) S.reason = PFX_SNES_CONVERGED_ITOL_UPDATE;
    }
  }
  S.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
  if (st) *st = S;
  return 0;
}

// One time step with the reference's cut and retry (timestepper_BE.F90:
// 344-399): a Newton solve that does not converge restores the old state
// (RTTimeCut, reactive_transport.F90:58-96: tran_xx <- tran_yy; here the
// free ions and the exchange state), scales dt by reduction_factor and
// tries again, until max_cuts cuts have been taken.
extern "C" int pfx_gi_solver_timestep(pfx_gi_solver *s, pfx_cells *cells, const pfx_grid *grid,
                                      const pfx_react_opts *o, const pfx_newton_opts *no, int32_t max_cuts,
                                      double reduction_factor, pfx_gi_step_stats *st) {
  if (!s || !cells || !grid || !o || !cells->pri_molal || !(o->dt > 0.0)) return PFX_E_ARG;
  if (cells->ncells != s->n) return PFX_E_ARG;
  if (max_cuts < 0) max_cuts = 16;                       // timestepper_base.F90:144
  if (!(reduction_factor > 0.0 && reduction_factor < 1.0)) reduction_factor = 0.5;  // :145
  const auto wall0 = std::chrono::steady_clock::now();
  KCHK(hipSetDevice(s->device));
  hipStream_t hs = (hipStream_t)o->stream;
  const size_t bm = (size_t)s->n * s->N * sizeof(double);
  const size_t bx = (size_t)s->n * s->R * sizeof(double);
  const bool ionx = s->R > 0 && cells->ionx_ref_sorbed;
  KCHK(hipMemcpyAsync(s->old_m, cells->pri_molal, bm, hipMemcpyDeviceToDevice, hs));
  if (ionx) KCHK(hipMemcpyAsync(s->old_ionx, cells->ionx_ref_sorbed, bx, hipMemcpyDeviceToDevice, hs));
  pfx_gi_step_stats S;
  memset(&S, 0, sizeof(S));
  pfx_react_opts oc = *o;
  for (;;) {
    pfx_newton_stats ns;
    int rc = pfx_gi_solver_step(s, cells, grid, &oc, no, &ns);
    if (rc) return rc;
    S.newton = ns;
    S.newton_iterations += ns.iterations;
    S.linear_iterations += ns.linear_iterations;
    if (ns.reason > 0) break;
    S.wasted_newton_iterations += ns.iterations;
    if (S.cuts >= max_cuts || oc.dt * reduction_factor < 1.0e-20) break;   // dt_min, :159
    S.cuts++;
    oc.dt *= reduction_factor;
    KCHK(hipMemcpyAsync(cells->pri_molal, s->old_m, bm, hipMemcpyDeviceToDevice, hs));
    if (ionx) KCHK(hipMemcpyAsync(cells->ionx_ref_sorbed, s->old_ionx, bx, hipMemcpyDeviceToDevice, hs));
  }
  KCHK(hipStreamSynchronize(hs));
  S.reason = S.newton.reason;
  S.dt = oc.dt;
  S.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
  if (st) *st = S;
  return 0;
}
