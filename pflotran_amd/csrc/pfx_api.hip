// Host side of libpfx_chem.so: the C-ABI of include/pfx_chem.h.
//
// pfx_tables_create flattens the reaction network into the device table
// blob (row-per-lane lists, see pfx_device.h); pfx_react_batch launches the
// persistent Newton kernel over a batch of cells and reduces the per-cell
// statistics on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pfx_device.h"

template <int G, int MODE, int W, int LUP, bool XM>
__global__ void pfx_react_kernel(const pfx_kernel_params P);
template <int W, bool XM>
__global__ void pfx_react_zones_kernel(const pfx_kernel_params *zone_params, const pfx_zone_launch Z,
                                       uint64_t *finish);
extern "C" int pfx_scr_doubles(int M, int G, int K, int xm);

struct pfx_tables {
  pfx_dev_layout L;
  unsigned char *d_blob = nullptr;
  unsigned int *d_queue = nullptr;
  // per-call scratch (grown on demand)
  int32_t *d_iters = nullptr, *d_status = nullptr;
  double *d_relres = nullptr;
  int64_t scratch_cells = 0;
  unsigned int *d_defer = nullptr;        // deferred-factorisation list: count, cells
  int64_t defer_cells = 0;
  unsigned long long *d_stats = nullptr;  // reduction buffer
  double *d_gi = nullptr;                 // global-implicit per-cell scratch
  pfx_constraint_rxn *d_crxn = nullptr;   // mineral / gas constraint reactions [N]
  int64_t gi_cells = 0;
  double last_ms[2] = {0.0, 0.0};         // device time of the last call's launches
  int variant[4] = {0, 0, 0, 0};          // register-cap variant per kernel mode (pfx_tables_tune)
  // zoned launches led by these tables: parameter buffer and retire stamps
  pfx_kernel_params *d_zone_params = nullptr;
  uint64_t *d_finish = nullptr;
  int finish_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // side stream of a zoned launch's pivoting follow-up for these tables'
  // zone (default priority), with its fork and join events
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // the last stream-ordered pfx_react_batch, whose statistics
  // pfx_tables_stats reads (-1: none pending)
  int64_t pending_n = -1;
  hipStream_t pending_stream = nullptr;
  bool pending_fallback = false;
  int device = 0;
  int num_cu = 0;
};

// ---------------------------------------------------------------------------
// Fixed scratch of pfx_react_batch_noalloc: the work-queue cursor and fallback
// counter, the statistics reduction and, for callers that pass no per-cell
// outputs, the per-cell iteration counts, status and scaled residuals.  Each
// GPU holds its own copy until the process ends; every call on a GPU shares
// it (overlapping calls on one GPU race).
constexpr size_t kNoallocQueueOff = 0;      // uint32 [3]
constexpr size_t kNoallocStatsOff = 256;    // uint64 [64]
constexpr size_t kNoallocCellsOff = 1024;   // int32 iters [cap], int32 status [cap], double rel_res [cap]
constexpr size_t kNoallocDeferOff = kNoallocCellsOff + (size_t)PFX_NOALLOC_MAX_CELLS * 16;  // uint32 [cap + 1]
static __device__ __align__(256) unsigned char scratch[PFX_NOALLOC_SCRATCH_BYTES];
static_assert(kNoallocDeferOff + ((size_t)PFX_NOALLOC_MAX_CELLS + 1) * 4 <= PFX_NOALLOC_SCRATCH_BYTES,
              "scratch too small for its per-cell slots");

namespace {

constexpr int kMaxDevices = 64;
unsigned char *g_scratch_ptr[kMaxDevices] = {nullptr};

constexpr int BLOCK = 256;
constexpr int64_t kMaxLds = 160 * 1024;   // gfx950 LDS per workgroup

int g_last_hip_error = 0;

#define HIPCHK(x)                                                   \
  do {                                                              \
    hipError_t e_ = (x);                                            \
    if (e_ != hipSuccess) {                                         \
      g_last_hip_error = (int)e_;                                   \
      fprintf(stderr, "pfx_chem: %s failed: %s (%s:%d)\n", #x,     \
              hipGetErrorString(e_), __FILE__, __LINE__);           \
      return PFX_E_DEVICE;                                          \
    }                                                               \
  } while (0)

// hipGetSymbolAddress of the scratch array on `device` (current), once
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return PFX_E_UNSUP;
  if (g_scratch_ptr[device]) return 0;
  void *ptr = nullptr;
  HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return PFX_E_DEVICE;
  g_scratch_ptr[device] = (unsigned char *)ptr;
  return 0;
}

int group_size(int N) {
  if (N <= 8) return 8;
  if (N <= 16) return 16;
  if (N <= 32) return 32;
  return 0;
}

struct BlobBuilder {
  std::vector<unsigned char> bytes;
  template <typename T>
  int64_t add(const std::vector<T> &v) {
    size_t off = (bytes.size() + 15) & ~size_t(15);
    bytes.resize(off + v.size() * sizeof(T) + 16);  // never zero-sized
    if (!v.empty()) memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
    bytes.resize(off + std::max<size_t>(v.size() * sizeof(T), 8));
    return (int64_t)off;
  }
};

// ---------------------------------------------------------------------------
__global__ void k_zero_u32(unsigned int *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}
// a reaction launch's counters: queue cursors and fallback count, and the
// count of the deferred list (if any)
__global__ void k_zero_launch(unsigned int *queue, unsigned int *defer) {
  if (threadIdx.x < 3) queue[threadIdx.x] = 0u;
  if (threadIdx.x == 0 && defer) defer[0] = 0u;
}
__global__ void k_zero_u64(unsigned long long *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// stats[0]=converged, [1]=sum iterations, [2]=max it, [3]=min it (inverted),
// [4]=max rel_res bits, [5..36]=histogram
__global__ void k_stats(const int32_t *it, const int32_t *st, const double *rr, int64_t n,
                        unsigned long long *out) {
  __shared__ unsigned long long s_conv, s_sum, s_max, s_min, s_rr;
  __shared__ unsigned long long s_hist[32];
  if (threadIdx.x == 0) { s_conv = 0; s_sum = 0; s_max = 0; s_min = 0; s_rr = 0; }
  if (threadIdx.x < 32) s_hist[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long conv = 0, sum = 0, mx = 0, mn = 0, rmax = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int k = it[i];
    conv += (st[i] == PFX_CELL_CONVERGED);
    sum += (unsigned long long)k;
    mx = max(mx, (unsigned long long)k);
    mn = max(mn, (unsigned long long)(0x7fffffff - k));
    double v = rr[i];
    unsigned long long b = (v == v) ? (unsigned long long)__double_as_longlong(fabs(v)) : 0x7ff8000000000000ull;
    rmax = max(rmax, b);
    atomicAdd(&s_hist[k < 31 ? (k < 0 ? 0 : k) : 31], 1ull);
  }
  atomicAdd(&s_conv, conv);
  atomicAdd(&s_sum, sum);
  atomicMax(&s_max, mx);
  atomicMax(&s_min, mn);
  atomicMax(&s_rr, rmax);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], s_conv);
    atomicAdd(&out[1], s_sum);
    atomicMax(&out[2], s_max);
    atomicMax(&out[3], s_min);
    atomicMax(&out[4], s_rr);
  }
  if (threadIdx.x < 32) atomicAdd(&out[5 + threadIdx.x], s_hist[threadIdx.x]);
}

// RUpdateKineticState (reaction.F90:5340-5366): vf += rate * molar volume
// * dt, clamped at zero; one thread per (cell, mineral)
__global__ void k_kinetic_state(const unsigned char *blob, int64_t o_molvol, int K, int64_t n, const double *rate,
                                double *volfrac, double dt) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n * K) return;
  const double mv = ((const double *)(blob + o_molvol))[i % K];
  const double v = volfrac[i] + rate[i] * mv * dt;
  volfrac[i] = v < 0.0 ? 0.0 : v;
}

// the deferred-factorisation list for n cells (count + n indices)
int ensure_defer(pfx_tables *t, int64_t n) {
  if (n <= t->defer_cells) return 0;
  if (t->d_defer) HIPCHK(hipFree(t->d_defer));
  t->d_defer = nullptr;
  HIPCHK(hipMalloc(&t->d_defer, (size_t)(n + 1) * sizeof(unsigned int)));
  t->defer_cells = n;
  return 0;
}

int ensure_scratch(pfx_tables *t, int64_t n) {
  if (n <= t->scratch_cells) return 0;
  if (t->d_iters) HIPCHK(hipFree(t->d_iters));
  if (t->d_status) HIPCHK(hipFree(t->d_status));
  if (t->d_relres) HIPCHK(hipFree(t->d_relres));
  t->d_iters = nullptr;
  t->d_status = nullptr;
  t->d_relres = nullptr;
  HIPCHK(hipMalloc(&t->d_iters, n * sizeof(int32_t)));
  HIPCHK(hipMalloc(&t->d_status, n * sizeof(int32_t)));
  HIPCHK(hipMalloc(&t->d_relres, n * sizeof(double)));
  t->scratch_cells = n;
  return 0;
}

template <int MODE, int W, int LUP, bool XM = false>
const void *kernel_for(int G) {
  return G == 8 ? (const void *)pfx_react_kernel<8, MODE, W, LUP, XM>
       : G == 16 ? (const void *)pfx_react_kernel<16, MODE, W, LUP, XM>
                 : (const void *)pfx_react_kernel<32, MODE, W, LUP, XM>;
}

// tables whose minerals have rate prefactors or scale factors run the
// kernels compiled with them (uncapped)
bool tables_xm(const pfx_tables *t) { return (t->L.mnrl_flags & 48) != 0; }

// variant v caps the registers at v + 1 wavefronts per SIMD (v = 0: none);
// only the Newton modes have the capped variants.  lup: the reaction
// kernel's factorisation policy (PFX_LU_*; the pivoting one has no capped
// variants, the other modes always pivot).  xm: the prefactor kernels
// (reaction and evaluation modes only; no capped variants).
const void *kernel_ptr(int G, int mode, int v, int lup = PFX_LU_DEFER, bool xm = false) {
  if (xm && mode == PFX_MODE_REACT)
    return lup == PFX_LU_PIVOT      ? kernel_for<PFX_MODE_REACT, 1, PFX_LU_PIVOT, true>(G)
         : lup == PFX_LU_ONDEMAND ? kernel_for<PFX_MODE_REACT, 1, PFX_LU_ONDEMAND, true>(G)
                                  : kernel_for<PFX_MODE_REACT, 1, PFX_LU_DEFER, true>(G);
  if (xm && mode == PFX_MODE_EVAL) return kernel_for<PFX_MODE_EVAL, 1, PFX_LU_PIVOT, true>(G);
  switch (mode) {
    case PFX_MODE_REACT:
      if (lup == PFX_LU_PIVOT) return kernel_for<PFX_MODE_REACT, 1, PFX_LU_PIVOT>(G);
      if (lup == PFX_LU_ONDEMAND)
        return v == 2 ? kernel_for<PFX_MODE_REACT, 3, PFX_LU_ONDEMAND>(G)
             : v == 1 ? kernel_for<PFX_MODE_REACT, 2, PFX_LU_ONDEMAND>(G)
                      : kernel_for<PFX_MODE_REACT, 1, PFX_LU_ONDEMAND>(G);
      return v == 2 ? kernel_for<PFX_MODE_REACT, 3, PFX_LU_DEFER>(G)
           : v == 1 ? kernel_for<PFX_MODE_REACT, 2, PFX_LU_DEFER>(G)
                    : kernel_for<PFX_MODE_REACT, 1, PFX_LU_DEFER>(G);
    case PFX_MODE_EVAL: return kernel_for<PFX_MODE_EVAL, 1, PFX_LU_PIVOT>(G);
    case PFX_MODE_EQUIL:
      return v == 2 ? kernel_for<PFX_MODE_EQUIL, 3, PFX_LU_PIVOT>(G)
           : v == 1 ? kernel_for<PFX_MODE_EQUIL, 2, PFX_LU_PIVOT>(G)
                    : kernel_for<PFX_MODE_EQUIL, 1, PFX_LU_PIVOT>(G);
    default: return kernel_for<PFX_MODE_AUX, 1, PFX_LU_PIVOT>(G);
  }
}

const void *zones_kernel_ptr(int v, bool xm) {
  if (xm) return (const void *)pfx_react_zones_kernel<1, true>;
  return v == 2 ? (const void *)pfx_react_zones_kernel<3, false>
       : v == 1 ? (const void *)pfx_react_zones_kernel<2, false> : (const void *)pfx_react_zones_kernel<1, false>;
}

// queue: the work-queue cursors and fallback counter of this launch (the
// tables' own, or the fixed scratch array of pfx_react_batch_noalloc).
// A reaction step with a deferred list (P.defer) is two launches: the
// unpivoted kernel over the batch, then the pivoting kernel over the cells
// it listed (an empty list ends that launch at once); otherwise a reaction
// step factors with on-demand pivoting in one launch.
int launch(pfx_tables *t, pfx_kernel_params &P, int mode, hipStream_t stream, double *ms,
           unsigned int *queue = nullptr) {
  if (!queue) queue = t->d_queue;
  const int G = t->L.G;
  const int groups = BLOCK / G;
  const int64_t lds = t->L.lds_table_bytes + (int64_t)groups * pfx_scr_doubles(t->L.M, G, t->L.K, tables_xm(t)) * 8;
  if (lds > kMaxLds) return PFX_E_UNSUP;
  int occ = 0;
  const bool deferring = mode == PFX_MODE_REACT && P.defer != nullptr;
  const int lup = mode != PFX_MODE_REACT ? PFX_LU_PIVOT : (deferring ? PFX_LU_DEFER : PFX_LU_ONDEMAND);
  const void *fn = kernel_ptr(G, mode, t->variant[mode & 3], lup, tables_xm(t));
  const void *fn_piv = deferring ? kernel_ptr(G, mode, 0, PFX_LU_PIVOT, tables_xm(t)) : nullptr;
  if (lds > 64 * 1024) {
    HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (fn_piv) HIPCHK(hipFuncSetAttribute(fn_piv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, BLOCK, (size_t)lds));
  if (occ < 1) occ = 1;
  int64_t need_blocks = (P.ncells + groups - 1) / groups;
  int64_t blocks = std::min<int64_t>(need_blocks, (int64_t)t->num_cu * occ);
  if (blocks < 1) blocks = 1;
  // chunk: enough cells per fetch to keep queue atomics rare, small enough
  // to balance the tail
  int64_t per_wave = std::max<int64_t>(1, P.ncells / (blocks * (BLOCK / 64)));
  P.chunk = (int32_t)std::max<int64_t>(64 / G, std::min<int64_t>(64, per_wave / 8 + 1));
  P.waves = (int32_t)(blocks * (BLOCK / 64));
  P.queue = queue;
  P.fallback_count = queue + 1;
  // the pivoting launch: its own cursor, one cell per fetch, one workgroup
  // per compute unit at most (the list is short)
  pfx_kernel_params P2 = P;
  P2.queue = queue + 2;
  P2.chunk = 1;
  P2.waves = 0;
  const int64_t blocks2 = std::max<int64_t>(1, std::min<int64_t>(need_blocks, (int64_t)t->num_cu));
  hipLaunchKernelGGL(k_zero_launch, dim3(1), dim3(64), 0, stream, queue, P.defer);
  HIPCHK(hipGetLastError());
  void *args[] = {(void *)&P};
  void *args2[] = {(void *)&P2};
  if (!ms) {
    // stream-ordered: no events, no wait (the launches may be graph-captured)
    HIPCHK(hipLaunchKernel(fn, dim3((unsigned)blocks), dim3(BLOCK), args, (size_t)lds, stream));
    if (fn_piv) HIPCHK(hipLaunchKernel(fn_piv, dim3((unsigned)blocks2), dim3(BLOCK), args2, (size_t)lds, stream));
    return 0;
  }
  HIPCHK(hipEventRecord(t->ev0, stream));
  HIPCHK(hipLaunchKernel(fn, dim3((unsigned)blocks), dim3(BLOCK), args, (size_t)lds, stream));
  if (fn_piv) HIPCHK(hipLaunchKernel(fn_piv, dim3((unsigned)blocks2), dim3(BLOCK), args2, (size_t)lds, stream));
  HIPCHK(hipEventRecord(t->ev1, stream));
  HIPCHK(hipEventSynchronize(t->ev1));
  float f = 0.f;
  HIPCHK(hipEventElapsedTime(&f, t->ev0, t->ev1));
  *ms = f;
  t->last_ms[0] = f;
  t->last_ms[1] = 0.0;
  return 0;
}

void fill_params(pfx_tables *t, pfx_cells *cells, const pfx_react_opts *o, pfx_kernel_params &P) {
  memset(&P, 0, sizeof(P));
  P.L = t->L;
  P.tables = t->d_blob;
  P.ncells = cells->ncells;
  P.total = cells->total;
  P.total_sorb = cells->total_sorb;
  P.pri_molal = cells->pri_molal;
  P.pri_act_coef = cells->pri_act_coef;
  P.sec_act_coef = cells->sec_act_coef;
  P.sec_molal = cells->sec_molal;
  P.ionx_ref = cells->ionx_ref_sorbed;
  P.mnrl_volfrac = cells->mnrl_volfrac;
  P.mnrl_area = cells->mnrl_area;
  P.mnrl_rate = cells->mnrl_rate;
  P.temperature = cells->temperature;
  P.porosity = cells->porosity;
  P.saturation = cells->saturation;
  P.volume = cells->volume;
  P.den_kg = cells->den_kg;
  P.guess = cells->guess;
  P.num_iterations = cells->num_iterations;
  P.status = cells->status;
  P.rel_res = cells->rel_res;
  P.dt = o->dt;
  P.rel_tol = o->rel_tol;
  P.default_guess = o->default_guess;
  P.def_porosity = o->porosity;
  P.def_saturation = o->saturation;
  P.def_volume = o->volume;
  P.def_den = o->den_kg;
  P.max_iterations = o->max_iterations;
  P.extra_iterations = o->extra_iterations;
  P.guess_mode = o->guess_mode;
  P.reference_stop = o->reference_stop;
  P.fallback_after = o->fallback_after > 0 ? o->fallback_after : o->max_iterations;
  P.jac_flags = o->jacobian_flags;
  P.rel_change_tol = o->rel_change_tol;
  P.fallback_previous = o->fallback_previous ? 1 : 0;
  P.fallback_count = t->d_queue + 1;
}

int check_cells(const pfx_tables *t, const pfx_cells *c, const pfx_react_opts *o) {
  if (!c || !o) return PFX_E_ARG;
  if (c->ncells < 0 || c->ncells > (int64_t)0x7fffffff) return PFX_E_ARG;
  if (c->ncells == 0) return 0;  // an empty batch needs no arrays (RTReact over no cells)
  if (!c->total || !c->pri_molal) return PFX_E_ARG;
  if (t->L.sorb_any && !c->total_sorb) return PFX_E_ARG;
  if (t->L.K > 0 && (!c->mnrl_volfrac || !c->mnrl_area)) return PFX_E_ARG;
  if ((o->guess_mode == PFX_GUESS_ARRAY || o->guess_mode == PFX_GUESS_ARRAY_FALLBACK) && !c->guess)
    return PFX_E_ARG;
  if (o->guess_mode < PFX_GUESS_PREVIOUS || o->guess_mode > PFX_GUESS_ARRAY_FALLBACK) return PFX_E_ARG;
  if (o->fallback_after < 0) return PFX_E_ARG;
  if (o->jacobian_flags & ~PFX_JAC_CAPPED_QK) return PFX_E_ARG;
  if (!(o->dt > 0.0)) return PFX_E_ARG;
  if (!(o->rel_change_tol >= 0.0)) return PFX_E_ARG;
  return 0;
}

// statistics reduction into t->d_stats (kernels only: capturable)
int stats_launch(pfx_tables *t, const pfx_kernel_params &P, int64_t n, hipStream_t stream,
                 unsigned long long *sbuf = nullptr) {
  if (!sbuf) sbuf = t->d_stats;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, sbuf, 64);
  int sblocks = (int)std::min<int64_t>(1024, (n + 255) / 256);
  hipLaunchKernelGGL(k_stats, dim3(sblocks), dim3(256), 0, stream, P.num_iterations, P.status, P.rel_res, n,
                     sbuf);
  HIPCHK(hipGetLastError());
  return 0;
}

int stats_read(pfx_tables *t, const pfx_kernel_params &P, int64_t n, hipStream_t stream, double ms,
               pfx_stats *stats, const unsigned long long *sbuf = nullptr, const unsigned int *queue = nullptr) {
  if (!sbuf) sbuf = t->d_stats;
  if (!queue) queue = t->d_queue;
  unsigned long long h[64];
  unsigned int nfb = 0;
  HIPCHK(hipMemcpyAsync(h, sbuf, sizeof(h), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipMemcpyAsync(&nfb, queue + 1, sizeof(nfb), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  const int64_t conv = (int64_t)h[0];
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->ncells = n;
    stats->converged = conv;
    stats->total_iterations = (int64_t)h[1];
    stats->max_iterations = (int32_t)h[2];
    stats->min_iterations = (int32_t)(0x7fffffff - (long long)h[3]);
    double rr;
    memcpy(&rr, &h[4], sizeof(rr));
    stats->max_rel_res = rr;
    stats->kernel_ms = ms;
    for (int i = 0; i < 32; i++) stats->iteration_histogram[i] = (int64_t)h[5 + i];
    stats->fallbacks = P.guess_mode == PFX_GUESS_ARRAY_FALLBACK ? (int64_t)nfb : 0;
  }
  return (int)(n - conv);
}

int finish_stats(pfx_tables *t, const pfx_kernel_params &P, int64_t n, hipStream_t stream, double ms,
                 pfx_stats *stats) {
  int rc = stats_launch(t, P, n, stream);
  if (rc) return rc;
  return stats_read(t, P, n, stream, ms, stats);
}

__global__ void k_stamp(uint64_t *p) {
  if (threadIdx.x == 0) *p = wall_clock64();
}

int use_scratch(pfx_tables *t, pfx_kernel_params &P, int64_t n);

// ---------------------------------------------------------------------------
// zoned launches
struct zone_setup {
  int nzones = 0;
  int variant = 0;
  const void *fn = nullptr;
  pfx_kernel_params P[PFX_MAX_ZONES];
  int64_t n[PFX_MAX_ZONES];
  pfx_zone_launch Z;
  int64_t lds = 0;
  int total_blocks = 0;
};

int ensure_side(pfx_tables *t) {
  if (!t->side) HIPCHK(hipStreamCreateWithFlags(&t->side, hipStreamNonBlocking));
  if (!t->fork_ev) HIPCHK(hipEventCreateWithFlags(&t->fork_ev, hipEventDisableTiming));
  if (!t->join_ev) HIPCHK(hipEventCreateWithFlags(&t->join_ev, hipEventDisableTiming));
  return 0;
}

int zone_lds(const pfx_tables *t) {
  const int G = t->L.G;
  return (int)(t->L.lds_table_bytes + (int64_t)(BLOCK / G) * pfx_scr_doubles(t->L.M, G, t->L.K, tables_xm(t)) * 8);
}

int zones_prepare(int32_t nzones, pfx_tables *const *tables, pfx_cells *cells, const pfx_react_opts *o,
                  const pfx_zone_plan *plan, zone_setup &zs) {
  if (nzones < 1 || nzones > PFX_MAX_ZONES || !tables || !cells || !o) return PFX_E_ARG;
  if (o->max_iterations < 0 || o->extra_iterations < 0) return PFX_E_ARG;
  const int dev = tables[0] ? tables[0]->device : 0;
  for (int z = 0; z < nzones; z++) {
    if (!tables[z] || tables[z]->device != dev) return PFX_E_ARG;
    for (int y = 0; y < z; y++)
      if (tables[y] == tables[z]) return PFX_E_ARG;  // queues and statistics are per tables
    int rc = check_cells(tables[z], &cells[z], o);
    if (rc) return rc;
  }
  HIPCHK(hipSetDevice(dev));
  for (int z = 0; z < nzones; z++) {
    int rc = ensure_side(tables[z]);
    if (rc) return rc;
  }
  zs.nzones = nzones;
  memset(&zs.Z, 0, sizeof(zs.Z));
  zs.Z.nzones = nzones;
  zs.lds = 0;
  double w[PFX_MAX_ZONES], wsum = 0.0;
  for (int z = 0; z < nzones; z++) {
    pfx_tables *t = tables[z];
    fill_params(t, &cells[z], o, zs.P[z]);
    int rc = use_scratch(t, zs.P[z], cells[z].ncells);
    if (rc) return rc;
    if ((rc = ensure_defer(t, cells[z].ncells))) return rc;
    zs.P[z].defer = t->d_defer;
    zs.P[z].queue = t->d_queue;
    zs.P[z].fallback_count = t->d_queue + 1;
    zs.n[z] = cells[z].ncells;
    zs.Z.G[z] = t->L.G;
    zs.lds = std::max<int64_t>(zs.lds, zone_lds(t));
    w[z] = (double)cells[z].ncells * t->L.N * t->L.N;
    wsum += w[z];
  }
  if (zs.lds > kMaxLds) return PFX_E_UNSUP;
  zs.variant = plan && plan->variant >= 0 && plan->variant <= 2 ? plan->variant : 0;
  bool xm = false;
  for (int z = 0; z < nzones; z++) xm = xm || tables_xm(tables[z]);
  zs.fn = zones_kernel_ptr(zs.variant, xm);
  if (zs.lds > 64 * 1024) HIPCHK(hipFuncSetAttribute(zs.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)zs.lds));
  int occ = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, zs.fn, BLOCK, (size_t)zs.lds));
  if (occ < 1) occ = 1;
  const int full = tables[0]->num_cu * occ;
  int blocks[PFX_MAX_ZONES];
  bool given = plan && plan->nzones == nzones;
  for (int z = 0; given && z < nzones; z++) given = plan->blocks[z] > 0;
  int total = 0;
  for (int z = 0; z < nzones; z++) {
    int b = given ? plan->blocks[z] : (int)(full * (wsum > 0 ? w[z] / wsum : 1.0 / nzones) + 0.5);
    const int64_t need = (zs.n[z] + (BLOCK / zs.Z.G[z]) - 1) / (BLOCK / zs.Z.G[z]);
    b = (int)std::max<int64_t>(1, std::min<int64_t>(b, std::max<int64_t>(need, 1)));
    blocks[z] = b;
    total += b;
  }
  int acc = 0;
  for (int z = 0; z < nzones; z++) {
    zs.Z.block_start[z] = acc;
    acc += blocks[z];
    // chunk: as in the single-batch launch, from the zone's own share
    const int64_t per_wave = std::max<int64_t>(1, zs.n[z] / ((int64_t)blocks[z] * (BLOCK / 64)));
    zs.P[z].chunk = (int32_t)std::max<int64_t>(64 / zs.Z.G[z], std::min<int64_t>(64, per_wave / 8 + 1));
    zs.P[z].waves = blocks[z] * (BLOCK / 64);
  }
  zs.Z.block_start[nzones] = acc;
  zs.total_blocks = total;
  return 0;
}

// queue resets, start stamp, zoned kernel, each zone's pivoting launch over
// the cells the zoned kernel deferred, statistics (all kernels)
int zones_enqueue(pfx_tables *const *tables, zone_setup &zs, const pfx_kernel_params *d_params, uint64_t *d_finish,
                  hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  for (int z = 0; z < zs.nzones; z++)
    hipLaunchKernelGGL(k_zero_launch, dim3(1), dim3(64), 0, s, tables[z]->d_queue, zs.P[z].defer);
  hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, d_finish + zs.total_blocks);
  HIPCHK(hipGetLastError());
  if (ev0) HIPCHK(hipEventRecord(ev0, s));
  void *args[] = {(void *)&d_params, (void *)&zs.Z, (void *)&d_finish};
  HIPCHK(hipLaunchKernel(zs.fn, dim3((unsigned)zs.total_blocks), dim3(BLOCK), args, (size_t)zs.lds, s));
  // the zones' pivoting launches run concurrently, each on its tables' side
  // stream forked from s and joined back into it (their tails overlap
  // instead of adding up)
  HIPCHK(hipEventRecord(tables[0]->fork_ev, s));
  for (int z = 0; z < zs.nzones; z++) {
    pfx_tables *t = tables[z];
    pfx_kernel_params P2 = zs.P[z];
    P2.queue = t->d_queue + 2;
    P2.chunk = 1;
    P2.waves = 0;
    const int64_t lds = zone_lds(t);
    const void *fn_piv = kernel_ptr(t->L.G, PFX_MODE_REACT, 0, PFX_LU_PIVOT, tables_xm(t));
    if (lds > 64 * 1024) HIPCHK(hipFuncSetAttribute(fn_piv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int64_t need = (zs.n[z] + (BLOCK / t->L.G) - 1) / (BLOCK / t->L.G);
    const int64_t blocks2 = std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)t->num_cu));
    void *args2[] = {(void *)&P2};
    HIPCHK(hipStreamWaitEvent(t->side, tables[0]->fork_ev, 0));
    HIPCHK(hipLaunchKernel(fn_piv, dim3((unsigned)blocks2), dim3(BLOCK), args2, (size_t)lds, t->side));
    HIPCHK(hipEventRecord(t->join_ev, t->side));
  }
  for (int z = 0; z < zs.nzones; z++) HIPCHK(hipStreamWaitEvent(s, tables[z]->join_ev, 0));
  if (ev1) HIPCHK(hipEventRecord(ev1, s));
  for (int z = 0; z < zs.nzones; z++) {
    int rc = stats_launch(tables[z], zs.P[z], zs.n[z], s);
    if (rc) return rc;
  }
  return 0;
}

int zones_finish(pfx_tables *const *tables, const zone_setup &zs, const uint64_t *d_finish, hipStream_t s,
                 double ms, pfx_zone_plan *plan, pfx_stats *stats) {
  int notconv = 0;
  for (int z = 0; z < zs.nzones; z++) {
    int rc = stats_read(tables[z], zs.P[z], zs.n[z], s, ms, stats ? &stats[z] : nullptr);
    if (rc < 0) return rc;
    notconv += rc;
  }
  if (plan) {
    std::vector<uint64_t> fin(zs.total_blocks + 1);
    HIPCHK(hipMemcpy(fin.data(), d_finish, fin.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    int rate_khz = 0;
    HIPCHK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, tables[0]->device));
    const uint64_t t0 = fin[zs.total_blocks];
    plan->nzones = zs.nzones;
    for (int z = 0; z < zs.nzones; z++) {
      uint64_t last = t0;
      for (int b = zs.Z.block_start[z]; b < zs.Z.block_start[z + 1]; b++) last = std::max(last, fin[b]);
      plan->blocks[z] = zs.Z.block_start[z + 1] - zs.Z.block_start[z];
      plan->zone_ms[z] = rate_khz > 0 ? (double)(last - t0) / rate_khz : 0.0;
    }
    plan->kernel_ms = ms;
    plan->variant = zs.variant;
  }
  return notconv;
}

int ensure_zone_buffers(pfx_tables *t, int blocks) {
  if (!t->d_zone_params) HIPCHK(hipMalloc(&t->d_zone_params, PFX_MAX_ZONES * sizeof(pfx_kernel_params)));
  if (blocks + 1 > t->finish_cap) {
    if (t->d_finish) HIPCHK(hipFree(t->d_finish));
    t->d_finish = nullptr;
    t->finish_cap = 0;
    HIPCHK(hipMalloc(&t->d_finish, (blocks + 1) * sizeof(uint64_t)));
    t->finish_cap = blocks + 1;
  }
  return 0;
}

int use_scratch(pfx_tables *t, pfx_kernel_params &P, int64_t n) {
  if (!P.num_iterations || !P.status || !P.rel_res) {
    int rc = ensure_scratch(t, n);
    if (rc) return rc;
    if (!P.num_iterations) P.num_iterations = t->d_iters;
    if (!P.status) P.status = t->d_status;
    if (!P.rel_res) P.rel_res = t->d_relres;
  }
  return 0;
}

}  // namespace

// per-cell scratch of the global-implicit assembly: non-flux residual [N],
// non-flux Jacobian [N*N], totals [N], dT/dm [N*N]
int pfx_internal::gi_scratch(pfx_tables *t, int64_t ncells, double **rnf, double **jnf, double **tot,
                             double **dtot) {
  const int64_t N = t->L.N, per = 2 * N + 2 * N * N;
  if (ncells > t->gi_cells) {
    HIPCHK(hipSetDevice(t->device));
    if (t->d_gi) HIPCHK(hipFree(t->d_gi));
    t->d_gi = nullptr;
    t->gi_cells = 0;
    HIPCHK(hipMalloc(&t->d_gi, (size_t)(ncells * per) * sizeof(double)));
    t->gi_cells = ncells;
  }
  *rnf = t->d_gi;
  *jnf = *rnf + ncells * N;
  *tot = *jnf + ncells * N * N;
  *dtot = *tot + ncells * N;
  return 0;
}

int pfx_internal::gi_evaluate_cells(pfx_tables *t, pfx_cells *cells, const pfx_react_opts *o, int col_scale,
                                    double *rnf, double *jnf, double *tot, double *dtot) {
  int rc = check_cells(t, cells, o);
  if (rc) return rc;
  if (cells->ncells == 0) return 0;
  HIPCHK(hipSetDevice(t->device));
  pfx_kernel_params P;
  fill_params(t, cells, o, P);
  P.res_out = rnf;
  P.jac_out = jnf;
  P.tot_eval_out = tot;
  P.dtot_out = dtot;
  P.col_scale = col_scale ? 1 : 0;
  double ms = 0.0;
  return launch(t, P, PFX_MODE_EVAL, (hipStream_t)o->stream, o->stream_ordered ? nullptr : &ms);
}

int pfx_internal::timed_launch_begin(pfx_tables *t, hipStream_t s) {
  HIPCHK(hipEventRecord(t->ev0, s));
  return 0;
}

int pfx_internal::timed_launch_end(pfx_tables *t, hipStream_t s, int slot) {
  HIPCHK(hipEventRecord(t->ev1, s));
  HIPCHK(hipEventSynchronize(t->ev1));
  float f = 0.f;
  HIPCHK(hipEventElapsedTime(&f, t->ev0, t->ev1));
  t->last_ms[slot] = f;
  return 0;
}

void pfx_internal::table_max_dlnC(const pfx_tables *t, double *max_dlnC) { *max_dlnC = t->L.max_dlnC; }

void pfx_internal::table_sizes(const pfx_tables *t, int *N, int *M, int *K, int *R) {
  *N = t->L.N;
  *M = t->L.M;
  *K = t->L.K;
  *R = t->L.R;
}

// ---------------------------------------------------------------------------
extern "C" int pfx_abi_version(void) { return PFX_ABI_VERSION; }

extern "C" int pfx_last_hip_error(void) { return g_last_hip_error; }

extern "C" int pfx_tables_create(const pfx_reaction_desc *d, pfx_tables **out) {
  if (!d || !out) return PFX_E_ARG;
  *out = nullptr;
  const int N = d->naqcomp, M = d->neqcplx, K = d->nkinmnrl, R = d->neqionxrxn;
  const int S0 = d->max_rxn_spec, C0 = d->max_cation;
  if (N < 1 || M < 0 || K < 0 || R < 0 || S0 < 1 || (R > 0 && C0 < 1)) return PFX_E_ARG;
  if (N > PFX_MAX_PRIMARY || M > PFX_MAX_CPLX || K > PFX_MAX_MINERALS || R > PFX_MAX_IONX_RXN ||
      S0 > PFX_MAX_RXN_SPEC || C0 > PFX_MAX_IONX_CATION)
    return PFX_E_UNSUP;
  if (!d->primary_spec_Z || !d->primary_spec_a0) return PFX_E_ARG;
  if (M > 0 && (!d->eqcplxspecid || !d->eqcplxstoich || !d->eqcplx_Z || !d->eqcplx_a0 || !d->eqcplx_logK))
    return PFX_E_ARG;
  if (K > 0 && (!d->kinmnrlspecid || !d->kinmnrlstoich || !d->kinmnrl_logK || !d->kinmnrl_rate_constant))
    return PFX_E_ARG;
  if (R > 0 && (!d->eqionx_rxn_cationid || !d->eqionx_rxn_k || !d->eqionx_rxn_CEC)) return PFX_E_ARG;
  if (d->act_coef_update_frequency < PFX_ACT_OFF || d->act_coef_update_frequency > PFX_ACT_NEWTON_ITER ||
      (d->act_coef_update_algorithm != PFX_ACT_ALG_LAG && d->act_coef_update_algorithm != PFX_ACT_ALG_NEWTON) ||
      (d->numerical_jacobian != 0 && d->numerical_jacobian != 1))
    return PFX_E_ARG;
  for (int i = 0; i < M; i++)
    if (d->eqcplxh2ostoich && d->eqcplxh2ostoich[i] != 0.0) {
      // water activity is 1 (ln a_w = 0, use_activity_h2o = F): the term vanishes
    }
  int G = group_size(N);
  if (d->group_lanes != 0) {
    if ((d->group_lanes != 8 && d->group_lanes != 16 && d->group_lanes != 32) || d->group_lanes < N)
      return PFX_E_ARG;
    G = d->group_lanes;
  }
  pfx_dev_layout L;
  memset(&L, 0, sizeof(L));
  L.N = N; L.M = M; L.K = K; L.R = R; L.G = G; L.S = S0;
  L.Q = (M + G - 1) / G;

  auto cid = [&](int i, int s) { return d->eqcplxspecid[(size_t)i * (S0 + 1) + 1 + s]; };
  auto cnt = [&](int i) { return d->eqcplxspecid[(size_t)i * (S0 + 1)]; };
  auto cnu = [&](int i, int s) { return d->eqcplxstoich[(size_t)i * S0 + s]; };
  for (int i = 0; i < M; i++) {
    if (cnt(i) < 0 || cnt(i) > S0) return PFX_E_ARG;
    for (int s = 0; s < cnt(i); s++)
      if (cid(i, s) < 0 || cid(i, s) >= N) return PFX_E_ARG;
  }

  BlobBuilder B;
  std::vector<double> Z(G, 0.0), a0(G, 0.0);
  for (int j = 0; j < N; j++) { Z[j] = d->primary_spec_Z[j]; a0[j] = d->primary_spec_a0[j]; }
  L.o_Z = B.add(Z);
  L.o_a0 = B.add(a0);
  std::vector<int32_t> cspec((size_t)M * S0, -1);
  std::vector<double> cnuv((size_t)M * S0, 0.0), cZ(M), ca0(M), clogK(M), ccoef((size_t)M * 5, 0.0);
  for (int i = 0; i < M; i++) {
    for (int s = 0; s < cnt(i); s++) { cspec[i * S0 + s] = cid(i, s); cnuv[i * S0 + s] = cnu(i, s); }
    cZ[i] = d->eqcplx_Z[i];
    ca0[i] = d->eqcplx_a0[i];
    clogK[i] = d->eqcplx_logK[i];
    if (d->eqcplx_logKcoef) for (int q = 0; q < 5; q++) ccoef[i * 5 + q] = d->eqcplx_logKcoef[i * 5 + q];
  }
  L.has_logk_coef = d->eqcplx_logKcoef != nullptr;
  L.o_cspec = B.add(cspec);
  L.o_cnu = B.add(cnuv);
  L.o_cZ = B.add(cZ);
  L.o_ca0 = B.add(ca0);
  L.o_clogK = B.add(clogK);

  // adjacency of each primary (its complexes) and the off-diagonal
  // co-occurrence lists of dT_r/dm_k, per column k, per lane r
  std::vector<std::vector<std::pair<int, double>>> adj(G);
  std::vector<std::vector<std::vector<std::pair<int, double>>>> off(G, std::vector<std::vector<std::pair<int, double>>>(G));
  for (int i = 0; i < M; i++)
    for (int s = 0; s < cnt(i); s++) {
      const int rr = cid(i, s);
      adj[rr].push_back({i, cnu(i, s)});
      for (int t2 = 0; t2 < cnt(i); t2++) {
        const int kk = cid(i, t2);
        if (kk == rr) continue;
        off[kk][rr].push_back({i, cnu(i, s) * cnu(i, t2)});
      }
    }
  // merge duplicate species within a complex would double-count: the
  // reference sums stoichiometry slots independently, so lists keep every slot
  int maxdeg = 0;
  for (int r = 0; r < G; r++) maxdeg = std::max<int>(maxdeg, (int)adj[r].size());
  L.maxdeg = maxdeg;
  std::vector<int32_t> adj_c((size_t)std::max(maxdeg, 1) * G, M);
  std::vector<double> adj_nu((size_t)std::max(maxdeg, 1) * G, 0.0);
  for (int r = 0; r < G; r++)
    for (size_t e = 0; e < adj[r].size(); e++) {
      adj_c[e * G + r] = adj[r][e].first;
      adj_nu[e * G + r] = adj[r][e].second;
    }
  L.o_adj_c = B.add(adj_c);
  L.o_adj_nu = B.add(adj_nu);
  // chunks of 4 columns, each with as many slots as its longest list; the
  // kernel reads a slot's 4 complexes (16-bit) and 4 coefficients with
  // one 8-byte and two 16-byte LDS loads per lane
  const int nq = G / 4;
  int rows = 0;
  for (int q = 0; q < nq; q++) {
    L.offq_start[q] = rows;
    int mx = 0;
    for (int j = 0; j < 4; j++)
      for (int r = 0; r < G; r++) mx = std::max<int>(mx, (int)off[4 * q + j][r].size());
    rows += mx;
  }
  L.offq_start[nq] = rows;
  for (int q = nq + 1; q <= PFX_MAX_PRIMARY / 4; q++) L.offq_start[q] = rows;
  std::vector<uint16_t> offq_c((size_t)std::max(rows, 1) * G * 4, (uint16_t)M);
  std::vector<double> offq_coef((size_t)std::max(rows, 1) * G * 4, 0.0);
  for (int q = 0; q < nq; q++)
    for (int j = 0; j < 4; j++)
      for (int r = 0; r < G; r++)
        for (size_t e = 0; e < off[4 * q + j][r].size(); e++) {
          const size_t row = (size_t)L.offq_start[q] + e;
          offq_c[(row * G + r) * 4 + j] = (uint16_t)off[4 * q + j][r][e].first;
          offq_coef[((row * 2 + j / 2) * G + r) * 2 + (j % 2)] = off[4 * q + j][r][e].second;
        }
  L.o_offq_c = B.add(offq_c);
  L.o_offq_coef = B.add(offq_coef);

  // minerals
  auto mid = [&](int k, int s) { return d->kinmnrlspecid[(size_t)k * (S0 + 1) + 1 + s]; };
  auto mcnt = [&](int k) { return d->kinmnrlspecid[(size_t)k * (S0 + 1)]; };
  std::vector<int32_t> mspec((size_t)std::max(K, 1) * S0, -1);
  std::vector<double> mnu((size_t)std::max(K, 1) * S0, 0.0), mnuG((size_t)std::max(K, 1) * G, 0.0);
  std::vector<double> mlogK(std::max(K, 1)), mcoef((size_t)std::max(K, 1) * 5, 0.0), mk(std::max(K, 1)),
      mmv(std::max(K, 1)), mEa(std::max(K, 1), 0.0), mthr(std::max(K, 1), 0.0), mlim(std::max(K, 1), 0.0),
      mpow(std::max(K, 1), 1.0), mtk(std::max(K, 1), 1.0);
  for (int k = 0; k < K; k++) {
    if (mcnt(k) < 0 || mcnt(k) > S0) return PFX_E_ARG;
    for (int s = 0; s < mcnt(k); s++) {
      const int sp = mid(k, s);
      if (sp < 0 || sp >= N) return PFX_E_ARG;
      mspec[k * S0 + s] = sp;
      mnu[k * S0 + s] = d->kinmnrlstoich[(size_t)k * S0 + s];
      mnuG[k * G + sp] += d->kinmnrlstoich[(size_t)k * S0 + s];
    }
    mlogK[k] = d->kinmnrl_logK[k];
    if (d->kinmnrl_logKcoef) for (int q = 0; q < 5; q++) mcoef[k * 5 + q] = d->kinmnrl_logKcoef[k * 5 + q];
    mk[k] = d->kinmnrl_rate_constant[k];
    mmv[k] = d->kinmnrl_molar_vol ? d->kinmnrl_molar_vol[k] : 0.0;
    if (d->kinmnrl_activation_energy) mEa[k] = d->kinmnrl_activation_energy[k];
    if (d->kinmnrl_affinity_threshold) mthr[k] = d->kinmnrl_affinity_threshold[k];
    if (d->kinmnrl_rate_limiter) mlim[k] = d->kinmnrl_rate_limiter[k];
    if (d->kinmnrl_affinity_power) mpow[k] = d->kinmnrl_affinity_power[k];
    if (d->kinmnrl_Temkin_const) mtk[k] = d->kinmnrl_Temkin_const[k];
  }
  // rate prefactors (reaction_mineral.F90:743-786) and min scale factor
  const int PP = d->kinmnrl_num_prefactors ? d->max_prefactors : 0;
  const int PQ = PP > 0 ? d->max_prefactor_species : 0;
  if (PP < 0 || PP > PFX_MAX_PREFACTORS || PQ < 0 || PQ > PFX_MAX_PREFACTOR_SPECIES) return PFX_E_UNSUP;
  if (PP > 0 && (PQ < 1 || !d->kinmnrl_prefactor_id || !d->kinmnrl_pref_alpha || !d->kinmnrl_pref_beta ||
                 !d->kinmnrl_pref_atten_coef || !d->kinmnrl_pref_rate))
    return PFX_E_ARG;
  const int PPa = std::max(PP, 1), PQa = std::max(PQ, 1);
  std::vector<int32_t> mpn(std::max(K, 1), 0), mpid((size_t)std::max(K, 1) * PPa * PQa, -1);
  std::vector<double> mpalpha((size_t)std::max(K, 1) * PPa * PQa, 0.0), mpbeta(mpalpha), mpatten(mpalpha);
  std::vector<double> mprate((size_t)std::max(K, 1) * PPa, 0.0), mpEa(mprate), mmsf(std::max(K, 1), 1.0);
  bool any_pref = false;
  for (int k = 0; k < K && PP > 0; k++) {
    const int np = d->kinmnrl_num_prefactors[k];
    if (np < 0 || np > PP) return PFX_E_ARG;
    mpn[k] = np;
    any_pref = any_pref || np > 0;
    for (int p = 0; p < np; p++) {
      const size_t kp = (size_t)k * PP + p;
      mprate[k * PPa + p] = d->kinmnrl_pref_rate[kp];
      mpEa[k * PPa + p] = d->kinmnrl_pref_activation_energy ? d->kinmnrl_pref_activation_energy[kp] : 0.0;
      const int ns = d->kinmnrl_prefactor_id[kp * (PQ + 1)];
      if (ns < 0 || ns > PQ) return PFX_E_ARG;
      for (int q = 0; q < ns; q++) {
        const int sp = d->kinmnrl_prefactor_id[kp * (PQ + 1) + 1 + q];
        if (sp < 0) return PFX_E_UNSUP;  // complexes as prefactor species
        if (sp >= N) return PFX_E_ARG;
        const size_t o = ((size_t)k * PPa + p) * PQa + q;
        mpid[o] = sp;
        mpalpha[o] = d->kinmnrl_pref_alpha[kp * PQ + q];
        mpbeta[o] = d->kinmnrl_pref_beta[kp * PQ + q];
        mpatten[o] = d->kinmnrl_pref_atten_coef[kp * PQ + q];
      }
    }
  }
  if (d->kinmnrl_min_scale_factor)
    for (int k = 0; k < K; k++) mmsf[k] = d->kinmnrl_min_scale_factor[k];
  L.npref = any_pref ? PPa : 0;
  L.npref_spec = any_pref ? PQa : 0;
  L.has_mnrl_coef = d->kinmnrl_logKcoef != nullptr;
  L.mnrl_flags = (d->kinmnrl_affinity_threshold ? 1 : 0) | (d->kinmnrl_rate_limiter ? 2 : 0) |
                 (d->kinmnrl_affinity_power ? 4 : 0) | (d->kinmnrl_Temkin_const ? 8 : 0) |
                 (any_pref ? 16 : 0) | (d->kinmnrl_min_scale_factor ? 32 : 0);
  L.o_mspec = B.add(mspec);
  L.o_mnu = B.add(mnu);
  L.o_mnuG = B.add(mnuG);
  L.o_mlogK = B.add(mlogK);
  L.o_mk = B.add(mk);
  L.o_mEa = B.add(mEa);
  L.o_mthr = B.add(mthr);
  L.o_mlim = B.add(mlim);
  L.o_mpow = B.add(mpow);
  L.o_mtemkin = B.add(mtk);
  // staged only where used (the LDS copy of the blob bounds the occupancy)
  if (d->kinmnrl_min_scale_factor) L.o_mmsf = B.add(mmsf);
  if (any_pref) {
    L.o_mpn = B.add(mpn);
    L.o_mpid = B.add(mpid);
    L.o_mpalpha = B.add(mpalpha);
    L.o_mpbeta = B.add(mpbeta);
    L.o_mpatten = B.add(mpatten);
    L.o_mprate = B.add(mprate);
    L.o_mpEa = B.add(mpEa);
  }
  // ion exchange
  std::vector<double> ionx_k((size_t)std::max(R, 1) * G, 0.0);
  std::vector<int32_t> ionx_role((size_t)std::max(R, 1) * G, -1);
  for (int q = 0; q < R; q++) {
    const int nc = d->eqionx_rxn_cationid[(size_t)q * (C0 + 1)];
    if (nc < 1 || nc > C0) return PFX_E_ARG;
    L.ionx_ncat[q] = nc;
    L.ionx_CEC[q] = d->eqionx_rxn_CEC[q];
    L.ionx_to_surf[q] = d->eqionx_rxn_to_surf ? d->eqionx_rxn_to_surf[q] : -1;
    if (L.ionx_to_surf[q] >= K) return PFX_E_ARG;
    for (int j = 0; j < nc; j++) {
      const int sp = d->eqionx_rxn_cationid[(size_t)q * (C0 + 1) + 1 + j];
      if (sp < 0 || sp >= N) return PFX_E_ARG;
      if (ionx_role[q * G + sp] >= 0) return PFX_E_ARG;  // cation listed twice
      ionx_role[q * G + sp] = j;
      ionx_k[q * G + sp] = d->eqionx_rxn_k[(size_t)q * C0 + j];
    }
    const int ref = d->eqionx_rxn_cationid[(size_t)q * (C0 + 1) + 1];
    L.ionx_ref_lane[q] = ref;
    L.ionx_ref_k[q] = d->eqionx_rxn_k[(size_t)q * C0];
    int zflag = 0;
    for (int i = 0; i < nc && !zflag; i++)
      for (int j = 0; j < nc; j++) {
        const int a = d->eqionx_rxn_cationid[(size_t)q * (C0 + 1) + 1 + i];
        const int b = d->eqionx_rxn_cationid[(size_t)q * (C0 + 1) + 1 + j];
        if (std::fabs(d->primary_spec_Z[a] - d->primary_spec_Z[b]) > 0.1) { zflag = 1; break; }
      }
    L.ionx_zflag[q] = zflag;
  }
  L.o_ionx_k = B.add(ionx_k);
  L.o_ionx_role = B.add(ionx_role);
  // isotherm sorption: one reaction per primary, per-lane parameters
  {
    const int D = d->neqkdrxn;
    if (D < 0 || D > PFX_MAX_KD_RXN) return PFX_E_UNSUP;
    if (D > 0 && (!d->eqkdspecid || !d->eqkdtype || !d->eqkddistcoef)) return PFX_E_ARG;
    std::vector<int32_t> kd_type(G, 0), kd_mnrl(G, -1);
    std::vector<double> kd_coef(G, 0.0), kd_b(G, 0.0), kd_invn(G, 1.0);
    for (int q = 0; q < D; q++) {
      const int sp = d->eqkdspecid[q];
      const int ty = d->eqkdtype[q];
      if (sp < 0 || sp >= N || ty < PFX_SORPTION_LINEAR || ty > PFX_SORPTION_FREUNDLICH) return PFX_E_ARG;
      if (kd_type[sp] != 0) return PFX_E_UNSUP;  // a second isotherm on one species
      if (ty == PFX_SORPTION_LANGMUIR && !d->eqkdlangmuirb) return PFX_E_ARG;
      if (ty == PFX_SORPTION_FREUNDLICH && !d->eqkdfreundlichn) return PFX_E_ARG;
      kd_type[sp] = ty;
      kd_coef[sp] = d->eqkddistcoef[q];
      kd_b[sp] = d->eqkdlangmuirb ? d->eqkdlangmuirb[q] : 0.0;
      kd_invn[sp] = (ty == PFX_SORPTION_FREUNDLICH) ? 1.0 / d->eqkdfreundlichn[q] : 1.0;
      kd_mnrl[sp] = d->eqkdmineral ? d->eqkdmineral[q] : -1;
      if (kd_mnrl[sp] >= K) return PFX_E_ARG;
    }
    L.kd_any = D > 0 ? 1 : 0;
    L.sorb_any = (R > 0 || D > 0) ? 1 : 0;
    L.soil_density = d->soil_particle_density;
    L.o_kd_type = B.add(kd_type);
    L.o_kd_mnrl = B.add(kd_mnrl);
    L.o_kd_coef = B.add(kd_coef);
    L.o_kd_b = B.add(kd_b);
    L.o_kd_invn = B.add(kd_invn);
  }
  // not needed inside the Newton loop: kept past the LDS prefix
  L.lds_table_bytes = ((int64_t)B.bytes.size() + 15) & ~int64_t(15);
  L.o_clogKcoef = B.add(ccoef);
  L.o_mlogKcoef = B.add(mcoef);
  L.o_mmolvol = B.add(mmv);
  L.o_ch2o = B.add(std::vector<double>(std::max(M, 1), 0.0));
  L.o_mh2o = B.add(std::vector<double>(std::max(K, 1), 0.0));
  // the coefficient tables are read per cell at load time, from the LDS copy
  // when they fit: stage the whole blob
  B.bytes.resize((B.bytes.size() + 15) & ~size_t(15));
  L.blob_bytes = (int64_t)B.bytes.size();
  L.lds_table_bytes = L.blob_bytes;
  L.debyeA = d->debyeA;
  L.debyeB = d->debyeB;
  L.debyeBdot = d->debyeBdot;
  L.use_log = d->use_log_formulation;
  L.act_freq = d->act_coef_update_frequency;
  L.act_alg = d->act_coef_update_algorithm;
  L.num_jac = d->numerical_jacobian;
  L.max_dlnC = d->max_dlnC;
  L.max_rel_change_tol = d->max_relative_change_tolerance;
  L.max_residual_tol = d->max_residual_tolerance;
  const int64_t lds = L.lds_table_bytes + (int64_t)(BLOCK / G) * pfx_scr_doubles(M, G, K, (L.mnrl_flags & 48) != 0) * 8;
  if (lds > kMaxLds) return PFX_E_UNSUP;

  pfx_tables *t = new pfx_tables();
  t->L = L;
  HIPCHK(hipGetDevice(&t->device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, t->device));
  t->num_cu = prop.multiProcessorCount;
  t->variant[PFX_MODE_REACT] = t->variant[PFX_MODE_EQUIL] = G <= 16 ? 2 : 1;
  if (lds > (int64_t)prop.sharedMemPerBlock) {
    delete t;
    return PFX_E_UNSUP;
  }
  HIPCHK(hipMalloc(&t->d_blob, (size_t)L.blob_bytes));
  HIPCHK(hipMemcpy(t->d_blob, B.bytes.data(), (size_t)L.blob_bytes, hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&t->d_queue, 256));
  HIPCHK(hipMalloc(&t->d_stats, 64 * sizeof(unsigned long long)));
  HIPCHK(hipEventCreate(&t->ev0));
  HIPCHK(hipEventCreate(&t->ev1));
  {
    // the device address of this GPU's fixed scratch array, once per device
    // (the array's name is a host stand-in, not a device address)
    int rc = noalloc_scratch_init(t->device);
    if (rc) {
      pfx_tables_destroy(t);
      return rc;
    }
  }
  *out = t;
  return 0;
}

extern "C" int pfx_tables_destroy(pfx_tables *t) {
  if (!t) return PFX_E_ARG;
  (void)hipFree(t->d_blob);
  (void)hipFree(t->d_queue);
  (void)hipFree(t->d_stats);
  if (t->d_gi) (void)hipFree(t->d_gi);
  if (t->d_crxn) (void)hipFree(t->d_crxn);
  if (t->d_zone_params) (void)hipFree(t->d_zone_params);
  if (t->d_finish) (void)hipFree(t->d_finish);
  if (t->d_iters) (void)hipFree(t->d_iters);
  if (t->d_status) (void)hipFree(t->d_status);
  if (t->d_relres) (void)hipFree(t->d_relres);
  if (t->d_defer) (void)hipFree(t->d_defer);
  if (t->ev0) (void)hipEventDestroy(t->ev0);
  if (t->ev1) (void)hipEventDestroy(t->ev1);
  if (t->fork_ev) (void)hipEventDestroy(t->fork_ev);
  if (t->join_ev) (void)hipEventDestroy(t->join_ev);
  if (t->side) (void)hipStreamDestroy(t->side);
  delete t;
  return 0;
}

extern "C" int pfx_react_batch(pfx_tables *t, pfx_cells *cells, const pfx_react_opts *o, pfx_stats *stats) {
  if (!t) return PFX_E_ARG;
  int rc = check_cells(t, cells, o);
  if (rc) return rc;
  if (o->max_iterations < 0 || o->extra_iterations < 0) return PFX_E_ARG;
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->ncells = cells->ncells;
  }
  if (cells->ncells == 0) return 0;
  HIPCHK(hipSetDevice(t->device));
  hipStream_t stream = (hipStream_t)o->stream;
  pfx_kernel_params P;
  fill_params(t, cells, o, P);
  if ((rc = use_scratch(t, P, cells->ncells))) return rc;
  if ((rc = ensure_defer(t, cells->ncells))) return rc;
  P.defer = t->d_defer;
  if (o->stream_ordered) {
    // enqueue the step and its statistics reduction; pfx_tables_stats
    // reads them later
    if ((rc = launch(t, P, PFX_MODE_REACT, stream, nullptr))) return rc;
    if ((rc = stats_launch(t, P, cells->ncells, stream))) return rc;
    t->pending_n = cells->ncells;
    t->pending_stream = stream;
    t->pending_fallback = P.guess_mode == PFX_GUESS_ARRAY_FALLBACK;
    return 0;
  }
  double ms = 0.0;
  if ((rc = launch(t, P, PFX_MODE_REACT, stream, &ms))) return rc;
  t->pending_n = -1;
  return finish_stats(t, P, cells->ncells, stream, ms, stats);
}

extern "C" int pfx_tables_stats(pfx_tables *t, pfx_stats *stats) {
  if (!t || !stats) return PFX_E_ARG;
  if (t->pending_n < 0) return PFX_E_ARG;  // no stream-ordered step since the last read
  HIPCHK(hipSetDevice(t->device));
  pfx_kernel_params P;
  memset(&P, 0, sizeof(P));
  P.guess_mode = t->pending_fallback ? PFX_GUESS_ARRAY_FALLBACK : PFX_GUESS_DEFAULT;
  const int64_t n = t->pending_n;
  t->pending_n = -1;
  int rc = stats_read(t, P, n, t->pending_stream, 0.0, stats);
  return rc;
}

extern "C" int pfx_react_batch_noalloc(pfx_tables *t, pfx_cells *cells, const pfx_react_opts *o,
                                       pfx_stats *stats) {
  if (!t) return PFX_E_ARG;
  int rc = check_cells(t, cells, o);
  if (rc) return rc;
  if (o->max_iterations < 0 || o->extra_iterations < 0) return PFX_E_ARG;
  if (o->stream_ordered) return PFX_E_ARG;   // the statistics are read back before returning
  const bool per_cell_scratch = !cells->num_iterations || !cells->status || !cells->rel_res;
  if (per_cell_scratch && cells->ncells > PFX_NOALLOC_MAX_CELLS) {
    fprintf(stderr,
            "pfx_react_batch_noalloc: %lld cells without per-cell outputs exceed the fixed scratch "
            "(at most %d; pass num_iterations, status and rel_res for larger batches)\n",
            (long long)cells->ncells, PFX_NOALLOC_MAX_CELLS);
    return PFX_E_ARG;
  }
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->ncells = cells->ncells;
  }
  if (cells->ncells == 0) return 0;
  if (t->device < 0 || t->device >= kMaxDevices || !g_scratch_ptr[t->device]) return PFX_E_DEVICE;
  HIPCHK(hipSetDevice(t->device));
  unsigned char *base = g_scratch_ptr[t->device];
  unsigned int *queue = (unsigned int *)(base + kNoallocQueueOff);
  unsigned long long *sbuf = (unsigned long long *)(base + kNoallocStatsOff);
  hipStream_t stream = (hipStream_t)o->stream;
  pfx_kernel_params P;
  fill_params(t, cells, o, P);
  if (per_cell_scratch) {
    // per-cell slots: every slot a launch uses is written by it before the
    // statistics read it, so nothing needs clearing
    const int64_t cap = PFX_NOALLOC_MAX_CELLS;
    if (!P.num_iterations) P.num_iterations = (int32_t *)(base + kNoallocCellsOff);
    if (!P.status) P.status = (int32_t *)(base + kNoallocCellsOff + cap * 4);
    if (!P.rel_res) P.rel_res = (double *)(base + kNoallocCellsOff + cap * 8);
  }
  // the deferred list where it fits (its count is cleared by the launch)
  if (cells->ncells <= PFX_NOALLOC_MAX_CELLS) P.defer = (unsigned int *)(base + kNoallocDeferOff);
  double ms = 0.0;
  if ((rc = launch(t, P, PFX_MODE_REACT, stream, &ms, queue))) return rc;
  if ((rc = stats_launch(t, P, cells->ncells, stream, sbuf))) return rc;
  return stats_read(t, P, cells->ncells, stream, ms, stats, sbuf, queue);
}

extern "C" int pfx_residual_jacobian(pfx_tables *t, pfx_cells *cells, const pfx_react_opts *o, double *residual,
                                     double *jacobian) {
  if (!t || !residual || !jacobian) return PFX_E_ARG;
  int rc = check_cells(t, cells, o);
  if (rc) return rc;
  if (cells->ncells == 0) return 0;
  HIPCHK(hipSetDevice(t->device));
  pfx_kernel_params P;
  fill_params(t, cells, o, P);
  P.res_out = residual;
  P.jac_out = jacobian;
  P.num_iterations = nullptr;
  P.status = nullptr;
  P.rel_res = nullptr;
  double ms = 0.0;
  return launch(t, P, PFX_MODE_EVAL, (hipStream_t)o->stream, &ms);
}

extern "C" int pfx_equilibrate_batch(pfx_tables *t, pfx_cells *cells, const int32_t *constraint_type,
                                     const double *constraint_conc, const pfx_constraint_rxn *constraint_rxn,
                                     int32_t molal, const pfx_react_opts *o, pfx_stats *stats) {
  if (!t || !cells || !o || !constraint_type) return PFX_E_ARG;
  if (cells->ncells < 0 || cells->ncells > (int64_t)0x7fffffff) return PFX_E_ARG;
  // an empty batch needs no device arrays
  if (cells->ncells > 0 && (!constraint_conc || !cells->pri_molal || !cells->total)) return PFX_E_ARG;
  if (cells->ncells > 0 && t->L.sorb_any && !cells->total_sorb) return PFX_E_ARG;
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->ncells = cells->ncells;
  }
  int nph = 0, nrxn = 0;
  for (int j = 0; j < t->L.N; j++) {
    const int c = constraint_type[j];
    if (c < PFX_CONSTRAINT_TOTAL || c > PFX_CONSTRAINT_GAS) return PFX_E_UNSUP;
    if (c == PFX_CONSTRAINT_TOTAL_SORB && !t->L.sorb_any) return PFX_E_ARG;
    nph += c == PFX_CONSTRAINT_PH;
    if (c == PFX_CONSTRAINT_MINERAL || c == PFX_CONSTRAINT_GAS) {
      if (!constraint_rxn) return PFX_E_ARG;
      const pfx_constraint_rxn &x = constraint_rxn[j];
      if (x.nspec < 1 || x.nspec > PFX_MAX_RXN_SPEC) return PFX_E_ARG;
      for (int s = 0; s < x.nspec; s++)
        if (x.specid[s] < 0 || x.specid[s] >= t->L.N) return PFX_E_ARG;
      nrxn++;
    }
  }
  if (nph > 1) return PFX_E_ARG;
  if (cells->ncells == 0) return 0;
  HIPCHK(hipSetDevice(t->device));
  hipStream_t stream = (hipStream_t)o->stream;
  pfx_kernel_params P;
  fill_params(t, cells, o, P);
  P.constraint_conc = constraint_conc;
  for (int j = 0; j < t->L.N; j++) P.ctype[j] = constraint_type[j];
  P.has_charge = 0;
  for (int j = 0; j < t->L.N; j++) P.has_charge |= constraint_type[j] == PFX_CONSTRAINT_CHARGE;
  P.has_sorb = 0;
  for (int j = 0; j < t->L.N; j++) P.has_sorb |= constraint_type[j] == PFX_CONSTRAINT_TOTAL_SORB;
  P.crxn = nullptr;
  if (nrxn > 0) {
    if (!t->d_crxn) HIPCHK(hipMalloc((void **)&t->d_crxn, PFX_MAX_PRIMARY * sizeof(pfx_constraint_rxn)));
    HIPCHK(hipMemcpyAsync(t->d_crxn, constraint_rxn, t->L.N * sizeof(pfx_constraint_rxn), hipMemcpyHostToDevice,
                          stream));
    P.crxn = t->d_crxn;
  }
  P.molal = molal;
  if (P.max_iterations <= 0) P.max_iterations = 10000;
  int rc = use_scratch(t, P, cells->ncells);
  if (rc) return rc;
  double ms = 0.0;
  if ((rc = launch(t, P, PFX_MODE_EQUIL, stream, &ms))) return rc;
  return finish_stats(t, P, cells->ncells, stream, ms, stats);
}

extern "C" int pfx_auxvar_batch(pfx_tables *t, pfx_cells *cells, int32_t activity_passes, const pfx_react_opts *o) {
  if (!t || !cells || !o || activity_passes < 0) return PFX_E_ARG;
  if (cells->ncells < 0 || cells->ncells > (int64_t)0x7fffffff) return PFX_E_ARG;
  if (cells->ncells == 0) return 0;  // an empty batch needs no device arrays
  if (!cells->pri_molal || !cells->total) return PFX_E_ARG;
  if (t->L.sorb_any && !cells->total_sorb) return PFX_E_ARG;
  HIPCHK(hipSetDevice(t->device));
  pfx_kernel_params P;
  fill_params(t, cells, o, P);
  P.aux_passes = activity_passes;
  P.num_iterations = nullptr;
  P.status = nullptr;
  P.rel_res = nullptr;
  double ms = 0.0;
  return launch(t, P, PFX_MODE_AUX, (hipStream_t)o->stream, &ms);
}

extern "C" int pfx_update_kinetic_state(pfx_tables *t, pfx_cells *cells, const pfx_react_opts *o) {
  if (!t || !cells || !o || cells->ncells < 0) return PFX_E_ARG;
  if (!(o->dt > 0.0)) return PFX_E_ARG;
  if (cells->ncells == 0 || t->L.K == 0) return 0;
  if (!cells->mnrl_rate || !cells->mnrl_volfrac) return PFX_E_ARG;
  HIPCHK(hipSetDevice(t->device));
  hipStream_t stream = (hipStream_t)o->stream;
  const int64_t n = cells->ncells * (int64_t)t->L.K;
  hipLaunchKernelGGL(k_kinetic_state, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, t->d_blob,
                     t->L.o_mmolvol, t->L.K, cells->ncells, cells->mnrl_rate, (double *)cells->mnrl_volfrac, o->dt);
  HIPCHK(hipGetLastError());
  if (!o->stream_ordered) HIPCHK(hipStreamSynchronize(stream));
  return 0;
}

namespace pfx_internal {
// the per-call buffers a reaction step of these cells needs, allocated
// ahead of a graph capture (no allocation happens inside it)
int tables_reserve(pfx_tables *t, const pfx_cells *cells) {
  HIPCHK(hipSetDevice(t->device));
  int rc = ensure_defer(t, cells->ncells);
  if (rc) return rc;
  if (!cells->num_iterations || !cells->status || !cells->rel_res) rc = ensure_scratch(t, cells->ncells);
  return rc;
}
// a captured stream-ordered pfx_react_batch of n cells: its statistics are
// readable by pfx_tables_stats after each replay
void tables_rearm_stats(pfx_tables *t, int64_t n, hipStream_t stream, bool fallback) {
  t->pending_n = n;
  t->pending_stream = stream;
  t->pending_fallback = fallback;
}
}  // namespace pfx_internal

extern "C" int pfx_tables_timing(pfx_tables *t, double *ms) {
  if (!t || !ms) return PFX_E_ARG;
  ms[0] = t->last_ms[0];
  ms[1] = t->last_ms[1];
  return 0;
}

extern "C" int pfx_react_zones(int32_t nzones, pfx_tables *const *tables, pfx_cells *cells, const pfx_react_opts *o,
                               pfx_zone_plan *plan, pfx_stats *stats) {
  zone_setup zs;
  int rc = zones_prepare(nzones, tables, cells, o, plan, zs);
  if (rc) return rc;
  if (stats)
    for (int z = 0; z < nzones; z++) {
      memset(&stats[z], 0, sizeof(pfx_stats));
      stats[z].ncells = cells[z].ncells;
    }
  pfx_tables *lead = tables[0];
  if ((rc = ensure_zone_buffers(lead, zs.total_blocks))) return rc;
  hipStream_t s = (hipStream_t)o->stream;
  HIPCHK(hipMemcpy(lead->d_zone_params, zs.P, nzones * sizeof(pfx_kernel_params), hipMemcpyHostToDevice));
  if ((rc = zones_enqueue(tables, zs, lead->d_zone_params, lead->d_finish, s, lead->ev0, lead->ev1))) return rc;
  HIPCHK(hipEventSynchronize(lead->ev1));
  float f = 0.f;
  HIPCHK(hipEventElapsedTime(&f, lead->ev0, lead->ev1));
  lead->last_ms[0] = f;
  lead->last_ms[1] = 0.0;
  return zones_finish(tables, zs, lead->d_finish, s, f, plan, stats);
}

extern "C" int pfx_zone_plan_rebalance(pfx_zone_plan *plan) {
  if (!plan || plan->nzones < 1 || plan->nzones > PFX_MAX_ZONES) return PFX_E_ARG;
  double work[PFX_MAX_ZONES], sum = 0.0;
  int total = 0;
  for (int z = 0; z < plan->nzones; z++) {
    if (plan->blocks[z] < 1) return PFX_E_ARG;
    work[z] = plan->blocks[z] * std::max(plan->zone_ms[z], 1e-6);
    sum += work[z];
    total += plan->blocks[z];
  }
  int acc = 0;
  for (int z = 0; z < plan->nzones; z++) {
    int b = (int)(total * work[z] / sum + 0.5);
    plan->blocks[z] = std::max(1, b);
    acc += plan->blocks[z];
  }
  // keep the total: adjust the largest share
  int zmax = 0;
  for (int z = 1; z < plan->nzones; z++)
    if (plan->blocks[z] > plan->blocks[zmax]) zmax = z;
  plan->blocks[zmax] = std::max(1, plan->blocks[zmax] + total - acc);
  return 0;
}

struct pfx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nzones = 0;
  pfx_tables *tables[PFX_MAX_ZONES];
  zone_setup zs;
  pfx_kernel_params *d_params = nullptr;
  uint64_t *d_finish = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipStream_t last = nullptr;
};

extern "C" int pfx_graph_destroy(pfx_graph *g) {
  if (!g) return PFX_E_ARG;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->d_params) (void)hipFree(g->d_params);
  if (g->d_finish) (void)hipFree(g->d_finish);
  if (g->ev0) (void)hipEventDestroy(g->ev0);
  if (g->ev1) (void)hipEventDestroy(g->ev1);
  delete g;
  return 0;
}

extern "C" int pfx_zones_graph_create(int32_t nzones, pfx_tables *const *tables, pfx_cells *cells,
                                      const pfx_react_opts *o, const pfx_zone_plan *plan, pfx_graph **out) {
  if (!out) return PFX_E_ARG;
  *out = nullptr;
  pfx_graph *g = new pfx_graph();
  int rc = zones_prepare(nzones, tables, cells, o, plan, g->zs);
  if (rc) {
    delete g;
    return rc;
  }
  g->nzones = nzones;
  for (int z = 0; z < nzones; z++) g->tables[z] = tables[z];
  hipStream_t cs = nullptr;
  bool ok = hipMalloc(&g->d_params, nzones * sizeof(pfx_kernel_params)) == hipSuccess &&
            hipMalloc(&g->d_finish, (g->zs.total_blocks + 1) * sizeof(uint64_t)) == hipSuccess &&
            hipEventCreate(&g->ev0) == hipSuccess && hipEventCreate(&g->ev1) == hipSuccess &&
            hipMemcpy(g->d_params, g->zs.P, nzones * sizeof(pfx_kernel_params), hipMemcpyHostToDevice) ==
                hipSuccess &&
            hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess &&
            hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) == hipSuccess;
  if (ok) {
    rc = zones_enqueue(tables, g->zs, g->d_params, g->d_finish, cs, nullptr, nullptr);
    hipGraph_t graph = nullptr;
    const bool ended = hipStreamEndCapture(cs, &graph) == hipSuccess;
    g->graph = graph;
    ok = rc == 0 && ended && graph && hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) == hipSuccess;
  }
  if (cs) (void)hipStreamDestroy(cs);
  if (!ok) {
    pfx_graph_destroy(g);
    return rc ? rc : PFX_E_DEVICE;
  }
  *out = g;
  return 0;
}

extern "C" int pfx_graph_launch(pfx_graph *g, void *stream) {
  if (!g || !g->exec) return PFX_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(hipEventRecord(g->ev0, s));
  HIPCHK(hipGraphLaunch(g->exec, s));
  HIPCHK(hipEventRecord(g->ev1, s));
  g->last = s;
  return 0;
}

extern "C" int pfx_graph_stats(pfx_graph *g, pfx_stats *stats, pfx_zone_plan *plan) {
  if (!g) return PFX_E_ARG;
  HIPCHK(hipEventSynchronize(g->ev1));
  float f = 0.f;
  HIPCHK(hipEventElapsedTime(&f, g->ev0, g->ev1));
  return zones_finish(g->tables, g->zs, g->d_finish, g->last, f, plan, stats);
}

extern "C" int pfx_tables_tune(pfx_tables *t, int32_t mode, int32_t variant) {
  if (!t || mode < PFX_MODE_REACT || mode > PFX_MODE_AUX) return PFX_E_ARG;
  if (variant < 0) return t->variant[mode];
  if (variant > 2 || ((mode == PFX_MODE_EVAL || mode == PFX_MODE_AUX) && variant != 0)) return PFX_E_ARG;
  t->variant[mode] = variant;
  return 0;
}
