// Global-implicit assembly on a structured grid (include/pfx_chem.h,
// pfx_gi_structure / pfx_gi_residual_jacobian).
//
// Two launches per evaluation:
//  1. the persistent per-cell kernel in EVAL mode writes, per cell, the
//     non-flux residual (accumulation/dt - old accumulation/dt + reaction),
//     the non-flux Jacobian block, the aqueous totals and dT/dm (column-
//     major blocks, columns already scaled by m for log unknowns);
//  2. k_gi_assemble, one wavefront per block row, adds the upwind flux
//     terms of the row's faces (TFluxCoef, TFlux, TFluxDerivative) and
//     writes the row's residual and its 7 (or fewer) BSR blocks.  With a
//     species-independent dispersion every flux block is a scalar times the
//     neighbour's dT/dm block, so the row is a handful of scaled block copies:
//     the kernel is HBM-bound (≈7 * N^2 * 8 B written per cell).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "pfx_device.h"
#include "../../include/pfx_gi_solver.h"

namespace {

#define GCHK(x)                                                               \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "pfx_gi: %s failed: %s (%s:%d)\n", #x,                  \
              hipGetErrorString(e_), __FILE__, __LINE__);                     \
      return PFX_E_DEVICE;                                                    \
    }                                                                         \
  } while (0)

struct gi_params {
  int32_t nx, ny, nz, N;
  double area[3];
  const double *q[3], *dd[3];
  const double *bc_q[2], *bc_dd[2], *bc_total[2];
  const int32_t *row_ptr;
  const double *rnf, *jnf, *tot, *dtot;
  double *res, *vals;
};

// TFluxCoef (transport.F90:756-819), upwinding, one liquid phase: L/s
__device__ __forceinline__ void flux_coef(double area, double q, double dd, double &Tup, double &Tdn) {
  double cup, cdn;
  if (q > 0.0) {
    cup = dd + q;
    cdn = -dd;
  } else {
    cup = dd;
    cdn = -dd + q;
  }
  Tup = cup * area * 1000.0;
  Tdn = cdn * area * 1000.0;
}

__global__ __launch_bounds__(256) void k_gi_assemble(const gi_params P) {
  const int lane = threadIdx.x & 63;
  const int64_t nx = P.nx, ny = P.ny, nz = P.nz, n = nx * ny * nz;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= n) return;  // whole wavefront
  const int N = P.N, NN = N * N;
  const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
  const bool comp = lane < N;
  const double tc = comp ? P.tot[c * N + lane] : 0.0;
  double res = comp ? P.rnf[c * N + lane] : 0.0;
  double s_diag = 0.0;  // scalar of dT_c in the diagonal block

  // neighbours in column order: z-, y-, x-, (self), x+, y+, z+
  int64_t nb[6];
  double tnb[6];
  bool has[6];
  const int64_t stride[3] = {1, nx, nx * ny};
  const int64_t idx[3] = {i, j, k};
  const int64_t ext[3] = {nx, ny, nz};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int64_t ex = nx - (d == 0), ey = ny - (d == 1);
    // lower face: (c - stride) is up, c is dn
    const int lo = 2 - d;  // slot: z- = 0, y- = 1, x- = 2
    has[lo] = idx[d] > 0;
    if (has[lo]) {
      const int64_t up = c - stride[d];
      const int64_t ui = up % nx, uj = (up / nx) % ny, uk = up / (nx * ny);
      const int64_t f = ui + ex * (uj + ey * uk);
      double Tup, Tdn;
      flux_coef(P.area[d], P.q[d][f], P.dd[d][f], Tup, Tdn);
      const double tn = comp ? P.tot[up * N + lane] : 0.0;
      res -= Tup * tn + Tdn * tc;
      s_diag -= Tdn;
      nb[lo] = up;
      tnb[lo] = -Tup;
    }
    // upper face: c is up, (c + stride) is dn
    const int hi = 3 + d;  // x+ = 3, y+ = 4, z+ = 5
    has[hi] = idx[d] < ext[d] - 1;
    if (has[hi]) {
      const int64_t dn = c + stride[d];
      const int64_t f = i + ex * (j + ey * k);
      double Tup, Tdn;
      flux_coef(P.area[d], P.q[d][f], P.dd[d][f], Tup, Tdn);
      const double tn = comp ? P.tot[dn * N + lane] : 0.0;
      res += Tup * tc + Tdn * tn;
      s_diag += Tup;
      nb[hi] = dn;
      tnb[hi] = Tdn;
    }
  }
  // boundary connections on the x faces (boundary = up, cell = dn)
#pragma unroll
  for (int side = 0; side < 2; side++) {
    if (P.bc_q[side] && i == (side == 0 ? 0 : nx - 1)) {
      const int64_t f = j + ny * k;
      double Tup, Tdn;
      flux_coef(P.area[0], P.bc_q[side][f], P.bc_dd[side][f], Tup, Tdn);
      const double tb = P.bc_total[side] ? (comp ? P.bc_total[side][lane] : 0.0) : tc;
      res -= Tup * tb + Tdn * tc;
      s_diag -= Tdn;
    }
  }
  if (comp) P.res[c * N + lane] = res;

  // blocks of the row, in column order.  All source words are loaded
  // before the first store (up to 8 x CH loads in flight per lane); the
  // row's blocks are written once and not read back here, so they bypass
  // the caches that hold the neighbours' dT/dm blocks.
  int64_t b = P.row_ptr[c];
  const int order[7] = {0, 1, 2, -1, 3, 4, 5};
  constexpr int CH = 4;  // 64-wide chunks of an N x N block, N <= 16
  if (NN <= CH * 64) {
    double v[7][CH], jn[CH];
#pragma unroll
    for (int s = 0; s < 7; s++) {
      const int o = order[s];
      const bool on = o < 0 || has[o];
      const double *src = P.dtot + (o < 0 ? c : (on ? nb[o] : c)) * NN;
#pragma unroll
      for (int q = 0; q < CH; q++) {
        const int e = lane + 64 * q;
        v[s][q] = (on && e < NN) ? src[e] : 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < CH; q++) {
      const int e = lane + 64 * q;
      jn[q] = e < NN ? __builtin_nontemporal_load(P.jnf + c * NN + e) : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 7; s++) {
      const int o = order[s];
      if (o < 0 || has[o]) {
        double *out = P.vals + b * NN;
#pragma unroll
        for (int q = 0; q < CH; q++) {
          const int e = lane + 64 * q;
          if (e < NN) __builtin_nontemporal_store(o < 0 ? jn[q] + s_diag * v[s][q] : tnb[o] * v[s][q], out + e);
        }
        b++;
      }
    }
    return;
  }
#pragma unroll
  for (int s = 0; s < 7; s++) {
    const int o = order[s];
    if (o < 0) {
      const double *Jn = P.jnf + c * NN;
      const double *Dc = P.dtot + c * NN;
      double *out = P.vals + b * NN;
      for (int e = lane; e < NN; e += 64) out[e] = Jn[e] + s_diag * Dc[e];
      b++;
    } else if (has[o]) {
      const double *Dn = P.dtot + nb[o] * NN;
      double *out = P.vals + b * NN;
      const double t = tnb[o];
      for (int e = lane; e < NN; e += 64) out[e] = t * Dn[e];
      b++;
    }
  }
}

struct disp_params {
  int32_t nx, ny, nz;
  const double *q[3], *bc_q[2];
  const double *phi, *sat, *tau;
  double alpha, dm, dx[3];
  int32_t bc_type[2];
  double *dd[3], *bc_dd[2];
};

__device__ __forceinline__ double hydro_disp(const disp_params &P, int64_t c, double q) {
  const double s = P.sat ? P.sat[c] : 1.0;
  const double t = P.tau ? P.tau[c] : 1.0;
  return fmax(P.alpha * fabs(q) + s * P.phi[c] * t * P.dm, 1e-40);
}

// one thread per face: x faces, then y faces, then z faces, then the
// x-min / x-max boundary faces
__global__ void k_gi_dispersion(const disp_params P) {
  const int64_t nx = P.nx, ny = P.ny, nz = P.nz;
  const int64_t nfx = (nx - 1) * ny * nz, nfy = nx * (ny - 1) * nz, nfz = nx * ny * (nz - 1), nb = ny * nz;
  int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int d = 0;
  if (f >= nfx) {
    f -= nfx;
    d = 1;
    if (f >= nfy) {
      f -= nfy;
      d = 2;
      if (f >= nfz) {
        f -= nfz;
        d = 3;
        if (f >= 2 * nb) return;
      }
    }
  }
  if (d < 3) {
    const int64_t ex = nx - (d == 0), ey = ny - (d == 1);
    const int64_t i = f % ex, j = (f / ex) % ey, k = f / (ex * ey);
    const int64_t up = i + nx * (j + ny * k);
    const int64_t dn = up + (d == 0 ? 1 : (d == 1 ? nx : nx * ny));
    const double q = P.q[d][f];
    const double Du = hydro_disp(P, up, q), Dd = hydro_disp(P, dn, q);
    const double half = 0.5 * P.dx[d];
    const double su = P.sat ? P.sat[up] : 1.0, sd = P.sat ? P.sat[dn] : 1.0;
    P.dd[d][f] = (su < 1e-40 || sd < 1e-40) ? 0.0 : (Du * Dd) / (Du * half + Dd * half);
  } else {
    const int side = (int)(f / nb);
    const int64_t jk = f % nb;
    if (!P.bc_dd[side]) return;
    double v = 0.0;
    if (P.bc_type[side] == 1) {
      const int64_t c = (side == 0 ? 0 : nx - 1) + nx * jk;
      const double q = P.bc_q[side] ? P.bc_q[side][jk] : 0.0;
      v = hydro_disp(P, c, q) / (0.5 * P.dx[0]);
    }
    P.bc_dd[side][jk] = v;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
static bool grid_ok(const pfx_grid *g) {
  if (!g || g->nx < 1 || g->ny < 1 || g->nz < 1) return false;
  if ((int64_t)g->nx * g->ny * g->nz > (int64_t)0x7fffffff / 8) return false;
  for (int d = 0; d < 3; d++) {
    const int64_t ext = d == 0 ? g->nx : (d == 1 ? g->ny : g->nz);
    if (ext > 1 && (!g->q[d] || !g->dd[d])) return false;
  }
  for (int s = 0; s < 2; s++)
    if (g->bc_q[s] && !g->bc_dd[s]) return false;
  return true;
}

// ---------------------------------------------------------------------------
// The transport half of an operator-split step: the scalar transport matrix
// of RTCalculateTransportMatrix (reactive_transport.F90:1379-1601), the same
// for every component, and the right-hand sides of RTCalculateRHS_t0/_t1
// (:1085-1375).  Rows in the order of pfx_gi_structure (block size 1).

struct os_params {
  int32_t nx, ny, nz, N;
  double area[3];
  const double *q[3], *dd[3];
  const double *bc_q[2], *bc_dd[2], *bc_total[2];
  const int32_t *row_ptr;
  const double *porosity, *saturation, *volume;
  double def_porosity, def_saturation, def_volume, dt;
  double *vals;        // [nnzb] matrix
  double *rhs, *x;     // [ncells] one component
  double *total;       // [ncells][N] in: old totals; out: transported
  int32_t comp;
};

__device__ __forceinline__ double os_accum(const os_params &P, int64_t c) {
  const double phi = P.porosity ? P.porosity[c] : P.def_porosity;
  const double sat = P.saturation ? P.saturation[c] : P.def_saturation;
  const double vol = P.volume ? P.volume[c] : P.def_volume;
  return phi * sat * 1000.0 * vol / P.dt;  // reactive_transport.F90:1526-1530
}

// one thread per cell: the row's values, neighbours in column order
__global__ void k_os_matrix(const os_params P) {
  const int64_t nx = P.nx, ny = P.ny, nz = P.nz, n = nx * ny * nz;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
  const int64_t idx[3] = {i, j, k}, ext[3] = {nx, ny, nz};
  double diag = os_accum(P, c);
  double off[6];
  bool has[6];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int64_t ex = nx - (d == 0), ey = ny - (d == 1);
    const int64_t stride = d == 0 ? 1 : (d == 1 ? nx : nx * ny);
    const int lo = 2 - d, hi = 3 + d;
    has[lo] = idx[d] > 0;
    off[lo] = 0.0;
    if (has[lo]) {  // c is the downstream cell of its lower face
      const int64_t up = c - stride;
      const int64_t f = (up % nx) + ex * (((up / nx) % ny) + ey * (up / (nx * ny)));
      double Tup, Tdn;
      flux_coef(P.area[d], P.q[d][f], P.dd[d][f], Tup, Tdn);
      diag -= Tdn;
      off[lo] = -Tup;
    }
    has[hi] = idx[d] < ext[d] - 1;
    off[hi] = 0.0;
    if (has[hi]) {  // c is the upstream cell of its upper face
      const int64_t f = i + ex * (j + ey * k);
      double Tup, Tdn;
      flux_coef(P.area[d], P.q[d][f], P.dd[d][f], Tup, Tdn);
      diag += Tup;
      off[hi] = Tdn;
    }
  }
  // boundary connections on the x faces (boundary up, cell down): the
  // diagonal takes -T_dn; a zero-gradient face (no boundary total) carries
  // the cell's own total, so its T_up joins the diagonal too
#pragma unroll
  for (int side = 0; side < 2; side++) {
    if (P.bc_q[side] && i == (side == 0 ? 0 : nx - 1)) {
      const int64_t f = j + ny * k;
      double Tup, Tdn;
      flux_coef(P.area[0], P.bc_q[side][f], P.bc_dd[side][f], Tup, Tdn);
      diag -= Tdn;
      if (!P.bc_total[side]) diag -= Tup;
    }
  }
  int64_t b = P.row_ptr[c];
  const int order[7] = {0, 1, 2, -1, 3, 4, 5};
#pragma unroll
  for (int s = 0; s < 7; s++) {
    const int o = order[s];
    if (o < 0) P.vals[b++] = diag;
    else if (has[o]) P.vals[b++] = off[o];
  }
}

// right-hand side of one component: accum * T_old plus T_up * T_bc on faces
// with a boundary total; the old totals start the solve's bookkeeping only
__global__ void k_os_rhs(const os_params P) {
  const int64_t nx = P.nx, ny = P.ny, n = nx * ny * (int64_t)P.nz;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
  double r = os_accum(P, c) * P.total[c * P.N + P.comp];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    if (P.bc_q[side] && P.bc_total[side] && i == (side == 0 ? 0 : nx - 1)) {
      const int64_t f = j + ny * k;
      double Tup, Tdn;
      flux_coef(P.area[0], P.bc_q[side][f], P.bc_dd[side][f], Tup, Tdn);
      r += Tup * P.bc_total[side][P.comp];
    }
  }
  P.rhs[c] = r;
}

__global__ void k_os_scatter(const os_params P) {
  const int64_t n = (int64_t)P.nx * P.ny * P.nz;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  P.total[c * P.N + P.comp] = P.x[c];
}

// every component's right-hand side at once, cell-major [n][N] (the
// layout of the totals and of the several-right-hand-side solve)
__global__ void k_os_rhs_all(const os_params P) {
  const int64_t nx = P.nx, ny = P.ny, n = nx * ny * (int64_t)P.nz;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * P.N) return;
  const int64_t c = e / P.N;
  const int comp = (int)(e % P.N);
  const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
  double r = os_accum(P, c) * P.total[e];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    if (P.bc_q[side] && P.bc_total[side] && i == (side == 0 ? 0 : nx - 1)) {
      const int64_t f = j + ny * k;
      double Tup, Tdn;
      flux_coef(P.area[0], P.bc_q[side][f], P.bc_dd[side][f], Tup, Tdn);
      r += Tup * P.bc_total[side][comp];
    }
  }
  P.rhs[e] = r;
}

__global__ void k_os_scatter_all(const os_params P) {
  const int64_t n = (int64_t)P.nx * P.ny * P.nz;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n * P.N) P.total[e] = P.x[e];
}

extern "C" int pfx_gi_structure(const pfx_grid *g, int32_t *row_ptr, int32_t *col_idx, int64_t *nnzb) {
  if (!g || !nnzb || g->nx < 1 || g->ny < 1 || g->nz < 1) return PFX_E_ARG;
  const int64_t nx = g->nx, ny = g->ny, nz = g->nz, n = nx * ny * nz;
  // 7 blocks per interior row, one fewer per box face the cell touches
  const int64_t total = 7 * n - 2 * (ny * nz + nx * nz + nx * ny);
  *nnzb = total;
  if (!row_ptr) return 0;
  if (!col_idx || total > (int64_t)0x7fffffff) return PFX_E_ARG;
  std::vector<int32_t> rp(n + 1), ci(total);
  int64_t p = 0;
  for (int64_t c = 0; c < n; c++) {
    const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    rp[c] = (int32_t)p;
    if (k > 0) ci[p++] = (int32_t)(c - nx * ny);
    if (j > 0) ci[p++] = (int32_t)(c - nx);
    if (i > 0) ci[p++] = (int32_t)(c - 1);
    ci[p++] = (int32_t)c;
    if (i < nx - 1) ci[p++] = (int32_t)(c + 1);
    if (j < ny - 1) ci[p++] = (int32_t)(c + nx);
    if (k < nz - 1) ci[p++] = (int32_t)(c + nx * ny);
  }
  rp[n] = (int32_t)p;
  if (p != total) return PFX_E_DEVICE;
  GCHK(hipMemcpy(row_ptr, rp.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  GCHK(hipMemcpy(col_idx, ci.data(), total * sizeof(int32_t), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int pfx_gi_residual_jacobian(pfx_tables *t, pfx_cells *cells, const pfx_grid *g, const pfx_react_opts *o,
                                        const int32_t *row_ptr, double *residual, double *bsr_values) {
  if (!t || !cells || !o || !row_ptr || !residual || !bsr_values || !grid_ok(g)) return PFX_E_ARG;
  const int64_t n = (int64_t)g->nx * g->ny * g->nz;
  if (cells->ncells != n) return PFX_E_ARG;
  int N, M, K, R;
  pfx_internal::table_sizes(t, &N, &M, &K, &R);
  double *rnf, *jnf, *tot, *dtot;
  int rc = pfx_internal::gi_scratch(t, n, &rnf, &jnf, &tot, &dtot);
  if (rc) return rc;
  rc = pfx_internal::gi_evaluate_cells(t, cells, o, g->log_formulation, rnf, jnf, tot, dtot);
  if (rc) return rc;
  gi_params P;
  memset(&P, 0, sizeof(P));
  P.nx = g->nx;
  P.ny = g->ny;
  P.nz = g->nz;
  P.N = N;
  for (int d = 0; d < 3; d++) {
    P.area[d] = g->area[d];
    P.q[d] = g->q[d];
    P.dd[d] = g->dd[d];
  }
  for (int s = 0; s < 2; s++) {
    P.bc_q[s] = g->bc_q[s];
    P.bc_dd[s] = g->bc_dd[s];
    P.bc_total[s] = g->bc_total[s];
  }
  P.row_ptr = row_ptr;
  P.rnf = rnf;
  P.jnf = jnf;
  P.tot = tot;
  P.dtot = dtot;
  P.res = residual;
  P.vals = bsr_values;
  hipStream_t s = (hipStream_t)o->stream;
  if ((rc = pfx_internal::timed_launch_begin(t, s))) return rc;
  hipLaunchKernelGGL(k_gi_assemble, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, P);
  GCHK(hipGetLastError());
  return pfx_internal::timed_launch_end(t, s, 1);
}

extern "C" int pfx_gi_dispersion(const pfx_grid *g, const double *porosity, const double *saturation,
                                 const double *tortuosity, double dispersivity, double diffusion, const double *dx,
                                 const int32_t *bc_type, double **dd_out, double **bc_dd_out, void *stream) {
  if (!g || !porosity || !dx || !dd_out || g->nx < 1 || g->ny < 1 || g->nz < 1) return PFX_E_ARG;
  disp_params P;
  memset(&P, 0, sizeof(P));
  P.nx = g->nx;
  P.ny = g->ny;
  P.nz = g->nz;
  const int64_t ext[3] = {g->nx, g->ny, g->nz};
  for (int d = 0; d < 3; d++) {
    P.q[d] = g->q[d];
    P.dd[d] = dd_out[d];
    P.dx[d] = dx[d];
    if (ext[d] > 1 && (!P.q[d] || !P.dd[d])) return PFX_E_ARG;
    if (!(dx[d] > 0.0)) return PFX_E_ARG;
  }
  for (int s = 0; s < 2; s++) {
    P.bc_q[s] = g->bc_q[s];
    P.bc_dd[s] = bc_dd_out ? bc_dd_out[s] : nullptr;
    P.bc_type[s] = bc_type ? bc_type[s] : 0;
  }
  P.phi = porosity;
  P.sat = saturation;
  P.tau = tortuosity;
  P.alpha = dispersivity;
  P.dm = diffusion;
  const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
  const int64_t nf = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1) + 2 * ny * nz;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gi_dispersion, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, P);
  GCHK(hipGetLastError());
  GCHK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int pfx_os_transport_step(pfx_bsr *h, const pfx_grid *g, pfx_cells *cells, int32_t ncomp,
                                     const pfx_react_opts *o, double *work, const pfx_krylov_opts *kopts,
                                     pfx_krylov_stats *stats) {
  if (!h || !cells || !o || !work || !cells->total || ncomp < 1 || !grid_ok(g) || !(o->dt > 0.0))
    return PFX_E_ARG;
  const int64_t n = (int64_t)g->nx * g->ny * g->nz;
  if (cells->ncells != n) return PFX_E_ARG;
  for (int side = 0; side < 2; side++)
    if (g->bc_q[side] && !g->bc_dd[side]) return PFX_E_ARG;
  int64_t nnzb = 0;
  int rc = pfx_gi_structure(g, nullptr, nullptr, &nnzb);
  if (rc) return rc;
  os_params P;
  memset(&P, 0, sizeof(P));
  P.nx = g->nx;
  P.ny = g->ny;
  P.nz = g->nz;
  P.N = ncomp;
  for (int d = 0; d < 3; d++) {
    P.area[d] = g->area[d];
    P.q[d] = g->q[d];
    P.dd[d] = g->dd[d];
  }
  for (int s = 0; s < 2; s++) {
    P.bc_q[s] = g->bc_q[s];
    P.bc_dd[s] = g->bc_dd[s];
    P.bc_total[s] = g->bc_total[s];
  }
  P.row_ptr = nullptr;
  P.porosity = cells->porosity;
  P.saturation = cells->saturation;
  P.volume = cells->volume;
  P.def_porosity = o->porosity;
  P.def_saturation = o->saturation;
  P.def_volume = o->volume;
  P.dt = o->dt;
  P.vals = work;
  P.rhs = work + nnzb;
  P.x = work + nnzb + n;
  P.total = cells->total;
  P.row_ptr = pfx_internal::bsr_row_ptr(h);
  // the handle must hold this grid's pattern at block size 1: the matrix
  // kernel writes nnzb values through its row pointers
  if (!P.row_ptr || pfx_internal::bsr_rows(h) != n || pfx_internal::bsr_block(h) != 1 ||
      pfx_internal::bsr_nnzb(h) != nnzb)
    return PFX_E_ARG;
  hipStream_t s = (hipStream_t)o->stream;
  const unsigned gb = (unsigned)((n + 255) / 256);
  // every component in one several-right-hand-side solve, cell-major as
  // the totals are (one pfx_bsr_solve per component has the same scalars
  // and tests per component, in ncomp times the launches)
  double *B = nullptr, *X = nullptr;
  if ((rc = pfx_internal::bsr_multi_io(h, ncomp, &B, &X))) return rc;
  P.rhs = B;
  P.x = X;
  const unsigned ge = (unsigned)((n * ncomp + 255) / 256);
  hipLaunchKernelGGL(k_os_matrix, dim3(gb), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_os_rhs_all, dim3(ge), dim3(256), 0, s, P);
  GCHK(hipGetLastError());
  pfx_krylov_stats acc;
  memset(&acc, 0, sizeof(acc));
  rc = pfx_bsr_solve_multi(h, P.vals, B, X, ncomp, kopts, o->stream, o->stream_ordered, &acc);
  if (rc) return rc;
  hipLaunchKernelGGL(k_os_scatter_all, dim3(ge), dim3(256), 0, s, P);
  GCHK(hipGetLastError());
  if (o->stream_ordered) return 0;   // the outcome: pfx_bsr_multi_stats
  GCHK(hipStreamSynchronize(s));
  if (stats) *stats = acc;
  return acc.reason < 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// one operator-split step as a graph
struct pfx_os_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  pfx_tables *tables = nullptr;
  pfx_bsr *h = nullptr;
  int64_t ncells = 0;
  bool fallback = false;
  int device = 0;
  hipStream_t last = nullptr;
};

extern "C" int pfx_os_graph_destroy(pfx_os_graph *g) {
  if (!g) return 0;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return 0;
}

extern "C" int pfx_os_graph_create(pfx_bsr *h, const pfx_grid *grid, pfx_tables *tables, pfx_cells *cells,
                                   int32_t ncomp, const pfx_react_opts *topts, const pfx_react_opts *ropts,
                                   double *work, const pfx_krylov_opts *kopts, pfx_os_graph **out) {
  if (!out || !h || !grid || !tables || !cells || !topts || !ropts || !work || !kopts) return PFX_E_ARG;
  *out = nullptr;
  if (kopts->max_iterations < 1 || kopts->max_iterations > PFX_KSP_MAX_ENQUEUED) return PFX_E_ARG;
  pfx_os_graph *g = new pfx_os_graph();
  GCHK(hipGetDevice(&g->device));
  g->tables = tables;
  g->h = h;
  g->ncells = cells->ncells;
  g->fallback = ropts->guess_mode == PFX_GUESS_ARRAY_FALLBACK;
  pfx_react_opts to = *topts, ro = *ropts;
  to.stream_ordered = 1;
  ro.stream_ordered = 1;
  // every buffer the three calls use is allocated before the capture
  double *B = nullptr, *X = nullptr;
  int rc0 = pfx_internal::bsr_multi_io(h, ncomp, &B, &X);
  if (rc0 == 0) rc0 = pfx_internal::tables_reserve(tables, cells);
  if (rc0) {
    delete g;
    return rc0;
  }
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    delete g;
    return PFX_E_DEVICE;
  }
  to.stream = (void *)cs;
  ro.stream = (void *)cs;
  int rc = 0;
  bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) == hipSuccess;
  if (ok) {
    rc = pfx_os_transport_step(h, grid, cells, ncomp, &to, work, kopts, nullptr);
    if (rc == 0) rc = pfx_react_batch(tables, cells, &ro, nullptr);
    if (rc == 0) rc = pfx_update_kinetic_state(tables, cells, &ro);
    hipGraph_t graph = nullptr;
    const bool ended = hipStreamEndCapture(cs, &graph) == hipSuccess;
    g->graph = graph;
    ok = rc == 0 && ended && graph && hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) == hipSuccess;
  }
  (void)hipStreamDestroy(cs);
  if (!ok) {
    (void)hipGetLastError();
    pfx_os_graph_destroy(g);
    return rc ? rc : PFX_E_DEVICE;
  }
  *out = g;
  return 0;
}

extern "C" int pfx_os_graph_launch(pfx_os_graph *g, void *stream) {
  if (!g || !g->exec) return PFX_E_ARG;
  GCHK(hipSetDevice(g->device));
  GCHK(hipGraphLaunch(g->exec, (hipStream_t)stream));
  g->last = (hipStream_t)stream;
  return 0;
}

extern "C" int pfx_os_graph_stats(pfx_os_graph *g, int32_t *rc, pfx_stats *rs, pfx_krylov_stats *ks) {
  if (!g || !g->exec) return PFX_E_ARG;
  GCHK(hipSetDevice(g->device));
  GCHK(hipStreamSynchronize(g->last));
  if (ks) {
    int r = pfx_bsr_multi_stats(g->h, (void *)g->last, ks);
    if (r) return r;
  }
  pfx_internal::tables_rearm_stats(g->tables, g->ncells, g->last, g->fallback);
  pfx_stats tmp;
  const int r = pfx_tables_stats(g->tables, rs ? rs : &tmp);
  if (r < 0) return r;
  if (rc) *rc = r;
  return 0;
}
