// Batched per-cell reaction step for MI355X (gfx950), FP64.
//
// One cell per group of G lanes (G = 8, 16 or 32 >= N primaries); lane r
// owns primary species r: its free ion, its total, row r of the local
// Jacobian and row r of the LU factors.  Reaction tables are staged in LDS
// once per workgroup; per-cell complex data (concentrations, activity
// coefficients, log K) live in a per-group LDS slice.  A persistent grid
// pulls cells from a device work queue in chunks, and a group whose cell
// converged refills from the queue while the other groups of its wavefront
// keep iterating, so cells with 2 or 25 Newton iterations share one launch.
//
// Reference path (src/pflotran/):
//   RReact                reaction.F90:3322-3511   (Newton loop, log update)
//   RTAuxVarCompute       reaction.F90:4969-5006
//   RTotal                reaction.F90:4057-4158   (phases A/B below)
//   RActivityCoefficients reaction.F90:3812-4053   (LAG and NEWTON)
//   RTotalSorbEqIonx      reaction.F90:4305-4535
//   RKineticMineral       reaction_mineral.F90:564-1000 (no prefactor species)
//   RSolve/ludcmp/lubksb  reaction.F90:4835-4880, utility.F90:393-523
//   RTAccumulation(Derivative), RAccumulationSorb(Derivative)
//                         reaction.F90:5072-5232, 4539-4603
// The residual is the dt-consistent per-cell residual of the global-implicit
// path (reactive_transport.F90:2539-2587, 2735-2758): (accum - accum_old)/dt
// + mineral rates, in mol/s.
#include <hip/hip_runtime.h>

#include "pfx_device.h"

namespace {

constexpr int BLOCK = 256;

// Profiling build (this file compiled with -DPFX_SECTION_TIMERS):
// shader-clock cycles per section of the Newton pass, summed per wavefront
// and added into a device array when the wavefront exits; the product build
// compiles every marker to nothing.  Sections: 0 refill/load, 1 activity,
// 2 complexes, 3 totals and dT/dm, 4 ion exchange, 5 residual, minerals and
// convergence, 6 constraint rows, 7 LU solve, 8 update and retire.
#ifdef PFX_SECTION_TIMERS
__device__ unsigned long long g_section_cycles[16];
#define PFX_PT(i)                                  \
  {                                                \
    const unsigned long long t_ = clock64();       \
    pt_acc[i] += t_ - pt_last;                     \
    pt_last = t_;                                  \
  }
#else
#define PFX_PT(i)
#endif

__device__ __forceinline__ double bperm_d(double v, int src_lane) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_ds_bpermute(src_lane << 2, lo);
  hi = __builtin_amdgcn_ds_bpermute(src_lane << 2, hi);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int bperm_i(int v, int src_lane) {
  return __builtin_amdgcn_ds_bpermute(src_lane << 2, v);
}

// Group-wide reductions.  Within a DPP row (16 lanes) the exchanges are
// DPP moves (VALU, a few cycles each) instead of LDS-crossbar permutes:
//   G = 16: row rotations by 8, 4, 2, 1;
//   G = 8 : half-row mirror, then quad permutations xor 2 and xor 1;
//   G = 32: the G = 16 steps, then one permute across the two rows.
// Every step combines a commutative pair, so all lanes of the group end
// with the bitwise-same value.
// Every lane of a row has a source lane under these controls, so no lane
// keeps an old value and the moves need no initialised destination.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}

constexpr int DPP_ROR8 = 0x128, DPP_ROR4 = 0x124, DPP_ROR2 = 0x122, DPP_ROR1 = 0x121;
constexpr int DPP_HALF_MIRROR = 0x141, DPP_XOR2 = 0x4E, DPP_XOR1 = 0xB1;

struct OpSum {
  __device__ __forceinline__ double operator()(double a, double b) const { return a + b; }
};
struct OpMax {  // NaN wins: a broken cell is seen as broken
  __device__ __forceinline__ double operator()(double a, double b) const {
    return (b > a || b != b) ? b : a;
  }
};
struct OpMin {
  __device__ __forceinline__ double operator()(double a, double b) const { return fmin(a, b); }
};

template <int G, class Op>
__device__ __forceinline__ double greduce(double v, int lane, Op op) {
  if constexpr (G == 8) {
    v = op(v, dpp_d<DPP_HALF_MIRROR>(v));
    v = op(v, dpp_d<DPP_XOR2>(v));
    v = op(v, dpp_d<DPP_XOR1>(v));
  } else {
    v = op(v, dpp_d<DPP_ROR8>(v));
    v = op(v, dpp_d<DPP_ROR4>(v));
    v = op(v, dpp_d<DPP_ROR2>(v));
    v = op(v, dpp_d<DPP_ROR1>(v));
    if constexpr (G == 32) v = op(v, bperm_d(v, lane ^ 16));
  }
  return v;
}
template <int G>
__device__ __forceinline__ double gsum(double v, int lane) { return greduce<G>(v, lane, OpSum()); }
template <int G>
__device__ __forceinline__ double gmax(double v, int lane) { return greduce<G>(v, lane, OpMax()); }
template <int G>
__device__ __forceinline__ double gmin(double v, int lane) { return greduce<G>(v, lane, OpMin()); }

// argmax with ties resolved to the larger row index
__device__ __forceinline__ void amax_step(double &v, int &idx, double w, int j) {
  const bool take = (w > v) | ((w == v) & (j > idx));  // no short circuit: selects
  v = take ? w : v;
  idx = take ? j : idx;
}
template <int G>
__device__ __forceinline__ void gargmax(double &v, int &idx, int lane) {
  if constexpr (G == 8) {
    amax_step(v, idx, dpp_d<DPP_HALF_MIRROR>(v), dpp_i<DPP_HALF_MIRROR>(idx));
    amax_step(v, idx, dpp_d<DPP_XOR2>(v), dpp_i<DPP_XOR2>(idx));
    amax_step(v, idx, dpp_d<DPP_XOR1>(v), dpp_i<DPP_XOR1>(idx));
  } else {
    amax_step(v, idx, dpp_d<DPP_ROR8>(v), dpp_i<DPP_ROR8>(idx));
    amax_step(v, idx, dpp_d<DPP_ROR4>(v), dpp_i<DPP_ROR4>(idx));
    amax_step(v, idx, dpp_d<DPP_ROR2>(v), dpp_i<DPP_ROR2>(idx));
    amax_step(v, idx, dpp_d<DPP_ROR1>(v), dpp_i<DPP_ROR1>(idx));
    if constexpr (G == 32) amax_step(v, idx, bperm_d(v, lane ^ 16), bperm_i(idx, lane ^ 16));
  }
}

struct OpMaxNum {  // a NaN candidate never wins the pivot search
  __device__ __forceinline__ double operator()(double a, double b) const { return fmax(a, b); }
};

// Pivot row of an elimination step: the group's largest candidate, ties to
// the larger row (the reference's `>=` scan keeps the last maximum,
// utility.F90:451-452).  A max reduction, then the highest lane of the
// group whose candidate equals it (one ballot); a group without a finite
// candidate keeps its own row (the cell is non-finite and stops).
template <int G>
__device__ __forceinline__ int gargmax_row(double cand, int lane, int gbase) {
  const double vmax = greduce<G>(cand, lane, OpMaxNum());
  const unsigned long long msk = __ballot(cand == vmax);
  const unsigned int gm = (unsigned int)(msk >> gbase) & (G == 32 ? 0xffffffffu : ((1u << G) - 1u));
  return gm ? 31 - __clz((int)gm) : lane - gbase;
}

// compile-time loop: f(Ic<J>) for J = B .. E-1
template <int V>
struct Ic {
  static constexpr int value = V;
};
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(Ic<B>{});
    static_for<B + 1, E>(f);
  }
}

// Lane j's value of a group, to every lane of the group.  G = 16: one DPP
// row broadcast (a VALU move, no LDS); G = 8: the two groups of a DPP row
// take lane j and lane 8 + j; G = 32: a lane permute (the group spans two
// DPP rows).  J is a compile-time lane index.
template <int G, int J>
__device__ __forceinline__ double gbcast(double v, int gbase) {
  if constexpr (G == 16) {
    return dpp_d<0x150 + J>(v);
  } else if constexpr (G == 8) {
    const double a = dpp_d<0x150 + J>(v);
    const double b = dpp_d<0x150 + 8 + J>(v);
    return (gbase & 8) ? b : a;
  } else {
    return bperm_d(v, gbase + J);
  }
}

__device__ __forceinline__ void wave_sync() {
  // LDS written by other lanes of this wavefront becomes visible to the
  // whole wavefront (lanes of one group never span wavefronts).
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS accessors: every table and scratch array is addressed by a 32-bit
// byte offset into the dynamic LDS segment, so all accesses stay ds_read /
// ds_write and no 64-bit pointers are held in registers.
extern __shared__ __align__(16) unsigned char smem[];
#define LD(off, i) (((const double *)(smem + (off)))[(i)])
#define LI(off, i) (((const int32_t *)(smem + (off)))[(i)])
#define SD(off, i) (((double *)(smem + (off)))[(i)])

// 5-term log K(T) (reaction_aux.F90:1461-1488) evaluated without fused
// multiply-add, in the reference's order: the fitted coefficients cancel to
// ~1e-3 of their terms, so contraction alone would move K by ~1e-12.
// ln_tk = log(tk), computed once per cell.
__device__ __forceinline__ double logk_5term(int cfo, double tk, double ln_tk) {
#pragma clang fp contract(off)
  double v = LD(cfo, 0) * ln_tk;
  v = v + LD(cfo, 1);
  v = v + LD(cfo, 2) * tk;
  v = v + LD(cfo, 3) / tk;
  v = v + LD(cfo, 4) / (tk * tk);
  return v;
}
// the same from a coefficient array in global memory
__device__ __forceinline__ double logk_5term_ptr(const double *cf, double tk, double ln_tk) {
#pragma clang fp contract(off)
  double v = cf[0] * ln_tk;
  v = v + cf[1];
  v = v + cf[2] * tk;
  v = v + cf[3] / tk;
  v = v + cf[4] / (tk * tk);
  return v;
}

// per-group scratch slice (doubles), as byte offsets from the slice base:
//   C[M+1] complex molalities (C[M] = 0, padding slot), gsec[M] complex
//   activity coefficients, lnK[M] = -log10 K * LOG_TO_LN of this cell,
//   la[G] ln activities, mv[G] free ions, im[G] inverse free ions,
//   xb[G] ion-exchange fractions, mlnK/mkeff/marea/mvf/mrate[K] per mineral,
//   cs[4] the cell's porosity, saturation, volume and water density (cs[6]
//   in the prefactor kernels: then also its temperature (C) and a spare),
//   ionx[PFX_MAX_IONX_RXN] the exchange state, fx[G] the fixed accumulation
//   of each row.  The last three hold cell state that lives across Newton
//   passes: in LDS it does not occupy registers for the whole loop.
struct Scr {
  int C, gsec, lnK, la, mv, im, xb, mlnK, mkeff, marea, mvf, mrate, cs, ionx, fx;
};

__host__ __device__ inline int scr_doubles(int M, int G, int K, bool xm) {
  int n = (M + 1) + 2 * M + 5 * G + 5 * K + (xm ? 6 : 4) + PFX_MAX_IONX_RXN;
  return (n + 1) & ~1;
}

// The fields whose size is fixed by the group size come first, at offsets
// known at compile time: their LDS accesses take the offset as an
// immediate of the slice base, and no per-field address stays live in a
// register.  The complex and mineral arrays (sizes M, K) follow.
template <bool XM>
__device__ __forceinline__ Scr make_scr(int base, int M, int G, int K) {
  Scr s;
  s.la = base;
  s.mv = base + 8 * G;
  s.im = base + 16 * G;
  s.xb = base + 24 * G;
  s.fx = base + 32 * G;
  s.cs = base + 40 * G;
  constexpr int cs_bytes = XM ? 48 : 32;
  s.ionx = base + 40 * G + cs_bytes;
  int p = base + 40 * G + cs_bytes + 8 * PFX_MAX_IONX_RXN;
  s.C = p; p += (M + 1) * 8;
  s.gsec = p; p += M * 8;
  s.lnK = p; p += M * 8;
  s.mlnK = p; p += K * 8;
  s.mkeff = p; p += K * 8;
  s.marea = p; p += K * 8;
  s.mvf = p; p += K * 8;
  s.mrate = p;
  return s;
}

}  // namespace

// ---------------------------------------------------------------------------
// x^z for the valence ratios of ion exchange: the common ratios 1, 2, 1/2
// and 3 without the log/exp of a general power (same value to the last bit
// or within one rounding)
__device__ __forceinline__ double pow_ratio(double x, double z) {
  if (z == 1.0) return x;
  if (z == 2.0) return x * x;
  if (z == 0.5) return sqrt(x);
  if (z == 3.0) return x * x * x;
  return pow(x, z);
}

// Kernel modes (template MODE):
//   PFX_MODE_REACT  reaction step, RReact (reaction.F90:3322-3511)
//   PFX_MODE_EVAL   residual + Jacobian at given free ions (reaction.F90:3419-3441)
//   PFX_MODE_EQUIL  constraint speciation, ReactionEquilibrateConstraint
//                   (reaction.F90:1308-2046), TOTAL/FREE/LOG/PH rows
//   PFX_MODE_AUX    aux-var refresh, RTUpdateAuxVars (reactive_transport.F90:3790-3846)
//
// Phases of a cell in REACT: 0 = RTotal at the previous state, 1 = activity
// update pending, 2 = Newton.  Other modes start in phase 2.
//
// The parameters live in constant memory (the kernarg segment for a single
// batch, a device buffer for a zoned launch) and are read through a pointer
// that is made opaque once per pass of the persistent loop, so the ~100
// uniform parameters are re-read through the scalar cache where they are
// used instead of all being held in SGPRs across the whole loop.
typedef const __attribute__((address_space(4))) pfx_kernel_params *KP;

// Arrhenius factor exp(Ea/R (1/298.15 - 1/T)) of an activation energy
// (reaction_mineral.F90:775-780, 786-792); 1 where Ea <= 0
__device__ __forceinline__ double arrhenius(double Ea, double tempc) {
  return Ea > 0.0 ? exp(Ea / PFX_IDEAL_GAS * (1.0 / (25.0 + 273.15) - 1.0 / (tempc + 273.15))) : 1.0;
}

// ln numerator alpha la and ln denominator ln(1 + exp(ln atten + beta la))
// of prefactor species slot o at ln activity la (reaction_mineral.F90:758-767)
__device__ __forceinline__ void pref_species_ln(KP Pp, int o, double la, double &num, double &den) {
  num = LD(Pp->L.o_mpalpha, o) * la;
  den = log(1.0 + exp(log(LD(Pp->L.o_mpatten, o)) + LD(Pp->L.o_mpbeta, o) * la));
}

// The rate constant of kinetic mineral k: sum over its prefactors of
// prod_s (monod factor) * k_p * Arrhenius_p (reaction_mineral.F90:743-786),
// or the mineral's own constant (with its Arrhenius factor) without
// prefactors.  jpert >= 0 takes the ln activity of that primary from
// SD(S.xb, jpert) (a forward difference).  Group-uniform.  XM: the kernel
// instantiation that carries prefactors and scale factors (without it the
// rate constant is the mineral's own).
template <bool XM>
__device__ __forceinline__ double mineral_rate_constant(KP Pp, const Scr &S, int k, int jpert) {
  if (!XM || !(Pp->L.mnrl_flags & 16)) return SD(S.mkeff, k);
  const int np = LI(Pp->L.o_mpn, k);
  if (np == 0) return SD(S.mkeff, k);
  const int PP = Pp->L.npref, PQ = Pp->L.npref_spec;
  double sum = 0.0;
  for (int p = 0; p < np; p++) {
    double lnp = 0.0;
    for (int q = 0; q < PQ; q++) {
      const int o = (k * PP + p) * PQ + q;
      const int sp = LI(Pp->L.o_mpid, o);
      if (sp < 0) continue;
      double num, den;
      pref_species_ln(Pp, o, sp == jpert ? SD(S.xb, sp) : SD(S.la, sp), num, den);
      lnp = lnp + num;
      lnp = lnp - den;
    }
    sum = sum + exp(lnp) * LD(Pp->L.o_mprate, k * PP + p) * arrhenius(LD(Pp->L.o_mpEa, k * PP + p), SD(S.cs, 4));
  }
  return sum;
}

// Rate (mol/s/m^3 bulk) of kinetic mineral k at ln(Q/K), 0 where the
// reference skips the mineral (reaction_mineral.F90:690-816).  Group-uniform.
template <bool XM>
__device__ __forceinline__ double mineral_rate(KP Pp, const Scr &S, int k, double lnQK, int jpert) {
  const double QK = lnQK <= 6.90776 ? exp(lnQK) : 1.0e3;
  const bool temkin = Pp->L.mnrl_flags & 8, msf_on = XM && (Pp->L.mnrl_flags & 32);
  const double msf = msf_on ? LD(Pp->L.o_mmsf, k) : 1.0;
  const double tk_ = temkin ? (msf_on ? msf * LD(Pp->L.o_mtemkin, k) : LD(Pp->L.o_mtemkin, k)) : msf;
  double aff = (temkin || msf_on) ? 1.0 - pow(QK, 1.0 / tk_) : 1.0 - QK;
  const double sgn = copysign(1.0, aff);
  if (!(SD(S.mvf, k) > 0.0 || sgn < 0.0)) return 0.0;
  const double thr = (Pp->L.mnrl_flags & 1) ? LD(Pp->L.o_mthr, k) : 0.0;
  if (thr > 0.0 && sgn < 0.0 && QK < thr) return 0.0;
  const double lim = (Pp->L.mnrl_flags & 2) ? LD(Pp->L.o_mlim, k) : 0.0;
  if (lim > 0.0) aff = aff / (1.0 + (1.0 - aff) / lim);
  double Im_const = -SD(S.marea, k);
  if (msf_on) Im_const = Im_const / msf;
  const double rate = mineral_rate_constant<XM>(Pp, S, k, jpert);
  return (Pp->L.mnrl_flags & 4) ? Im_const * sgn * pow(fabs(aff), LD(Pp->L.o_mpow, k)) * rate
                             : Im_const * sgn * fabs(aff) * rate;
}

// Derivative of a prefactor mineral's rate through its prefactors
// (reaction_mineral.F90:887-944; primary prefactor species): for each
// species s of prefactor p, dIm/dm_s = (Im / sum) (prefactor_p / monod_s)
// d monod_s / d a_s * gamma_s * k_p * Arrhenius_p, added to column s of the
// lane's row with the lane's stoichiometric coefficient nur.  dIm_dsum is
// Im (mol/s) over the summed rate constant.
template <int G>
__device__ __forceinline__ void prefactor_jacobian(KP Pp, const Scr &S, int k, double dIm_dsum, double nur,
                                                   double *row) {
  const int np = LI(Pp->L.o_mpn, k);
  const int PP = Pp->L.npref, PQ = Pp->L.npref_spec;
  for (int p = 0; p < np; p++) {
    const double arr = arrhenius(LD(Pp->L.o_mpEa, k * PP + p), SD(S.cs, 4));
    double lnp = 0.0;
    for (int q = 0; q < PQ; q++) {
      const int o = (k * PP + p) * PQ + q;
      if (LI(Pp->L.o_mpid, o) < 0) continue;
      double num, den;
      pref_species_ln(Pp, o, SD(S.la, LI(Pp->L.o_mpid, o)), num, den);
      lnp = lnp + num;
      lnp = lnp - den;
    }
    for (int q = 0; q < PQ; q++) {
      const int o = (k * PP + p) * PQ + q;
      const int sp = LI(Pp->L.o_mpid, o);
      if (sp < 0) continue;
      const double la = SD(S.la, sp);
      double num, den;
      pref_species_ln(Pp, o, la, num, den);
      const double lps = num - den;
      const double dpref = exp(lnp - lps);
      const double dnum = LD(Pp->L.o_mpalpha, o) * exp(lps - la);
      const double atten = LD(Pp->L.o_mpatten, o), beta = LD(Pp->L.o_mpbeta, o);
      const double lgb = beta * la;
      const double denom = 1.0 + exp(log(atten) + lgb);
      const double dden = -1.0 * exp(lps) / denom * atten * beta * exp(lgb - la);
      const double gam_s = exp(la) * SD(S.im, sp);  // activity over molality
      const double dspec = (dnum + dden) * gam_s;
      const double dIm = dIm_dsum * dpref * dspec * LD(Pp->L.o_mprate, k * PP + p) * arr;
#pragma unroll
      for (int kk = 0; kk < G; kk++) row[kk] += (kk == sp) ? nur * dIm : 0.0;
    }
  }
}

// Reaction part of the Jacobian by forward differences (NUMERICAL_JACOBIAN,
// RReactionDerivative reaction.F90:3625-3684): column j is
// (Res(m + e_j m_j 1e-5) - Res(m)) / (m_j 1e-5), Res = sum_k nu_ik Im_k V in
// mineral order from zero; entries below 1e-40 in magnitude become 0.  The
// perturbed ln activities use the unperturbed coefficients.  Adds to row[].
template <int G, bool XM>
__device__ __forceinline__ void mineral_fd_jacobian(KP Pp, const Scr &S, int r, bool live, double m, double lng,
                                                    double vol, double *row) {
  const int K = Pp->L.K, nS = Pp->L.S;
  {
#pragma clang fp contract(off)
    if (live) SD(S.xb, r) = log(m + m * 1.0e-5) + lng;  // perturbed ln a_r
  }
  wave_sync();
  double ro = 0.0;
  for (int k = 0; k < K; k++) {
    double lnQK = SD(S.mlnK, k);
    for (int s = 0; s < nS; s++) {
      const int sp = LI(Pp->L.o_mspec, k * nS + s);
      if (sp >= 0) lnQK += LD(Pp->L.o_mnu, k * nS + s) * SD(S.la, sp);
    }
    const double nur = live ? LD(Pp->L.o_mnuG, k * G + r) : 0.0;
    ro += nur * (mineral_rate<XM>(Pp, S, k, lnQK, -1) * vol);
  }
#pragma unroll
  for (int j = 0; j < G; j++) {
    if (j < Pp->L.N) {
      double rp = 0.0;
      for (int k = 0; k < K; k++) {
        double lnQK = SD(S.mlnK, k);
        for (int s = 0; s < nS; s++) {
          const int sp = LI(Pp->L.o_mspec, k * nS + s);
          if (sp >= 0) lnQK += LD(Pp->L.o_mnu, k * nS + s) * (sp == j ? SD(S.xb, j) : SD(S.la, sp));
        }
        const double nur = live ? LD(Pp->L.o_mnuG, k * G + r) : 0.0;
        rp += nur * (mineral_rate<XM>(Pp, S, k, lnQK, j) * vol);
      }
      double pert;
      {
#pragma clang fp contract(off)
        pert = SD(S.mv, j) * 1.0e-5;
      }
      double d = (rp - ro) / pert;
      if (fabs(d) < 1.0e-40) d = 0.0;
      if (live) row[j] += d;
    }
  }
  wave_sync();
}

// A load issued whether or not the array is given: a missing array reads
// `dummy` (any valid device address) and the caller selects its default.
__device__ __forceinline__ double ld_or(const double *a, long long i, const unsigned char *dummy) {
  return *(a ? a + i : (const double *)dummy);
}

template <int G, int MODE, int LUP, bool XM>
__device__ __forceinline__ void react_body(uint64_t praw) {
  const KP P0 = (KP)praw;
  const int N = P0->L.N, M = P0->L.M, K = P0->L.K, R = P0->L.R;
  // a pivoting reaction launch given the deferred list works through it
  constexpr bool kCanDefer = MODE == PFX_MODE_REACT && LUP == PFX_LU_DEFER;
  const bool from_list = MODE == PFX_MODE_REACT && LUP == PFX_LU_PIVOT && P0->defer != nullptr;
  const long long ncl = from_list ? (long long)P0->defer[0] : P0->ncells;
  const int lds_tab = (int)P0->L.lds_table_bytes;

  // stage the reaction tables into LDS (16-byte words)
  {
    const int4 *src = (const int4 *)P0->tables;
    int4 *dst = (int4 *)smem;
    const int nw = lds_tab >> 4;
    for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int r = threadIdx.x & (G - 1);
  const int gbase = lane - r;
  const bool live = r < N;  // padding lanes carry an identity row
  const int sdoubles = scr_doubles(M, G, K, XM);
  // The group's scratch offsets, and the lane's charge and ion size, are
  // rebuilt every pass from opaque copies of their base values: held across
  // the persistent loop they cost ~15 registers that the register-capped
  // variants spilled to scratch and reloaded on every pass.
  int sbase = lds_tab + (int)(threadIdx.x / G) * sdoubles * 8;
  int r_opq = r;
  if (r == 0) SD(make_scr<XM>(sbase, M, G, K).C, M) = 0.0;

  // wave-level work queue state (identical in all lanes of the wave)
  long long q_next = 0, q_end = 0;
  bool exhausted = false;

#ifdef PFX_SECTION_TIMERS
  unsigned long long pt_acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long pt_last = clock64();
#endif
  // per-group cell state
  long long cell = -1;
  int phase = 0, it = 0, extra = 0;
  int it_cap = 0;           // iteration at which the current attempt gives up
  bool fell_back = false;   // warm attempt abandoned, restarted from the default guess
  bool stop_next = false;   // reference relative-change stop: finish at the next evaluation
  bool reload = false;      // load the current cell again (fallback restart)
  bool lu_pivot = false;    // this cell factors with row pivoting (its unpivoted LU broke down)
  bool deferred = false;    // left to the pivoting launch (factorisation broke down, or not converged)
  // step test at a stagnating candidate: this pass solves for the FP64
  // resolution of the free ions at the candidate iterate, whose step
  // (m_cand, its size rel_cand) waits
  bool floor_pass = false;
  double rel_cand = 0.0, m_cand = 0.0;
  double last_change = INFINITY;  // max|dm/m| of the attempt's last Newton update
  int cell_flag = PFX_CELL_CONVERGED;
  double m = 1.0, gam = 1.0, lng = 0.0;
  // constraint speciation state
  int ctype = PFX_CONSTRAINT_FREE, nit_on = 0;
  bool compute_act = false;
  double tcon = 0.0, free_conc = 1e-9;
  double clnK = 0.0;  // -ln K of the lane's mineral / gas constraint reaction

  for (;;) {
    asm volatile("" : "+s"(praw));
    PFX_PT(8);
    KP Pp = (KP)praw;
#define P (*Pp)
#define L (Pp->L)
    asm volatile("" : "+v"(sbase), "+v"(r_opq));
    const Scr S = make_scr<XM>(sbase, L.M, G, L.K);
    const double Zr = live ? LD(L.o_Z, r_opq) : 0.0;
    const double a0r = live ? LD(L.o_a0, r_opq) : 0.0;
    double tot_out = 0.0, sorb_out = 0.0;  // totals and sorbed totals of this pass
    // ---------------------------------------------------------------
    // refill: groups without a cell take the next ones from the queue
    bool load_now = reload;
    reload = false;
    {
      const bool need = (cell < 0) && !exhausted;
      unsigned long long mask = __ballot(need && r == 0);
      int nneed = __popcll(mask);
      int rank = __popcll(mask & ((1ull << lane) - 1ull));
      long long mycell = -1;
      while (nneed > 0) {
        long long avail = q_end - q_next;
        if (avail <= 0) {
          // guided sizes: near the end of the queue a fetch takes less, so
          // the last wavefronts do not hold long private runs of cells
          unsigned int base = 0, take_n = (unsigned int)P.chunk;
          if (lane == 0) {
            if (P.waves > 0) {
              const long long left =
                  ncl - (long long)__hip_atomic_load(P.queue, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              const long long share = left / (4LL * P.waves);
              const long long lo = 64 / G;
              take_n = (unsigned int)(share < lo ? lo : (share < (long long)take_n ? share : (long long)take_n));
            }
            base = atomicAdd(P.queue, take_n);
          }
          base = __shfl(base, 0);
          take_n = __shfl(take_n, 0);
          q_next = base;
          q_end = (long long)base + take_n;
          if (q_end > ncl) q_end = ncl;
          if (q_next >= ncl) {
            exhausted = true;
            break;
          }
          continue;
        }
        long long take = nneed < avail ? nneed : avail;
        if (need && r == 0 && rank >= 0 && rank < take) mycell = q_next + rank;
        rank -= (int)take;
        nneed -= (int)take;
        q_next += take;
      }
      // a listed cell: its index, and bit 31 set where it failed after its
      // warm attempt had been given up (the pivoting launch then repeats the
      // plain attempt, counting on from the iterations already spent)
      bool listed_fb = false;
      if (from_list && mycell >= 0) mycell = (long long)P.defer[1 + mycell];
      if (need) {
        int lo = bperm_i((int)(mycell & 0xffffffff), gbase);
        int hi = bperm_i((int)(mycell >> 32), gbase);
        long long c = ((long long)hi << 32) | (unsigned int)lo;
        if (from_list && c >= 0) {
          listed_fb = (c >> 31) & 1;
          c &= 0x7fffffff;
        }
        if (c >= 0) {
          cell = c;
          it = listed_fb ? P.num_iterations[c] : 0;
          fell_back = listed_fb;
          lu_pivot = false;
          deferred = false;
          floor_pass = false;
          load_now = true;
        }
      }
      // ---- load the cell: a new one, or the same one again when its warm
      // attempt failed (the restart then repeats the plain method from the
      // step's start: activity coefficients, complexes and exchange state
      // are read again from the unchanged inputs)
      if (load_now) {
        {
          const long long c = cell;
          phase = (MODE == PFX_MODE_REACT) ? 0 : 2;
          extra = 0;
          it_cap = (MODE == PFX_MODE_REACT && P.guess_mode == PFX_GUESS_ARRAY_FALLBACK && !fell_back)
                       ? P.fallback_after
                       : it + P.max_iterations;
          stop_next = false;
          last_change = INFINITY;
          cell_flag = PFX_CELL_CONVERGED;
          compute_act = false;
          nit_on = 0;
          // ---- load the cell.  Every global load is issued before any is
          // used: a missing optional array reads the table blob instead and
          // its default is selected afterwards, so the loads are in flight
          // together instead of one round trip per array.
          const unsigned char *dmy = P.tables;
          const int rr = live ? r : 0;
          const long long ci = c * N + rr;
          const double v_por = ld_or(P.porosity, c, dmy);
          const double v_sat = ld_or(P.saturation, c, dmy);
          const double v_vol = ld_or(P.volume, c, dmy);
          const double v_den = ld_or(P.den_kg, c, dmy);
          const double v_tmp = ld_or(P.temperature, c, dmy);
          double v_m = 1.0, v_gam = 1.0, v_tot = 0.0, v_sorb = 0.0, v_conc = 0.0;
          double v_fguess = 1.0e-9;
          if (MODE == PFX_MODE_EQUIL) {
            v_conc = ld_or(P.constraint_conc, ci, dmy);
            v_fguess = ld_or(P.guess_mode == PFX_GUESS_ARRAY ? P.guess : nullptr, ci, dmy);
          } else {
            v_m = ld_or(P.pri_molal, ci, dmy);
            v_gam = ld_or(P.pri_act_coef, ci, dmy);
            if (MODE != PFX_MODE_AUX) {
              v_tot = ld_or(P.total, ci, dmy);
              v_sorb = ld_or(L.sorb_any ? P.total_sorb : nullptr, ci, dmy);
            }
          }
          double mguess = 1.0;
          if (MODE == PFX_MODE_REACT)
            mguess = ld_or((P.guess_mode == PFX_GUESS_ARRAY || P.guess_mode == PFX_GUESS_ARRAY_FALLBACK) && !fell_back
                               ? P.guess
                               : nullptr,
                           ci, dmy);
          // the fallback's starting point: the previous free ions or the default guess
          if (MODE == PFX_MODE_REACT && fell_back) mguess = P.fallback_previous ? v_m : P.default_guess;
          // the warm-start guess waits in the group's exchange slot (unused
          // until the first ion-exchange pass) instead of a register
          SD(S.xb, r) = mguess;
          double v_ionx[PFX_MAX_IONX_RXN];
#pragma unroll
          for (int q = 0; q < PFX_MAX_IONX_RXN; q++)
            v_ionx[q] = ld_or((MODE != PFX_MODE_EQUIL && q < R) ? P.ionx_ref : nullptr, c * R + q, dmy);
          const double porosity = P.porosity ? v_por : P.def_porosity;
          const double sat = P.saturation ? v_sat : P.def_saturation;
          const double vol = P.volume ? v_vol : P.def_volume;
          const double den = P.den_kg ? v_den : P.def_den;
          if (r == 0) {
            SD(S.cs, 0) = porosity;
            SD(S.cs, 1) = sat;
            SD(S.cs, 2) = vol;
            SD(S.cs, 3) = den;
          }
          const double tempc = P.temperature ? v_tmp : 25.0;
          if constexpr (XM) {
            if (r == 0) SD(S.cs, 4) = tempc;
          }
          const double tk = tempc + 273.15;
          const double ln_tk = log(tk);
          if (MODE == PFX_MODE_EQUIL) {
            gam = 1.0;
            for (int cc = r; cc < M; cc += G) SD(S.gsec, cc) = 1.0;
            const double molal_to_molar = P.molal ? den / 1000.0 : 1.0;
            const double molar_to_molal = P.molal ? 1.0 : 1000.0 / den;
            ctype = live ? P.ctype[r] : PFX_CONSTRAINT_FREE;
            const double conc = live ? v_conc : 0.0;
            // the starting free ions: a FREE_ION_GUESS (cells->guess with
            // PFX_GUESS_ARRAY) or 1e-9 (reaction.F90:1476-1482)
            free_conc = (P.guess_mode == PFX_GUESS_ARRAY && P.guess) ? v_fguess : 1.0e-9;
            tcon = 0.0;
            if (ctype == PFX_CONSTRAINT_TOTAL) tcon = conc * molal_to_molar;
            else if (ctype == PFX_CONSTRAINT_TOTAL_SORB) tcon = conc;  // mol/m^3 bulk (reaction.F90:1490-1492)
            else if (ctype == PFX_CONSTRAINT_FREE) free_conc = conc * molar_to_molal;
            else if (ctype == PFX_CONSTRAINT_LOG) free_conc = pow(10.0, conc) * molar_to_molal;
            else if (ctype == PFX_CONSTRAINT_PH) { free_conc = pow(10.0, -conc); tcon = conc; }
            else if (ctype == PFX_CONSTRAINT_CHARGE) free_conc = conc * molar_to_molal;
            else if (ctype == PFX_CONSTRAINT_MINERAL) free_conc = conc * molar_to_molal;
            else if (ctype == PFX_CONSTRAINT_GAS) tcon = conc <= 0.0 ? pow(10.0, conc) : conc;  // bar
            if ((ctype == PFX_CONSTRAINT_MINERAL || ctype == PFX_CONSTRAINT_GAS) && P.crxn) {
              const pfx_constraint_rxn &X = P.crxn[r];
              const double lk = (X.has_logK_coef && P.temperature) ? logk_5term_ptr(X.logK_coef, tk, ln_tk) : X.logK;
              clnK = -lk * PFX_LOG_TO_LN;
            }
            m = live ? free_conc : 1.0;
          } else {
            m = live ? v_m : 1.0;
            gam = (live && P.pri_act_coef) ? v_gam : 1.0;
            if (MODE != PFX_MODE_AUX) {
              const double sorb_old = (L.sorb_any && P.total_sorb) ? v_sorb : 0.0;
              SD(S.fx, r) = live ? porosity * sat * 1000.0 * vol * v_tot + (L.sorb_any ? sorb_old * vol : 0.0) : 0.0;
            }
            // up to 4 complexes per pass with their loads issued together
#pragma unroll 4
            for (int cc = r; cc < M; cc += G) {
              const double g = ld_or(P.sec_act_coef, c * M + cc, dmy);
              SD(S.gsec, cc) = P.sec_act_coef ? g : 1.0;
            }
          }
          for (int cc = r; cc < M; cc += G) {
            double lk = LD(L.o_clogK, cc);
            if (P.temperature && L.has_logk_coef) {
              const int cfo = (int)L.o_clogKcoef + 5 * cc * 8;
              lk = logk_5term(cfo, tk, ln_tk);
            }
            SD(S.lnK, cc) = -lk * PFX_LOG_TO_LN;
          }
          for (int k = r; k < K; k += G) {
            const double v_area = ld_or(P.mnrl_area, c * K + k, dmy);
            const double v_vf = ld_or(P.mnrl_volfrac, c * K + k, dmy);
            double lk = LD(L.o_mlogK, k);
            if (P.temperature && L.has_mnrl_coef) {
              const int cfo = (int)L.o_mlogKcoef + 5 * k * 8;
              lk = logk_5term(cfo, tk, ln_tk);
            }
            SD(S.mlnK, k) = -lk * PFX_LOG_TO_LN;
            double arr = 1.0;
            const double Ea = LD(L.o_mEa, k);
            if (Ea > 0.0) arr = exp(Ea / PFX_IDEAL_GAS * (1.0 / (25.0 + 273.15) - 1.0 / (tempc + 273.15)));
            SD(S.mkeff, k) = LD(L.o_mk, k) * arr;
            SD(S.marea, k) = P.mnrl_area ? v_area : 0.0;
            SD(S.mvf, k) = P.mnrl_volfrac ? v_vf : 0.0;
            SD(S.mrate, k) = 0.0;
          }
#pragma unroll
          for (int q = 0; q < PFX_MAX_IONX_RXN; q++)
            SD(S.ionx, q) = (MODE != PFX_MODE_EQUIL && q < R && P.ionx_ref) ? v_ionx[q] : 1e-9;
          lng = log(gam);
          if (MODE == PFX_MODE_REACT && L.act_freq == PFX_ACT_OFF) {
            phase = 2;
            if (live) {
              if (P.guess_mode == PFX_GUESS_ARRAY || P.guess_mode == PFX_GUESS_ARRAY_FALLBACK)
                m = SD(S.xb, r);
              else if (P.guess_mode == PFX_GUESS_DEFAULT) m = P.default_guess;
            }
          }
        }
      }
    }
    if (__ballot(cell >= 0) == 0ull) break;
    PFX_PT(0);
    wave_sync();
    const bool active = cell >= 0;

    // ---------------------------------------------------------------
    // activity coefficients from the current free ions and the complex
    // molalities of the previous RTotal (RActivityCoefficients,
    // reaction.F90:3812-4053)
    bool act_now = false;
    if (MODE == PFX_MODE_REACT)
      act_now = phase == 1 || (phase == 2 && L.act_freq == PFX_ACT_NEWTON_ITER && !stop_next && !floor_pass);
    else if (MODE == PFX_MODE_EQUIL)
      act_now = compute_act && L.act_freq != PFX_ACT_OFF;
    else if (MODE == PFX_MODE_AUX)
      act_now = it >= 1;
    if (active && act_now) {
      // with updates every Newton iteration, the step's first pass updates
      // once at the previous state (reaction.F90:3407-3409) and once more at
      // the initial iterate, as iteration 1 does (reaction.F90:3415-3418)
      const int npass = (MODE == PFX_MODE_REACT && phase == 1 && L.act_freq == PFX_ACT_NEWTON_ITER) ? 2 : 1;
      for (int pass = 0; pass < npass; pass++) {
        if (MODE == PFX_MODE_REACT && pass == 1) {
          phase = 2;
          if (live) {
            if (P.guess_mode == PFX_GUESS_ARRAY || P.guess_mode == PFX_GUESS_ARRAY_FALLBACK)
              m = SD(S.xb, r);
            else if (P.guess_mode == PFX_GUESS_DEFAULT) m = P.default_guess;
          }
        }
        const double A = L.debyeA, B = L.debyeB, Bd = L.debyeBdot;
        if (L.act_alg == PFX_ACT_ALG_NEWTON) {
          // Newton on the ionic strength that also updates the complex
          // molalities at fixed primary activities (reaction.F90:3846-3991)
          SD(S.la, r) = live ? log(m) + lng : 0.0;
          const double fpri = gsum<G>(live ? m * Zr * Zr : 0.0, lane);
          double II = 0.0;
          int ait = 0;
          wave_sync();
          for (;;) {
            if (++ait > 50) {  // reference: NaN poisoning (reaction.F90:3866-3877)
              cell_flag = PFX_CELL_NONFINITE;
              break;
            }
            double part = 0.0;
            for (int cc = r; cc < M; cc += G) part += SD(S.C, cc) * LD(L.o_cZ, cc) * LD(L.o_cZ, cc);
            double I = 0.5 * (fpri + gsum<G>(part, lane));
            const double f = I;
            if (fabs(I - II) < 1.0e-6 * I) break;
            if (M > 0) {
              double sqI = sqrt(I);
              // d log10(gamma_j)/dI of the primaries (0 for neutral ones)
              if (live) {
                const double tj = 1.0 + B * a0r * sqI;
                SD(S.im, r) = fabs(Zr) > 0.0 ? -0.5 * A * Zr * Zr / (sqI * tj * tj) + Bd : 0.0;
              }
              wave_sync();
              double dpart = 0.0;
              for (int cc = r; cc < M; cc += G) {
                const double z = LD(L.o_cZ, cc);
                if (fabs(z) > 0.0) {
                  const double t = 1.0 + B * LD(L.o_ca0, cc) * sqI;
                  double sum = 0.5 * A * z * z / (sqI * t * t) - Bd;
                  const int nS = L.S;
                  for (int s = 0; s < nS; s++) {
                    const int sp = LI(L.o_cspec, cc * nS + s);
                    if (sp >= 0 && fabs(LD(L.o_Z, sp)) > 0.0) sum += LD(L.o_cnu, cc * nS + s) * SD(S.im, sp);
                  }
                  const double dcdi = SD(S.C, cc) * PFX_LOG_TO_LN * sum;
                  dpart += 0.5 * z * z * dcdi;
                }
              }
              const double didi = gsum<G>(dpart, lane);
              const double dd = 1.0 - didi;
              II = fabs(dd) > 0.0 ? (f - I * didi) / dd : f;
            } else {
              II = f;
            }
            I = II;
            const double sqI = sqrt(I);
            if (live) {
              gam = fabs(Zr) > 0.0 ? exp((-Zr * Zr * sqI * A / (1.0 + a0r * B * sqI) + Bd * I) * PFX_LOG_TO_LN)
                                   : 1.0;
              lng = log(gam);
            }
            for (int cc = r; cc < M; cc += G) {
              const double z = LD(L.o_cZ, cc);
              const double gs =
                  fabs(z) > 0.0 ? exp((-z * z * sqI * A / (1.0 + LD(L.o_ca0, cc) * B * sqI) + Bd * I) * PFX_LOG_TO_LN)
                                : 1.0;
              SD(S.gsec, cc) = gs;
              double lnQK = SD(S.lnK, cc);
              const int nS = L.S;
              for (int s = 0; s < nS; s++) {
                const int sp = LI(L.o_cspec, cc * nS + s);
                if (sp >= 0) lnQK += LD(L.o_cnu, cc * nS + s) * SD(S.la, sp);
              }
              SD(S.C, cc) = exp(lnQK) / gs;
            }
            wave_sync();
          }
        } else {
          // LAG: one pass (reaction.F90:3992-4051)
          double part = live ? m * Zr * Zr : 0.0;
          for (int cc = r; cc < M; cc += G) part += SD(S.C, cc) * LD(L.o_cZ, cc) * LD(L.o_cZ, cc);
          const double I = 0.5 * gsum<G>(part, lane);
          const double sqI = sqrt(I);
          if (live) {
            gam = fabs(Zr) > 1e-10 ? exp((-Zr * Zr * sqI * A / (1.0 + a0r * B * sqI) + Bd * I) * PFX_LOG_TO_LN)
                                   : 1.0;
            lng = log(gam);
          }
          for (int cc = r; cc < M; cc += G) {
            const double z = LD(L.o_cZ, cc);
            SD(S.gsec, cc) = fabs(z) > 1e-10
                                 ? exp((-z * z * sqI * A / (1.0 + LD(L.o_ca0, cc) * B * sqI) + Bd * I) * PFX_LOG_TO_LN)
                                 : 1.0;
          }
        }
      }
      if (MODE == PFX_MODE_REACT && phase == 1) {
        phase = 2;
        if (live) {
          if (P.guess_mode == PFX_GUESS_ARRAY || P.guess_mode == PFX_GUESS_ARRAY_FALLBACK)
            m = SD(S.xb, r);
          else if (P.guess_mode == PFX_GUESS_DEFAULT) m = P.default_guess;
        }
      }
    }
    PFX_PT(1);
    if (MODE == PFX_MODE_EQUIL && active && live &&
        (ctype == PFX_CONSTRAINT_FREE || ctype == PFX_CONSTRAINT_LOG))
      m = free_conc;  // reaction.F90:1552-1558

    // ---------------------------------------------------------------
    // RTotal, phase A: complex molalities, complex c computed by lane c % G
    if (active) {
      SD(S.la, r) = live ? log(m) + lng : 0.0;
      SD(S.mv, r) = m;
      SD(S.im, r) = 1.0 / m;
    }
    wave_sync();
    if (active) {
      for (int cc = r; cc < M; cc += G) {
        double lnQK = SD(S.lnK, cc);
        const int nS = L.S;
#pragma unroll
        for (int s = 0; s < PFX_MAX_RXN_SPEC; s++) {
          if (s < nS) {
            const int sp = LI(L.o_cspec, cc * nS + s);
            if (sp >= 0) lnQK += LD(L.o_cnu, cc * nS + s) * SD(S.la, sp);
          }
        }
        SD(S.C, cc) = exp(lnQK) / SD(S.gsec, cc);
      }
    }
    wave_sync();

    PFX_PT(2);
    // phase B: totals, |contributions|, and row r of dT/dm (kg/L)
    // (not needed at the previous state, whose complexes only feed the
    // activity update)
    double row[G];
    double Tsum = 0.0, absum = 0.0;
    const double den_kg_per_L = SD(S.cs, 3) * 1.0e-3;
    if (active && !(MODE == PFX_MODE_REACT && phase == 0)) {
      double diag = 0.0;
      const int maxdeg = L.maxdeg;
#pragma unroll 2
      for (int e = 0; e < maxdeg; e++) {
        const int cc = LI(L.o_adj_c, e * G + r);
        const double nu = LD(L.o_adj_nu, e * G + r);
        const double Cc = SD(S.C, cc);
        Tsum += nu * Cc;
        absum += fabs(nu) * Cc;
        diag += nu * nu * Cc;
      }
      // off-diagonal sums, 4 columns per chunk: one slot is one 8-byte load
      // of 4 complex numbers (16-bit), two 16-byte loads of coefficients and
      // 4 independent gathers; a column with a shorter list reads the zero
      // slot C[M] with coefficient 0
#pragma unroll
      for (int q = 0; q < G / 4; q++) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        const int s1 = L.offq_start[q + 1];
        for (int s = L.offq_start[q]; s < s1; s++) {
          const uint2 w = *(const uint2 *)(smem + L.o_offq_c + (s * G + r) * 8);
          const double2 fa = *(const double2 *)(smem + L.o_offq_coef + ((2 * s) * G + r) * 16);
          const double2 fb = *(const double2 *)(smem + L.o_offq_coef + ((2 * s + 1) * G + r) * 16);
          a0 += fa.x * SD(S.C, (int)(w.x & 0xffffu));
          a1 += fa.y * SD(S.C, (int)(w.x >> 16));
          a2 += fb.x * SD(S.C, (int)(w.y & 0xffffu));
          a3 += fb.y * SD(S.C, (int)(w.y >> 16));
        }
        row[4 * q] = a0;
        row[4 * q + 1] = a1;
        row[4 * q + 2] = a2;
        row[4 * q + 3] = a3;
      }
#pragma unroll
      for (int k = 0; k < G; k++)
        row[k] = (k == r) ? (live ? 1.0 + diag * SD(S.im, k) : 1.0) : row[k] * SD(S.im, k);
#pragma unroll
      for (int k = 0; k < G; k++) row[k] = live ? row[k] * den_kg_per_L : row[k];
      tot_out = (m + Tsum) * den_kg_per_L;
    }
    PFX_PT(3);
    if (MODE == PFX_MODE_EVAL && active && live) {
      // aqueous totals and dT/dm of the cell for the flux terms of the
      // global-implicit assembly (RTAuxVarCompute's total and dtotal)
      if (P.tot_eval_out) P.tot_eval_out[cell * N + r] = tot_out;
      if (P.dtot_out) {
#pragma unroll
        for (int k = 0; k < G; k++)
          if (k < N) P.dtot_out[cell * N * N + (long long)k * N + r] = P.col_scale ? row[k] * SD(S.mv, k) : row[k];
      }
    }
    if (MODE == PFX_MODE_REACT && active && phase == 0) {
      phase = 1;  // complexes of the previous state are in LDS: update activity next pass
      wave_sync();
      continue;
    }
    if (MODE == PFX_MODE_AUX && active && it < P.aux_passes) {
      it++;
      wave_sync();
      continue;
    }

    // ---------------------------------------------------------------
    // equilibrium ion exchange (RTotalSorbEqIonx): sorbed total of row r and
    // d(sorbed)/dm added to row[] with weight ionx_w
    const bool need_ionx = (MODE == PFX_MODE_REACT || MODE == PFX_MODE_EVAL) ? true : false;
    const bool ionx_final = (MODE == PFX_MODE_AUX || MODE == PFX_MODE_EQUIL);
    double Sr = 0.0;
    // d(sorbed total of row r)/dm for a total-sorbed constraint row
    // (reaction.F90:1608-1614); only the constraint kernel keeps it
    double srow[G];
#pragma unroll
    for (int k = 0; k < G; k++) srow[k] = 0.0;
    const double dt = P.dt;
    const double vol = SD(S.cs, 2);
    const double psv_t = SD(S.cs, 0) * SD(S.cs, 1) * 1000.0 * vol;
    if (active && R > 0 && (need_ionx || ionx_final)) {
      const double ionx_w = need_ionx ? vol / dt : 0.0;
      for (int q = 0; q < R; q++) {
        const int role = live ? LI(L.o_ionx_role, q * G + r) : -1;
        const double kq = live ? LD(L.o_ionx_k, q * G + r) : 0.0;
        double omega = L.ionx_CEC[q];
        if (L.ionx_to_surf[q] >= 0) omega = fmax(L.ionx_CEC[q] * SD(S.mvf, L.ionx_to_surf[q]), 1e-40);
        double X = 0.0;
        if (L.ionx_zflag[q]) {
          const int rl = L.ionx_ref_lane[q];
          const double ref_conc = bperm_d(m * gam, gbase + rl);
          const double ref_Z = LD(L.o_Z, rl);
          const double ref_k = L.ionx_ref_k[q];
          double ref_X = ref_Z * SD(S.ionx, q) / omega;
          double KDj = ref_X / (ref_k * ref_conc);
          bool one_more = false;
          const double zratio = Zr / ref_Z;
          int iit = 0;
          for (;;) {
            if (++iit > 20000) {  // reference: error stop (reaction.F90:4382-4385)
              cell_flag = PFX_CELL_MAX_ITERATIONS;
              break;
            }
            ref_X = KDj * (ref_k * ref_conc);
            if (role > 0) X = kq * m * gam * pow_ratio(KDj, zratio);
            else X = role == 0 ? ref_X : 0.0;
            const double total = gsum<G>(X, lane);
            double dres = gsum<G>(role > 0 ? X / KDj * Zr : 0.0, lane);
            dres = dres / ref_Z + (ref_k * ref_conc);
            const double res = 1.0 - total;
            if (one_more) break;
            const double delta = res / dres;
            KDj = KDj + delta;
            KDj = fmax(KDj, 1e-40);
            if (fabs(delta / KDj) < 1e-12) one_more = true;
          }
          SD(S.ionx, q) = ref_X * omega / ref_Z;
        } else {
          X = role >= 0 ? m * gam * kq : 0.0;
          const double sumkm = gsum<G>(X, lane);
          X = X / sumkm;
        }
        const double sumZX = gsum<G>(Zr * X, lane);
        const double t1 = role >= 0 ? X * omega / Zr : 0.0;
        const double t2 = Zr / sumZX;
        Sr += t1;
        if (need_ionx) {
          const double wscale = ionx_w / (live ? psv_t / dt : 1.0);
          SD(S.xb, r) = X;
          wave_sync();
#pragma unroll
          for (int k = 0; k < G; k++) {
            const double Xk = SD(S.xb, k);
            const double d = (k == r) ? t1 * (1.0 - (t2 * Xk)) : (-t1) * t2 * Xk;
            row[k] += d * SD(S.im, k) * wscale;
          }
          wave_sync();
        } else if (MODE == PFX_MODE_EQUIL && P.has_sorb) {
          SD(S.xb, r) = X;
          wave_sync();
#pragma unroll
          for (int k = 0; k < G; k++) {
            const double Xk = SD(S.xb, k);
            const double d = (k == r) ? t1 * (1.0 - (t2 * Xk)) : (-t1) * t2 * Xk;
            srow[k] += d * SD(S.im, k);
          }
          wave_sync();
        }
      }
    }
    // isotherm sorption (RTotalSorbKD, reaction.F90:4220-4301): the lane's
    // own primary only, so its derivative is diagonal
    if (active && L.kd_any && (need_ionx || ionx_final)) {
      const int kty = live ? LI(L.o_kd_type, r) : 0;
      if (kty > 0) {
        double kd = LD(L.o_kd_coef, r);
        const int km = LI(L.o_kd_mnrl, r);
        // K_D in mL water/g soil scaled by the mineral's volume fraction
        if (km >= 0) kd = kd * SD(S.cs, 3) * (1.0 - SD(S.cs, 0)) * L.soil_density * 1.0e-3 * SD(S.mvf, km);
        double res, dres;
        if (kty == PFX_SORPTION_LINEAR) {
          res = kd * m;
          dres = kd;
        } else if (kty == PFX_SORPTION_LANGMUIR) {
          const double tr = kd * m;
          res = tr * LD(L.o_kd_b, r) / (1.0 + tr);
          dres = res / m - res / (1.0 + tr) * tr / m;
        } else {
          const double inv_n = LD(L.o_kd_invn, r);
          res = kd * pow(m, inv_n);
          dres = res / m * inv_n;
        }
        Sr += res;
        const double jw = need_ionx ? vol / psv_t : 1.0;  // row[] is scaled by psv_t/dt later
#pragma unroll
        for (int k = 0; k < G; k++) {
          if (k == r) {
            if (need_ionx) row[k] += dres * jw;
            else if (MODE == PFX_MODE_EQUIL && P.has_sorb) srow[k] += dres;
          }
        }
      }
    }
    sorb_out = Sr;
    PFX_PT(4);

    // ---------------------------------------------------------------
    // residual row, Jacobian row, convergence test
    double Rr = 0.0, scale = 1.0;
    bool finish = false;
    bool candidate = false;  // rel_res passed; the step test below decides
    int status = PFX_CELL_CONVERGED;
    double rr = 0.0;
    int use_log = L.use_log;
    if (MODE == PFX_MODE_AUX && active) finish = true;
    if ((MODE == PFX_MODE_REACT || MODE == PFX_MODE_EVAL) && active) {
      // accumulation (mol/s) and its derivative: the dt-consistent form
#pragma unroll
      for (int k = 0; k < G; k++) row[k] = live ? row[k] * (psv_t / dt) : row[k];
      if (live) {
        const double fixed = SD(S.fx, r);
        Rr = (psv_t * tot_out + (L.sorb_any ? Sr * vol : 0.0)) / dt - fixed / dt;
        const double psv = SD(S.cs, 0) * SD(S.cs, 1) * SD(S.cs, 3) * vol;
        scale = psv * (m + absum) / dt + fabs(Sr) * vol / dt + fabs(fixed) / dt;
      }
      // kinetic minerals (group-uniform rates, row-local residual/Jacobian)
      const bool fd_jac = MODE == PFX_MODE_EVAL && L.num_jac;
      // the capped-Q/K derivative for the attempt's first iterations only
      // (the attempt started at it_cap less its budget)
      const int att_it = it - (it_cap - ((MODE == PFX_MODE_REACT && P.guess_mode == PFX_GUESS_ARRAY_FALLBACK &&
                                          !fell_back) ? P.fallback_after : P.max_iterations));
      const bool capped_jac = (P.jac_flags & PFX_JAC_CAPPED_QK) && (MODE == PFX_MODE_EVAL || att_it < PFX_JAC_CAPPED_ITERS);
      for (int k = 0; k < K; k++) {
        double lnQK = SD(S.mlnK, k);
        for (int s = 0; s < L.S; s++) {
          const int sp = LI(L.o_mspec, k * L.S + s);
          if (sp >= 0) lnQK += LD(L.o_mnu, k * L.S + s) * SD(S.la, sp);
        }
        const double QK = lnQK <= 6.90776 ? exp(lnQK) : 1.0e3;
        const bool temkin = L.mnrl_flags & 8, msf_on = XM && (L.mnrl_flags & 32);
        const double msf = msf_on ? LD(L.o_mmsf, k) : 1.0;
        const double tk_ = temkin ? (msf_on ? msf * LD(L.o_mtemkin, k) : LD(L.o_mtemkin, k)) : msf;
        double aff = (temkin || msf_on) ? 1.0 - pow(QK, 1.0 / tk_) : 1.0 - QK;
        const double sgn = copysign(1.0, aff);
        double Im = 0.0;
        if (SD(S.mvf, k) > 0.0 || sgn < 0.0) {
          const double thr = (L.mnrl_flags & 1) ? LD(L.o_mthr, k) : 0.0;
          if (!(thr > 0.0 && sgn < 0.0 && QK < thr)) {
            const double lim = (L.mnrl_flags & 2) ? LD(L.o_mlim, k) : 0.0;
            if (lim > 0.0) aff = aff / (1.0 + (1.0 - aff) / lim);
            const double rate = mineral_rate_constant<XM>(Pp, S, k, -1);
            double Im_const = -SD(S.marea, k);
            if (msf_on) Im_const = Im_const / msf;
            const bool pw = L.mnrl_flags & 4;
            Im = pw ? Im_const * sgn * pow(fabs(aff), LD(L.o_mpow, k)) * rate : Im_const * sgn * fabs(aff) * rate;
            if (r == 0) SD(S.mrate, k) = Im;
            Im_const *= vol;
            const double ImV = Im * vol;
            const double nur = live ? LD(L.o_mnuG, k * G + r) : 0.0;
            Rr += nur * ImV;
            double dIm_dQK = pw ? -ImV * LD(L.o_mpow, k) / fabs(aff) : -Im_const * rate;
            if (temkin || msf_on) dIm_dQK = dIm_dQK * (1.0 / tk_) / QK * (1.0 - aff);
            double f = nur * dIm_dQK * QK * SD(S.cs, 3) * 1.0e-3;
            if (lim > 0.0) {
              const double dd = 1.0 + (1.0 - aff) / lim;
              f = f * (1.0 + QK / lim / dd) / dd;
            }
            if (!fd_jac && !(capped_jac && lnQK > 6.90776)) {
#pragma unroll
              for (int kk = 0; kk < G; kk++) row[kk] += f * LD(L.o_mnuG, k * G + kk) * SD(S.im, kk);
            }
            if constexpr (XM) {
              if (!fd_jac && (L.mnrl_flags & 16) && LI(L.o_mpn, k) > 0)
                prefactor_jacobian<G>(Pp, S, k, ImV / rate, nur, row);
            }
          } else if (r == 0) {
            SD(S.mrate, k) = 0.0;
          }
        } else if (r == 0) {
          SD(S.mrate, k) = 0.0;
        }
      }
      if (MODE == PFX_MODE_EVAL && fd_jac && K > 0) mineral_fd_jacobian<G, XM>(Pp, S, r, live, m, lng, vol, row);
      rr = gmax<G>(live ? fabs(Rr) / fmax(scale, 1e-300) : 0.0, lane);
      if (MODE == PFX_MODE_EVAL) {
        finish = true;
      } else if (cell_flag != PFX_CELL_CONVERGED) {
        finish = true;
        status = cell_flag;
      } else if (!(rr == rr) || rr == INFINITY) {
        finish = true;
        status = PFX_CELL_NONFINITE;
      } else if (stop_next) {
        finish = true;
      } else {
        if (rr <= P.rel_tol) {
          if (extra >= P.extra_iterations) {
            // with a step test the Newton step of this iterate decides
            if (MODE == PFX_MODE_REACT && P.rel_change_tol > 0.0) candidate = true;
            else finish = true;
          } else {
            extra++;
          }
        }
        if (!finish && P.reference_stop) {
          const double mx = gmax<G>(live ? fabs(Rr * dt) : 0.0, lane);
          if (mx < L.max_residual_tol) finish = true;
        }
        if (!finish && !candidate && it >= it_cap) {
          finish = true;
          status = PFX_CELL_MAX_ITERATIONS;
        }
      }
    }
    PFX_PT(5);
    if (MODE == PFX_MODE_EQUIL && active) {
      // constraint rows (reaction.F90:1597-1660)
      if (P.has_charge) {
        // charge balance (reaction.F90:1624-1651): R = sum_j z_j T_j and
        // row sum_j z_j dT_j/dm_k, reduced over the group from the rows
        // of dT/dm before any row is replaced
        const double zc = live ? Zr : 0.0;
        const double Rc = gsum<G>(zc * tot_out, lane);
        double crow[G];
#pragma unroll
        for (int k = 0; k < G; k++) crow[k] = gsum<G>(zc * row[k], lane);
        if (ctype == PFX_CONSTRAINT_CHARGE && live) {
          Rr = Rc;
#pragma unroll
          for (int k = 0; k < G; k++) row[k] = crow[k];
          // a vanishing counter-ion restarts at 1e-3 when the imbalance
          // needs more of it (reaction.F90:1635-1650)
          if (m < 1e-20 && ((Rr > 0.0 && Zr > 0.0) || (Rr < 0.0 && Zr < 0.0))) m = 1.0e-3;
        }
      }
      if (ctype == PFX_CONSTRAINT_TOTAL) {
        Rr = tot_out - tcon;
      } else if (ctype == PFX_CONSTRAINT_TOTAL_SORB) {
        Rr = sorb_out - tcon;
#pragma unroll
        for (int k = 0; k < G; k++) row[k] = srow[k];
      } else if (ctype == PFX_CONSTRAINT_MINERAL || ctype == PFX_CONSTRAINT_GAS) {
        // ln(Q/K) of the mineral, or of the gas less ln p (reaction.F90:1694-1745)
        const pfx_constraint_rxn &X = P.crxn[r];
        double lnQK = clnK;
        for (int s = 0; s < X.nspec; s++) lnQK += X.stoich[s] * SD(S.la, X.specid[s]);
        Rr = ctype == PFX_CONSTRAINT_MINERAL ? lnQK : lnQK - log(tcon);
#pragma unroll
        for (int k = 0; k < G; k++) row[k] = 0.0;
        for (int s = 0; s < X.nspec; s++) {
          const int sp = X.specid[s];
          const double v = X.stoich[s];
#pragma unroll
          for (int k = 0; k < G; k++)
            if (k == sp) row[k] += v / SD(S.mv, k);
        }
      } else if (ctype == PFX_CONSTRAINT_CHARGE) {
        // row set above
      } else {
        Rr = 0.0;
        if (ctype == PFX_CONSTRAINT_PH && live) m = pow(10.0, -tcon) / gam;
#pragma unroll
        for (int k = 0; k < G; k++) row[k] = (k == r) ? 1.0 : 0.0;
      }
      if (ctype == PFX_CONSTRAINT_PH || ctype == PFX_CONSTRAINT_CHARGE) SD(S.mv, r) = m;
      if (it > 3 && it < 9) use_log = use_log ? (it % 2 == 0) : 0;  // reaction.F90:1797-1804
      rr = gmax<G>(live ? fabs(Rr) : 0.0, lane);  // max |Res| (mol/L)
      if (cell_flag != PFX_CELL_CONVERGED) {
        finish = true;
        status = cell_flag;
      } else if (!(rr == rr) || rr == INFINITY) {
        finish = true;
        status = PFX_CELL_NONFINITE;
      } else if (stop_next) {
        finish = true;
      }
    }
    wave_sync();

    // ---------------------------------------------------------------
    PFX_PT(6);
    // RSolve: row equilibration, ln-column scaling, pivoted LU, solve
    double upd = 0.0;
    bool redo = false;  // the unpivoted LU broke down
    if ((MODE == PFX_MODE_REACT || MODE == PFX_MODE_EQUIL) && active && !finish) {
      // A floor pass solves for the rounding level of each residual row
      // (unit roundoff times the magnitude of the row's terms) instead of
      // the residual: its solution is the resolution of the free ions in
      // FP64.
      // Straight-line code: the steps run over all G rows with selects
      // instead of branches (a divergent branch, or a step skipped for
      // j >= N, makes the compiler copy the whole row at every merge).
      // Padding lanes (r >= N) hold identity rows with a zero right-hand
      // side; they are no pivot candidates before step N and pivot on
      // themselves after it, where every multiplier is zero, so the live
      // rows see exactly the N-step elimination.
      double b = 0.0;
      {
        double mx = 0.0;
#pragma unroll
        for (int k = 0; k < G; k++) mx = fmax(mx, fabs(row[k]));
        const double norm = live ? 1.0 / fmax(1.0, mx) : 1.0;
        b = live ? (floor_pass ? 2.220446049250313e-16 * scale : Rr) * norm : 0.0;
        const bool lg = live && use_log;
#pragma unroll
        for (int k = 0; k < G; k++) {
          // a multiply by 1 when the columns are not scaled (exact): a
          // select of the product made the compiler branch around the load
          const double sk = SD(S.mv, k);
          row[k] = row[k] * norm;
          row[k] = row[k] * (lg ? sk : 1.0);
        }
      }
      double vv = 0.0;
      {
        double mx = 0.0;
#pragma unroll
        for (int k = 0; k < G; k++) mx = fmax(mx, fabs(row[k]));
        vv = mx > 0.0 ? 1.0 / mx : 0.0;
        if (live && !(mx > 0.0)) status = PFX_CELL_SINGULAR;
      }
      // Elimination with the pivot row of step j in lane j, broadcast across
      // the group.  The reaction step's Jacobian in ln m is the
      // accumulation's diagonal plus positive semi-definite complexation,
      // exchange and mineral terms (nu^T diag(C) nu, diag(X) - X X^T,
      // nu nu^T dI/dlnQ) with rows scaled by positive factors, so it needs
      // no row exchanges, and the reaction kernels factor without a pivot
      // search.  A zero pivot or a non-finite multiplier (far from the
      // solution, or where a rate-law option makes a mineral term
      // indefinite) breaks the unpivoted factorisation down: under
      // PFX_LU_DEFER the cell is listed and left untouched for the pivoting
      // launch, which redoes its step from the start; under
      // PFX_LU_ONDEMAND its wave pivots from the cell's next pass on, at
      // the same iterate.  Pivoting (the constraint solve always, and
      // PFX_LU_PIVOT): before step j the largest scaled candidate of the
      // rows not yet eliminated (implicit row scaling vv, ties to the larger
      // row as the reference's `>=` scan, utility.F90:451-452) trades places
      // with row j, lane to lane, with its right-hand side and scale.
      // Padding lanes (r >= N) hold identity rows with a zero right-hand
      // side and zeros in the live rows' columns: they are never exchanged
      // and leave the live rows' elimination unchanged.
      constexpr bool kCanPivot = MODE != PFX_MODE_REACT || LUP != PFX_LU_DEFER;
      const bool pivoting = MODE != PFX_MODE_REACT || LUP == PFX_LU_PIVOT ||
                            (LUP == PFX_LU_ONDEMAND && __ballot(active && !finish && lu_pivot) != 0ull);
      {
        bool lbad = false;
        double inv_diag = 0.0;  // 1 / U_rr, kept by lane r for the back substitution
        static_for<0, G>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          if (kCanPivot && pivoting && j < N) {
            const double cand = (r >= j && live) ? vv * fabs(row[j]) : -1.0;
            const int pr = gargmax_row<G>(cand, lane, gbase);
            const int src = gbase + ((r == j) ? pr : ((r == pr) ? j : r));
#pragma unroll
            for (int k = 0; k < G; k++) {
              row[k] = bperm_d(row[k], src);
              if ((k & 7) == 7) __builtin_amdgcn_sched_barrier(0);
            }
            b = bperm_d(b, src);
            vv = bperm_d(vv, src);
          }
          double piv = gbcast<G, j>(row[j], gbase);
          if (pivoting) piv = piv == 0.0 ? 1.0e-20 : piv;  // ludcmp's TINY (utility.F90:462)
          const double ip = 1.0 / piv;
          const bool elim = r > j;
          // pivoting divides, as the reference's elimination does; the
          // unpivoted path multiplies by the reciprocal
          const double f = elim ? (pivoting ? row[j] / piv : row[j] * ip) : 0.0;
          lbad = lbad || (r == j && !(fabs(piv) > 0.0)) || !(fabs(f) <= 1.0e300);
          inv_diag = (r == j) ? ip : inv_diag;
          row[j] = (r == j) ? piv : (elim ? f : row[j]);
#pragma unroll
          for (int k = j + 1; k < G; k++) {
            const double pk = gbcast<G, j>(row[k], gbase);
            row[k] -= f * pk;
            if ((k & 7) == 7) __builtin_amdgcn_sched_barrier(0);
          }
          const double pb = gbcast<G, j>(b, gbase);
          b -= f * pb;
        });
        static_for<0, G>([&](auto jc) {
          constexpr int j = G - 1 - decltype(jc)::value;
          const double xj = gbcast<G, j>(pivoting ? b / row[j] : b * inv_diag, gbase);
          upd = (r == j && live) ? xj : upd;
          b -= row[j] * xj;
        });
        if (!pivoting) {
          redo = gmax<G>(lbad ? 1.0 : 0.0, lane) > 0.0;
          if (LUP == PFX_LU_ONDEMAND) lu_pivot = lu_pivot || redo;
          if (kCanDefer && redo) {
            // a warm attempt is given up as one that fails (the plain
            // attempt follows in this launch); otherwise the cell is left as
            // it came in, for the pivoting launch to redo
            finish = true;
            if (P.guess_mode == PFX_GUESS_ARRAY_FALLBACK && !fell_back) status = PFX_CELL_SINGULAR;
            else deferred = true;
          }
        }
      }
      const double bad = gmax<G>(status == PFX_CELL_SINGULAR ? 1.0 : 0.0, lane);
      if (bad > 0.0) {
        finish = true;  // singular Jacobian row (reference: MPI_Abort)
        status = PFX_CELL_SINGULAR;
      }
    }

    // ---------------------------------------------------------------
    PFX_PT(7);
    // Newton update (reaction.F90:3454-3470; constraint form :1806-1826)
    double mnew = m, rel = 0.0;
    if (MODE == PFX_MODE_REACT && active && !finish && !redo && floor_pass) {
      // the solve gave the resolution of the free ions at the candidate: a
      // waiting step within twice it is rounding noise, and the iterate is
      // as good as FP64 makes it; otherwise the step is taken now
      floor_pass = false;
      const double frel = gmax<G>(live ? fabs(use_log ? upd : upd / m) : 0.0, lane);
      if (rel_cand <= 2.0 * frel) {
        finish = true;
      } else if (it >= it_cap) {
        finish = true;
        status = PFX_CELL_MAX_ITERATIONS;
      } else {
        mnew = m_cand;
        rel = rel_cand;
      }
    } else if ((MODE == PFX_MODE_REACT || MODE == PFX_MODE_EQUIL) && active && !finish && !redo) {
      const double mprev = m;
      if (use_log) {
        const double u = copysign(1.0, upd) * fmin(fabs(upd), L.max_dlnC);
        if (live) mnew = m * exp(-u);
      } else {
        double ratio = (live && m <= upd) ? fabs(m / upd) : 1e20;
        const double min_ratio = gmin<G>(ratio, lane);
        if (MODE == PFX_MODE_EQUIL ? (min_ratio <= 1.0) : (min_ratio < 1.0)) upd = upd * min_ratio * 0.99;
        if (live) mnew = m - upd;
      }
      rel = gmax<G>(live ? fabs((mnew - mprev) / mprev) : 0.0, lane);
      if (MODE == PFX_MODE_REACT && candidate) {
        // step test: the change this step would make, enlarged by
        // 1/(1 - rho), rho = this step over the last one (at most 10x),
        // bounds the distance to the root also where the iteration
        // contracts only linearly; the iterate itself is returned (its
        // residual passed), the step is not taken.  Where the steps
        // stagnate (rho >= 1/2) a step within twice the FP64 resolution of
        // the free ions is rounding noise: the next pass, at the same
        // iterate, solves for that resolution and decides.
        const double rho = fmin(last_change > 0.0 && last_change < INFINITY ? rel / last_change : 0.0, 0.9);
        if (rel / (1.0 - rho) <= P.rel_change_tol) {
          finish = true;
        } else if (rho >= 0.5) {
          floor_pass = true;
          rel_cand = rel;
          m_cand = mnew;
        } else if (it >= it_cap) {
          finish = true;
          status = PFX_CELL_MAX_ITERATIONS;
        }
      }
    }
    if ((MODE == PFX_MODE_REACT || MODE == PFX_MODE_EQUIL) && active && !finish && !redo && !floor_pass) {
      m = mnew;
      it++;
      if (MODE == PFX_MODE_REACT) {
        last_change = rel;
        if (P.reference_stop && rel < L.max_rel_change_tol) stop_next = true;
      } else {
        // constraint convergence (reaction.F90:1952-1966)
        const double mn = gmin<G>(live ? m : 1.0, lane);
        if (!(mn > 0.0)) {
          finish = true;
          status = PFX_CELL_NONFINITE;
        } else if (it >= 10000 || it >= P.max_iterations) {
          finish = true;
          status = PFX_CELL_MAX_ITERATIONS;
        } else if (rr < L.max_residual_tol && rel < L.max_rel_change_tol) {
          // activity switches on after the first convergence; at least two
          // more iterations follow (reaction.F90:1952-1966).  With activity
          // off the switch is still taken, only the update is skipped.
          if (compute_act && it - nit_on > 1) {
            finish = true;
          } else {
            if (!compute_act) nit_on = it;
            compute_act = true;
          }
        }
      }
    }

    // ---------------------------------------------------------------
    // warm start abandoned: the cell is loaded again at the next pass and
    // repeats the plain method from the step's start (previous state,
    // activity update, then Newton from the fallback guess) with the full
    // iteration budget
    if (MODE == PFX_MODE_REACT && active && finish && status != PFX_CELL_CONVERGED &&
        P.guess_mode == PFX_GUESS_ARRAY_FALLBACK && !fell_back) {
      fell_back = true;
      finish = false;
      reload = true;
      if (r == 0) atomicAdd(P.fallback_count, 1u);
    }

    // ---------------------------------------------------------------
    // a cell the unpivoted kernel did not converge is redone from its
    // inputs by the pivoting launch (the reference's factorisation), so a
    // failure is only reported where partial pivoting fails too
    if (kCanDefer && active && finish && status != PFX_CELL_CONVERGED) deferred = true;
    if (kCanDefer && active && finish && deferred && r == 0) {
      const unsigned int slot = atomicAdd(P.defer, 1u);
      P.defer[1 + slot] = (unsigned int)cell | (fell_back ? 0x80000000u : 0u);
      if (fell_back) P.num_iterations[cell] = it_cap - P.max_iterations;  // the warm attempt's iterations
    }

    // ---------------------------------------------------------------
    // retire the cell: write the state at this iterate
    if (active && finish) {
      const long long c = cell;
      if (deferred) {
        // nothing written: the inputs stay for the pivoting launch
      } else if (MODE == PFX_MODE_EVAL) {
        if (live) {
          P.res_out[c * N + r] = Rr;
#pragma unroll
          for (int k = 0; k < G; k++)
            if (k < N) P.jac_out[c * N * N + (long long)k * N + r] = P.col_scale ? row[k] * SD(S.mv, k) : row[k];
        }
        // the mineral rates at this iterate (the last evaluation of a
        // global-implicit solve leaves those of its solution)
        if (P.mnrl_rate)
          for (int k = r; k < K; k += G) P.mnrl_rate[c * K + k] = SD(S.mrate, k);
      } else {
        if (MODE == PFX_MODE_EQUIL && R > 0) {
          // sorbed totals at the final free ions (reaction.F90:1970-1975)
          // are produced by the aux-var refresh; keep the exchange state
        }
        if (live) {
          P.pri_molal[c * N + r] = m;
          P.total[c * N + r] = tot_out;
          if (P.pri_act_coef) P.pri_act_coef[c * N + r] = gam;
          if (L.sorb_any && P.total_sorb) P.total_sorb[c * N + r] = sorb_out;
        }
        for (int cc = r; cc < M; cc += G) {
          if (P.sec_act_coef) P.sec_act_coef[c * M + cc] = SD(S.gsec, cc);
          if (P.sec_molal) P.sec_molal[c * M + cc] = SD(S.C, cc);
        }
        if (MODE == PFX_MODE_REACT && P.mnrl_rate)
          for (int k = r; k < K; k += G) P.mnrl_rate[c * K + k] = SD(S.mrate, k);
        if (r == 0) {
          for (int q = 0; q < R; q++)
            if (P.ionx_ref) P.ionx_ref[c * R + q] = SD(S.ionx, q);
          if (MODE != PFX_MODE_AUX) {
            if (P.num_iterations) P.num_iterations[c] = it;
            if (P.status) P.status[c] = status;
            if (P.rel_res) P.rel_res[c] = rr;
          }
        }
      }
      cell = -1;
    }
    wave_sync();
#undef P
#undef L
  }
#ifdef PFX_SECTION_TIMERS
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 10; i++) atomicAdd(&g_section_cycles[i], pt_acc[i]);
    atomicAdd(&g_section_cycles[15], 1ull);
  }
#endif
}

// W: minimum wavefronts per SIMD the register allocation must allow (1 =
// as many registers as the code wants; 2, 3 spill what does not fit).  The
// fastest choice depends on the group size and mode and is picked by
// measurement (pfx_tables_tune).  XM: mineral rate prefactors and scale
// factors compiled in (tables that use them launch these instantiations,
// uncapped; the others keep the smaller register footprint).
template <int G, int MODE, int W, int LUP, bool XM>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(W))) void pfx_react_kernel(
    const pfx_kernel_params Pk) {
  (void)Pk;
  react_body<G, MODE, LUP, XM>((uint64_t)(uintptr_t)__builtin_amdgcn_kernarg_segment_ptr());
}

// Several chemistry zones in one persistent launch: workgroups
// [block_start[z], block_start[z+1]) serve zone z, whose parameters (its own
// tables, cells and queue) are zone_params[z]; each workgroup stamps the
// wall clock when it retires, so the host can rebalance the split.
template <bool XM>
__device__ __forceinline__ void zones_body(const pfx_kernel_params *zone_params, const pfx_zone_launch &Z,
                                           uint64_t *finish) {
  int z = 0;
  while (z + 1 < Z.nzones && (int)blockIdx.x >= Z.block_start[z + 1]) z++;
  const uint64_t praw = (uint64_t)(uintptr_t)(zone_params + z);
  switch (Z.G[z]) {
    case 8: react_body<8, PFX_MODE_REACT, PFX_LU_DEFER, XM>(praw); break;
    case 16: react_body<16, PFX_MODE_REACT, PFX_LU_DEFER, XM>(praw); break;
    default: react_body<32, PFX_MODE_REACT, PFX_LU_DEFER, XM>(praw); break;
  }
  __syncthreads();
  if (threadIdx.x == 0 && finish) finish[blockIdx.x] = wall_clock64();
}

// W as for pfx_react_kernel: 1 leaves the registers to the 32-lane path (one
// wavefront per SIMD when such a zone is present), 2 and 3 cap them; XM as
// for pfx_react_kernel (uncapped only)
template <int W, bool XM>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(W))) void pfx_react_zones_kernel(
    const pfx_kernel_params *zone_params, const pfx_zone_launch Z, uint64_t *finish) {
  zones_body<XM>(zone_params, Z, finish);
}
template __global__ void pfx_react_zones_kernel<1, false>(const pfx_kernel_params *, const pfx_zone_launch,
                                                          uint64_t *);
template __global__ void pfx_react_zones_kernel<2, false>(const pfx_kernel_params *, const pfx_zone_launch,
                                                          uint64_t *);
template __global__ void pfx_react_zones_kernel<3, false>(const pfx_kernel_params *, const pfx_zone_launch,
                                                          uint64_t *);
template __global__ void pfx_react_zones_kernel<1, true>(const pfx_kernel_params *, const pfx_zone_launch,
                                                         uint64_t *);

#define PFX_INST_LU(G, MODE, W, LUP, XM) \
  template __global__ void pfx_react_kernel<G, MODE, W, LUP, XM>(const pfx_kernel_params);
#define PFX_INST(G)                                        \
  PFX_INST_LU(G, PFX_MODE_REACT, 1, PFX_LU_DEFER, false)    \
  PFX_INST_LU(G, PFX_MODE_REACT, 2, PFX_LU_DEFER, false)    \
  PFX_INST_LU(G, PFX_MODE_REACT, 3, PFX_LU_DEFER, false)    \
  PFX_INST_LU(G, PFX_MODE_REACT, 1, PFX_LU_ONDEMAND, false) \
  PFX_INST_LU(G, PFX_MODE_REACT, 2, PFX_LU_ONDEMAND, false) \
  PFX_INST_LU(G, PFX_MODE_REACT, 3, PFX_LU_ONDEMAND, false) \
  PFX_INST_LU(G, PFX_MODE_REACT, 1, PFX_LU_PIVOT, false)    \
  PFX_INST_LU(G, PFX_MODE_EVAL, 1, PFX_LU_PIVOT, false)     \
  PFX_INST_LU(G, PFX_MODE_EQUIL, 1, PFX_LU_PIVOT, false)    \
  PFX_INST_LU(G, PFX_MODE_EQUIL, 2, PFX_LU_PIVOT, false)    \
  PFX_INST_LU(G, PFX_MODE_EQUIL, 3, PFX_LU_PIVOT, false)    \
  PFX_INST_LU(G, PFX_MODE_AUX, 1, PFX_LU_PIVOT, false)      \
  PFX_INST_LU(G, PFX_MODE_REACT, 1, PFX_LU_DEFER, true)     \
  PFX_INST_LU(G, PFX_MODE_REACT, 1, PFX_LU_ONDEMAND, true)  \
  PFX_INST_LU(G, PFX_MODE_REACT, 1, PFX_LU_PIVOT, true)     \
  PFX_INST_LU(G, PFX_MODE_EVAL, 1, PFX_LU_PIVOT, true)
PFX_INST(8)
PFX_INST(16)
PFX_INST(32)

extern "C" int pfx_scr_doubles(int M, int G, int K, int xm) { return scr_doubles(M, G, K, xm != 0); }

#ifdef PFX_SECTION_TIMERS
// Profiling build only (no counterpart in the reference): read and clear the
// per-section cycle sums of the current device; out[0..9] the sections,
// out[15] the number of wavefronts that added to them.
extern "C" int pfx_section_cycles(unsigned long long *out, int32_t clear) {
  if (!out) return PFX_E_ARG;
  if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_section_cycles), sizeof(unsigned long long) * 16, 0,
                          hipMemcpyDeviceToHost) != hipSuccess)
    return PFX_E_DEVICE;
  if (clear) {
    unsigned long long z[16] = {0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_section_cycles), z, sizeof(z), 0, hipMemcpyHostToDevice) != hipSuccess)
      return PFX_E_DEVICE;
  }
  return 0;
}
#endif
