// Device-side table layout and kernel parameters shared by the HIP kernels
// and the host code of libpfx_chem.so.
#pragma once
#include <stdint.h>

#include "../../include/pfx_chem.h"

#define PFX_LOG_TO_LN 2.30258509299       /* pflotran_constants.F90:48 */
#define PFX_IDEAL_GAS 8.31446             /* pflotran_constants.F90:53 */

// Tables derived once per reaction network by pfx_tables_create.  All arrays
// live in one device allocation; offsets are in bytes from its base.
//
// Row-per-lane layout: a cell occupies a group of G lanes, lane r owns
// primary species r (row r of the local Jacobian).  Per-lane tables are
// stored [entry][G] so the G lanes of a group read consecutive words.
struct pfx_dev_layout {
  int32_t N, M, K, R, G, S;     // sizes; S = species slots per complex/mineral
  int32_t Q;                    // complex rounds per lane = ceil(M/G)
  int32_t maxdeg;               // max complexes touching one primary
  // off-diagonal co-occurrence lists of dT_r/dm_k in chunks of 4 columns:
  // chunk q owns rows [offq_start[q], offq_start[q+1]) of the chunked
  // tables, one row per slot, as many slots as its longest list
  int32_t offq_start[PFX_MAX_PRIMARY / 4 + 1];
  int32_t has_logk_coef, has_mnrl_coef;
  int32_t ionx_ncat[PFX_MAX_IONX_RXN];
  int32_t ionx_zflag[PFX_MAX_IONX_RXN];
  int32_t ionx_to_surf[PFX_MAX_IONX_RXN];
  double ionx_CEC[PFX_MAX_IONX_RXN];
  double ionx_ref_k[PFX_MAX_IONX_RXN];
  int32_t ionx_ref_lane[PFX_MAX_IONX_RXN];
  double debyeA, debyeB, debyeBdot;
  int32_t use_log, act_freq, act_alg, num_jac;
  double max_dlnC, max_rel_change_tol, max_residual_tol;
  int32_t mnrl_flags;           // bit0 threshold, bit1 limiter, bit2 power, bit3 Temkin,
                                // bit4 rate prefactors, bit5 min scale factor
  int32_t npref, npref_spec;    // P, Q of the prefactor tables (0 without prefactors)
  int32_t kd_any;               // some primary carries isotherm (K_D) sorption
  int32_t sorb_any;             // sorbed totals present: ion exchange or isotherms
  double soil_density;          // kg/m^3, for K_D scaled by a mineral's volume fraction

  // byte offsets into the table blob (doubles unless noted)
  int64_t o_Z, o_a0;                 // [G]
  int64_t o_cspec;                   // int32 [M][S]  (-1 = empty slot)
  int64_t o_cnu;                     // [M][S]
  int64_t o_cZ, o_ca0;               // [M]
  int64_t o_clogK;                   // [M] log10 K at the reference temperature
  int64_t o_clogKcoef;               // [M][5]
  int64_t o_ch2o;                    // [M]
  int64_t o_adj_c;                   // int32 [maxdeg][G], -> M = zero slot
  int64_t o_adj_nu;                  // [maxdeg][G]
  int64_t o_offq_c;                  // uint16 [rows][G][4] complex (M = zero slot)
  int64_t o_offq_coef;               // [rows][2][G][2] nu_r * nu_k
  int64_t o_mspec;                   // int32 [K][S]
  int64_t o_mnu;                     // [K][S]
  int64_t o_mnuG;                    // [K][G] dense stoichiometry per lane
  int64_t o_mlogK, o_mlogKcoef, o_mh2o;
  int64_t o_mk, o_mmolvol, o_mEa, o_mthr, o_mlim, o_mpow, o_mtemkin;  // [K]
  int64_t o_mmsf;                    // [K] min scale factor (1 where unset)
  int64_t o_mpn;                     // int32 [K] rate prefactors of the mineral
  int64_t o_mpid;                    // int32 [K][P][Q] primary id of each prefactor species, -1 empty
  int64_t o_mpalpha, o_mpbeta, o_mpatten;  // [K][P][Q]
  int64_t o_mprate, o_mpEa;          // [K][P] rate constant (mol/m^2/s), activation energy (J/mol)
  int64_t o_ionx_k;                  // [R][G] selectivity of the lane's cation (0 if none)
  int64_t o_ionx_role;               // int32 [R][G] cation slot (0 = reference) or -1
  int64_t o_kd_type;                 // int32 [G] PFX_SORPTION_* of the lane's primary (0 = none)
  int64_t o_kd_mnrl;                 // int32 [G] mineral scaling the lane's K_D, or -1
  int64_t o_kd_coef, o_kd_b, o_kd_invn;  // [G] K_D, Langmuir b, 1/n (Freundlich)
  int64_t blob_bytes;
  int64_t lds_table_bytes;           // prefix of the blob staged into LDS
};

enum { PFX_MODE_REACT = 0, PFX_MODE_EVAL = 1, PFX_MODE_EQUIL = 2, PFX_MODE_AUX = 3 };

// block split of a zoned launch (pfx_react_zones)
struct pfx_zone_launch {
  int32_t nzones;
  int32_t block_start[PFX_MAX_ZONES + 1];
  int32_t G[PFX_MAX_ZONES];
};

struct pfx_kernel_params {
  pfx_dev_layout L;
  const unsigned char *tables;    // device blob
  // cells
  int64_t ncells;
  double *total, *total_sorb, *pri_molal, *pri_act_coef, *sec_act_coef, *sec_molal, *ionx_ref;
  const double *mnrl_volfrac, *mnrl_area;
  double *mnrl_rate;
  const double *temperature, *porosity, *saturation, *volume, *den_kg, *guess;
  int32_t *num_iterations, *status;
  double *rel_res;
  // options
  double dt, rel_tol, default_guess;
  double def_porosity, def_saturation, def_volume, def_den;
  int32_t max_iterations, extra_iterations, guess_mode, reference_stop;
  int32_t fallback_after;       // warm-attempt iteration cap (PFX_GUESS_ARRAY_FALLBACK)
  int32_t jac_flags;            // PFX_JAC_*
  double rel_change_tol;        // step test at the accepted iterate (0 = off)
  int32_t fallback_previous;    // a failed warm attempt restarts from the previous free ions
  int32_t reserved0;
  unsigned int *fallback_count; // cells restarted by the fallback
  double *res_out, *jac_out;    // residual/Jacobian outputs
  double *dtot_out;             // EVAL: dT/dm blocks [ncells][N][N] column-major (kg/L), may be NULL
  double *tot_eval_out;         // EVAL: totals at the free ions [ncells][N] (mol/L), may be NULL
  int32_t col_scale;            // EVAL: scale Jacobian and dT/dm columns by m (log unknowns)
  // constraint speciation
  const double *constraint_conc;   // [ncells*N]
  int32_t ctype[PFX_MAX_PRIMARY];  // PFX_CONSTRAINT_* per component
  int32_t has_charge;              // some component carries the charge balance
  int32_t has_sorb;                // some component is constrained by its sorbed total
  const pfx_constraint_rxn *crxn;  // [N] mineral / gas constraint reactions (device), or null
  int32_t molal;                   // constraint concentrations in mol/kg
  int32_t aux_passes;              // activity passes of the aux-var refresh
  // work queue
  unsigned int *queue;          // global cell cursor
  int32_t chunk;                // cells taken per fetch (at most)
  int32_t waves;                // wavefronts draining the queue: a fetch takes
                                // a quarter of their fair share of what is
                                // left, at most chunk (0 = always chunk)
  // cells whose unpivoted factorisation broke down: defer[0] their count,
  // defer[1..] their indices (written by the unpivoted reaction kernel,
  // worked through by the pivoting one); null where the kernel pivots
  // itself
  unsigned int *defer;
};

// Factorisation policy of a reaction-step kernel (template argument):
//   PFX_LU_DEFER     unpivoted; a cell whose factorisation breaks down, or
//                    that does not converge, is left untouched and listed in
//                    defer[] for a pivoting launch
//   PFX_LU_PIVOT     partial pivoting; with defer[] given the launch works
//                    through the listed cells instead of the batch
//   PFX_LU_ONDEMAND  unpivoted; such a cell pivots from its next pass on
#define PFX_LU_DEFER 0
#define PFX_LU_PIVOT 1
#define PFX_LU_ONDEMAND 2

// Sizes of a network held by an opaque pfx_tables (defined in pfx_api.hip)
// for the other translation units of the library.
struct pfx_bsr;  // include/pfx_gi_solver.h
namespace pfx_internal {
void table_sizes(const pfx_tables *t, int *N, int *M, int *K, int *R);
void table_max_dlnC(const pfx_tables *t, double *max_dlnC);
int gi_scratch(pfx_tables *t, int64_t ncells, double **rnf, double **jnf, double **tot, double **dtot);
int gi_evaluate_cells(pfx_tables *t, pfx_cells *cells, const pfx_react_opts *o, int col_scale, double *rnf,
                      double *jnf, double *tot, double *dtot);
// shape of a pfx_bsr handle (defined in pfx_krylov.hip)
const int32_t *bsr_row_ptr(const ::pfx_bsr *h);
int64_t bsr_rows(const ::pfx_bsr *h);
int32_t bsr_block(const ::pfx_bsr *h);
int64_t bsr_nnzb(const ::pfx_bsr *h);
// right-hand-side and solution vectors [nrows][C] of a handle's
// several-right-hand-side solve (allocated on first use)
int bsr_multi_io(::pfx_bsr *h, int C, double **B, double **X);
// allocate what a reaction step of these cells needs (before a graph capture)
int tables_reserve(pfx_tables *t, const pfx_cells *cells);
// arm pfx_tables_stats for a replayed stream-ordered reaction step of n cells
void tables_rearm_stats(pfx_tables *t, int64_t n, hipStream_t stream, bool fallback);
// event-timed launch bracket; the time goes to the tables' last_ms[slot]
int timed_launch_begin(pfx_tables *t, hipStream_t s);
int timed_launch_end(pfx_tables *t, hipStream_t s, int slot);
}
