/*
 * pfx_chem.h -- C-ABI of the MI355X per-cell geochemistry library.
 *
 * This is the drop-in boundary for PFLOTRAN's per-cell reaction step.  Every
 * type is plain C (fixed-width integers, doubles, pointers), so the Fortran
 * host binds it with ISO_C_BINDING (see pflotran_amd/fortran/pfx_chem_mod.F90).
 *
 * Conventions (shared by every entry point):
 *   - return 0 on success, >0 = number of cells that did not converge
 *     (details in the per-cell status array), <0 = hard error (PFX_E_*).
 *     Nothing ever aborts the process (reference: ludcmp calls MPI_Abort,
 *     src/pflotran/utility.F90:423-432; RReact can `stop`, reaction.F90:3487).
 *   - species / complex / mineral ids are 0-based (the reference's are
 *     1-based with count in row 0, reaction_aux.F90:230-231).
 *   - per-cell vectors are cell-major ("AoS"): value j of cell c lives at
 *     [c*len + j], exactly the layout of the transport Vec tran_xx
 *     (reactive_transport.F90:1701-1709).  Jacobian blocks are column-major
 *     N x N per cell, the layout of the Fortran J(ncomp,ncomp) and of PETSc
 *     BAIJ blocks with MAT_ROW_ORIENTED=FALSE (discretization.F90:857).
 *   - all arithmetic is FP64.
 *   - cell arrays are DEVICE pointers (hipMalloc'd or torch CUDA tensors);
 *     the call is ordered on `stream` (NULL = default stream) and returns
 *     after the stream has drained, with the statistics filled in.
 */
#ifndef PFX_CHEM_H
#define PFX_CHEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFX_ABI_VERSION 6

/* largest primary-species count handled by one cell group (row per lane) */
#define PFX_MAX_PRIMARY 32
/* largest number of aqueous complexes per table */
#define PFX_MAX_CPLX 256
/* largest species count in one complex or mineral reaction (H2O excluded) */
#define PFX_MAX_RXN_SPEC 8
#define PFX_MAX_MINERALS 16
#define PFX_MAX_IONX_RXN 4
#define PFX_MAX_IONX_CATION 8
/* rate prefactors per kinetic mineral, and species per prefactor */
#define PFX_MAX_PREFACTORS 4
#define PFX_MAX_PREFACTOR_SPECIES 4
#define PFX_MAX_ZONES 8

/* error codes */
#define PFX_E_ARG      (-1)   /* bad argument / inconsistent sizes */
#define PFX_E_DEVICE   (-2)   /* HIP runtime error */
#define PFX_E_NOMEM    (-3)
#define PFX_E_UNSUP    (-4)   /* table shape beyond the compiled limits */

/* isotherm (K_D) sorption types: pflotran_constants.F90:94-96 SORPTION_* */
#define PFX_SORPTION_LINEAR     1   /* S = K_D m                          */
#define PFX_SORPTION_LANGMUIR   2   /* S = K_D m b / (1 + K_D m)          */
#define PFX_SORPTION_FREUNDLICH 3   /* S = K_D m^(1/n)                    */
#define PFX_MAX_KD_RXN 32

/* activity-coefficient update frequency: reaction_aux.F90 ACT_COEF_FREQUENCY_* */
#define PFX_ACT_OFF          0
#define PFX_ACT_TIMESTEP     1   /* once per reaction step (LAG), reaction.F90:3407 */
#define PFX_ACT_NEWTON_ITER  2   /* every Newton iteration, reaction.F90:3415-3418 */
/* activity-coefficient algorithm: ACT_COEF_ALGORITHM_* */
#define PFX_ACT_ALG_LAG      0   /* reaction.F90:3992-4051 */
#define PFX_ACT_ALG_NEWTON   1   /* reaction.F90:3846-3991 */

/* per-cell status */
#define PFX_CELL_CONVERGED      0
#define PFX_CELL_MAX_ITERATIONS 1
#define PFX_CELL_NONFINITE      2
#define PFX_CELL_SINGULAR       3

/* constraint types of pfx_equilibrate_batch (transport_constraint.F90
   CONSTRAINT_* meanings) */
#define PFX_CONSTRAINT_TOTAL 0   /* total component (mol/L or mol/kg)      */
#define PFX_CONSTRAINT_FREE  1   /* free ion concentration                 */
#define PFX_CONSTRAINT_LOG   2   /* log10 free ion concentration           */
#define PFX_CONSTRAINT_PH    3   /* pH (component must be H+)              */
#define PFX_CONSTRAINT_CHARGE 4  /* charge balance: sum z_j T_j = 0; the value
                                    is the starting free concentration      */
#define PFX_CONSTRAINT_TOTAL_SORB 5 /* sorbed (ion-exchange) total, mol/m^3 bulk
                                       (reaction.F90:1490-1492, 1608-1614);
                                       needs an exchange site                */
#define PFX_CONSTRAINT_MINERAL 6 /* equilibrium with a mineral, ln(Q/K) = 0
                                    (reaction.F90:1694-1718); the value is the
                                    starting free concentration; the reaction
                                    is constraint_rxn[j]                      */
#define PFX_CONSTRAINT_GAS     7 /* equilibrium with a gas at the value's
                                    partial pressure (bar; a value <= 0 is
                                    log10 of it), ln(Q/K) = ln p
                                    (reaction.F90:1531-1534, 1720-1745); the
                                    reaction is constraint_rxn[j]             */

/* The reaction of a mineral or gas constraint row: the primaries it holds
   (0-based ids; water activity is 1 and drops out), their stoichiometry,
   log10 K at the reference temperature and, with has_logK_coef, the 5
   Maier-Kelly coefficients used for cells that carry a temperature
   (mineral_reaction%mnrl_logK / gas%paseqlogK after
   RUpdateTempDependentCoefs, reaction.F90:1472-1474). */
typedef struct pfx_constraint_rxn {
  int32_t nspec;
  int32_t has_logK_coef;
  int32_t specid[PFX_MAX_RXN_SPEC];
  double stoich[PFX_MAX_RXN_SPEC];
  double logK;
  double logK_coef[5];
} pfx_constraint_rxn;

/* initial Newton iterate */
#define PFX_GUESS_PREVIOUS  0   /* previous free ions (RReact's guess, reaction.F90:3370) */
#define PFX_GUESS_ARRAY     1   /* caller array cells->guess (e.g. learned warm start) */
#define PFX_GUESS_DEFAULT   2   /* constant opts->default_guess (1e-9 molal, reaction.F90:1481) */
#define PFX_GUESS_ARRAY_FALLBACK 3 /* cells->guess; a cell not converged after
                                     opts->fallback_after iterations (or that
                                     fails) restarts from opts->default_guess
                                     with the full max_iterations budget */

/*
 * Flattened reaction_type (reaction_aux.F90:142-335, mineral_type
 * reaction_mineral_aux.F90:77-128).  All pointers are HOST pointers read once
 * by pfx_tables_create; the library keeps its own device copy.
 * Row-major 2-D tables with the stated leading dimensions.
 */
typedef struct pfx_reaction_desc {
  int32_t naqcomp;        /* N: primary aqueous species                       */
  int32_t neqcplx;        /* M: aqueous complexes                             */
  int32_t nkinmnrl;       /* K: kinetic minerals                              */
  int32_t neqionxrxn;     /* R: ion-exchange reactions (sites)                */
  int32_t max_rxn_spec;   /* leading dim of *specid (+1) and *stoich tables   */
  int32_t max_cation;     /* leading dim of eqionx tables (+1 for specid)     */

  const double *primary_spec_Z;     /* [N]                                    */
  const double *primary_spec_a0;    /* [N] Debye-Hueckel ion size (Angstrom)  */

  /* aqueous complexes: reaction%eqcplx* */
  const int32_t *eqcplxspecid;      /* [M][max_rxn_spec+1]; [i][0] = count    */
  const double *eqcplxstoich;       /* [M][max_rxn_spec]                      */
  const double *eqcplxh2ostoich;    /* [M] (0 when water absent)              */
  const double *eqcplx_Z;           /* [M]                                    */
  const double *eqcplx_a0;          /* [M]                                    */
  const double *eqcplx_logK;        /* [M] log10 K at the reference temperature */
  const double *eqcplx_logKcoef;    /* [M][5] Maier-Kelly fit, NULL if unused  */

  /* kinetic minerals: mineral%kinmnrl* */
  const int32_t *kinmnrlspecid;     /* [K][max_rxn_spec+1]                    */
  const double *kinmnrlstoich;      /* [K][max_rxn_spec]                      */
  const double *kinmnrlh2ostoich;   /* [K]                                    */
  const double *kinmnrl_logK;       /* [K]                                    */
  const double *kinmnrl_logKcoef;   /* [K][5] or NULL                         */
  const double *kinmnrl_rate_constant;      /* [K] mol/m^2/s                  */
  const double *kinmnrl_molar_vol;          /* [K] m^3/mol                    */
  const double *kinmnrl_activation_energy;  /* [K] J/mol, <=0: none; or NULL  */
  const double *kinmnrl_affinity_threshold; /* [K] or NULL                    */
  const double *kinmnrl_rate_limiter;       /* [K] or NULL                    */
  const double *kinmnrl_affinity_power;     /* [K] or NULL (beta, else 1)     */
  const double *kinmnrl_Temkin_const;       /* [K] or NULL (sigma, else 1)    */

  /* equilibrium ion exchange: reaction%eqionx_rxn_* */
  const int32_t *eqionx_rxn_cationid; /* [R][max_cation+1]; [r][0] = count,
                                         [r][1] = reference cation            */
  const double *eqionx_rxn_k;         /* [R][max_cation]                      */
  const double *eqionx_rxn_CEC;       /* [R] eq/m^3 bulk                      */
  const int32_t *eqionx_rxn_to_surf;  /* [R] kinetic mineral id or -1; NULL = none */

  /* Debye-Hueckel parameters, fixed at the reference temperature
     (reaction_database.F90:896-990) */
  double debyeA, debyeB, debyeBdot;

  /* Newton controls (reaction_aux.F90:569-571) */
  int32_t use_log_formulation;
  int32_t act_coef_update_frequency;  /* PFX_ACT_*     */
  int32_t act_coef_update_algorithm;  /* PFX_ACT_ALG_* */
  int32_t numerical_jacobian;         /* 1: the reaction part of the per-cell
                                         Jacobian by forward differences,
                                         perturbation 1e-5 m_j (NUMERICAL_JACOBIAN,
                                         RReactionDerivative reaction.F90:3625-3684);
                                         applies to pfx_residual_jacobian and the
                                         global-implicit assembly, as in the
                                         reference; the reaction step's Newton
                                         stays analytic (reaction.F90:3438-3439) */
  double max_dlnC;                    /* 5                         */
  double max_relative_change_tolerance;  /* 1e-6                   */
  double max_residual_tolerance;         /* 1e-12                  */

  /* equilibrium isotherm sorption: reaction%eqkd* (reaction.F90:505-600),
     evaluated by RTotalSorbKD (:4220-4301); at most one reaction per
     primary species.  The sorbed totals add to those of ion exchange. */
  int32_t neqkdrxn;                   /* D: isotherm reactions (0: none)      */
  int32_t reserved_kd;
  const int32_t *eqkdspecid;          /* [D] primary id                        */
  const int32_t *eqkdtype;            /* [D] PFX_SORPTION_*                    */
  const double *eqkddistcoef;         /* [D] K_D: kg water/m^3 bulk, or with a
                                         mineral mL water/g soil (L/kg)       */
  const double *eqkdlangmuirb;        /* [D] Langmuir b (mol/m^3 bulk) or NULL */
  const double *eqkdfreundlichn;      /* [D] Freundlich n or NULL              */
  const int32_t *eqkdmineral;         /* [D] kinetic mineral whose volume
                                         fraction scales K_D, -1 none; NULL = none */
  double soil_particle_density;       /* kg/m^3 (ROCK_DENSITY), with eqkdmineral */

  /* Rate prefactors of the kinetic minerals (mineral%kinmnrl_num_prefactors,
     kinmnrl_prefactor_id, kinmnrl_pref_*; RKineticMineral
     reaction_mineral.F90:743-786 and its derivative :887-944).  A mineral
     with prefactors takes as its rate constant
       sum_p k_p A_p(T) prod_s a_s^alpha_ps / (1 + atten_ps a_s^beta_ps),
     a_s the activity of primary species s and A_p the Arrhenius factor of
     the prefactor's activation energy (none where <= 0); a mineral without
     uses kinmnrl_rate_constant.  Prefactor species must be primaries (the
     reference's negative ids for complexes are not supported). */
  int32_t max_prefactors;             /* P: leading dim; 0 = no prefactors    */
  int32_t max_prefactor_species;      /* Q: leading dim                        */
  const int32_t *kinmnrl_num_prefactors;        /* [K]                         */
  const int32_t *kinmnrl_prefactor_id;          /* [K][P][Q+1]; [k][p][0] = count,
                                                   then 0-based primary ids   */
  const double *kinmnrl_pref_alpha;             /* [K][P][Q]                   */
  const double *kinmnrl_pref_beta;              /* [K][P][Q]                   */
  const double *kinmnrl_pref_atten_coef;        /* [K][P][Q]                   */
  const double *kinmnrl_pref_rate;              /* [K][P] mol/m^2/s            */
  const double *kinmnrl_pref_activation_energy; /* [K][P] J/mol, or NULL       */
  /* kinmnrl_min_scale_factor f (reaction_mineral.F90:704-716, 800-802,
     848-858): affinity 1 - (Q/K)^(1/(f sigma)) and the rate divided by f;
     NULL = none */
  const double *kinmnrl_min_scale_factor;       /* [K] or NULL                 */

  /* Lanes per cell group in the per-cell kernels (no counterpart in the
     reference; a tuning knob): 0 picks the smallest of 8, 16, 32 that holds
     N; 8, 16 or 32 (>= N) forces it: a group is the LU's row tile, so a wider
     group runs fewer cells per wavefront with idle padding rows. */
  int32_t group_lanes;
  int32_t reserved_lanes;
} pfx_reaction_desc;

/* Per-cell state, device pointers, cell-major.  In = read, Out = written. */
typedef struct pfx_cells {
  int64_t ncells;
  double *total;          /* [ncells*N] In: transported totals tran_xx (mol/L);
                             Out: totals at the solution                     */
  double *total_sorb;     /* [ncells*N] In: sorbed totals of the previous state
                             (mol/m^3 bulk); Out: at the solution. NULL if R==0 */
  double *pri_molal;      /* [ncells*N] In: previous free ions (molal);
                             Out: converged free ions                         */
  double *pri_act_coef;   /* [ncells*N] In/Out; NULL -> 1 (activity off)       */
  double *sec_act_coef;   /* [ncells*M] In/Out; NULL -> 1                      */
  double *sec_molal;      /* [ncells*M] Out, may be NULL                       */
  double *ionx_ref_sorbed;/* [ncells*R] In/Out eqionx_ref_cation_sorbed_conc   */
  const double *mnrl_volfrac; /* [ncells*K]                                  */
  const double *mnrl_area;    /* [ncells*K] m^2/m^3 bulk                      */
  double *mnrl_rate;      /* [ncells*K] Out mol/m^3/s at the returned state (also
                             from the residual evaluations), may be NULL     */
  const double *temperature;  /* [ncells] deg C; NULL -> table log K (isothermal) */
  const double *porosity;     /* [ncells] or NULL -> opts default             */
  const double *saturation;   /* [ncells] or NULL                              */
  const double *volume;       /* [ncells] m^3 or NULL                          */
  const double *den_kg;       /* [ncells] kg/m^3 or NULL                       */
  const double *guess;        /* [ncells*N] used with PFX_GUESS_ARRAY         */
  int32_t *num_iterations;    /* [ncells] Out, may be NULL                     */
  int32_t *status;            /* [ncells] Out, may be NULL                     */
  double *rel_res;            /* [ncells] Out final scaled residual, may be NULL */
} pfx_cells;

typedef struct pfx_react_opts {
  double dt;              /* reaction step (s)                                 */
  double rel_tol;         /* stop when rel_res <= rel_tol (1e-10)             */
  int32_t max_iterations; /* Newton cap per cell                               */
  int32_t extra_iterations; /* Newton updates taken after rel_tol is met       */
  int32_t guess_mode;     /* PFX_GUESS_*                                      */
  int32_t reference_stop; /* also stop on max|R|<max_residual_tolerance or
                             max|dm/m|<max_relative_change_tolerance (RReact) */
  double default_guess;   /* PFX_GUESS_DEFAULT value (molal)                  */
  double porosity, saturation, volume, den_kg; /* used when arrays are NULL   */
  void *stream;           /* hipStream_t, NULL = default stream               */
  int32_t fallback_after; /* PFX_GUESS_ARRAY_FALLBACK: iterations of the warm
                             attempt (0 = max_iterations)                     */
  int32_t jacobian_flags; /* PFX_JAC_*; 0 = the reference's Jacobian          */
  double rel_change_tol;  /* > 0: a cell converges only when, besides
                             rel_res <= rel_tol, the last Newton update moved
                             no free ion by more than this relative amount
                             (max|dm/m|, the quantity of the reference's
                             relative-change stop, reaction.F90:3473-3476,
                             tightened); an attempt that has taken no update
                             yet takes one.  0 = off                         */
  int32_t fallback_previous; /* PFX_GUESS_ARRAY_FALLBACK: 1 restarts a failed
                             warm attempt from the previous free ions (RReact's
                             own starting point, reaction.F90:3370), 0 from
                             opts->default_guess                             */
  int32_t stream_ordered; /* 1: pfx_react_batch enqueues its launches on
                             `stream` and returns without waiting; the stats
                             argument is ignored and pfx_tables_stats reads
                             the statistics of the last call later.  No
                             host synchronisation, no events: such a call
                             can be captured into a HIP graph.  0: returns
                             after the stream has drained, stats filled    */
} pfx_react_opts;

/* pfx_react_opts.jacobian_flags.  PFX_JAC_CAPPED_QK: where ln(Q/K) of a
   kinetic mineral exceeds the reference's cap 6.90776 (Q/K held at 1e3,
   reaction_mineral.F90:699-703) the rate does not change with m, and its
   derivative is left out.  The reference keeps the uncapped derivative there
   (reaction_mineral.F90:860-875), with which Newton converges only linearly
   for such cells.  The residual, and so the solution, are unchanged.  No
   counterpart in the reference. */
#define PFX_JAC_CAPPED_QK 1
/* The capped derivative is used for the first PFX_JAC_CAPPED_ITERS Newton
   iterations of an attempt, the reference's after them: the residual has
   a kink at the cap, across which Newton with the capped derivative can
   cycle (rare cells); the reference's derivative, larger there, does not. */
#ifndef PFX_JAC_CAPPED_ITERS
#define PFX_JAC_CAPPED_ITERS 12
#endif

typedef struct pfx_stats {
  int64_t ncells;
  int64_t converged;
  int64_t total_iterations;
  int32_t max_iterations;
  int32_t min_iterations;
  double max_rel_res;
  double kernel_ms;       /* device time of the batch (HIP events)             */
  int64_t iteration_histogram[32]; /* [min(it,31)]                             */
  int64_t fallbacks;      /* cells restarted from the default guess            */
} pfx_stats;

typedef struct pfx_tables pfx_tables;   /* opaque, owns device copies */

/* Library version/ABI check. */
int pfx_abi_version(void);

/* Flattened reaction_type -> device tables.
   Replaces the reaction_type consumed by RReact/RTAuxVarCompute
   (reaction_aux.F90:142-335; setup BasisInit reaction_database.F90:772). */
int pfx_tables_create(const pfx_reaction_desc *desc, pfx_tables **out);
int pfx_tables_destroy(pfx_tables *tables);

/* Register-cap variant of the per-cell kernels (no counterpart in the
   reference; an autotuning knob): kernel = PFX_KERNEL_*; variant 0 gives
   the kernel the registers its code wants, 1 and 2 cap them at 2 and 3
   wavefronts per SIMD (the compiler spills the rest).  Only the Newton
   kernels (REACT, EQUIL) have capped variants.  variant < 0 queries.
   Returns the variant (query) or 0, or PFX_E_ARG.  Defaults, from
   measurement on MI355X: 2 for groups of up to 16 lanes, 1 for 32. */
#define PFX_KERNEL_REACT 0   /* pfx_react_batch        */
#define PFX_KERNEL_EVAL  1   /* pfx_residual_jacobian  */
#define PFX_KERNEL_EQUIL 2   /* pfx_equilibrate_batch  */
#define PFX_KERNEL_AUX   3   /* pfx_auxvar_batch       */
int pfx_tables_tune(pfx_tables *tables, int32_t kernel, int32_t variant);

/* Device time (ms, HIP events) of the launches of the last call made with
   these tables: ms[0] the per-cell kernel, ms[1] the global-implicit
   assembly kernel (0 if the call had none).  Profiling aid; no counterpart
   in the reference (its PetscLogEvent timers play this role). */
int pfx_tables_timing(pfx_tables *tables, double *ms);

/* Batched operator-split reaction step: RTReact's cell loop
   (reactive_transport.F90:1697-1724) around RReact (reaction.F90:3322-3511),
   solving the dt-consistent per-cell residual of the global-implicit path
   (reactive_transport.F90:2539-2587, 2735-2758).  One persistent launch. */
int pfx_react_batch(pfx_tables *tables, pfx_cells *cells,
                    const pfx_react_opts *opts, pfx_stats *stats);

/* No counterpart in the reference.  pfx_react_batch for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc): the work-queue counters, the
   statistics reduction and, when cells->num_iterations, status or rel_res is
   NULL, the per-cell iteration counts, status and scaled residuals live in
   a fixed device array of the library (PFX_NOALLOC_SCRATCH_BYTES on each
   GPU, held until the process ends; its address is taken at
   pfx_tables_create), never in memory allocated by the call, and so does
   the list of cells left to the pivoting launch.  Limit: with any of those
   per-cell outputs NULL, at most PFX_NOALLOC_MAX_CELLS cells (PFX_E_ARG
   beyond); larger batches (per-cell outputs given) factor with pivoting
   switched on per cell inside one launch instead of a follow-up launch,
   which gives the same converged cells up to rounding.  Every call on one
   GPU shares the array, which keeps the last call's contents: calls that
   overlap on one GPU race.  opts->stream_ordered must be 0 (the statistics
   are read back before the call returns).  Otherwise identical to
   pfx_react_batch. */
#define PFX_NOALLOC_MAX_CELLS (1 << 20)
#define PFX_NOALLOC_SCRATCH_BYTES (1280 + 20 * PFX_NOALLOC_MAX_CELLS)
int pfx_react_batch_noalloc(pfx_tables *tables, pfx_cells *cells, const pfx_react_opts *opts, pfx_stats *stats);

/* Statistics of the last pfx_react_batch made with opts->stream_ordered = 1
   on these tables: waits for that call's stream, fills stats (kernel_ms is 0:
   a stream-ordered call records no events) and returns the number of cells
   that did not converge, as pfx_react_batch does.  PFX_E_ARG when no such
   call is pending (each is read once).  The reference's counterpart is the
   per-cell num_iterations that RTReact sums after its loop
   (reactive_transport.F90:1720-1724). */
int pfx_tables_stats(pfx_tables *tables, pfx_stats *stats);

/* Per-cell residual and analytic Jacobian at the given free ions, without a
   Newton solve: RTAuxVarCompute + RTAccumulation(Derivative) +
   RAccumulationSorb(Derivative) + RReaction(derivative=TRUE)
   (reaction.F90:3419-3441), divided by dt as in RTResidualNonFlux /
   RTJacobianNonFlux (reactive_transport.F90:2564, 3342-3465).
   residual: [ncells*N] mol/s; jacobian: [ncells*N*N] column-major blocks
   (dR/dm, kg water/s).  The activity coefficients in `cells` are used as
   given (no update).  fixed_accum: [ncells*N] mol, the accumulation of the
   old state, may be NULL (then computed from cells->total/total_sorb). */
int pfx_residual_jacobian(pfx_tables *tables, pfx_cells *cells,
                          const pfx_react_opts *opts,
                          double *residual, double *jacobian);

/* Batched constraint speciation: ReactionEquilibrateConstraint
   (reaction.F90:1308-2046) for every cell, as called per cell by
   CondControlAssignTranInitCond (condition_control.F90:725-750).
   constraint_type[N] (PFX_CONSTRAINT_*) is shared by all cells;
   constraint_conc[ncells*N] (device) holds each cell's values;
   constraint_rxn[N] (host; NULL when no component has a mineral or gas
   constraint) the reactions of the mineral and gas rows.  Starts from
   1e-9 molal, or with opts->guess_mode == PFX_GUESS_ARRAY from
   cells->guess (a FREE_ION_GUESS, reaction.F90:1476-1482; rows fixed by
   their own value still start from it), with unit activity coefficients,
   switches activity on after the
   first convergence (LAG, every iteration), and stops on the reference
   criteria max|R| < max_residual_tolerance and max|dm/m| <
   max_relative_change_tolerance.  Out: pri_molal, pri_act_coef,
   sec_act_coef, sec_molal, total (at the last evaluation), total_sorb and
   ionx_ref_sorbed at the solution, num_iterations, status.
   molal != 0: concentrations in mol/kg water (CHEMISTRY MOLAL). */
int pfx_equilibrate_batch(pfx_tables *tables, pfx_cells *cells, const int32_t *constraint_type,
                          const double *constraint_conc, const pfx_constraint_rxn *constraint_rxn,
                          int32_t molal, const pfx_react_opts *opts, pfx_stats *stats);

/* Auxiliary-variable refresh at the given free ions: RTUpdateAuxVars's cell
   loop (reactive_transport.F90:3790-3846): RTAuxVarCompute (totals, sorbed
   totals, complex molalities), preceded `activity_passes` times by
   RActivityCoefficients + RTAuxVarCompute (the initial-condition refresh
   of condition_control.F90:929-935 uses 2). */
int pfx_auxvar_batch(pfx_tables *tables, pfx_cells *cells, int32_t activity_passes,
                     const pfx_react_opts *opts);

/*
 * Structured grid of the global-implicit assembly (a DMDA box, natural
 * ordering: cell c = i + nx*(j + ny*k)).  Internal faces connect a cell
 * ("up") with its +x/+y/+z neighbour ("dn"); their Darcy fluxes are
 * positive from up to dn.  Face arrays are device pointers:
 *   x faces [nz][ny][nx-1], y faces [nz][ny-1][nx], z faces [nz-1][ny][nx].
 * dd is the harmonic mean of saturation * porosity * (mechanical dispersion
 * + tortuosity * molecular diffusion) over the distance (TDispersion,
 * transport.F90:65-234), m/s, the same for every component.  The x-min and
 * x-max faces of the box may carry boundary connections (the boundary is
 * "up", the cell "dn", flux positive into the domain, [nz][ny] arrays);
 * their totals are bc_total (mol/L, [N], device) or, when NULL, the cell's
 * own (zero gradient).  Other boundary faces are closed.
 */
typedef struct pfx_grid {
  int32_t nx, ny, nz;
  int32_t log_formulation;   /* 1: Jacobian columns scaled by m (unknown ln m) */
  double area[3];            /* face area per direction (m^2)                */
  const double *q[3];        /* internal-face Darcy flux (m/s)                */
  const double *dd[3];       /* internal-face dispersion / distance (m/s)     */
  const double *bc_q[2];     /* x-min, x-max boundary flux (NULL = closed)    */
  const double *bc_dd[2];    /* x-min, x-max boundary dispersion / distance   */
  const double *bc_total[2]; /* x-min, x-max boundary totals (NULL = zero gradient) */
} pfx_grid;

/* Block sparsity of the 7-point Jacobian: row_ptr[ncells+1] and
   col_idx[nnzb] (DEVICE, int32), columns sorted within a row, the layout of
   PETSc BAIJ from DMCreateMatrix on a star stencil
   (discretization.F90:849-857).  With row_ptr == NULL only *nnzb is set. */
int pfx_gi_structure(const pfx_grid *grid, int32_t *row_ptr, int32_t *col_idx, int64_t *nnzb);

/* Global-implicit residual and Jacobian of reactive transport at the free
   ions in cells->pri_molal: RTResidual = RTResidualFlux + RTResidualNonFlux
   (reactive_transport.F90:2057-2782) and RTJacobian = RTJacobianFlux +
   RTJacobianNonFlux (:2903-3525), upwind fluxes (TFluxCoef,
   transport.F90:756-819; TFlux :368-439; TFluxDerivative :529-622).
   cells->total / total_sorb hold the accumulation of the OLD state.
   residual: [ncells*N] mol/s; bsr_values: [nnzb][N][N] column-major blocks
   in the row_ptr order of pfx_gi_structure (kg water/s, or per unit ln m
   with log_formulation).  Activity coefficients are used as given. */
int pfx_gi_residual_jacobian(pfx_tables *tables, pfx_cells *cells, const pfx_grid *grid,
                             const pfx_react_opts *opts, const int32_t *row_ptr,
                             double *residual, double *bsr_values);

/* Transport coefficients of the grid's faces: RTUpdateTransportCoefs
   (reactive_transport.F90:852-1020) -> TDispersion / TDispersionBC
   (transport.F90:65-379), liquid phase, longitudinal dispersion only
   (no transverse tensor), uniform spacing dx[3] (host):
     D = max(alpha_L |q| + s phi tau D_m, 1e-40) per side,
     dd = D_up D_dn / (D_up dist_dn + D_dn dist_up), dist = dx/2 each side;
   boundary faces: bc_type 1 (Dirichlet) dd = D_cell / (dx/2), 2 (zero
   gradient or closed) dd = 0.  Cell arrays [ncells] (device; saturation and
   tortuosity may be NULL = 1) and face fluxes grid->q / grid->bc_q in;
   dd_out[3] (internal faces) and bc_dd_out[2] (device, may be NULL) out. */
int pfx_gi_dispersion(const pfx_grid *grid, const double *porosity, const double *saturation,
                      const double *tortuosity, double dispersivity, double diffusion,
                      const double *dx, const int32_t *bc_type, double **dd_out,
                      double **bc_dd_out, void *stream);

/*
 * Chemistry zones in one persistent launch.  No counterpart in the
 * reference, where RTReact loops over every cell with one reaction_type:
 * here each zone has its own tables (its own primaries, complexes and
 * minerals) and its own batch of cells, and one launch reacts them all.
 * The workgroups are split between the zones; each zone drains its own
 * work queue, and every workgroup stamps the wall clock when it retires, so
 * the split can follow the measured work (pfx_zone_plan_rebalance).
 */
typedef struct pfx_zone_plan {
  int32_t nzones;
  int32_t blocks[PFX_MAX_ZONES];   /* In: workgroups per zone (all 0 = split by
                                      cells * N^2); Out: the split used       */
  double zone_ms[PFX_MAX_ZONES];   /* Out: when the zone's last workgroup retired */
  double kernel_ms;                /* Out: the launch (HIP events)            */
  int32_t variant;                 /* In/Out: 0 registers as needed (1 wavefront
                                      per SIMD with 32-lane zones), 1 and 2
                                      capped at 2 and 3 wavefronts per SIMD   */
  int32_t reserved;
} pfx_zone_plan;

/* pfx_react_batch for nzones (tables[z], cells[z]) pairs in one launch;
   opts are shared (dt, tolerances, guess mode).  plan may be NULL (split
   by cells * N^2).  stats: [nzones], may be NULL.  Returns the number of
   cells that did not converge, or a PFX_E_* code. */
int pfx_react_zones(int32_t nzones, pfx_tables *const *tables, pfx_cells *cells, const pfx_react_opts *opts,
                    pfx_zone_plan *plan, pfx_stats *stats);

/* New split from the last launch's zone times: blocks_z proportional to
   blocks_z * zone_ms_z (the work each zone took), same total. */
int pfx_zone_plan_rebalance(pfx_zone_plan *plan);

/* The same zoned step captured as a HIP graph (queue resets, the zoned
   kernel and the statistics reductions as kernel nodes, no copies), for
   time-step loops that replay it on the same buffers. */
typedef struct pfx_graph pfx_graph;
int pfx_zones_graph_create(int32_t nzones, pfx_tables *const *tables, pfx_cells *cells,
                           const pfx_react_opts *opts, const pfx_zone_plan *plan, pfx_graph **out);
int pfx_graph_launch(pfx_graph *graph, void *stream);
/* statistics of the last replay (waits for the stream it was launched on) */
int pfx_graph_stats(pfx_graph *graph, pfx_stats *stats, pfx_zone_plan *plan);
int pfx_graph_destroy(pfx_graph *graph);

/* Mineral volume-fraction update after a converged step: RUpdateKineticState
   (reaction.F90:5340-5366), called per cell by RTUpdateKineticState
   (reactive_transport.F90:642).  vf_k += mnrl_rate_k * molar_vol_k * dt,
   clamped at 0, with the rates that pfx_react_batch returned at the
   solution (the RKineticMineral evaluation the reference repeats there).
   Writes cells->mnrl_volfrac (device).  With opts->stream_ordered it
   returns without waiting for the stream. */
int pfx_update_kinetic_state(pfx_tables *tables, pfx_cells *cells, const pfx_react_opts *opts);

#ifdef __cplusplus
}
#endif

#endif /* PFX_CHEM_H */
