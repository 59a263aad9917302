/*
 * pfx_gi_solver.h -- the global-implicit time step closed on the device:
 * Newton's method on the residual/Jacobian of pfx_gi_residual_jacobian with
 * a preconditioned BiCGStab linear solve, so a transport step needs no host
 * round trip of the 7-point block system.
 *
 * Reference interfaces replaced (src/pflotran/):
 *   SNESSolve of the transport process model   timestepper_BE.F90:326
 *     (PETSc SNES NEWTONLS, basic line search, SNESConvergedDefault through
 *      ConvergenceTest, convergence.F90:63-200)
 *   KSPBCGS, the transport linear solver       solver.F90:161, 237
 *   PMRTCheckUpdatePre (max_dlnC clip, or the
 *     0.99 min-ratio scaling in linear form)   pm_rt.F90:734-840
 * PETSc itself is not part of the reference tree (pinned commit 1a9d3c3c50);
 * its published algorithms are restated here: SNESConvergedDefault's order
 * of tests and KSPBCGS (van der Vorst's BiCGStab).  The preconditioner is the
 * point-block Jacobi of the N x N diagonal blocks (PETSc PCPBJACOBI), not
 * PETSc's serial default ILU(0): it is exact to set up in parallel and costs
 * one block product per application; ILU(0)'s triangular solves serialise
 * along the grid's hyperplanes.
 *
 * All pointers are DEVICE pointers unless marked (host).  Return codes as in
 * pfx_chem.h (0 success, <0 PFX_E_*); solver outcomes are in the stats.
 */
#ifndef PFX_GI_SOLVER_H
#define PFX_GI_SOLVER_H

#include "pfx_chem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* KSPSetTolerances meanings (solver.F90:237); zeros select PETSc's defaults
   rtol 1e-5, atol 1e-50, dtol 1e5, max_iterations 10000.  Convergence:
   ||r||_2 <= max(rtol ||b||_2, atol), r the true (unpreconditioned)
   residual, the preconditioner applied on the right. */
typedef struct pfx_krylov_opts {
  double rtol;
  double atol;
  double dtol;
  int32_t max_iterations;
  int32_t reserved;
} pfx_krylov_opts;

/* KSPConvergedReason values */
#define PFX_KSP_CONVERGED_RTOL       2
#define PFX_KSP_CONVERGED_ATOL       3
#define PFX_KSP_DIVERGED_ITS       (-3)
#define PFX_KSP_DIVERGED_DTOL      (-4)
#define PFX_KSP_DIVERGED_BREAKDOWN (-5)
#define PFX_KSP_DIVERGED_NANORINF  (-9)
#define PFX_KSP_DIVERGED_PC_FAILED (-11)  /* a singular diagonal block */

typedef struct pfx_krylov_stats {
  int32_t iterations;
  int32_t reason;         /* PFX_KSP_* */
  double rnorm0;          /* ||b||_2 */
  double rnorm;           /* final ||r||_2 */
  double ms;              /* device time of the solve (events) */
} pfx_krylov_stats;

/* A block-sparse (BSR) system with column-major N x N blocks, the layout of
   pfx_gi_structure / PETSc BAIJ (MAT_ROW_ORIENTED = FALSE); the handle owns
   the solver's work vectors and the inverted diagonal blocks.  row_ptr
   [nrows+1] and col_idx [nnzb] (device, int32) must stay valid while the
   handle lives; every row must hold its diagonal block. */
typedef struct pfx_bsr pfx_bsr;
int pfx_bsr_create(int64_t nrows, int32_t N, int64_t nnzb, const int32_t *row_ptr, const int32_t *col_idx,
                   pfx_bsr **out);
int pfx_bsr_destroy(pfx_bsr *h);

/* y = A x (values [nnzb][N][N] column-major). */
int pfx_bsr_spmv(pfx_bsr *h, const double *values, const double *x, double *y, void *stream);

/* Solve A x = b by BiCGStab with point-block Jacobi preconditioning from
   x = 0.  The diagonal blocks are inverted (partial pivoting) first. */
int pfx_bsr_solve(pfx_bsr *h, const double *values, const double *b, double *x, const pfx_krylov_opts *opts,
                  void *stream, pfx_krylov_stats *stats);

/* Several right-hand sides of one scalar matrix (a handle of block size
   1): B, X [nrows][ncomp] (device, cell-major: component c of row i at
   i*ncomp + c).  Each column is solved by its own BiCGStab with Jacobi
   preconditioning, as pfx_bsr_solve would (the same scalars and tests per
   column), but every scalar stays on the device: the columns share each
   launch, and a decided column skips its work.  stream_ordered != 0
   enqueues exactly opts->max_iterations iterations (at most
   PFX_KSP_MAX_ENQUEUED) and returns without waiting, so the solve can be
   captured in a graph; pfx_bsr_multi_stats then reads the outcome.
   Otherwise the host reads one flag every 4 iterations and stats are
   filled on return (iterations summed over the columns, the worst reason,
   the largest norms).  No counterpart in the reference: its transport
   components are separate KSP solves. */
#define PFX_KSP_MAX_ENQUEUED 1000
int pfx_bsr_solve_multi(pfx_bsr *h, const double *values, const double *B, double *X, int32_t ncomp,
                        const pfx_krylov_opts *opts, void *stream, int32_t stream_ordered, pfx_krylov_stats *stats);
/* The outcome of the handle's last several-right-hand-side solve once
   `stream` has run it (waits for the stream). */
int pfx_bsr_multi_stats(pfx_bsr *h, void *stream, pfx_krylov_stats *stats);

/* SNESSetTolerances meanings (solver.F90:256); zeros select PETSc's
   defaults atol 1e-50, rtol 1e-8, stol 1e-8, max_iterations 50. */
typedef struct pfx_newton_opts {
  double atol;
  double rtol;
  double stol;
  int32_t max_iterations;
  int32_t reserved;
  pfx_krylov_opts krylov;
  /* PFLOTRAN's own tests, each off at 0 (NEWTON_SOLVER ITOL, ITOL_UPDATE,
     solver.F90:861-867; the transport process model's ITOL_RELATIVE_UPDATE,
     pm_rt.F90:171-174): ||F||_inf < itol_res, ||dx||_inf < itol_update and
     max |dx / x_old| < itol_rel_update (x = ln m in log form) converge */
  double itol_res;
  double itol_update;
  double itol_rel_update;
} pfx_newton_opts;

/* SNESConvergedReason values */
#define PFX_SNES_CONVERGED_FNORM_ABS        2
#define PFX_SNES_CONVERGED_FNORM_RELATIVE   3
#define PFX_SNES_CONVERGED_SNORM_RELATIVE   4
#define PFX_SNES_DIVERGED_LINEAR_SOLVE    (-3)
#define PFX_SNES_DIVERGED_FNORM_NAN       (-4)
#define PFX_SNES_DIVERGED_MAX_IT          (-5)
/* PFLOTRAN's ConvergenceTest: residual infinity norm above max_norm 1e20
   (convergence.F90:230-232, solver.F90:137) */
#define PFX_SNES_DIVERGED_MAX_NORM       (-10)
/* PFLOTRAN's own convergence reasons: ||F||_inf below ITOL (convergence.F90:
   215-217), ||dx||_inf below ITOL_UPDATE (:221-223), and the transport
   process model's ITOL_RELATIVE_UPDATE, max |dx / x| below it
   (pm_rt.F90:901-925 with convergence.F90:194-198) */
#define PFX_SNES_CONVERGED_ITOL_RES        10
#define PFX_SNES_CONVERGED_ITOL_UPDATE     11
#define PFX_SNES_CONVERGED_RELATIVE_UPDATE 12

typedef struct pfx_newton_stats {
  int32_t iterations;         /* Newton updates taken */
  int32_t reason;             /* PFX_SNES_* */
  int32_t linear_iterations;  /* BiCGStab iterations, all Newton steps */
  int32_t reserved;
  double fnorm0, fnorm;       /* ||F||_2 at the start and at the end */
  double ms_assembly;         /* device time in residual/Jacobian evaluations */
  double ms_linear;           /* device time in the linear solves */
  double ms_total;            /* host wall time of the step */
} pfx_newton_stats;

/* One global-implicit time step (the SNESSolve of a transport step).
   The grid's BSR structure and matrix are held by the handle; create it
   once per grid shape and network. */
typedef struct pfx_gi_solver pfx_gi_solver;
int pfx_gi_solver_create(pfx_tables *tables, const pfx_grid *grid, pfx_gi_solver **out);
int pfx_gi_solver_destroy(pfx_gi_solver *s);

/* Newton from the free ions in cells->pri_molal (in: initial iterate, the
   old state as timestepper_BE starts it; out: the last iterate).
   cells->total / total_sorb hold the OLD state's accumulation, as for
   pfx_gi_residual_jacobian; activity coefficients are held fixed (the
   once-per-step update).  Each iteration: F, J at the iterate; J y = F by
   pfx_bsr_solve; y clipped as PMRTCheckUpdatePre does (|y| <= max_dlnC in
   log form, else scaled by 0.99 min(m / y) where that is below 1); x -= y
   (m <- m exp(-y) in log form); tests in SNESConvergedDefault's order:
   NaN, ||F|| < atol, ||F|| <= rtol ||F0||, ||y|| < stol ||x||, the
   iteration limit, then PFLOTRAN's PMRTCheckUpdatePost and ConvergenceTest
   (pm_rt.F90:885-925, convergence.F90:170-245): the itol tests of
   pfx_newton_opts converge (reasons 12, 10, 11 in that order of override),
   ||F||_inf > 1e20 diverges, and no test but NaN applies to the initial
   residual (newton_min_iterations 1: one Newton iteration at
   least, also from a zero residual).  Totals at the solution come from
   pfx_auxvar_batch. */
int pfx_gi_solver_step(pfx_gi_solver *s, pfx_cells *cells, const pfx_grid *grid, const pfx_react_opts *opts,
                       const pfx_newton_opts *nopts, pfx_newton_stats *stats);

/* One time step of the global-implicit solve with the reference's time-step
   cut (timestepper_BE.F90:344-399): pfx_gi_solver_step with opts->dt; a
   step whose Newton solve does not converge (reason <= 0) restores the old
   free ions and exchange state (RTTimeCut, reactive_transport.F90:58-96),
   multiplies dt by reduction_factor and is solved again, at most max_cuts
   times (< 0: the reference's 16; a factor outside (0,1): its 0.5,
   timestepper_base.F90:144-145).  The step taken is stats->dt; the caller
   advances time by it (the reference's target time moves with the cut). */
typedef struct pfx_gi_step_stats {
  pfx_newton_stats newton;          /* the last Newton solve                 */
  int32_t reason;                   /* its PFX_SNES_* reason                 */
  int32_t cuts;                     /* time-step cuts taken                  */
  int32_t newton_iterations;        /* over every attempt                    */
  int32_t wasted_newton_iterations; /* in the attempts that were cut         */
  int32_t linear_iterations;
  int32_t reserved;
  double dt;                        /* the step taken (s)                    */
  double ms_total;                  /* host wall time                        */
} pfx_gi_step_stats;
int pfx_gi_solver_timestep(pfx_gi_solver *s, pfx_cells *cells, const pfx_grid *grid, const pfx_react_opts *opts,
                           const pfx_newton_opts *nopts, int32_t max_cuts, double reduction_factor,
                           pfx_gi_step_stats *stats);

/* The transport half of an operator-split step (the reference's
   deprecated/timestepper.F90:2156-2245 sequence): one implicit step of the
   totals of every component, A T_new = b, with A the scalar transport
   matrix of RTCalculateTransportMatrix (reactive_transport.F90:1379-1601):
   upwind face coefficients (TFluxCoef, transport.F90:756-819), boundary
   faces on the diagonal, accumulation phi*s*1000*V/dt (phi, s, V from
   cells or the opts defaults, dt = opts->dt); and b of RTCalculateRHS_t0
   (:1085-1146) plus T_up * T_bc on boundary faces with a boundary total
   (RTCalculateRHS_t1, :1232-1258).  A zero-gradient boundary face (no
   boundary total) carries the cell's new total, on the diagonal (the
   reference's RTCalculateRHS_t1 stops with "must be refactored", so it
   fixes no other choice).  A is the same for all components: it is
   assembled once and each component solved by pfx_bsr_solve.
   h: a handle of block size 1 on pfx_gi_structure's pattern for the grid;
   cells->total [ncells][ncomp] (device): in the old totals, out the
   transported ones; work: device doubles [nnzb + 2 ncells] (the matrix
   values in the first nnzb).  The components are solved together by
   pfx_bsr_solve_multi.  stats sum the components' iterations and keep the
   worst reason.  Returns 0, 1 if a component's solve did not converge, or a
   negative PFX_E_* code.  With opts->stream_ordered the step is enqueued
   without waiting (kopts->max_iterations iterations, graph-capturable) and
   pfx_bsr_multi_stats gives the outcome.
   No counterpart in the current reference build (its operator split is
   disabled). */
int pfx_os_transport_step(pfx_bsr *h, const pfx_grid *grid, pfx_cells *cells, int32_t ncomp,
                          const pfx_react_opts *opts, double *work, const pfx_krylov_opts *kopts,
                          pfx_krylov_stats *stats);

/* One whole operator-split step captured as a HIP graph: the transport
   half (pfx_os_transport_step with transport_opts, every component in one
   several-right-hand-side solve of kopts->max_iterations enqueued
   iterations), the reaction half (pfx_react_batch with react_opts) and the
   volume-fraction update (pfx_update_kinetic_state with react_opts), all
   stream-ordered, in the order of the reference's operator-split step
   (deprecated/timestepper.F90:2156-2245 transport, RTReact, then
   RTUpdateKineticState).  The arguments are those of the three calls; the
   graph holds their pointers and options (dt included), so a replay
   advances the same cells by one more step of the same size.  Replay with
   pfx_os_graph_launch; after a replay has run, pfx_os_graph_stats reads the
   reaction statistics (pfx_react_batch's return value in rc) and the
   transport solve's outcome.  No counterpart in the reference (its steps are
   host loops); the captured step is the device-resident form of one
   RTReact step. */
typedef struct pfx_os_graph pfx_os_graph;
int pfx_os_graph_create(pfx_bsr *transport, const pfx_grid *grid, pfx_tables *tables, pfx_cells *cells,
                        int32_t ncomp, const pfx_react_opts *transport_opts, const pfx_react_opts *react_opts,
                        double *work, const pfx_krylov_opts *kopts, pfx_os_graph **out);
int pfx_os_graph_launch(pfx_os_graph *graph, void *stream);
int pfx_os_graph_stats(pfx_os_graph *graph, int32_t *rc, pfx_stats *react_stats, pfx_krylov_stats *transport_stats);
int pfx_os_graph_destroy(pfx_os_graph *graph);

#ifdef __cplusplus
}
#endif

#endif /* PFX_GI_SOLVER_H */
