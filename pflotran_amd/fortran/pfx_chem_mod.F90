! ISO_C_BINDING interface of libpfx_chem.so (include/pfx_chem.h) for the
! PFLOTRAN Fortran host.  The derived types mirror the C structs field by
! field; reaction_type and reactive_transport_auxvar_type stay unchanged and
! are packed into these contiguous arrays by the caller.
module pfx_chem_module

  use, intrinsic :: iso_c_binding

  implicit none

  private

  integer(c_int), parameter, public :: PFX_ABI_EXPECTED = 5
  integer(c_int), parameter, public :: PFX_E_ARG = -1, PFX_E_DEVICE = -2, &
                                       PFX_E_NOMEM = -3, PFX_E_UNSUP = -4
  integer(c_int), parameter, public :: PFX_ACT_OFF = 0, PFX_ACT_TIMESTEP = 1, &
                                       PFX_ACT_NEWTON_ITER = 2
  integer(c_int), parameter, public :: PFX_ACT_ALG_LAG = 0, PFX_ACT_ALG_NEWTON = 1
  integer(c_int), parameter, public :: PFX_CELL_CONVERGED = 0, &
                                       PFX_CELL_MAX_ITERATIONS = 1, &
                                       PFX_CELL_NONFINITE = 2, PFX_CELL_SINGULAR = 3
  integer(c_int), parameter, public :: PFX_CONSTRAINT_TOTAL = 0, &
                                       PFX_CONSTRAINT_FREE = 1, &
                                       PFX_CONSTRAINT_LOG = 2, PFX_CONSTRAINT_PH = 3, &
                                       PFX_CONSTRAINT_CHARGE = 4, PFX_CONSTRAINT_TOTAL_SORB = 5, &
                                       PFX_CONSTRAINT_MINERAL = 6, PFX_CONSTRAINT_GAS = 7
  integer(c_int), parameter, public :: PFX_GUESS_PREVIOUS = 0, PFX_GUESS_ARRAY = 1, &
                                       PFX_GUESS_DEFAULT = 2, PFX_GUESS_ARRAY_FALLBACK = 3

  type, bind(c), public :: pfx_reaction_desc
    integer(c_int32_t) :: naqcomp, neqcplx, nkinmnrl, neqionxrxn
    integer(c_int32_t) :: max_rxn_spec, max_cation
    type(c_ptr) :: primary_spec_Z, primary_spec_a0
    type(c_ptr) :: eqcplxspecid, eqcplxstoich, eqcplxh2ostoich
    type(c_ptr) :: eqcplx_Z, eqcplx_a0, eqcplx_logK, eqcplx_logKcoef
    type(c_ptr) :: kinmnrlspecid, kinmnrlstoich, kinmnrlh2ostoich
    type(c_ptr) :: kinmnrl_logK, kinmnrl_logKcoef, kinmnrl_rate_constant
    type(c_ptr) :: kinmnrl_molar_vol, kinmnrl_activation_energy
    type(c_ptr) :: kinmnrl_affinity_threshold, kinmnrl_rate_limiter
    type(c_ptr) :: kinmnrl_affinity_power, kinmnrl_Temkin_const
    type(c_ptr) :: eqionx_rxn_cationid, eqionx_rxn_k, eqionx_rxn_CEC
    type(c_ptr) :: eqionx_rxn_to_surf
    real(c_double) :: debyeA, debyeB, debyeBdot
    integer(c_int32_t) :: use_log_formulation
    integer(c_int32_t) :: act_coef_update_frequency
    integer(c_int32_t) :: act_coef_update_algorithm
    integer(c_int32_t) :: numerical_jacobian
    real(c_double) :: max_dlnC
    real(c_double) :: max_relative_change_tolerance
    real(c_double) :: max_residual_tolerance
    integer(c_int32_t) :: neqkdrxn, reserved_kd
    type(c_ptr) :: eqkdspecid, eqkdtype, eqkddistcoef, eqkdlangmuirb
    type(c_ptr) :: eqkdfreundlichn, eqkdmineral
    real(c_double) :: soil_particle_density
    integer(c_int32_t) :: max_prefactors, max_prefactor_species
    type(c_ptr) :: kinmnrl_num_prefactors, kinmnrl_prefactor_id
    type(c_ptr) :: kinmnrl_pref_alpha, kinmnrl_pref_beta, kinmnrl_pref_atten_coef
    type(c_ptr) :: kinmnrl_pref_rate, kinmnrl_pref_activation_energy
    type(c_ptr) :: kinmnrl_min_scale_factor
    integer(c_int32_t) :: group_lanes, reserved_lanes
  end type pfx_reaction_desc

  type, bind(c), public :: pfx_cells
    integer(c_int64_t) :: ncells
    type(c_ptr) :: total, total_sorb, pri_molal, pri_act_coef, sec_act_coef
    type(c_ptr) :: sec_molal, ionx_ref_sorbed, mnrl_volfrac, mnrl_area
    type(c_ptr) :: mnrl_rate, temperature, porosity, saturation, volume
    type(c_ptr) :: den_kg, guess, num_iterations, status, rel_res
  end type pfx_cells

  type, bind(c), public :: pfx_react_opts
    real(c_double) :: dt, rel_tol
    integer(c_int32_t) :: max_iterations, extra_iterations
    integer(c_int32_t) :: guess_mode, reference_stop
    real(c_double) :: default_guess
    real(c_double) :: porosity, saturation, volume, den_kg
    type(c_ptr) :: stream
    integer(c_int32_t) :: fallback_after, jacobian_flags
    real(c_double) :: rel_change_tol
    integer(c_int32_t) :: fallback_previous, stream_ordered
  end type pfx_react_opts

  ! the reaction of a mineral or gas constraint row
  type, bind(c), public :: pfx_constraint_rxn
    integer(c_int32_t) :: nspec, has_logK_coef
    integer(c_int32_t) :: specid(8)
    real(c_double) :: stoich(8)
    real(c_double) :: logK
    real(c_double) :: logK_coef(5)
  end type pfx_constraint_rxn

  type, bind(c), public :: pfx_grid
    integer(c_int32_t) :: nx, ny, nz, log_formulation
    real(c_double) :: area(3)
    type(c_ptr) :: q(3), dd(3)
    type(c_ptr) :: bc_q(2), bc_dd(2), bc_total(2)
  end type pfx_grid

  type, bind(c), public :: pfx_zone_plan
    integer(c_int32_t) :: nzones
    integer(c_int32_t) :: blocks(8)
    real(c_double) :: zone_ms(8)
    real(c_double) :: kernel_ms
    integer(c_int32_t) :: variant, reserved
  end type pfx_zone_plan

  type, bind(c), public :: pfx_stats
    integer(c_int64_t) :: ncells, converged, total_iterations
    integer(c_int32_t) :: max_iterations, min_iterations
    real(c_double) :: max_rel_res, kernel_ms
    integer(c_int64_t) :: iteration_histogram(32)
    integer(c_int64_t) :: fallbacks
  end type pfx_stats

  ! global-implicit step on the device (include/pfx_gi_solver.h)
  type, bind(c), public :: pfx_krylov_opts
    real(c_double) :: rtol, atol, dtol
    integer(c_int32_t) :: max_iterations, reserved
  end type pfx_krylov_opts

  type, bind(c), public :: pfx_krylov_stats
    integer(c_int32_t) :: iterations, reason
    real(c_double) :: rnorm0, rnorm, ms
  end type pfx_krylov_stats

  type, bind(c), public :: pfx_newton_opts
    real(c_double) :: atol, rtol, stol
    integer(c_int32_t) :: max_iterations, reserved
    type(pfx_krylov_opts) :: krylov
    real(c_double) :: itol_res, itol_update, itol_rel_update
  end type pfx_newton_opts

  type, bind(c), public :: pfx_newton_stats
    integer(c_int32_t) :: iterations, reason, linear_iterations, reserved
    real(c_double) :: fnorm0, fnorm, ms_assembly, ms_linear, ms_total
  end type pfx_newton_stats

  public :: pfx_bsr_create, pfx_bsr_destroy, pfx_bsr_spmv, pfx_bsr_solve, &
            pfx_gi_solver_create, pfx_gi_solver_destroy, pfx_gi_solver_step, &
            pfx_os_transport_step

  public :: pfx_abi_version, pfx_tables_create, pfx_tables_destroy, &
            pfx_react_batch, pfx_tables_stats, pfx_residual_jacobian, pfx_equilibrate_batch, &
            pfx_auxvar_batch, pfx_update_kinetic_state, pfx_tables_timing, &
            pfx_gi_structure, pfx_gi_residual_jacobian, pfx_gi_dispersion, &
            pfx_react_zones, pfx_zone_plan_rebalance, pfx_zones_graph_create, &
            pfx_graph_launch, pfx_graph_stats, pfx_graph_destroy, pfx_tables_tune
  ! learned warm start (include/pfx_surrogate.h)
  public :: pfx_mlp_create, pfx_mlp_destroy, pfx_mlp_params, pfx_mlp_forward, &
            pfx_mlp_gradients, pfx_mlp_adam, pfx_mlp_train_step, &
            pfx_surrogate_features, pfx_surrogate_guess, pfx_surrogate_guess_from

  interface
    function pfx_abi_version() bind(c, name='pfx_abi_version')
      import :: c_int
      integer(c_int) :: pfx_abi_version
    end function pfx_abi_version

    function pfx_tables_create(desc, tables) bind(c, name='pfx_tables_create')
      import :: c_int, c_ptr, pfx_reaction_desc
      type(pfx_reaction_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: tables
      integer(c_int) :: pfx_tables_create
    end function pfx_tables_create

    function pfx_tables_destroy(tables) bind(c, name='pfx_tables_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: tables
      integer(c_int) :: pfx_tables_destroy
    end function pfx_tables_destroy

    function pfx_react_batch(tables, cells, opts, stats) &
        bind(c, name='pfx_react_batch')
      import :: c_int, c_ptr, pfx_cells, pfx_react_opts, pfx_stats
      type(c_ptr), value :: tables
      type(pfx_cells), intent(inout) :: cells
      type(pfx_react_opts), intent(in) :: opts
      type(pfx_stats), intent(out) :: stats
      integer(c_int) :: pfx_react_batch
    end function pfx_react_batch

    function pfx_tables_stats(tables, stats) bind(c, name='pfx_tables_stats')
      import :: c_int, c_ptr, pfx_stats
      type(c_ptr), value :: tables
      type(pfx_stats), intent(out) :: stats
      integer(c_int) :: pfx_tables_stats
    end function pfx_tables_stats

    function pfx_residual_jacobian(tables, cells, opts, residual, jacobian) &
        bind(c, name='pfx_residual_jacobian')
      import :: c_int, c_ptr, pfx_cells, pfx_react_opts
      type(c_ptr), value :: tables
      type(pfx_cells), intent(inout) :: cells
      type(pfx_react_opts), intent(in) :: opts
      type(c_ptr), value :: residual, jacobian
      integer(c_int) :: pfx_residual_jacobian
    end function pfx_residual_jacobian

    function pfx_equilibrate_batch(tables, cells, constraint_type, &
                                   constraint_conc, constraint_rxn, molal, opts, stats) &
        bind(c, name='pfx_equilibrate_batch')
      import :: c_int, c_int32_t, c_ptr, pfx_cells, pfx_react_opts, pfx_stats
      type(c_ptr), value :: tables
      type(pfx_cells), intent(inout) :: cells
      integer(c_int32_t), intent(in) :: constraint_type(*)
      type(c_ptr), value :: constraint_conc
      type(c_ptr), value :: constraint_rxn   ! pfx_constraint_rxn(N) or c_null_ptr
      integer(c_int32_t), value :: molal
      type(pfx_react_opts), intent(in) :: opts
      type(pfx_stats), intent(out) :: stats
      integer(c_int) :: pfx_equilibrate_batch
    end function pfx_equilibrate_batch

    function pfx_auxvar_batch(tables, cells, activity_passes, opts) &
        bind(c, name='pfx_auxvar_batch')
      import :: c_int, c_int32_t, c_ptr, pfx_cells, pfx_react_opts
      type(c_ptr), value :: tables
      type(pfx_cells), intent(inout) :: cells
      integer(c_int32_t), value :: activity_passes
      type(pfx_react_opts), intent(in) :: opts
      integer(c_int) :: pfx_auxvar_batch
    end function pfx_auxvar_batch

    function pfx_tables_tune(tables, kernel, variant) bind(c, name='pfx_tables_tune')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: tables
      integer(c_int32_t), value :: kernel, variant
      integer(c_int) :: pfx_tables_tune
    end function pfx_tables_tune

    function pfx_tables_timing(tables, ms) bind(c, name='pfx_tables_timing')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: tables
      real(c_double), intent(out) :: ms(2)
      integer(c_int) :: pfx_tables_timing
    end function pfx_tables_timing

    function pfx_gi_structure(grid, row_ptr, col_idx, nnzb) bind(c, name='pfx_gi_structure')
      import :: c_int, c_int64_t, c_ptr, pfx_grid
      type(pfx_grid), intent(in) :: grid
      type(c_ptr), value :: row_ptr, col_idx
      integer(c_int64_t), intent(out) :: nnzb
      integer(c_int) :: pfx_gi_structure
    end function pfx_gi_structure

    function pfx_gi_residual_jacobian(tables, cells, grid, opts, row_ptr, residual, values) &
        bind(c, name='pfx_gi_residual_jacobian')
      import :: c_int, c_ptr, pfx_cells, pfx_grid, pfx_react_opts
      type(c_ptr), value :: tables
      type(pfx_cells), intent(inout) :: cells
      type(pfx_grid), intent(in) :: grid
      type(pfx_react_opts), intent(in) :: opts
      type(c_ptr), value :: row_ptr, residual, values
      integer(c_int) :: pfx_gi_residual_jacobian
    end function pfx_gi_residual_jacobian

    function pfx_gi_dispersion(grid, porosity, saturation, tortuosity, dispersivity, &
                               diffusion, dx, bc_type, dd_out, bc_dd_out, stream) &
        bind(c, name='pfx_gi_dispersion')
      import :: c_int, c_int32_t, c_double, c_ptr, pfx_grid
      type(pfx_grid), intent(in) :: grid
      type(c_ptr), value :: porosity, saturation, tortuosity
      real(c_double), value :: dispersivity, diffusion
      real(c_double), intent(in) :: dx(3)
      integer(c_int32_t), intent(in) :: bc_type(2)
      type(c_ptr), intent(in) :: dd_out(3), bc_dd_out(2)
      type(c_ptr), value :: stream
      integer(c_int) :: pfx_gi_dispersion
    end function pfx_gi_dispersion

    function pfx_react_zones(nzones, tables, cells, opts, plan, stats) bind(c, name='pfx_react_zones')
      import :: c_int, c_int32_t, c_ptr, pfx_cells, pfx_react_opts, pfx_zone_plan, pfx_stats
      integer(c_int32_t), value :: nzones
      type(c_ptr), intent(in) :: tables(*)
      type(pfx_cells), intent(inout) :: cells(*)
      type(pfx_react_opts), intent(in) :: opts
      type(pfx_zone_plan), intent(inout) :: plan
      type(pfx_stats), intent(out) :: stats(*)
      integer(c_int) :: pfx_react_zones
    end function pfx_react_zones

    function pfx_zone_plan_rebalance(plan) bind(c, name='pfx_zone_plan_rebalance')
      import :: c_int, pfx_zone_plan
      type(pfx_zone_plan), intent(inout) :: plan
      integer(c_int) :: pfx_zone_plan_rebalance
    end function pfx_zone_plan_rebalance

    function pfx_zones_graph_create(nzones, tables, cells, opts, plan, graph) &
        bind(c, name='pfx_zones_graph_create')
      import :: c_int, c_int32_t, c_ptr, pfx_cells, pfx_react_opts, pfx_zone_plan
      integer(c_int32_t), value :: nzones
      type(c_ptr), intent(in) :: tables(*)
      type(pfx_cells), intent(inout) :: cells(*)
      type(pfx_react_opts), intent(in) :: opts
      type(pfx_zone_plan), intent(in) :: plan
      type(c_ptr), intent(out) :: graph
      integer(c_int) :: pfx_zones_graph_create
    end function pfx_zones_graph_create

    function pfx_graph_launch(graph, stream) bind(c, name='pfx_graph_launch')
      import :: c_int, c_ptr
      type(c_ptr), value :: graph, stream
      integer(c_int) :: pfx_graph_launch
    end function pfx_graph_launch

    function pfx_graph_stats(graph, stats, plan) bind(c, name='pfx_graph_stats')
      import :: c_int, c_ptr, pfx_stats, pfx_zone_plan
      type(c_ptr), value :: graph
      type(pfx_stats), intent(out) :: stats(*)
      type(pfx_zone_plan), intent(out) :: plan
      integer(c_int) :: pfx_graph_stats
    end function pfx_graph_stats

    function pfx_bsr_create(nrows, n, nnzb, row_ptr, col_idx, h) bind(c, name='pfx_bsr_create')
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      integer(c_int64_t), value :: nrows
      integer(c_int32_t), value :: n
      integer(c_int64_t), value :: nnzb
      type(c_ptr), value :: row_ptr, col_idx
      type(c_ptr), intent(out) :: h
      integer(c_int) :: pfx_bsr_create
    end function pfx_bsr_create

    function pfx_bsr_destroy(h) bind(c, name='pfx_bsr_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: pfx_bsr_destroy
    end function pfx_bsr_destroy

    function pfx_bsr_spmv(h, values, x, y, stream) bind(c, name='pfx_bsr_spmv')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, values, x, y, stream
      integer(c_int) :: pfx_bsr_spmv
    end function pfx_bsr_spmv

    function pfx_bsr_solve(h, values, b, x, opts, stream, stats) bind(c, name='pfx_bsr_solve')
      import :: c_int, c_ptr, pfx_krylov_opts, pfx_krylov_stats
      type(c_ptr), value :: h, values, b, x
      type(pfx_krylov_opts), intent(in) :: opts
      type(c_ptr), value :: stream
      type(pfx_krylov_stats), intent(out) :: stats
      integer(c_int) :: pfx_bsr_solve
    end function pfx_bsr_solve

    function pfx_gi_solver_create(tables, grid, solver) bind(c, name='pfx_gi_solver_create')
      import :: c_int, c_ptr, pfx_grid
      type(c_ptr), value :: tables
      type(pfx_grid), intent(in) :: grid
      type(c_ptr), intent(out) :: solver
      integer(c_int) :: pfx_gi_solver_create
    end function pfx_gi_solver_create

    function pfx_gi_solver_destroy(solver) bind(c, name='pfx_gi_solver_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
      integer(c_int) :: pfx_gi_solver_destroy
    end function pfx_gi_solver_destroy

    function pfx_gi_solver_step(solver, cells, grid, opts, nopts, stats) bind(c, name='pfx_gi_solver_step')
      import :: c_int, c_ptr, pfx_cells, pfx_grid, pfx_react_opts, pfx_newton_opts, pfx_newton_stats
      type(c_ptr), value :: solver
      type(pfx_cells), intent(inout) :: cells
      type(pfx_grid), intent(in) :: grid
      type(pfx_react_opts), intent(in) :: opts
      type(pfx_newton_opts), intent(in) :: nopts
      type(pfx_newton_stats), intent(out) :: stats
      integer(c_int) :: pfx_gi_solver_step
    end function pfx_gi_solver_step

    ! transport half of an operator-split step (include/pfx_gi_solver.h)
    function pfx_os_transport_step(bsr, grid, cells, ncomp, opts, work, kopts, stats) &
        bind(c, name='pfx_os_transport_step')
      import :: c_int, c_int32_t, c_ptr, pfx_cells, pfx_grid, pfx_react_opts, pfx_krylov_opts, pfx_krylov_stats
      type(c_ptr), value :: bsr
      type(pfx_grid), intent(in) :: grid
      type(pfx_cells), intent(inout) :: cells
      integer(c_int32_t), value :: ncomp
      type(pfx_react_opts), intent(in) :: opts
      type(c_ptr), value :: work
      type(pfx_krylov_opts), intent(in) :: kopts
      type(pfx_krylov_stats), intent(out) :: stats
      integer(c_int) :: pfx_os_transport_step
    end function pfx_os_transport_step

    function pfx_graph_destroy(graph) bind(c, name='pfx_graph_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: graph
      integer(c_int) :: pfx_graph_destroy
    end function pfx_graph_destroy

    function pfx_update_kinetic_state(tables, cells, opts) &
        bind(c, name='pfx_update_kinetic_state')
      import :: c_int, c_ptr, pfx_cells, pfx_react_opts
      type(c_ptr), value :: tables
      type(pfx_cells), intent(inout) :: cells
      type(pfx_react_opts), intent(in) :: opts
      integer(c_int) :: pfx_update_kinetic_state
    end function pfx_update_kinetic_state

    function pfx_mlp_create(nlayers, dims, batch_max, seed, mlp) &
        bind(c, name='pfx_mlp_create')
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      integer(c_int32_t), value :: nlayers
      integer(c_int32_t), intent(in) :: dims(*)
      integer(c_int32_t), value :: batch_max
      integer(c_int64_t), value :: seed
      type(c_ptr), intent(out) :: mlp
      integer(c_int) :: pfx_mlp_create
    end function pfx_mlp_create

    function pfx_mlp_destroy(mlp) bind(c, name='pfx_mlp_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: mlp
      integer(c_int) :: pfx_mlp_destroy
    end function pfx_mlp_destroy

    function pfx_mlp_params(mlp, which, layer, host, to_device) &
        bind(c, name='pfx_mlp_params')
      import :: c_int, c_int32_t, c_float, c_ptr
      type(c_ptr), value :: mlp
      integer(c_int32_t), value :: which, layer, to_device
      real(c_float), intent(inout) :: host(*)
      integer(c_int) :: pfx_mlp_params
    end function pfx_mlp_params

    function pfx_mlp_forward(mlp, x, b, y, stream) bind(c, name='pfx_mlp_forward')
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: mlp, x, y, stream
      integer(c_int64_t), value :: b
      integer(c_int) :: pfx_mlp_forward
    end function pfx_mlp_forward

    function pfx_mlp_gradients(mlp, x, t, b, loss, stream) &
        bind(c, name='pfx_mlp_gradients')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: mlp, x, t, loss, stream
      integer(c_int32_t), value :: b
      integer(c_int) :: pfx_mlp_gradients
    end function pfx_mlp_gradients

    function pfx_mlp_adam(mlp, lr, beta1, beta2, eps, stream) bind(c, name='pfx_mlp_adam')
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: mlp, stream
      real(c_float), value :: lr, beta1, beta2, eps
      integer(c_int) :: pfx_mlp_adam
    end function pfx_mlp_adam

    function pfx_mlp_train_step(mlp, x, t, b, lr, loss, stream) &
        bind(c, name='pfx_mlp_train_step')
      import :: c_int, c_int32_t, c_float, c_ptr
      type(c_ptr), value :: mlp, x, t, loss, stream
      integer(c_int32_t), value :: b
      real(c_float), value :: lr
      integer(c_int) :: pfx_mlp_train_step
    end function pfx_mlp_train_step

    function pfx_surrogate_features(tables, mlp, cells, x, stream) &
        bind(c, name='pfx_surrogate_features')
      import :: c_int, c_ptr, pfx_cells
      type(c_ptr), value :: tables, mlp, x, stream
      type(pfx_cells), intent(in) :: cells
      integer(c_int) :: pfx_surrogate_features
    end function pfx_surrogate_features

    function pfx_surrogate_guess(tables, mlp, cells, stream) &
        bind(c, name='pfx_surrogate_guess')
      import :: c_int, c_ptr, pfx_cells
      type(c_ptr), value :: tables, mlp, stream
      type(pfx_cells), intent(inout) :: cells
      integer(c_int) :: pfx_surrogate_guess
    end function pfx_surrogate_guess

    function pfx_surrogate_guess_from(mlp, y, ncells, naqcomp, prev, guess, stream) &
        bind(c, name='pfx_surrogate_guess_from')
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: mlp, y, prev, guess, stream
      integer(c_int64_t), value :: ncells
      integer(c_int32_t), value :: naqcomp
      integer(c_int) :: pfx_surrogate_guess_from
    end function pfx_surrogate_guess_from
  end interface

end module pfx_chem_module
