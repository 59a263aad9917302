! Host-side probe of the Fortran binding: ABI version and argument checks
! (no device work), plus the byte sizes of the bound types for comparison
! with the C structs.
program pfx_abi_probe
  use, intrinsic :: iso_c_binding
  use pfx_chem_module
  implicit none
  type(pfx_reaction_desc) :: desc
  type(pfx_cells) :: cells
  type(pfx_react_opts) :: opts
  type(pfx_stats) :: stats
  type(pfx_grid) :: grid
  type(pfx_zone_plan) :: plan
  type(pfx_newton_opts) :: nopts
  type(pfx_newton_stats) :: nstats
  type(pfx_krylov_stats) :: kstats
  type(pfx_constraint_rxn) :: crxn
  type(c_ptr) :: tables, solver
  integer(c_int) :: rc

  print '(a,i0)', 'abi ', pfx_abi_version()
  desc%naqcomp = 0
  rc = pfx_tables_create(desc, tables)
  print '(a,i0)', 'create_bad ', rc
  rc = pfx_react_batch(c_null_ptr, cells, opts, stats)
  print '(a,i0)', 'react_null ', rc
  rc = pfx_mlp_destroy(c_null_ptr)
  print '(a,i0)', 'mlp_null ', rc
  print '(a,i0)', 'sizeof_desc ', c_sizeof(desc)
  print '(a,i0)', 'sizeof_cells ', c_sizeof(cells)
  print '(a,i0)', 'sizeof_opts ', c_sizeof(opts)
  print '(a,i0)', 'sizeof_stats ', c_sizeof(stats)
  print '(a,i0)', 'sizeof_grid ', c_sizeof(grid)
  print '(a,i0)', 'sizeof_plan ', c_sizeof(plan)
  rc = pfx_gi_solver_create(c_null_ptr, grid, solver)
  print '(a,i0)', 'solver_null ', rc
  print '(a,i0)', 'sizeof_newton_opts ', c_sizeof(nopts)
  print '(a,i0)', 'sizeof_newton_stats ', c_sizeof(nstats)
  print '(a,i0)', 'sizeof_krylov_stats ', c_sizeof(kstats)
  print '(a,i0)', 'sizeof_constraint_rxn ', c_sizeof(crxn)
end program pfx_abi_probe
