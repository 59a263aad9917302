"""ctypes mirror of include/pfx_chem.h (the C-ABI of libpfx_chem.so).

The structures here are byte-for-byte the C structs; tests/test_abi.py checks
the field offsets against the header compiled with the host compiler.
"""
import ctypes as C

PFX_ABI_VERSION = 6
PFX_MAX_PRIMARY = 32
PFX_MAX_CPLX = 256
PFX_MAX_RXN_SPEC = 8
PFX_MAX_MINERALS = 16
PFX_MAX_IONX_RXN = 4
PFX_MAX_IONX_CATION = 8
PFX_MAX_PREFACTORS = 4
PFX_MAX_PREFACTOR_SPECIES = 4
PFX_MAX_ZONES = 8

PFX_E_ARG, PFX_E_DEVICE, PFX_E_NOMEM, PFX_E_UNSUP = -1, -2, -3, -4

PFX_ACT_OFF, PFX_ACT_TIMESTEP, PFX_ACT_NEWTON_ITER = 0, 1, 2
PFX_KERNEL_REACT, PFX_KERNEL_EVAL, PFX_KERNEL_EQUIL, PFX_KERNEL_AUX = 0, 1, 2, 3
PFX_ACT_ALG_LAG, PFX_ACT_ALG_NEWTON = 0, 1

PFX_CELL_CONVERGED, PFX_CELL_MAX_ITERATIONS, PFX_CELL_NONFINITE, PFX_CELL_SINGULAR = 0, 1, 2, 3
PFX_GUESS_PREVIOUS, PFX_GUESS_ARRAY, PFX_GUESS_DEFAULT, PFX_GUESS_ARRAY_FALLBACK = 0, 1, 2, 3
PFX_CONSTRAINT_TOTAL, PFX_CONSTRAINT_FREE, PFX_CONSTRAINT_LOG, PFX_CONSTRAINT_PH, PFX_CONSTRAINT_CHARGE = 0, 1, 2, 3, 4
PFX_CONSTRAINT_TOTAL_SORB = 5
PFX_CONSTRAINT_MINERAL, PFX_CONSTRAINT_GAS = 6, 7
PFX_JAC_CAPPED_QK = 1
PFX_SORPTION_LINEAR, PFX_SORPTION_LANGMUIR, PFX_SORPTION_FREUNDLICH = 1, 2, 3
PFX_NOALLOC_MAX_CELLS = 1 << 20

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class ReactionDesc(C.Structure):
    _fields_ = [
        ("naqcomp", C.c_int32), ("neqcplx", C.c_int32), ("nkinmnrl", C.c_int32),
        ("neqionxrxn", C.c_int32), ("max_rxn_spec", C.c_int32), ("max_cation", C.c_int32),
        ("primary_spec_Z", _dp), ("primary_spec_a0", _dp),
        ("eqcplxspecid", _ip), ("eqcplxstoich", _dp), ("eqcplxh2ostoich", _dp),
        ("eqcplx_Z", _dp), ("eqcplx_a0", _dp), ("eqcplx_logK", _dp), ("eqcplx_logKcoef", _dp),
        ("kinmnrlspecid", _ip), ("kinmnrlstoich", _dp), ("kinmnrlh2ostoich", _dp),
        ("kinmnrl_logK", _dp), ("kinmnrl_logKcoef", _dp), ("kinmnrl_rate_constant", _dp),
        ("kinmnrl_molar_vol", _dp), ("kinmnrl_activation_energy", _dp),
        ("kinmnrl_affinity_threshold", _dp), ("kinmnrl_rate_limiter", _dp),
        ("kinmnrl_affinity_power", _dp), ("kinmnrl_Temkin_const", _dp),
        ("eqionx_rxn_cationid", _ip), ("eqionx_rxn_k", _dp), ("eqionx_rxn_CEC", _dp),
        ("eqionx_rxn_to_surf", _ip),
        ("debyeA", C.c_double), ("debyeB", C.c_double), ("debyeBdot", C.c_double),
        ("use_log_formulation", C.c_int32), ("act_coef_update_frequency", C.c_int32),
        ("act_coef_update_algorithm", C.c_int32), ("numerical_jacobian", C.c_int32),
        ("max_dlnC", C.c_double), ("max_relative_change_tolerance", C.c_double),
        ("max_residual_tolerance", C.c_double),
        ("neqkdrxn", C.c_int32), ("reserved_kd", C.c_int32), ("eqkdspecid", _ip), ("eqkdtype", _ip),
        ("eqkddistcoef", _dp), ("eqkdlangmuirb", _dp), ("eqkdfreundlichn", _dp), ("eqkdmineral", _ip),
        ("soil_particle_density", C.c_double),
        ("max_prefactors", C.c_int32), ("max_prefactor_species", C.c_int32),
        ("kinmnrl_num_prefactors", _ip), ("kinmnrl_prefactor_id", _ip), ("kinmnrl_pref_alpha", _dp),
        ("kinmnrl_pref_beta", _dp), ("kinmnrl_pref_atten_coef", _dp), ("kinmnrl_pref_rate", _dp),
        ("kinmnrl_pref_activation_energy", _dp), ("kinmnrl_min_scale_factor", _dp),
        ("group_lanes", C.c_int32), ("reserved_lanes", C.c_int32),
    ]


class Cells(C.Structure):
    _fields_ = [
        ("ncells", C.c_int64),
        ("total", C.c_void_p), ("total_sorb", C.c_void_p), ("pri_molal", C.c_void_p),
        ("pri_act_coef", C.c_void_p), ("sec_act_coef", C.c_void_p), ("sec_molal", C.c_void_p),
        ("ionx_ref_sorbed", C.c_void_p), ("mnrl_volfrac", C.c_void_p), ("mnrl_area", C.c_void_p),
        ("mnrl_rate", C.c_void_p), ("temperature", C.c_void_p), ("porosity", C.c_void_p),
        ("saturation", C.c_void_p), ("volume", C.c_void_p), ("den_kg", C.c_void_p),
        ("guess", C.c_void_p), ("num_iterations", C.c_void_p), ("status", C.c_void_p),
        ("rel_res", C.c_void_p),
    ]


class ReactOpts(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("rel_tol", C.c_double), ("max_iterations", C.c_int32),
        ("extra_iterations", C.c_int32), ("guess_mode", C.c_int32), ("reference_stop", C.c_int32),
        ("default_guess", C.c_double), ("porosity", C.c_double), ("saturation", C.c_double),
        ("volume", C.c_double), ("den_kg", C.c_double), ("stream", C.c_void_p),
        ("fallback_after", C.c_int32), ("jacobian_flags", C.c_int32),
        ("rel_change_tol", C.c_double), ("fallback_previous", C.c_int32), ("stream_ordered", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("ncells", C.c_int64), ("converged", C.c_int64), ("total_iterations", C.c_int64),
        ("max_iterations", C.c_int32), ("min_iterations", C.c_int32), ("max_rel_res", C.c_double),
        ("kernel_ms", C.c_double), ("iteration_histogram", C.c_int64 * 32),
        ("fallbacks", C.c_int64),
    ]

    def as_dict(self):
        return {
            "ncells": self.ncells, "converged": self.converged,
            "total_iterations": self.total_iterations, "max_iterations": self.max_iterations,
            "min_iterations": self.min_iterations, "max_rel_res": self.max_rel_res,
            "kernel_ms": self.kernel_ms, "iteration_histogram": list(self.iteration_histogram),
            "fallbacks": self.fallbacks,
        }


class Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("log_formulation", C.c_int32),
        ("area", C.c_double * 3), ("q", C.c_void_p * 3), ("dd", C.c_void_p * 3),
        ("bc_q", C.c_void_p * 2), ("bc_dd", C.c_void_p * 2), ("bc_total", C.c_void_p * 2),
    ]


class ZonePlan(C.Structure):
    _fields_ = [
        ("nzones", C.c_int32), ("blocks", C.c_int32 * PFX_MAX_ZONES), ("zone_ms", C.c_double * PFX_MAX_ZONES),
        ("kernel_ms", C.c_double), ("variant", C.c_int32), ("reserved", C.c_int32),
    ]


class ConstraintRxn(C.Structure):
    """pfx_constraint_rxn: the reaction of a mineral or gas constraint row."""
    _fields_ = [("nspec", C.c_int32), ("has_logK_coef", C.c_int32), ("specid", C.c_int32 * PFX_MAX_RXN_SPEC),
                ("stoich", C.c_double * PFX_MAX_RXN_SPEC), ("logK", C.c_double), ("logK_coef", C.c_double * 5)]


class KrylovOpts(C.Structure):
    """pfx_krylov_opts (include/pfx_gi_solver.h); zeros = PETSc defaults."""
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("dtol", C.c_double),
                ("max_iterations", C.c_int32), ("reserved", C.c_int32)]


class KrylovStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("reason", C.c_int32), ("rnorm0", C.c_double), ("rnorm", C.c_double),
                ("ms", C.c_double)]


class NewtonOpts(C.Structure):
    """pfx_newton_opts; zeros = PETSc SNES defaults."""
    _fields_ = [("atol", C.c_double), ("rtol", C.c_double), ("stol", C.c_double),
                ("max_iterations", C.c_int32), ("reserved", C.c_int32), ("krylov", KrylovOpts),
                ("itol_res", C.c_double), ("itol_update", C.c_double), ("itol_rel_update", C.c_double)]


class NewtonStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("reason", C.c_int32), ("linear_iterations", C.c_int32),
                ("reserved", C.c_int32), ("fnorm0", C.c_double), ("fnorm", C.c_double),
                ("ms_assembly", C.c_double), ("ms_linear", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class GiStepStats(C.Structure):
    _fields_ = [("newton", NewtonStats), ("reason", C.c_int32), ("cuts", C.c_int32),
                ("newton_iterations", C.c_int32), ("wasted_newton_iterations", C.c_int32),
                ("linear_iterations", C.c_int32), ("reserved", C.c_int32), ("dt", C.c_double),
                ("ms_total", C.c_double)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in ("reserved", "newton")}
        d["newton"] = self.newton.as_dict()
        return d


PFX_KSP_CONVERGED_RTOL, PFX_KSP_CONVERGED_ATOL = 2, 3
PFX_KSP_DIVERGED_ITS, PFX_KSP_DIVERGED_DTOL, PFX_KSP_DIVERGED_BREAKDOWN = -3, -4, -5
PFX_KSP_DIVERGED_NANORINF, PFX_KSP_DIVERGED_PC_FAILED = -9, -11
PFX_SNES_CONVERGED_FNORM_ABS, PFX_SNES_CONVERGED_FNORM_RELATIVE, PFX_SNES_CONVERGED_SNORM_RELATIVE = 2, 3, 4
PFX_SNES_DIVERGED_LINEAR_SOLVE, PFX_SNES_DIVERGED_FNORM_NAN, PFX_SNES_DIVERGED_MAX_IT = -3, -4, -5
PFX_SNES_DIVERGED_MAX_NORM = -10
PFX_SNES_CONVERGED_ITOL_RES, PFX_SNES_CONVERGED_ITOL_UPDATE, PFX_SNES_CONVERGED_RELATIVE_UPDATE = 10, 11, 12


CELL_FIELDS = [f for f, _ in Cells._fields_ if f != "ncells"]


def default_opts(dt=1.0, rel_tol=1e-10, max_iterations=100, extra_iterations=0,
                 guess_mode=PFX_GUESS_PREVIOUS, reference_stop=0, default_guess=1e-9,
                 porosity=0.25, saturation=1.0, volume=1.0, den_kg=997.0, stream=None, fallback_after=0,
                 jacobian_flags=0, rel_change_tol=0.0, fallback_previous=0, stream_ordered=0):
    o = ReactOpts()
    o.dt = dt
    o.rel_tol = rel_tol
    o.max_iterations = max_iterations
    o.extra_iterations = extra_iterations
    o.guess_mode = guess_mode
    o.reference_stop = reference_stop
    o.default_guess = default_guess
    o.porosity, o.saturation, o.volume, o.den_kg = porosity, saturation, volume, den_kg
    o.stream = stream
    o.fallback_after = fallback_after
    o.jacobian_flags = jacobian_flags
    o.rel_change_tol = rel_change_tol
    o.fallback_previous = fallback_previous
    o.stream_ordered = stream_ordered
    return o
