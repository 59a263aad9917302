"""The ISO_C_BINDING module compiles with the image's Fortran compiler,
links against libpfx_chem.so and binds types of the same size and field
offsets as the C structs; the probe calls only host-side argument checks."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from pflotran_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "pflotran_amd", "fortran")
LIBDIR = os.path.join(ROOT, "pflotran_amd")
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)


@pytest.mark.skipif(FC is None, reason="no Fortran compiler in this image")
def test_fortran_module_binds_the_c_abi(tmp_path):
    subprocess.run([FC, "-c", os.path.join(FDIR, "pfx_chem_mod.F90")], cwd=tmp_path, check=True)
    subprocess.run([FC, "-o", "probe", os.path.join(FDIR, "pfx_abi_probe.F90"), "pfx_chem_mod.o",
                    f"-L{LIBDIR}", "-lpfx_chem", f"-Wl,-rpath,{LIBDIR}"], cwd=tmp_path, check=True)
    out = subprocess.run([str(tmp_path / "probe")], check=True, capture_output=True, text=True).stdout
    lines = out.strip().splitlines()
    vals = dict(line.split() for line in lines if not line.startswith("off "))
    offsets = {line.split()[1]: int(line.split()[2]) for line in lines if line.startswith("off ")}
    assert int(vals["abi"]) == abi.PFX_ABI_VERSION
    assert int(vals["create_bad"]) == abi.PFX_E_ARG
    assert int(vals["react_null"]) == abi.PFX_E_ARG
    assert int(vals["mlp_null"]) == abi.PFX_E_ARG
    assert int(vals["sizeof_desc"]) == C.sizeof(abi.ReactionDesc)
    assert int(vals["sizeof_cells"]) == C.sizeof(abi.Cells)
    assert int(vals["sizeof_opts"]) == C.sizeof(abi.ReactOpts)
    assert int(vals["sizeof_stats"]) == C.sizeof(abi.Stats)
    assert int(vals["sizeof_grid"]) == C.sizeof(abi.Grid)
    assert int(vals["sizeof_plan"]) == C.sizeof(abi.ZonePlan)
    assert int(vals["solver_null"]) == abi.PFX_E_ARG
    assert int(vals["sizeof_newton_opts"]) == C.sizeof(abi.NewtonOpts)
    assert int(vals["sizeof_newton_stats"]) == C.sizeof(abi.NewtonStats)
    assert int(vals["sizeof_krylov_stats"]) == C.sizeof(abi.KrylovStats)
    assert int(vals["sizeof_constraint_rxn"]) == C.sizeof(abi.ConstraintRxn)
    # every field of every bound type sits at the C offset (tests/test_abi.py
    # holds the ctypes mirror to the header's offsetof), so two swapped
    # fields of one size are caught as well
    structs = [abi.ReactionDesc, abi.Cells, abi.ReactOpts, abi.Stats, abi.Grid, abi.ZonePlan, abi.ConstraintRxn,
               abi.KrylovOpts, abi.KrylovStats, abi.NewtonOpts, abi.NewtonStats]
    expected = {f"{st.__name__}.{f}": getattr(st, f).offset for st in structs for f, _ in st._fields_}
    assert offsets == expected
