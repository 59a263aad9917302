"""The C-ABI library loads and exports every entry point declared in
include/pfx_chem.h, and the ctypes mirror matches the C struct layouts."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

from pflotran_amd import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfx_chem.h")
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_functions():
    names = set()
    for h in HEADERS:
        names |= set(re.findall(r"^\s*int\s+(pfx_\w+)\s*\(", open(h).read(), flags=re.M))
    return sorted(names)


def test_header_declares_expected_entry_points():
    assert declared_functions() == sorted(runtime.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = runtime.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.pfx_abi_version() == abi.PFX_ABI_VERSION


def test_null_arguments_rejected_without_gpu_work():
    lib = runtime.load_library()
    assert lib.pfx_tables_create(None, None) == abi.PFX_E_ARG
    assert lib.pfx_react_batch(None, None, None, None) == abi.PFX_E_ARG
    assert lib.pfx_tables_destroy(None) == abi.PFX_E_ARG
    assert lib.pfx_mlp_create(0, None, 0, 0, None) == abi.PFX_E_ARG
    assert lib.pfx_mlp_destroy(None) == abi.PFX_E_ARG
    assert lib.pfx_mlp_forward(None, None, 0, None, None) == abi.PFX_E_ARG
    assert lib.pfx_surrogate_guess(None, None, None, None) == abi.PFX_E_ARG
    assert lib.pfx_bsr_create(0, 15, 0, None, None, None) == abi.PFX_E_ARG
    assert lib.pfx_bsr_solve(None, None, None, None, None, None, None) == abi.PFX_E_ARG
    assert lib.pfx_gi_solver_create(None, None, None) == abi.PFX_E_ARG
    assert lib.pfx_gi_solver_step(None, None, None, None, None, None) == abi.PFX_E_ARG
    assert lib.pfx_gi_solver_timestep(None, None, None, None, None, -1, 0.5, None) == abi.PFX_E_ARG
    assert lib.pfx_os_transport_step(None, None, None, 0, None, None, None, None) == abi.PFX_E_ARG


def test_every_header_compiles_as_c(tmp_path):
    src = tmp_path / "inc.c"
    src.write_text("".join(f'#include "{h}"\n' for h in HEADERS) + "int main(void){return 0;}\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-o", str(tmp_path / "inc"), str(src)], check=True)


@pytest.mark.parametrize("struct,cname", [(abi.ReactionDesc, "pfx_reaction_desc"), (abi.Cells, "pfx_cells"),
                                          (abi.ReactOpts, "pfx_react_opts"), (abi.Stats, "pfx_stats"),
                                          (abi.Grid, "pfx_grid"), (abi.ZonePlan, "pfx_zone_plan"),
                                          (abi.KrylovOpts, "pfx_krylov_opts"), (abi.KrylovStats, "pfx_krylov_stats"),
                                          (abi.NewtonOpts, "pfx_newton_opts"), (abi.NewtonStats, "pfx_newton_stats"),
                                          (abi.GiStepStats, "pfx_gi_step_stats"),
                                          (abi.ConstraintRxn, "pfx_constraint_rxn")])
def test_struct_layout_matches_header(tmp_path, struct, cname):
    fields = [f for f, _ in struct._fields_]
    src = ['#include <stddef.h>', '#include <stdio.h>'] + [f'#include "{h}"' for h in HEADERS] + ['int main(void){',
           f'printf("%zu\\n", sizeof({cname}));']
    src += [f'printf("%zu\\n", offsetof({cname}, {f}));' for f in fields]
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == C.sizeof(struct)
    for f, off in zip(fields, out[1:]):
        assert getattr(struct, f).offset == off, f
