"""C-ABI export check and host-side argument validation (no GPU needed)."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from helpers import cube
from newton_solver_fem_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "fem_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fem_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    # No undeclared binding either.
    assert set(_lib.SIGNATURES) == set(names)


def test_struct_layouts_match_header():
    text = (ROOT / "include" / "fem_hip.h").read_text()
    for cname, pyt in (("fem_model_desc", _lib.ModelDesc), ("fem_params", _lib.Params)):
        body = re.search(rf"typedef struct {cname} \{{(.*?)\}} {cname};", text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            names = decl.split()[-1] if "," not in decl else decl
            for n in re.findall(r"(\w+)(?:\[\d+\])?\s*(?:,|$)", decl.split(None, 1)[1] if "," in decl else names):
                fields.append(n)
        assert [f[0] for f in pyt._fields_] == fields, cname


def test_last_error_is_readable():
    assert isinstance(_lib.load().fem_last_error(), bytes)


def test_null_arguments_are_rejected_without_a_device():
    lib = _lib.load()
    assert lib.fem_create(None, None) == _lib.FEM_OK + 1
    assert b"null" in lib.fem_last_error()
    assert lib.fem_step(None, *([None] * 7), ctypes.c_float(0.01), None) == 1
    assert lib.fem_set_params(None, None) == 1


def _solver(model, **kw):
    from newton_solver_fem_amd.solver import SolverFEM

    base = dict(solver_type="cg", preconditioner_type="block_diag", self_contact_ke=0.0)
    base.update(kw)
    return SolverFEM(model, **base)


@pytest.mark.parametrize("tol", [0.0, -1e-5, np.nan, np.inf])
def test_tolerance_validation(tol):
    """Reference test_solver.py:136-144."""
    with pytest.raises(ValueError, match="finite and positive"):
        _solver(cube(), linear_solver_tolerance=tol)


@pytest.mark.parametrize("indices", [(-1,), (27,), (1.0,), ((1, 2),)])
def test_fixed_index_validation(indices):
    """Reference test_fixed_particles.py:136-147."""
    with pytest.raises(ValueError):
        _solver(cube(), fixed_particle_indices=indices)


def test_fixed_with_differentiable_rejected():
    with pytest.raises(ValueError, match="differentiable"):
        _solver(cube(), fixed_particle_indices=(0,), differentiable=True)


def test_unknown_krylov_methods_and_preconditioners_raise():
    """Reference _make_linear_solver solver.py:1222-1224 ('Invalid solver type')
    and warp's preconditioner() for an unknown ptype."""
    with pytest.raises(ValueError, match="Invalid solver type"):
        _solver(cube(), solver_type="minres")
    with pytest.raises(ValueError, match="preconditioner type"):
        _solver(cube(), preconditioner_type="ilu")


def test_host_model_is_rejected_loudly():
    # The HIP library is the only compute path: a CPU-resident model is an error.
    with pytest.raises(ValueError, match="HIP device"):
        _solver(cube(device="cpu"))


def test_public_names():
    import newton_solver_fem_amd as pkg

    assert pkg.SolverFEM is pkg.SolverLinearFEM
