"""CPU checks of the HIP library: it loads, exports every entry point the C
header declares, and its host-side validation reports the reference's fatal
conditions as error codes (no compute calls, so no GPU is needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from fr_learn_amd import _lib, cases, mesh
from fr_learn_amd.solver import Solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, f"binding missing for {s}"


def _expect(code, cfg, m):
    with pytest.raises(_lib.FrError) as ei:
        Solver(cfg, m)
    assert ei.value.code == code, str(ei.value)
    return str(ei.value)


def test_validation_errors():
    m = mesh.box(3, (3, 3, 3), (0, 0, 0), (1, 1, 1))
    assert "Roe not implemented in 3D" in _expect(-1, cases.config(nDims=3, riemannType=1), m)
    _expect(-1, cases.config(nDims=3, order=7), m)
    _expect(-1, cases.config(nDims=3, timeType=3), m)
    _expect(-1, cases.config(nDims=3, vcjhScheme=3, order=1), m)
    # unmatched boundary: a wall box without boundary sets
    m2 = mesh.box(3, (3, 3, 3), (0, 0, 0), (1, 1, 1), periodic=[False, True, True], bc="slip_wall")
    m2.bounds = []
    _expect(-2, cases.config(nDims=3), m2)


def test_device_error_without_gpu():
    """Valid input reaches the device stage; without a GPU that is FR_EHIP."""
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    m = mesh.box(3, (3, 3, 3), (0, 0, 0), (2 * np.pi,) * 3)
    _expect(-4, cases.config(nDims=3), m)


def test_mesh_requires_three_periodic_cells():
    with pytest.raises(ValueError):
        mesh.box(3, (2, 3, 3), (0, 0, 0), (1, 1, 1))


def test_gmsh_writer_roundtrip(tmp_path):
    m = mesh.box(3, (3, 3, 3), (0, 0, 0), (1, 1, 1), periodic=[False] * 3, bc="slip_wall")
    p = tmp_path / "box.msh"
    mesh.write_gmsh(m, str(p), names=["walls"] * len(m.bounds))
    text = p.read_text()
    assert "$Elements" in text and text.count("\n") > 27
