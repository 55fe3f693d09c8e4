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


def test_batch_argument_checks():
    """The batch calls' host-side checks (no device work): the tile switch
    takes -1, 0 or 1; a null batch is an invalid argument for every batch
    call."""
    lib = _lib.load()
    assert lib.fr_batch_force_tile(2) == -1
    assert lib.fr_batch_force_tile(-2) == -1
    assert lib.fr_batch_force_tile(1) == 0
    assert lib.fr_batch_force_tile(-1) == 0
    i = ctypes.c_int()
    d = ctypes.c_double()
    assert lib.fr_batch_group_tuning(None, 1, ctypes.byref(i), ctypes.byref(i), ctypes.byref(i),
                                     ctypes.byref(i), ctypes.byref(d)) == -1
    assert lib.fr_batch_tuning(None, ctypes.byref(i), ctypes.byref(i), ctypes.byref(d)) == -1
    assert lib.fr_batch_update(None, 1) == -1
    assert lib.fr_batch_create(None, 0, None) == -1
    assert lib.fr_batch_destroy(None) == 0


def test_device_error_without_gpu():
    """Valid input reaches the device stage; without a GPU that is FR_EHIP.
    The GPU is detected by the kernel driver's device node, not by a HIP
    runtime of another library in this process."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
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


def test_gmsh_reader_roundtrip(tmp_path):
    """write_gmsh -> read_gmsh gives back vertices, cells and boundary vertex
    sets, in 3D (hexes) and 2D (quads), with the meshBounds name mapping."""
    for D, spec in ((3, {"xmin": "char", "xmax": "sup_out", "ymin": "slip_wall", "ymax": "sub_out",
                         "zmin": "sub_in", "zmax": "adiabatic_noslip"}),
                    (2, {"xmin": "char", "xmax": "sub_out", "ymin": "slip_wall", "ymax": "sup_in"})):
        n = (3,) * D
        m = mesh.box(D, n, (0,) * D, (1,) * D, periodic=[False] * D, bc=spec, perturb=0.1, seed=2)
        inv = {v: k for k, v in mesh.BC.items()}
        names = [f"Surf_{i}" for i in range(len(m.bounds))]
        p = tmp_path / f"box{D}.msh"
        mesh.write_gmsh(m, str(p), names=names)
        bmap = {nm.lower(): inv[bc] for nm, (bc, _) in zip(names, m.bounds)}
        bmap["fluid"] = "fluid"
        r = mesh.read_gmsh(str(p), mesh_bounds=bmap)
        assert r.nDims == D
        assert np.array_equal(r.xv, m.xv) and np.array_equal(r.c2v, m.c2v)
        assert [bc for bc, _ in r.bounds] == [bc for bc, _ in m.bounds]
        for (_, a), (_, b) in zip(r.bounds, m.bounds):
            assert np.array_equal(a, b)


def test_gmsh_triangles(tmp_path):
    """Triangle cells (Gmsh type 2) round-trip; clockwise triangles come back
    counter-clockwise (the library's corner order)."""
    spec = {"xmin": "char", "xmax": "sub_out", "ymin": "slip_wall", "ymax": "sup_in"}
    m = mesh.tri_box((3, 3), (0, 0), (1, 1), periodic=[False, False], bc=spec, perturb=0.1, seed=3)
    inv = {v: k for k, v in mesh.BC.items()}
    names = [f"S{i}" for i in range(len(m.bounds))]
    p = tmp_path / "tri.msh"
    mesh.write_gmsh(m, str(p), names=names)
    bmap = {nm.lower(): inv[bc] for nm, (bc, _) in zip(names, m.bounds)} | {"fluid": "fluid"}
    r = mesh.read_gmsh(str(p), mesh_bounds=bmap)
    assert r.nNodes == 3 and np.array_equal(r.c2v, m.c2v) and np.array_equal(r.xv, m.xv)
    cw = mesh.Mesh(2, m.xv, m.c2v[:, [0, 2, 1]], None, m.bounds)
    mesh.write_gmsh(cw, str(p), names=names)
    r = mesh.read_gmsh(str(p), mesh_bounds=bmap)
    X = r.xv[r.c2v]
    area = (X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) - (X[:, 1, 1] - X[:, 0, 1]) * (X[:, 2, 0] - X[:, 0, 0])
    assert (area > 0).all()


def test_gmsh_reader_names_and_errors(tmp_path):
    """Physical names are unquoted and lower-cased; unknown names and cell
    types raise the reference's errors (geo.cpp:494-510, 636-641)."""
    text = """$MeshFormat
2.2 0 8
$EndMeshFormat
$PhysicalNames
2
1 1 "Wall"
2 2 "FLUID"
$EndPhysicalNames
$Nodes
6
1 0 0 0
2 1 0 0
3 2 0 0
4 0 1 0
5 1 1 0
6 2 1 0
$EndNodes
$Elements
4
1 1 2 1 1 1 2
2 1 2 1 1 2 3
3 3 2 2 2 1 2 5 4
4 3 2 2 2 2 3 6 5
$EndElements
"""
    p = tmp_path / "q.msh"
    p.write_text(text)
    r = mesh.read_gmsh(str(p), mesh_bounds={"wall": "slip_wall", "fluid": "fluid"})
    assert r.nDims == 2 and r.nEles == 2
    assert r.c2v.tolist() == [[0, 1, 4, 3], [1, 2, 5, 4]]
    assert r.bounds[0][0] == mesh.BC["slip_wall"] and r.bounds[0][1].tolist() == [0, 1, 2]
    with pytest.raises(ValueError, match="Unrecognized mesh boundary"):
        mesh.read_gmsh(str(p), mesh_bounds={"inlet": "sup_in", "fluid": "fluid"})
    with pytest.raises(ValueError, match="Unrecognized boundary condition"):
        mesh.read_gmsh(str(p), mesh_bounds={"wall": "no_such_bc", "fluid": "fluid"})
    p.write_text(text.replace("3 3 2 2 2 1 2 5 4", "3 4 2 2 2 1 2 5 4"))
    with pytest.raises(ValueError, match="element type not recognized"):
        mesh.read_gmsh(str(p), mesh_bounds={"wall": "slip_wall", "fluid": "fluid"})
    p.write_text(text.replace("3 3 2 2 2 1 2 5 4", "3 2 2 2 2 1 2 5"))
    with pytest.raises(ValueError, match="mixed cell types"):
        mesh.read_gmsh(str(p), mesh_bounds={"wall": "slip_wall", "fluid": "fluid"})
