"""Curved hexes on the GPU (20-node serendipity, 27 / 64 / 125 / 216-node
Lagrange, as geo::readGmsh reads them, geo.cpp:597-647) against the oracle,
whose shape functions and metrics tests/test_oracle_curved.py pins.

The library builds the metrics from the shape-function derivatives of every
node (solver::calcTransforms, solver.cpp:1227-1356).  Bar: FP64 per-field
max|gpu - oracle| / max|oracle| <= 1e-12 (helpers.TOL_FP64) for the
geometry, the residual and RK4 steps; the Gmsh reader's round trip gives the
same mesh."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases, mesh
from oracle import Oracle

from curved import curved_box
from helpers import MIXED_BCS, TOL_FP64, euler_cfg, perturbed_state, rel_err

pytestmark = pytest.mark.gpu


def _run(cfg, m, ic, steps=2, amp=1e-2):
    print(f"curved case nNodes={m.nNodes} P={cfg['order']} nEles={m.nEles} bounds={len(m.bounds)}", flush=True)
    o = Oracle(cfg, m)
    g = Solver(cfg, m)
    # curved elements are neither affine nor linear: per-point geometry only
    assert g.geometry_mode() == 0 and g.geometry_mode(2) == 0
    for which in (0, 1, 2, 3, 4):
        err = rel_err(g.geometry(which), o.geometry(which), axis=None)
        assert err <= TOL_FP64, (which, err)
    U = perturbed_state(o.initialize(ic), seed=2, amp=amp)
    o.set_state(U)
    g.set_state(U)
    o.calc_residual(0)
    g.calc_residual(0)
    err = rel_err(g.residual(0), o.residual(0))
    print("residual rel err", err, flush=True)
    assert err.max() <= TOL_FP64
    o.update(steps)
    g.update(steps)
    err = rel_err(g.get_state(), o.get_state())
    print("state rel err", err, flush=True)
    assert err.max() <= TOL_FP64
    g.close()
    o.close()


@pytest.mark.parametrize("nNodes,order", [(20, 2), (27, 3), (64, 3), (125, 4), (216, 3)])
def test_periodic_tgv(nNodes, order):
    m = curved_box(3, nNodes, L=2 * np.pi, amp=0.15)
    _run(euler_cfg(3, order, dt=1e-3), m, cases.ic_tgv)


@pytest.mark.parametrize("nNodes", [20, 27, 64])
def test_boundaries(nNodes):
    m = curved_box(3, nNodes, L=1.0, amp=0.02, periodic=False, bc=MIXED_BCS)
    ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4)
    _run(euler_cfg(3, 3, dt=1e-3), m, ic, amp=2e-2)


def test_gmsh_round_trip(tmp_path):
    """A 27-node and a 20-node mesh written as Gmsh 2.2 (types 12 and 17) and
    read back give the residual of the original mesh."""
    for nNodes in (27, 20):
        m = curved_box(3, nNodes, L=1.0, amp=0.02, periodic=False, bc=MIXED_BCS)
        path = tmp_path / f"curved{nNodes}.msh"
        names = ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
        mesh.write_gmsh(m, path, names)
        inv = {v: k for k, v in mesh.BC.items()}
        m2 = mesh.read_gmsh(path, dict({nm: inv[m.bounds[i][0]] for i, nm in enumerate(names)}, fluid="fluid"))
        assert m2.nNodes == nNodes and np.array_equal(m2.c2v, m.c2v)
        cfg = euler_cfg(3, 2, dt=1e-3)
        ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4)
        a, b = Solver(cfg, m), Solver(cfg, m2)
        U = perturbed_state(a.initialize(ic), seed=4)
        for s in (a, b):
            s.set_state(U)
            s.calc_residual(0)
        assert np.array_equal(a.residual(0), b.residual(0))
        a.close()
        b.close()
