"""The residual's intermediate arrays on the GPU (fr_get_array: F_spts,
Fn_fpts, disFn_fpts, dU_spts, dU_fpts -- the solver data members
include/solver.hpp:57-60 that ele.cpp:68-97 reads) against the oracle's, after
a residual evaluation of the same state (solver::calcResidual,
solver.cpp:252-305).

Bar: per-field max|gpu - oracle| / max|oracle| <= 1e-12 (helpers.TOL_FP64);
the hex P3 cases run the matrix-core stage kernel, the others the VALU one,
inviscid and viscous, with and without boundary faces, every geometry
layout of the inviscid linear hexes."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases
from oracle import Oracle

from helpers import MIXED_BCS, QUAD_BCS, TOL_FP64, WALL_BCS, euler_cfg, hex_box, perturbed_state, quad_box, rel_err
from test_gpu_viscous import ns_cfg

pytestmark = pytest.mark.gpu

INVISCID = ("F_spts", "Fn_fpts", "disFn_fpts")
VISCOUS = INVISCID + ("dU_spts", "dU_fpts")
ORACLE = {"F_spts": 2, "Fn_fpts": 3, "disFn_fpts": 4, "dU_spts": 5, "dU_fpts": 6}


def _check(cfg, m, ic, names, seed=0, amp=1e-2, modes=(None,)):
    print(f"arrays nDims={cfg['nDims']} P={cfg['order']} viscous={cfg['viscous']} nEles={m.nEles}", flush=True)
    o = Oracle(cfg, m)
    g = Solver(cfg, m)
    U = perturbed_state(o.initialize(ic), seed=seed, amp=amp)
    o.set_state(U)
    o.calc_residual(0)
    for mode in modes:
        if mode is not None:
            g.geometry_mode(mode)
        g.set_state(U)
        for name in names:
            a, b = g.array(name), o.array(ORACLE[name])
            err = rel_err(a, b, axis=tuple(range(a.ndim - 1)))
            print(name, "mode", g.geometry_mode(), "rel err", err, flush=True)
            assert err.max() <= TOL_FP64, (name, mode, err)
        # the call evaluated the residual of the state, as calc_residual(0) does
        assert rel_err(g.residual(0), o.residual(0)).max() <= TOL_FP64
    # the debug outputs are off again: further steps follow the oracle
    o.update(1)
    g.update(1)
    assert rel_err(g.get_state(), o.get_state()).max() <= TOL_FP64
    g.close()
    o.close()


def test_hex_p3_periodic_every_geometry_layout():
    _check(euler_cfg(3, 3, dt=1e-3), hex_box(3), cases.ic_tgv, INVISCID, modes=(0, 2, 3))


def test_hex_p3_affine_layout():
    from fr_learn_amd import mesh
    m = mesh.box(3, (3, 3, 3), (0, 0, 0), (2 * np.pi,) * 3, periodic=[True] * 3)
    _check(euler_cfg(3, 3, dt=1e-3), m, cases.ic_tgv, INVISCID, modes=(1,))


@pytest.mark.parametrize("order", [2, 3])
def test_hex_boundaries(order):
    m = hex_box(3, periodic=False, bc=MIXED_BCS, L=1.0)
    ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4)
    _check(euler_cfg(3, order, dt=1e-3), m, ic, INVISCID, amp=2e-2)


def test_quad_boundaries():
    m = quad_box(4, periodic=False, bc=QUAD_BCS)
    _check(euler_cfg(2, 3, dt=1e-3), m, lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4), INVISCID)


def test_hex_navier_stokes_walls():
    m = hex_box(3, periodic=False, bc=WALL_BCS, L=1.0)
    ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4)
    _check(ns_cfg(3, 3, dt=1e-4), m, ic, VISCOUS, amp=2e-2)


def test_quad_navier_stokes_periodic():
    _check(ns_cfg(2, 2, dt=1e-4), quad_box(4), lambda x, c: cases.ic_vortex(x, c, beta=5.0, mach=0.5), VISCOUS)


def test_refusals():
    from fr_learn_amd import _lib
    cfg = euler_cfg(3, 2, dt=1e-3)
    m = hex_box(3)
    g = Solver(cfg, m)
    g.set_state(g.initialize(cases.ic_tgv))
    for name in ("dU_spts", "dU_fpts"):  # no gradient on the inviscid path
        with pytest.raises(_lib.FrError):
            g.array(name)
    buf = np.zeros(8 * g.nSpts * g.nEles * g.nFields)
    for which in (1, 7):  # out of range, through the C ABI itself
        with pytest.raises(_lib.FrError):
            _lib.check(g._lib.fr_get_array(g._h, which, _lib.dptr(buf)))
    g.close()
    g32 = Solver(cfg, m, precision="fp32")
    g32.set_state(g32.initialize(cases.ic_tgv))
    with pytest.raises(_lib.FrError):
        g32.array("F_spts")
    g32.close()
