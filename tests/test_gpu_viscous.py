"""GPU (HIP, gfx950) viscous path (LDG gradients and viscous fluxes) vs the
CPU oracle on the same mesh and state.

Bar: per-field max|gpu - oracle| / max|oracle| <= 1e-12 in FP64
(helpers.TOL_FP64)."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases
from oracle import Oracle

from helpers import MIXED_BCS, QUAD_BCS, TOL_FP64, WALL_BCS, hex_box, perturbed_state, quad_box, rel_err

pytestmark = pytest.mark.gpu


def ns_cfg(nDims=3, order=3, **kw):
    base = dict(nDims=nDims, order=order, equation=1, viscous=1, mu_inf=0.05, prandtl=0.72, penFact=0.5, tau=1.0,
                fixVis=0, c_sth=0.4, rt_inf=0.3, uBound=0.3, vBound=0.1, wBound=-0.05, pBound=1.0 / 1.4,
                TWall=1.0, RGas=1.0 / 1.4, uWall=0.1)
    base.update(kw)
    return cases.config(**base)


def ad_cfg(nDims=3, order=3, **kw):
    base = dict(nDims=nDims, order=order, equation=0, viscous=1, diffD=0.1, penFact=0.5, advectVx=1.0,
                advectVy=0.5, advectVz=-0.25)
    base.update(kw)
    return cases.config(**base)


def _pair(cfg, m, ic, seed=0, amp=1e-2, precision="fp64"):
    print(f"viscous case nDims={cfg['nDims']} P={cfg['order']} eq={cfg['equation']} viscFix={cfg['viscFix']} "
          f"fixVis={cfg['fixVis']} nEles={m.nEles} prec={precision}", flush=True)
    o = Oracle(cfg, m)
    g = Solver(cfg, m, precision=precision)
    U = perturbed_state(o.initialize(ic), seed=seed, amp=amp)
    o.set_state(U)
    g.set_state(U)
    return o, g, U


def _check_residual(o, g):
    o.calc_residual(0)
    g.calc_residual(0)
    err = rel_err(g.residual(0), o.residual(0))
    print("rel err", err)
    assert err.max() <= TOL_FP64


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("viscFix", [0, 1])
def test_hex_ns_residual_periodic(order, viscFix):
    cfg = ns_cfg(3, order, viscFix=viscFix)
    o, g, _ = _pair(cfg, hex_box(3), cases.ic_tgv)
    _check_residual(o, g)


@pytest.mark.parametrize("fixVis", [0, 1])
def test_hex_ns_residual_constant_viscosity(fixVis):
    cfg = ns_cfg(3, 3, fixVis=fixVis, penFact=-0.3, tau=0.2)
    o, g, _ = _pair(cfg, hex_box(3, seed=4), cases.ic_tgv, seed=3)
    _check_residual(o, g)


@pytest.mark.parametrize("bcs", [MIXED_BCS, WALL_BCS])
def test_hex_ns_residual_boundaries(bcs):
    cfg = ns_cfg(3, 3)
    m = hex_box(3, periodic=False, bc=bcs, L=1.0)
    o, g, _ = _pair(cfg, m, lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4), amp=2e-2)
    _check_residual(o, g)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_quad_ns_residual(order):
    cfg = ns_cfg(2, order, riemannType=order % 2)
    o, g, _ = _pair(cfg, quad_box(4), lambda x, c: cases.ic_vortex(x, c, beta=5.0, mach=0.5))
    _check_residual(o, g)


def test_quad_ns_residual_boundaries():
    cfg = ns_cfg(2, 3, riemannType=1)
    bcs = dict(QUAD_BCS, ymin="adiabatic_noslip", ymax="isothermal_noslip_moving")
    o, g, _ = _pair(cfg, quad_box(4, periodic=False, bc=bcs), lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4))
    _check_residual(o, g)


@pytest.mark.parametrize("D", [2, 3])
def test_advection_diffusion_residual(D):
    cfg = ad_cfg(D, 3)
    m = hex_box(3, periodic=False, bc="sup_out", L=2.0) if D == 3 else quad_box(5, periodic=False, bc="sup_out")
    shift = 1.0 if D == 3 else 0.0
    o, g, _ = _pair(cfg, m, lambda x, c: cases.ic_gaussian(x - shift, c))
    _check_residual(o, g)
    cfg = ad_cfg(D, 4)
    m = hex_box(3) if D == 3 else quad_box(5)
    o, g, _ = _pair(cfg, m, cases.ic_gaussian)
    _check_residual(o, g)


@pytest.mark.parametrize("timeType", [4, 5])
def test_ns_rk_update(timeType):
    cfg = ns_cfg(3, 3, timeType=timeType, dt=1e-3)
    o, g, _ = _pair(cfg, hex_box(3, periodic=False, bc=WALL_BCS, L=1.0),
                    lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4), amp=2e-2)
    o.update(2)
    g.update(2)
    err = rel_err(g.get_state(), o.get_state())
    print("state rel err", err)
    assert err.max() <= TOL_FP64
    for s in range(o.nRKSteps):
        assert rel_err(g.residual(s), o.residual(s)).max() <= TOL_FP64


def test_ns_fp32_steps():
    cfg = ns_cfg(3, 3, dt=1e-3)
    o, g, _ = _pair(cfg, hex_box(4), cases.ic_tgv, precision="fp32")
    o.calc_residual(0)
    g.calc_residual(0)
    r_err = rel_err(g.residual(0), o.residual(0))
    print("fp32 residual rel err", r_err)
    assert r_err.max() < 1e-3
    o.update(3)
    g.update(3)
    Ug, Uo = g.get_state(), o.get_state()
    l2 = np.linalg.norm(Ug - Uo) / np.linalg.norm(Uo)
    print("fp32 state rel L2", l2)
    assert l2 < 1e-6
