"""GPU (HIP, gfx950) FP64 path vs the CPU oracle on the same mesh and state.

Bar: per-field max|gpu - oracle| / max|oracle| <= 1e-12 for residuals, states
after RK steps, traces and geometry (helpers.TOL_FP64)."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases
from oracle import Oracle

from helpers import (MIXED_BCS, QUAD_BCS, TOL_FP64, WALL_BCS, advection_cfg, euler_cfg, hex_box,
                     perturbed_state, quad_box, rel_err)

pytestmark = pytest.mark.gpu


def _pair(cfg, m, ic, seed=0, amp=1e-2, precision="fp64"):
    print(f"case nDims={cfg['nDims']} P={cfg['order']} eq={cfg['equation']} nEles={m.nEles} prec={precision}",
          flush=True)
    o = Oracle(cfg, m)
    g = Solver(cfg, m, precision=precision)
    U = perturbed_state(o.initialize(ic), seed=seed, amp=amp)
    o.set_state(U)
    g.set_state(U)
    return o, g, U


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_hex_euler_residual_periodic(order):
    cfg = euler_cfg(3, order)
    o, g, _ = _pair(cfg, hex_box(3), cases.ic_tgv)
    o.calc_residual(0)
    g.calc_residual(0)
    err = rel_err(g.residual(0), o.residual(0))
    print("rel err", err)
    assert err.max() <= TOL_FP64


@pytest.mark.parametrize("bcs", [MIXED_BCS, WALL_BCS])
def test_hex_euler_residual_boundaries(bcs):
    cfg = euler_cfg(3, 3)
    m = hex_box(3, periodic=False, bc=bcs, L=1.0)
    o, g, _ = _pair(cfg, m, lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4), amp=2e-2)
    o.calc_residual(0)
    g.calc_residual(0)
    err = rel_err(g.residual(0), o.residual(0))
    print("rel err", err)
    assert err.max() <= TOL_FP64


@pytest.mark.parametrize("order", [2, 3])
def test_hex_advection_residual(order):
    cfg = advection_cfg(3, order)
    m = hex_box(3, periodic=False, bc="sup_out", L=2.0)
    o, g, _ = _pair(cfg, m, lambda x, c: cases.ic_gaussian(x - 1.0, c))
    o.calc_residual(0)
    g.calc_residual(0)
    assert rel_err(g.residual(0), o.residual(0)).max() <= TOL_FP64


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("riemann", [0, 1])
def test_quad_euler_residual(order, riemann):
    cfg = euler_cfg(2, order, riemannType=riemann)
    o, g, _ = _pair(cfg, quad_box(4), lambda x, c: cases.ic_vortex(x, c, beta=5.0, mach=0.5))
    o.calc_residual(0)
    g.calc_residual(0)
    err = rel_err(g.residual(0), o.residual(0))
    print("rel err", err)
    assert err.max() <= TOL_FP64


def test_quad_boundaries_and_advection():
    cfg = euler_cfg(2, 3, riemannType=1)
    o, g, _ = _pair(cfg, quad_box(4, periodic=False, bc=QUAD_BCS),
                    lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4))
    o.calc_residual(0)
    g.calc_residual(0)
    assert rel_err(g.residual(0), o.residual(0)).max() <= TOL_FP64
    cfg = advection_cfg(2, 4)
    o, g, _ = _pair(cfg, quad_box(5, periodic=True), cases.ic_gaussian)
    o.calc_residual(0)
    g.calc_residual(0)
    assert rel_err(g.residual(0), o.residual(0)).max() <= TOL_FP64


@pytest.mark.parametrize("timeType", [0, 2, 4, 5, 6])
def test_rk_update(timeType):
    cfg = euler_cfg(3, 3, timeType=timeType, dt=2e-3)
    o, g, _ = _pair(cfg, hex_box(3), cases.ic_tgv)
    o.update(3)
    g.update(3)
    err = rel_err(g.get_state(), o.get_state())
    print("state rel err", err)
    assert err.max() <= TOL_FP64
    for s in range(o.nRKSteps):
        assert rel_err(g.residual(s), o.residual(s)).max() <= TOL_FP64


def test_stage_verbs_match():
    """timeStepA / timeStepB / copies as separate verbs (solver.cpp:307-399)."""
    cfg = euler_cfg(3, 2, dt=1e-3)
    o, g, _ = _pair(cfg, hex_box(3), cases.ic_tgv)
    for s in (o, g):
        s.copy_u_u0()
        s.calc_residual(0)
        s.time_step_a(0, 0.5)
        s.calc_residual(1)
        s.copy_u0_u()
        s.time_step_b(0)
        s.time_step_b(1)
    assert rel_err(g.get_state(), o.get_state()).max() <= TOL_FP64


def test_traces_geometry_operators():
    cfg = euler_cfg(3, 3)
    o, g, _ = _pair(cfg, hex_box(3), cases.ic_tgv)
    o.calc_residual(0)
    g.calc_residual(0)
    assert rel_err(g.traces(), o.array(1)).max() <= TOL_FP64
    for w in range(5):
        assert np.array_equal(g.operator(w), o.operator(w)), w
    for w in range(5):
        a, b = g.geometry(w), o.geometry(w)
        ax = tuple(range(a.ndim))
        assert rel_err(a, b, axis=ax) <= TOL_FP64, w
    go, _ = o.faces()
    assert np.array_equal(g.faces(), go)


def test_norms_and_kinetic_energy():
    cfg = euler_cfg(3, 3)
    o, g, U = _pair(cfg, hex_box(4, perturb=0.0, rotate=False, shuffle=False), cases.ic_tgv, amp=0.0)
    o.calc_residual(0)
    g.calc_residual(0)
    for kind in (1, 2, 3):
        assert rel_err(g.norm_residual(kind), o.norm_residual(kind), axis=0) <= 1e-12
    ke, vol = g.kinetic_energy()
    assert abs(vol - (2 * np.pi) ** 3) <= 1e-12 * vol
    # TGV at t=0: KE / Vol = rho0 V0^2 / 8 up to the quadrature error of a 4^3 P3 mesh
    assert abs(ke / vol - 0.125) < 1e-4


def test_fp32_residual_and_steps():
    cfg = euler_cfg(3, 3, dt=2e-3)
    m = hex_box(4)
    o, g, U = _pair(cfg, m, cases.ic_tgv, precision="fp32")
    o.calc_residual(0)
    g.calc_residual(0)
    r_err = rel_err(g.residual(0), o.residual(0))
    print("fp32 residual rel err", r_err)
    assert r_err.max() < 1e-3
    o.update(5)
    g.update(5)
    Ug, Uo = g.get_state(), o.get_state()
    l2 = np.linalg.norm(Ug - Uo) / np.linalg.norm(Uo)
    print("fp32 state rel L2", l2)
    assert l2 < 1e-6


@pytest.mark.parametrize("timeType", [0, 2, 4, 5, 6])
@pytest.mark.parametrize("D", [2, 3])
def test_fp32_rk_schemes(timeType, D):
    """FP32 stage kernels with every RK scheme (stage states formed from the
    base and the previous residual in FP32, final combination in FP64) against
    the FP64 oracle: state rel L2 < 1e-6 after 4 steps."""
    cfg = euler_cfg(D, 3, timeType=timeType, dt=2e-3 if D == 3 else 1e-3)
    m = hex_box(3) if D == 3 else quad_box(5, periodic=True)
    ic = cases.ic_tgv if D == 3 else (lambda x, c: cases.ic_vortex(x, c, beta=5.0, mach=0.5))
    o, g, _ = _pair(cfg, m, ic, precision="fp32")
    o.update(4)
    g.update(4)
    Ug, Uo = g.get_state(), o.get_state()
    l2 = np.linalg.norm(Ug - Uo) / np.linalg.norm(Uo)
    print("fp32 state rel L2", l2)
    assert l2 < 1e-6


@pytest.mark.parametrize("dtType", [1, 2])
def test_cfl_time_step_and_local_update(dtType):
    """CFL dt (global min or per element) and RK4 steps with it."""
    cfg = euler_cfg(3, 3, timeType=4)
    o, g, _ = _pair(cfg, hex_box(3, perturb=0.2, seed=3), cases.ic_tgv, amp=2e-2)
    dto = o.calc_dt(dtType, 0.7)
    dtg = g.calc_dt(dtType, 0.7)
    assert abs(dtg - dto) <= 1e-12 * dto
    dto_e = np.array([o.dt(e if dtType == 2 else -1) for e in range(o.nEles)])
    assert np.abs(g.dt() - dto_e).max() <= 1e-12 * dto
    o.update(2)
    g.update(2)
    err = rel_err(g.get_state(), o.get_state())
    print("state rel err", err)
    assert err.max() <= TOL_FP64
    g.set_dt(1e-3)
    assert np.all(g.dt() == 1e-3)


def test_residual_on_gmsh_mesh(tmp_path):
    """A mesh read back from a Gmsh 2.2 file (geo::readGmsh path)."""
    from fr_learn_amd import mesh as fmesh
    m = hex_box(3, periodic=False, bc=WALL_BCS, L=1.0, seed=7)
    inv = {v: k for k, v in fmesh.BC.items()}
    names = [f"b{i}" for i in range(len(m.bounds))]
    p = tmp_path / "walls.msh"
    fmesh.write_gmsh(m, str(p), names=names)
    r = fmesh.read_gmsh(str(p), mesh_bounds={nm: inv[bc] for nm, (bc, _) in zip(names, m.bounds)} |
                        {"fluid": "fluid"})
    cfg = euler_cfg(3, 2)
    o, g, _ = _pair(cfg, r, lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4), amp=2e-2)
    o.calc_residual(0)
    g.calc_residual(0)
    assert rel_err(g.residual(0), o.residual(0)).max() <= TOL_FP64


@pytest.mark.parametrize("scheme", [1, 2, 3])
@pytest.mark.parametrize("D", [2, 3])
def test_vcjh_schemes_residual(scheme, D):
    """VCJH correction functions other than DG (SD, HU, C+; polynomials.cpp
    392-475) through the same kernels."""
    P = 3
    cfg = euler_cfg(D, P, vcjhScheme=scheme)
    m, ic = (hex_box(3), cases.ic_tgv) if D == 3 else (quad_box(4), lambda x, c: cases.ic_vortex(x, c))
    o, g, _ = _pair(cfg, m, ic)
    o.calc_residual(0)
    g.calc_residual(0)
    assert rel_err(g.residual(0), o.residual(0)).max() <= TOL_FP64


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("order", [2, 4])
def test_lobatto_points_residual_and_step(D, order):
    """Gauss-Lobatto solution points (sptsType 1, points.cpp:254-330)."""
    cfg = euler_cfg(D, order, sptsType=1, dt=1e-3)
    m, ic = (hex_box(3), cases.ic_tgv) if D == 3 else (quad_box(4), lambda x, c: cases.ic_vortex(x, c))
    o, g, _ = _pair(cfg, m, ic)
    o.calc_residual(0)
    g.calc_residual(0)
    assert rel_err(g.residual(0), o.residual(0)).max() <= TOL_FP64
    o.update(2)
    g.update(2)
    assert rel_err(g.get_state(), o.get_state()).max() <= TOL_FP64


@pytest.mark.parametrize("case", ["hex_mixed", "hex_walls", "quad_roe"])
def test_fp32_boundaries_and_traces(case):
    """FP32 states carry the energy less a reference energy (the mean
    pressure over gamma - 1); every boundary ghost state, the Roe flux and
    the traces returned to the caller take it back.  Per-field relative L2
    against the FP64 oracle after 4 RK4 steps (< 1e-6), traces to FP32
    rounding."""
    if case == "quad_roe":
        cfg = euler_cfg(2, 3, riemannType=1, dt=1e-3)
        m, ic = quad_box(5, periodic=False, bc=QUAD_BCS), (lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4))
    else:
        cfg = euler_cfg(3, 2, dt=1e-3)
        m = hex_box(3, periodic=False, bc=MIXED_BCS if case == "hex_mixed" else WALL_BCS, L=1.0)
        ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4)
    o, g, U = _pair(cfg, m, ic, amp=2e-2, precision="fp32")
    o.calc_residual(0)
    tr_err = rel_err(g.traces(), o.array(1))
    print("fp32 traces rel err", tr_err)
    assert tr_err.max() < 1e-6
    o.update(4)
    g.update(4)
    Ug, Uo = g.get_state(), o.get_state()
    l2 = np.array([np.linalg.norm(Ug[..., k] - Uo[..., k]) / np.linalg.norm(Uo[..., k]) for k in range(g.nFields)])
    print("fp32 per-field rel L2", l2)
    assert (l2 < 1e-6).all()
