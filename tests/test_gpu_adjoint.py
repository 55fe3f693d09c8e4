"""Adjoint (vector-Jacobian products) of the residual and of RK steps.

Checks: exact transposition for the linear advection residual (dot-product
test to rounding); the FP64 CPU reference gradient J^T lam from the oracle's
dense Jacobian (central differences, column by column) on small meshes;
dot-product tests against central finite differences of the GPU residual and
update for Euler with periodic and wall boundaries; checkpointed multi-step
adjoints against the unrolled chain and against finite differences of a
kinetic-energy objective (bar 1e-5)."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases
from oracle import Oracle

from helpers import MIXED_BCS, QUAD_BCS, WALL_BCS, advection_cfg, euler_cfg, hex_box, perturbed_state, quad_box

pytestmark = pytest.mark.gpu


def _solver(cfg, m, ic, seed=0, amp=1e-2):
    print(f"adjoint case nDims={cfg['nDims']} P={cfg['order']} eq={cfg['equation']} nEles={m.nEles}", flush=True)
    g = Solver(cfg, m)
    U = perturbed_state(g.initialize(ic), seed=seed, amp=amp) if amp else g.initialize(ic)
    g.set_state(U)
    return g, U


def _residual(g, U):
    g.set_state(U)
    g.calc_residual(0)
    return g.residual(0)


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("periodic", [True, False])
def test_vjp_residual_linear_advection(D, periodic):
    """The advection residual is linear in U, so <lam, R v> = <R^T lam, v> to rounding."""
    cfg = advection_cfg(D, 3)
    if D == 3:
        m = hex_box(3, periodic=periodic, bc=None if periodic else "sup_out", L=2.0)
    else:
        m = quad_box(4, periodic=periodic, bc=None if periodic else QUAD_BCS)
    g, U = _solver(cfg, m, lambda x, c: cases.ic_gaussian(x - (1.0 if D == 3 else 0.0), c), amp=0)
    rng = np.random.default_rng(1)
    lam = rng.standard_normal(U.shape)
    v = rng.standard_normal(U.shape)
    lhs = float(np.sum(lam * _residual(g, v)))
    g.set_state(U)
    rhs = float(np.sum(g.vjp_residual(lam) * v))
    print("dot test", lhs, rhs)
    assert _rel(lhs, rhs) < 1e-12


def _dense_jacobian_T_lam(cfg, m, U, lam, h=1e-6):
    """FP64 CPU reference: J^T lam with J from central differences of the oracle
    residual, one column per state entry (small meshes only)."""
    o = Oracle(cfg, m)
    flat = U.reshape(-1)
    out = np.zeros_like(flat)
    for i in range(flat.size):
        step = h * max(abs(flat[i]), 1.0)
        up, dn = flat.copy(), flat.copy()
        up[i] += step
        dn[i] -= step
        o.set_state(up.reshape(U.shape))
        o.calc_residual(0)
        rp = o.residual(0)
        o.set_state(dn.reshape(U.shape))
        o.calc_residual(0)
        rm = o.residual(0)
        out[i] = np.sum(lam * (rp - rm)) / (2 * step)
    return out.reshape(U.shape)


@pytest.mark.parametrize("case", ["hex_periodic", "hex_walls", "quad_roe_bcs"])
def test_vjp_residual_vs_cpu_reference_gradient(case):
    if case == "hex_periodic":
        cfg, m, ic = euler_cfg(3, 1), hex_box(3), cases.ic_tgv
    elif case == "hex_walls":
        cfg = euler_cfg(3, 1)
        m = hex_box(3, periodic=False, bc=WALL_BCS, L=1.0)
        ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4)
    else:
        # a free stream oblique to every boundary: finite differences need the
        # normal velocity away from 0, where |vn| and the characteristic
        # condition's in/outflow switch have their kinks
        cfg = euler_cfg(2, 2, riemannType=1)
        m = quad_box(3, periodic=False, bc=QUAD_BCS)
        ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.2], p=1.0 / 1.4)
    g, U = _solver(cfg, m, ic, amp=2e-2)
    lam = np.random.default_rng(2).standard_normal(U.shape)
    got = g.vjp_residual(lam)
    want = _dense_jacobian_T_lam(cfg, m, U, lam)
    err = np.abs(got - want).max() / np.abs(want).max()
    print("max rel err vs CPU reference gradient", err)
    assert err < 1e-6


@pytest.mark.parametrize("bcs", [None, MIXED_BCS])
@pytest.mark.parametrize("order", [2, 3])
def test_vjp_residual_dot_fd(bcs, order):
    cfg = euler_cfg(3, order)
    if bcs is None:
        m, ic = hex_box(3), cases.ic_tgv
    else:
        m = hex_box(3, periodic=False, bc=bcs, L=1.0)
        ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4)
    g, U = _solver(cfg, m, ic, amp=2e-2)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(U.shape)
    dU = U * rng.standard_normal(U.shape)
    h = 1e-6
    fd = float(np.sum(lam * (_residual(g, U + h * dU) - _residual(g, U - h * dU)))) / (2 * h)
    g.set_state(U)
    adj = float(np.sum(g.vjp_residual(lam) * dU))
    print("fd", fd, "adjoint", adj)
    assert _rel(fd, adj) < 1e-7


@pytest.mark.parametrize("timeType", [0, 4, 5])
def test_vjp_step_dot_fd(timeType):
    cfg = euler_cfg(3, 2, timeType=timeType, dt=2e-3)
    g, U = _solver(cfg, hex_box(3), cases.ic_tgv, amp=2e-2)
    rng = np.random.default_rng(4)
    lam = rng.standard_normal(U.shape)
    dU = U * rng.standard_normal(U.shape)
    h = 1e-6

    def step(V):
        g.set_state(V)
        g.update(1)
        return g.get_state()

    fd = float(np.sum(lam * (step(U + h * dU) - step(U - h * dU)))) / (2 * h)
    g.set_state(U)
    adj = float(np.sum(g.vjp_step(lam) * dU))
    print("fd", fd, "adjoint", adj)
    assert _rel(fd, adj) < 1e-7
    assert np.array_equal(g.get_state(), U)  # the step adjoint leaves the state alone


def test_vjp_steps_checkpointing_and_ke_gradient():
    """7 RK4 steps with a checkpoint every 3 equal the unrolled chain of
    one-step adjoints; the gradient of J = KE(U_N) matches finite differences."""
    cfg = euler_cfg(3, 2, timeType=4, dt=2e-3)
    g, U0 = _solver(cfg, hex_box(3), cases.ic_tgv, amp=2e-2)
    n = 7
    states = [U0]
    for _ in range(n):
        g.update(1)
        states.append(g.get_state())
    lamN = g.kinetic_energy_gradient()
    lam = lamN
    for i in range(n - 1, -1, -1):
        g.set_state(states[i])
        lam = g.vjp_step(lam)
    g.set_state(U0)
    got = g.vjp_steps(n, 3, lamN)
    assert np.abs(got - lam).max() <= 1e-12 * np.abs(lam).max()
    assert np.abs(g.get_state() - states[-1]).max() <= 1e-13 * np.abs(U0).max()
    # J(U0) = KE(U_N) along a direction
    dU = U0 * np.random.default_rng(5).standard_normal(U0.shape)
    h = 1e-6

    def J(V):
        g.set_state(V)
        g.update(n)
        return g.kinetic_energy()[0]

    fd = (J(U0 + h * dU) - J(U0 - h * dU)) / (2 * h)
    adj = float(np.sum(got * dU))
    print("fd", fd, "adjoint", adj)
    assert _rel(fd, adj) < 1e-5


# ---------------------------------------------------------------------------
# viscous (LDG) residual
# ---------------------------------------------------------------------------
def _ns_cfg(D, P, **kw):
    base = dict(viscous=1, mu_inf=0.05, prandtl=0.72, penFact=0.3, tau=0.7, fixVis=0, c_sth=0.4, rt_inf=0.3)
    base.update(kw)
    return euler_cfg(D, P, **base)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("periodic", [True, False])
def test_vjp_residual_linear_advection_diffusion(D, periodic):
    """Advection-diffusion with LDG is linear in U: exact transposition."""
    cfg = advection_cfg(D, 3, viscous=1, diffD=0.1, penFact=0.4)
    if D == 3:
        m = hex_box(3, periodic=periodic, bc=None if periodic else "sup_out", L=2.0)
    else:
        m = quad_box(4, periodic=periodic, bc=None if periodic else QUAD_BCS)
    g, U = _solver(cfg, m, lambda x, c: cases.ic_gaussian(x - (1.0 if D == 3 else 0.0), c), amp=0)
    rng = np.random.default_rng(12)
    lam = rng.standard_normal(U.shape)
    v = rng.standard_normal(U.shape)
    lhs = float(np.sum(lam * _residual(g, v)))
    g.set_state(U)
    rhs = float(np.sum(g.vjp_residual(lam) * v))
    print("dot test", lhs, rhs)
    assert _rel(lhs, rhs) < 1e-12


@pytest.mark.parametrize("case", ["hex_periodic", "hex_walls", "quad_bcs"])
def test_vjp_viscous_residual_vs_cpu_reference_gradient(case):
    if case == "hex_periodic":
        cfg, m, ic = _ns_cfg(3, 1, viscFix=0), hex_box(3), cases.ic_tgv
    elif case == "hex_walls":
        cfg = _ns_cfg(3, 1, viscFix=1, fixVis=1)
        m = hex_box(3, periodic=False, bc=WALL_BCS, L=1.0)
        ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4)
    else:
        cfg = _ns_cfg(2, 2, riemannType=1)
        m = quad_box(3, periodic=False, bc=dict(QUAD_BCS, ymin="adiabatic_noslip"))
        ic = lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.2], p=1.0 / 1.4)
    g, U = _solver(cfg, m, ic, amp=2e-2)
    lam = np.random.default_rng(13).standard_normal(U.shape)
    got = g.vjp_residual(lam)
    want = _dense_jacobian_T_lam(cfg, m, U, lam)
    err = np.abs(got - want).max() / np.abs(want).max()
    print("max rel err vs CPU reference gradient", err)
    assert err < 1e-6


@pytest.mark.parametrize("viscFix", [0, 1])
def test_vjp_viscous_residual_and_step_dot_fd(viscFix):
    cfg = _ns_cfg(3, 2, viscFix=viscFix, dt=1e-3)
    g, U = _solver(cfg, hex_box(3, periodic=False, bc=MIXED_BCS, L=1.0),
                   lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4), amp=2e-2)
    rng = np.random.default_rng(14)
    lam = rng.standard_normal(U.shape)
    dU = U * rng.standard_normal(U.shape)
    h = 1e-6
    fd = float(np.sum(lam * (_residual(g, U + h * dU) - _residual(g, U - h * dU)))) / (2 * h)
    g.set_state(U)
    adj = float(np.sum(g.vjp_residual(lam) * dU))
    print("residual: fd", fd, "adjoint", adj)
    assert _rel(fd, adj) < 1e-7

    def step(V):
        g.set_state(V)
        g.update(1)
        return g.get_state()

    fd = float(np.sum(lam * (step(U + h * dU) - step(U - h * dU)))) / (2 * h)
    g.set_state(U)
    adj = float(np.sum(g.vjp_step(lam) * dU))
    print("step: fd", fd, "adjoint", adj)
    assert _rel(fd, adj) < 1e-7


def test_ns_tgv_adjoint_workload_small():
    """The Navier-Stokes TGV adjoint workload of bench.py --adjoint at 4^3 and
    12 steps (checkpoint every 5): gradient of KE(T) with respect to the
    initial velocity against a central difference (bar 1e-5)."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    r = bench.tgv_ns_adjoint(n=4, steps=12, every=5)
    print(r)
    assert r["rel_diff_vs_fd"] < 1e-5
