"""CFL time stepping over several updates (solver::update recomputes dt at
every step, solver.cpp:218-219, 236-240; solver::calcDt 293-305): the GPU's
per-element time steps and states after each of three updates against the
oracle (bar 1e-12 relative), the adjoint of CFL-stepped runs, and the residual
norms' NaN check (output.cpp:833-839)."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases
from fr_learn_amd._lib import FrError
from oracle import Oracle

from helpers import (QUAD_BCS, TOL_FP64, WALL_BCS, advection_cfg, euler_cfg, hex_box, perturbed_state, quad_box,
                     rel_err)

pytestmark = pytest.mark.gpu

CASES = {
    "hex_rusanov_rk4": lambda: (euler_cfg(3, 3, timeType=4), hex_box(3, perturb=0.2, seed=3), "tgv"),
    "hex_walls_rk4": lambda: (euler_cfg(3, 2, timeType=4), hex_box(3, periodic=False, bc=WALL_BCS, L=1.0, seed=5),
                              "uniform"),
    "quad_roe_bc_rk4": lambda: (euler_cfg(2, 3, riemannType=1, timeType=4),
                                quad_box(4, periodic=False, bc=QUAD_BCS, seed=2), "vortex"),
    "quad_jameson": lambda: (euler_cfg(2, 4, riemannType=0, timeType=5), quad_box(4, seed=4), "vortex"),
    "advection_rk4": lambda: (advection_cfg(3, 2, timeType=4), hex_box(3, perturb=0.2, seed=3), "gauss"),
    "euler_fe": lambda: (euler_cfg(3, 2, timeType=0), hex_box(3, perturb=0.2, seed=3), "tgv"),
}


def _ic(kind):
    if kind == "tgv":
        return cases.ic_tgv
    if kind == "vortex":
        return lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4)
    if kind == "gauss":
        return cases.ic_gaussian
    return lambda x, c: cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05][:c["nDims"]], p=1.0 / 1.4)


@pytest.mark.parametrize("dtType", [1, 2])
@pytest.mark.parametrize("case", sorted(CASES))
def test_cfl_time_step_every_update(case, dtType):
    cfg, m, ic = CASES[case]()
    print(f"cfl case {case} dtType {dtType} nEles {m.nEles}", flush=True)
    o, g = Oracle(cfg, m), Solver(cfg, m)
    U = perturbed_state(o.initialize(_ic(ic)), seed=1, amp=2e-2)
    o.set_state(U)
    g.set_state(U)
    dto, dtg = o.calc_dt(dtType, 0.3), g.calc_dt(dtType, 0.3)
    assert abs(dtg - dto) <= 1e-12 * dto
    dts = []
    for step in range(3):
        o.update(1)
        g.update(1)
        want = np.array([o.dt(e if dtType == 2 else -1) for e in range(o.nEles)])
        got = g.dt()
        err_dt = np.abs(got - want).max() / want.max()
        err = rel_err(g.get_state(), o.get_state())
        print("step", step, "dt", want.min(), "dt err", err_dt, "state err", err.max())
        assert err_dt <= 1e-12
        assert err.max() <= TOL_FP64
        dts.append(want.min())
    if ic != "gauss":
        # the time step follows the wave speeds of the flow
        assert abs(dts[-1] - dto) > 1e-9 * dto
    # a residual between updates stores new wave speeds (face.cpp:354), which
    # the next update's time step takes
    o.calc_residual(0)
    g.calc_residual(0)
    o.update(1)
    g.update(1)
    want = np.array([o.dt(e if dtType == 2 else -1) for e in range(o.nEles)])
    assert np.abs(g.dt() - want).max() <= 1e-12 * want.max()
    assert rel_err(g.get_state(), o.get_state()).max() <= TOL_FP64


def test_cfl_adjoint_steps_against_replayed_time_steps():
    """fr_vjp_steps with a global CFL time step: the adjoint keeps each step's
    dt as a constant, so it equals the derivative of the run that replays the
    recorded time steps (fr_set_dt before each step), by central differences."""
    cfg = euler_cfg(3, 2, timeType=4)
    m = hex_box(3, perturb=0.2, seed=3)
    g = Solver(cfg, m)
    U0 = perturbed_state(g.initialize(cases.ic_tgv), seed=2, amp=2e-2)
    nSteps, every = 7, 3
    g.set_state(U0)
    g.calc_dt(1, 0.3)
    dts = []
    for _ in range(nSteps):
        g.update(1)
        dts.append(g.dt()[0])
    assert len(set(dts)) > 1
    UN = g.get_state()

    def replay(U):
        g.set_state(U)
        for dt in dts:
            g.set_dt(dt)
            g.update(1)
        return g.get_state()

    assert rel_err(replay(U0), UN).max() <= 1e-14
    rng = np.random.default_rng(3)
    lamN, v = rng.standard_normal(U0.shape), U0 * rng.standard_normal(U0.shape)
    g.set_state(U0)
    g.calc_dt(1, 0.3)
    lam0 = g.vjp_steps(nSteps, every, lamN)
    # the run itself: the same final state and time step as the forward updates
    assert rel_err(g.get_state(), UN).max() <= 1e-14
    assert abs(g.dt()[0] - dts[-1]) <= 1e-15 * dts[-1]
    h = 1e-6
    fd = float(np.sum(lamN * (replay(U0 + h * v) - replay(U0 - h * v)))) / (2 * h)
    adj = float(np.sum(lam0 * v))
    print("fd", fd, "adjoint", adj)
    assert abs(fd - adj) <= 1e-7 * abs(fd)


@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("where", ["all", "one"])
def test_residual_norm_nan(kind, where):
    """A NaN in the residual makes every norm kind report FR_ENAN ("NaN
    Encountered in Solution Residual!", output.cpp:833-839, 859-866)."""
    cfg = euler_cfg(3, 2)
    g = Solver(cfg, hex_box(3))
    U = g.initialize(cases.ic_tgv)
    if where == "all":
        U[:] = np.nan
    else:
        U[5, 7, 2] = np.nan
    g.set_state(U)
    g.calc_residual(0)
    with pytest.raises(FrError) as ei:
        g.norm_residual(kind)
    assert ei.value.code == -5
