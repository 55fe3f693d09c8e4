"""Error norms on the GPU (fr_integrate_error) against the oracle's
restatement of solver::integrateError (solver.cpp:1507-1574) and calcError
(funcs.cpp:1875-2056), whose values tests/test_oracle_error.py pins.

Bar: 1e-12 per field, relative to the norm or to the same norm of the
solution itself, whichever is larger (the error is a difference of terms of
the solution's size); every analytic solution of calcError, L1, L2 and the
errorNorm 0 difference of integrals, quads and hexes, after a few RK steps
so that the state differs from the sampled solution."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases, mesh
from oracle import Oracle

from helpers import advection_cfg, euler_cfg, rel_err

pytestmark = pytest.mark.gpu


def _pair(cfg, m, U):
    o = Oracle(cfg, m)
    g = Solver(cfg, m)
    o.set_state(U)
    g.set_state(U)
    return o, g


def _scale(o, norm):
    """The same norm of the solution itself, per field: the error U - exact is
    formed from terms of the solution's size, so rounding in U at the
    quadrature points (1e-16 of U) moves the L1 norm and the errorNorm 0
    difference by at most 1e-16 * int |U| and the L2 norm by at most
    1e-16 * sqrt(int U^2) (Cauchy-Schwarz)."""
    U = o.get_state()
    wJ = o.operator(8)[:, None] * o.geometry(1)
    if norm == 2:
        return np.sqrt(np.einsum("se,sek->k", wJ, U ** 2))
    return np.einsum("se,sek->k", wJ, np.abs(U))


def _compare(o, g, **kw):
    """Bar: 1e-12 of the norm itself or of the solution's own norm (_scale),
    whichever is larger."""
    a, b = g.integrate_error(**kw), o.integrate_error(**kw)
    ref = np.maximum(np.abs(b), _scale(o, kw.get("errorNorm", 1)))
    err = np.abs(a - b) / np.maximum(ref, 1e-300)
    print(kw, "gpu", a, "oracle", b, "rel", err, flush=True)
    assert (err <= 1e-12).all()


def _vortex_state(x, gamma=1.4, eps=5.0):
    f = 1.0 - (x[..., 0] ** 2 + x[..., 1] ** 2)
    rho = (1.0 - eps ** 2 * (gamma - 1) / (8 * gamma * np.pi ** 2) * np.exp(f)) ** (1 / (gamma - 1))
    vx = 1 - eps * x[..., 1] / (2 * np.pi) * np.exp(f / 2)
    vy = 1 + eps * x[..., 0] / (2 * np.pi) * np.exp(f / 2)
    return cases.conservative(rho, [vx, vy], rho ** gamma, gamma)


@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("norm", [0, 1, 2])
def test_vortex_quads(order, norm):
    cfg = euler_cfg(2, order, riemannType=1, dt=1e-3, gamma=1.4)
    m = mesh.box(2, (8, 8), (-5.0, -5.0), (5.0, 5.0), perturb=0.1, seed=3)
    o = Oracle(cfg, m)
    U = _vortex_state(o.positions())
    o, g = _pair(cfg, m, U)
    o.update(3)
    g.update(3)
    for quad in ([4, 10] if norm else [10]):
        _compare(o, g, icType=1, errorNorm=norm, quadOrder=quad, time=3e-3)
    _compare(o, g, icType=2, errorNorm=norm, quadOrder=6, time=0.5, box=(-6.0, 6.0, -6.0, 6.0))
    _compare(o, g, testCase=0, errorNorm=norm, quadOrder=5)


@pytest.mark.parametrize("norm", [1, 2])
def test_hex_entropy_and_tgv(norm):
    cfg = euler_cfg(3, 3, dt=1e-3, gamma=1.4)
    m = mesh.box(3, (3, 3, 3), (0, 0, 0), (2 * np.pi,) * 3, perturb=0.1, seed=2)
    o = Oracle(cfg, m)
    U = o.initialize(cases.ic_tgv)
    o, g = _pair(cfg, m, U)
    o.update(2)
    g.update(2)
    _compare(o, g, icType=0, errorNorm=norm, quadOrder=7, rhoIC=1.0, pIC=71.42857142857143)
    _compare(o, g, icType=1, errorNorm=norm, quadOrder=4, box=(0.0, 2 * np.pi, 0.0, 2 * np.pi))


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("ic", [0, 1, 2])
def test_advection(D, ic):
    cfg = advection_cfg(D, 3, dt=1e-3)
    m = (mesh.box(2, (6, 6), (-5.0, -5.0), (5.0, 5.0), perturb=0.1, seed=1) if D == 2
         else mesh.box(3, (3, 3, 3), (-5.0, -5.0, -5.0), (5.0, 5.0, 5.0), perturb=0.1, seed=1))
    o = Oracle(cfg, m)
    U = o.initialize(cases.ic_gaussian)
    o, g = _pair(cfg, m, U)
    o.update(2)
    g.update(2)
    for norm in (0, 1, 2):
        _compare(o, g, icType=ic, errorNorm=norm, quadOrder=8, time=0.37)


def test_fp32_context_uses_the_fp64_state():
    cfg = euler_cfg(2, 3, dt=1e-3, gamma=1.4)
    m = mesh.box(2, (6, 6), (-5.0, -5.0), (5.0, 5.0))
    g = Solver(cfg, m, precision="fp32")
    o = Oracle(cfg, m)
    U = _vortex_state(o.positions())
    g.set_state(U)
    o.set_state(U)
    _compare(o, g, icType=1, errorNorm=2)


def test_bad_arguments():
    cfg = advection_cfg(2, 2)
    g = Solver(cfg, mesh.box(2, (3, 3), (-1.0, -1.0), (1.0, 1.0)))
    for kw in (dict(errorNorm=3), dict(quadOrder=11), dict(testCase=2)):
        with pytest.raises(RuntimeError):
            g.integrate_error(**kw)
