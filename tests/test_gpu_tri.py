"""Triangle elements on the GPU (fr_tri_ops.cpp operators, k_tri_stage) vs the
numpy triangle checker (oracle/tri.py, pinned in tests/test_tri_oracle.py).

Bar: operators, geometry and face tables to 1e-12; FP64 residuals and RK
states per-field max|gpu - checker| / max|checker| <= 1e-12 (helpers.TOL_FP64);
FP32 states within 1e-6 relative L2."""
import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

from fr_learn_amd import Solver, cases, mesh
from fr_learn_amd._lib import FrError
from oracle import tri

from helpers import QUAD_BCS, TOL_FP64, advection_cfg, euler_cfg, perturbed_state, rel_err

pytestmark = pytest.mark.gpu


def _mesh(n=4, periodic=True, bc=None, perturb=0.15, seed=1):
    return mesh.tri_box((n, n), (-5.0, -5.0), (5.0, 5.0), periodic=[periodic] * 2, bc=bc, perturb=perturb,
                        rotate=True, shuffle=True, seed=seed)


def _checker(g, cfg, m):
    P = cfg["order"]
    x1d, w1d = leggauss(P + 1)
    ops = tri.operators(P, x1d, w1d)
    return ops, g.geometry(0), g.geometry(1), g.geometry(2), g.geometry(3), [tuple(r) for r in g.faces()]


def _state(g, cfg, seed=0, amp=2e-2):
    if cfg["equation"] == 0:
        ic = lambda x, c: cases.ic_gaussian(x, c)
    else:
        ic = lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4)
    U = perturbed_state(g.initialize(ic), seed=seed, amp=amp)
    g.set_state(U)
    return U


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_operators_geometry_faces(P):
    cfg = euler_cfg(2, P)
    m = _mesh(3)
    g = Solver(cfg, m)
    ops, JG, detJ, nrm, dA, faces = _checker(g, cfg, m)
    nS, nF = g.nSpts, g.nFpts
    assert (nS, nF) == ((P + 1) * (P + 2) // 2, 3 * (P + 1))
    assert np.abs(g.operator(0) - ops["E"]).max() < 1e-12
    G = g.operator(1)
    assert np.abs(G[0] - ops["Dr"]).max() < 1e-12 * np.abs(ops["Dr"]).max()
    assert np.abs(G[1] - ops["Ds"]).max() < 1e-12 * np.abs(ops["Ds"]).max()
    assert np.abs(g.operator(3) - ops["L"]).max() < 1e-12 * np.abs(ops["L"]).max()
    JGo, detJo, nrmo, dAo = tri.geometry(m, ops)
    assert np.abs(JG - JGo).max() < 1e-13 and np.abs(detJ - detJo).max() < 1e-13
    assert np.abs(nrm - nrmo).max() < 1e-13 and np.abs(dA - dAo).max() < 1e-13
    assert [tuple(int(v) for v in f) for f in faces] == [tuple(int(v) for v in f) for f in tri.faces(m, P)]


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("riemann", [0, 1])
def test_euler_residual_periodic(P, riemann):
    cfg = euler_cfg(2, P, riemannType=riemann)
    m = _mesh(4)
    g = Solver(cfg, m)
    U = _state(g, cfg)
    print(f"tri case P={P} riemann={riemann} nEles={m.nEles}", flush=True)
    ops, JG, detJ, nrm, dA, faces = _checker(g, cfg, m)
    g.calc_residual(0)
    err = rel_err(g.residual(0), tri.residual(U, cfg, ops, JG, nrm, dA, faces))
    print("rel err", err)
    assert err.max() <= TOL_FP64


@pytest.mark.parametrize("P", [2, 4])
def test_euler_residual_boundaries(P):
    cfg = euler_cfg(2, P, riemannType=1)
    m = _mesh(4, periodic=False, bc=QUAD_BCS)
    g = Solver(cfg, m)
    U = _state(g, cfg)
    ops, JG, detJ, nrm, dA, faces = _checker(g, cfg, m)
    g.calc_residual(0)
    assert rel_err(g.residual(0), tri.residual(U, cfg, ops, JG, nrm, dA, faces)).max() <= TOL_FP64


@pytest.mark.parametrize("P", [1, 3])
def test_advection_residual(P):
    cfg = advection_cfg(2, P)
    m = _mesh(4, periodic=False, bc="sup_out")
    g = Solver(cfg, m)
    U = _state(g, cfg)
    ops, JG, detJ, nrm, dA, faces = _checker(g, cfg, m)
    g.calc_residual(0)
    assert rel_err(g.residual(0), tri.residual(U, cfg, ops, JG, nrm, dA, faces)).max() <= TOL_FP64


@pytest.mark.parametrize("timeType", [4, 5])
def test_rk_steps(timeType):
    """Classical RK4 and a Jameson scheme (input.cpp:424-461) through the
    fused stage kernel against the checker's stage-by-stage update."""
    cfg = euler_cfg(2, 3, timeType=timeType, dt=2e-3)
    m = _mesh(4)
    g = Solver(cfg, m)
    U = _state(g, cfg)
    ops, JG, detJ, nrm, dA, faces = _checker(g, cfg, m)
    R = lambda V: tri.residual(V, cfg, ops, JG, nrm, dA, faces)
    dt, dJ = cfg["dt"], detJ[..., None]
    for _ in range(2):
        if timeType == 4:
            a, b = [0.5, 0.5, 1.0], [1 / 6, 1 / 3, 1 / 3, 1 / 6]
            Us, rs = U, []
            for s in range(4):
                rs.append(R(Us))
                if s < 3:
                    Us = U - a[s] * rs[-1] / dJ * dt
            for s in range(4):
                U = U - b[s] * rs[s] / dJ * dt
        else:
            a = [1 / 4, 1 / 3, 1 / 2, 1.0]
            Us = U
            for s in range(4):
                Us = U - a[s] * R(Us) / dJ * dt
            U = Us
    g.update(2)
    err = rel_err(g.get_state(), U)
    print("state rel err", err)
    assert err.max() <= TOL_FP64


def test_fp32_and_free_stream():
    cfg = euler_cfg(2, 3, dt=2e-3)
    m = _mesh(6)
    g64, g32 = Solver(cfg, m), Solver(cfg, m, precision="fp32")
    U = _state(g64, cfg)
    g32.set_state(U)
    g64.update(5)
    g32.update(5)
    a, b = g32.get_state(), g64.get_state()
    l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
    print("fp32 rel L2", l2)
    assert l2 < 1e-6
    Uf = np.zeros_like(U)
    Uf[..., 0], Uf[..., 1], Uf[..., 2], Uf[..., 3] = 1.1, 0.2, -0.3, 2.4
    g64.set_state(Uf)
    g64.calc_residual(0)
    assert np.abs(g64.residual(0)).max() < 1e-12


def test_unsupported_verbs_raise():
    cfg = euler_cfg(2, 2)
    g = Solver(cfg, _mesh(3))
    _state(g, cfg)
    with pytest.raises(FrError):
        g.calc_dt(1, 0.5)
    with pytest.raises(FrError):
        g.vjp_residual(np.zeros(g.state_shape()))
