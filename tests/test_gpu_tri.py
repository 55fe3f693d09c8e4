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
@pytest.mark.parametrize("geo", [0, 1])
def test_euler_residual_periodic(P, riemann, geo):
    """geo 0: the per-point metrics and normals; 1 (the default): one metric
    set per element and one normal per edge (every linear triangle is
    affine)."""
    cfg = euler_cfg(2, P, riemannType=riemann)
    m = _mesh(4)
    g = Solver(cfg, m)
    assert g.geometry_mode(geo) == geo
    U = _state(g, cfg)
    print(f"tri case P={P} riemann={riemann} geometry={geo} nEles={m.nEles}", flush=True)
    ops, JG, detJ, nrm, dA, faces = _checker(g, cfg, m)
    g.calc_residual(0)
    err = rel_err(g.residual(0), tri.residual(U, cfg, ops, JG, nrm, dA, faces))
    print("rel err", err)
    assert err.max() <= TOL_FP64


@pytest.mark.parametrize("P", [2, 4])
@pytest.mark.parametrize("geo", [0, 1])
def test_euler_residual_boundaries(P, geo):
    cfg = euler_cfg(2, P, riemannType=1)
    m = _mesh(4, periodic=False, bc=QUAD_BCS)
    g = Solver(cfg, m)
    assert g.geometry_mode(geo) == geo
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


@pytest.mark.parametrize("P", [1, 3, 5])
@pytest.mark.parametrize("timeType", [4, 5])
def test_rk_steps(timeType, P):
    """Classical RK4 and a Jameson scheme (input.cpp:424-461) through the
    fused stage kernel against the checker's stage-by-stage update (the final
    RK4 stage combines the three earlier residuals it loads up front)."""
    cfg = euler_cfg(2, P, timeType=timeType, dt=2e-3)
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


CFL_LIMIT = [1.393, 0.464, 0.235, 0.1453, 0.100, 0.0736]


def _edges(m):
    """Per element and edge (v0v1, v1v2, v2v0): physical length and outward
    unit normal of the counter-clockwise triangle, and the element's area."""
    X = m.xv[m.c2v][:, :, :2]
    a, b = X, np.roll(X, -1, axis=1)
    d = b - a
    L = np.linalg.norm(d, axis=-1)
    n = np.stack([d[..., 1], -d[..., 0]], axis=-1) / L[..., None]
    e1, e2 = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]
    area = 0.5 * (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    return L, n, area


@pytest.mark.parametrize("P", [1, 3, 5])
@pytest.mark.parametrize("equation", [0, 1])
def test_cfl_dt_uniform_state(P, equation):
    """fr_calc_dt on triangles (dtType 2).  A uniform state has one wave speed
    per straight edge, so dt_e = CFL lim(P) 2 area_e / sum_edges ws L has a
    closed form in the vertices (ele::calcDt's formula with the physical edge
    measure)."""
    cfg = euler_cfg(2, P) if equation == 1 else advection_cfg(2, P)
    m = _mesh(5, perturb=0.2, seed=4)
    g = Solver(cfg, m)
    L, n, area = _edges(m)
    if equation == 1:
        rho, u, v, p = 1.1, 0.3, -0.2, 1.7
        gam = cfg["gamma"]
        U = np.zeros(g.state_shape())
        U[..., 0], U[..., 1], U[..., 2] = rho, rho * u, rho * v
        U[..., 3] = p / (gam - 1) + 0.5 * rho * (u * u + v * v)
        ws = np.abs(n[..., 0] * u + n[..., 1] * v) + np.sqrt(gam * p / rho)
    else:
        U = np.ones(g.state_shape())
        ws = np.abs(n[..., 0] * cfg["advectVx"] + n[..., 1] * cfg["advectVy"])
    g.set_state(U)
    CFL = 0.7
    dtmin = g.calc_dt(2, CFL)
    want = CFL * CFL_LIMIT[P] * 2 * area / (ws * L).sum(axis=1)
    got = g.dt()
    assert np.abs(got - want).max() <= 1e-12 * want.max()
    assert abs(dtmin - want.min()) <= 1e-12 * want.min()
    assert abs(g.calc_dt(1, CFL) - want.min()) <= 1e-12 * want.min()
    assert np.all(g.dt() == g.dt()[0])


@pytest.mark.parametrize("P", [2, 4])
def test_cfl_dt_varying_state(P):
    """A varying state: the wave speeds at the flux points from the checker's
    traces E U, own normals and dA, weighted by the edge Gauss weights and the
    reference edge's half length (sqrt 2 on the hypotenuse)."""
    cfg = euler_cfg(2, P)
    m = _mesh(4, perturb=0.15, seed=6)
    g = Solver(cfg, m)
    U = _state(g, cfg, amp=5e-2)
    ops, JG, detJ, nrm, dA, faces = _checker(g, cfg, m)
    x1d, w1d = leggauss(P + 1)
    Uf = np.einsum("qs,sek->qek", ops["E"], U)
    gam = cfg["gamma"]
    rho = Uf[..., 0]
    vx, vy = Uf[..., 1] / rho, Uf[..., 2] / rho
    pr = (gam - 1) * (Uf[..., 3] - 0.5 * rho * (vx * vx + vy * vy))
    ws = np.abs(vx * nrm[..., 0] + vy * nrm[..., 1]) + np.sqrt(np.maximum(gam * pr / rho, 0))
    q = np.arange(3 * (P + 1))
    wq = w1d[q % (P + 1)] * np.where(q // (P + 1) == 1, np.sqrt(2.0), 1.0)
    _, _, area = _edges(m)
    want = 0.5 * CFL_LIMIT[P] * 2 * area / np.einsum("qe,q,qe->e", ws, wq, dA)
    g.calc_dt(2, 0.5)
    assert np.abs(g.dt() - want).max() <= 1e-12 * want.max()


def test_unsupported_verbs_raise():
    cfg = euler_cfg(2, 2)
    g = Solver(cfg, _mesh(3))
    _state(g, cfg)
    with pytest.raises(FrError):
        g.vjp_residual(np.zeros(g.state_shape()))
