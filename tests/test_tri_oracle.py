"""The triangle FR checker (oracle/tri.py: the build's own triangle operators
and residual -- FlurryPP has none), pinned by what the mathematics fixes:
orthonormal basis, exact differentiation and extrapolation of polynomials of
degree <= P, the lifting identity of the DG correction, free-stream
preservation, discrete conservation on periodic meshes, exact residuals for
polynomial advection, and the residual's convergence order."""
import math

import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

from fr_learn_amd import cases, mesh
from oracle import tri

from helpers import advection_cfg, euler_cfg


def _ops(P):
    x1d, w1d = leggauss(P + 1)
    return tri.operators(P, x1d, w1d)


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_operators_exact_for_polynomials(P):
    o = _ops(P)
    S, F = o["spts"], o["fpts"]
    n = P + 3
    xa, wa = leggauss(n)
    A, B = np.meshgrid(xa, xa, indexing="ij")
    WA = (np.outer(wa, wa) * (1 - B) / 2).ravel()
    r, s = ((1 + A) * (1 - B) / 2 - 1).ravel(), B.ravel()
    Vq = tri.vandermonde(P, r, s)
    assert np.abs(Vq.T @ (WA[:, None] * Vq) - np.eye(len(S))).max() < 1e-13
    V = tri.vandermonde(P, S[:, 0], S[:, 1])
    rng = np.random.default_rng(P)
    for a in range(P + 1):
        for b in range(P + 1 - a):
            f = S[:, 0] ** a * S[:, 1] ** b
            dfr = a * S[:, 0] ** max(a - 1, 0) * S[:, 1] ** b
            dfs = b * S[:, 0] ** a * S[:, 1] ** max(b - 1, 0)
            assert np.abs(o["Dr"] @ f - dfr).max() < 1e-12
            assert np.abs(o["Ds"] @ f - dfs).max() < 1e-12
            assert np.abs(o["E"] @ f - F[:, 0] ** a * F[:, 1] ** b).max() < 1e-13
            g = rng.standard_normal(len(F))
            vals = Vq @ np.linalg.solve(V, o["L"] @ g)
            lhs = np.sum(WA * r ** a * s ** b * vals)
            assert abs(lhs - np.sum(o["W"] * F[:, 0] ** a * F[:, 1] ** b * g)) < 1e-12
    assert abs(o["wts"].sum() - 2.0) < 1e-14


def _setup(cfg, m, P):
    o = _ops(P)
    JG, detJ, nrm, dA = tri.geometry(m, o)
    X = m.xv[m.c2v][:, :, :2]
    pos = np.einsum("sn,end->sed", np.stack([-(o["spts"][:, 0] + o["spts"][:, 1]) / 2, (1 + o["spts"][:, 0]) / 2,
                                             (1 + o["spts"][:, 1]) / 2], axis=1), X)
    return o, JG, detJ, nrm, dA, tri.faces(m, P), pos


@pytest.mark.parametrize("P", [1, 3])
def test_free_stream_and_conservation(P):
    cfg = euler_cfg(2, P)
    m = mesh.tri_box((4, 4), (0, 0), (1, 1), perturb=0.15, rotate=True, shuffle=True, seed=3)
    o, JG, detJ, nrm, dA, fc, pos = _setup(cfg, m, P)
    U = np.zeros((len(o["spts"]), m.nEles, 4))
    U[..., 0], U[..., 1], U[..., 2], U[..., 3] = 1.2, 0.3, -0.2, 2.5
    r = tri.residual(U, cfg, o, JG, nrm, dA, fc)
    assert np.abs(r).max() < 1e-13
    U = U * (1 + 0.05 * np.random.default_rng(1).standard_normal(U.shape))
    r = tri.residual(U, cfg, o, JG, nrm, dA, fc)
    tot = np.einsum("s,sek->k", o["wts"], r)
    assert np.abs(tot).max() < 1e-13 * np.abs(r).sum()


@pytest.mark.parametrize("P", [2, 3])
def test_polynomial_advection_is_exact_inside(P):
    cfg = advection_cfg(2, P)
    m = mesh.tri_box((5, 5), (0, 0), (1, 1), periodic=[False, False], bc="sup_out", perturb=0.1, seed=4)
    o, JG, detJ, nrm, dA, fc, pos = _setup(cfg, m, P)
    x, y = pos[..., 0], pos[..., 1]
    u = 0.3 + x ** P - 0.5 * x * y ** (P - 1) + y
    ux = P * x ** (P - 1) - 0.5 * y ** (P - 1)
    uy = -0.5 * (P - 1) * x * y ** (P - 2) + 1
    r = tri.residual(u[..., None], cfg, o, JG, nrm, dA, fc)[..., 0] / detJ
    exact = cfg["advectVx"] * ux + cfg["advectVy"] * uy
    bnd = {e for e, f, e2, f2, rot, bc in fc if bc >= 0}
    inner = [e for e in range(m.nEles) if e not in bnd]
    assert len(inner) > 10
    assert np.abs(r[:, inner] - exact[:, inner]).max() < 1e-11


def test_residual_converges_at_order_p():
    P = 2
    cfg = advection_cfg(2, P)
    errs = []
    for n in (6, 12):
        m = mesh.tri_box((n, n), (0, 0), (1, 1), perturb=0.0)
        o, JG, detJ, nrm, dA, fc, pos = _setup(cfg, m, P)
        x, y = pos[..., 0], pos[..., 1]
        k = 2 * math.pi
        u = np.sin(k * x) * np.cos(k * y)
        exact = cfg["advectVx"] * k * np.cos(k * x) * np.cos(k * y) - cfg["advectVy"] * k * np.sin(k * x) * np.sin(k * y)
        r = tri.residual(u[..., None], cfg, o, JG, nrm, dA, fc)[..., 0] / detJ
        errs.append(np.abs(r - exact).max())
    rate = math.log2(errs[0] / errs[1])
    print("errors", errs, "rate", rate)
    assert rate > P - 0.5
