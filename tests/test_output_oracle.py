"""The numpy restatement of the reference's solution files (oracle/output.py),
pinned by what the mathematics fixes: plot-point interpolation reproduces
polynomials of the element's degree, plot-point positions of a sheared box
are the linear map of the reference grid, the restart reader inverts the
writer, and the inter-order interpolation of a restart is exact for
polynomials of the file's degree."""
import numpy as np
import pytest

from fr_learn_amd import cases, mesh
from oracle import Oracle
from oracle import output as oo

from helpers import euler_cfg, hex_box

SHEAR3 = np.array([[1.0, 0.3, -0.1], [0.05, 0.9, 0.2], [0.0, -0.15, 1.1]])


def sheared_box(n=2, L=1.0):
    m = hex_box(n, periodic=False, bc="sup_out", L=L, perturb=0.0, rotate=True, shuffle=True)
    xv = m.xv.copy()
    xv[:, :3] = xv[:, :3] @ SHEAR3.T
    return mesh.Mesh(3, xv, m.c2v, m.vmap, m.bounds)


def prim_poly(x):
    """Primitive fields of total degree <= 2 in the physical coordinates."""
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    return np.stack([1.0 + 0.1 * X + 0.05 * Y * Z, 0.3 + 0.1 * Y * Y, -0.2 + 0.05 * X * Z, 0.1 * X - 0.2 * Y,
                     1.0 + 0.2 * X * Z - 0.1 * Y], axis=-1)


def cons_of(V, gamma=1.4):
    U = np.empty_like(V)
    U[..., 0] = V[..., 0]
    U[..., 1:4] = V[..., :1] * V[..., 1:4]
    U[..., 4] = V[..., 4] / (gamma - 1) + 0.5 * V[..., 0] * (V[..., 1:4] ** 2).sum(axis=-1)
    return U


@pytest.mark.parametrize("order", [2, 3])
def test_plot_points_reproduce_polynomials(order):
    cfg = euler_cfg(3, order)
    m = sheared_box()
    o = Oracle(cfg, m)
    x1d = o.operator(7)
    U = cons_of(prim_poly(o.positions()))
    Vp, pos = oo.plot_data(U, cfg, m, x1d)
    # positions: the shear of the reference grid mapped into each box cell
    assert np.abs(Vp - prim_poly(pos)).max() < 1e-12
    # corner plot points are the element's vertices, in the writer's order
    NP = order + 3
    for e in range(m.nEles):
        X = m.xv[m.c2v[e]]
        for q, c in [(0, 0), (NP - 1, 1), (NP * NP - 1, 2), (NP * (NP - 1), 3)]:
            assert np.allclose(pos[q, e], X[c], atol=1e-14)
        assert np.allclose(pos[NP ** 3 - 1, e], X[6], atol=1e-14)


def test_positions_are_the_shear_of_the_grid():
    cfg = euler_cfg(3, 2)
    m = hex_box(1, periodic=False, bc="sup_out", L=2.0, perturb=0.0, rotate=False, shuffle=False)
    xv = m.xv.copy()
    xv[:, :3] = xv[:, :3] @ SHEAR3.T
    ms = mesh.Mesh(3, xv, m.c2v, m.vmap, m.bounds)
    o = Oracle(cfg, ms)
    _, pos = oo.plot_data(np.ones((o.nSpts, 1, 5)), cfg, ms, o.operator(7))
    ref = oo.loc_ppts(o.operator(7), 3)  # the element is [0,2]^3 = reference + 1
    assert np.abs(pos[:, 0, :] - (ref + 1.0) @ SHEAR3.T).max() < 1e-14


def test_restart_inverts_the_writer():
    cfg = euler_cfg(3, 3)
    m = sheared_box()
    o = Oracle(cfg, m)
    U = o.initialize(cases.ic_tgv)
    text = oo.paraview_text(U, cfg, m, o.operator(7), iter=7, time=0.125)
    U2, t = oo.read_restart(text, cfg, o.operator(7), None)
    assert t == 0.125
    assert np.abs(U2 - U).max() <= 1e-14 * np.abs(U).max()
    head, pieces = oo.parse_vtu(text)
    assert len(pieces) == m.nEles
    assert head[2] == "<!-- TIME 0.125 -->" and head[3] == "<!-- ITER 7 -->"
    conn = pieces[0]["connectivity"].astype(int)
    assert conn.min() == 0 and conn.max() == 6 ** 3 - 1 and len(conn) == 8 * 5 ** 3


def test_restart_interpolates_between_orders():
    cfg2, cfg3 = euler_cfg(3, 2), euler_cfg(3, 3)
    m = sheared_box()
    o2, o3 = Oracle(cfg2, m), Oracle(cfg3, m)
    # constant density, linear velocity and pressure: every conservative
    # field has degree <= 2, which the P2 -> P3 interpolation keeps exactly
    def low(x):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        return np.stack([1.2 + 0 * X, 0.3 + 0.1 * X, -0.2 + 0.05 * Y, 0.1 * Z - 0.1 * X, 1.0 + 0.2 * X - 0.1 * Z],
                        axis=-1)
    U2 = cons_of(low(o2.positions()))
    text = oo.paraview_text(U2, cfg2, m, o2.operator(7))
    U3, _ = oo.read_restart(text, cfg3, o3.operator(7), lambda P: Oracle(euler_cfg(3, P), m).operator(7))
    assert np.abs(U3 - cons_of(low(o3.positions()))).max() < 1e-13


def test_csv_rows_follow_the_state():
    cfg = euler_cfg(3, 2)
    m = sheared_box()
    o = Oracle(cfg, m)
    U = o.initialize(cases.ic_tgv)
    text = oo.csv_text(U, cfg, m, o.operator(5), o.operator(6), o.operator(0))
    rows = text.splitlines()
    assert rows[0] == "x,y,z,rho,u,v,p"
    data = np.array([[float(t) for t in r.split(",")] for r in rows[1:]])
    per = o.nSpts + o.nFpts
    assert data.shape == (m.nEles * per, 8)
    pos = o.positions()
    V = oo.primitives(U, cfg)
    for e in range(m.nEles):
        blk = data[e * per:e * per + o.nSpts]
        assert np.abs(blk[:, :3] - pos[:, e, :]).max() < 2e-14
        assert np.abs(blk[:, 3:] - V[:, e, :]).max() < 1e-13
