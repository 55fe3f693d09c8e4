"""Pins of the oracle's curved-hex shape functions (shape_hex / dshape_hex,
funcs.cpp:178-1607) and of the metrics it builds from them
(solver::calcTransforms, solver.cpp:1227-1356), by what the mathematics and
Gmsh's published numbering fix:

- the 27-node order is Gmsh's Hexahedron27 (corners, then the midpoints of
  the edges 0-1 0-3 0-4 1-2 1-5 2-3 2-6 3-7 4-5 4-7 5-6 6-7, then the centres
  of the faces z=-1, y=-1, x=-1, x=+1, y=+1, z=+1, then the centre);
- every shape function is 1 at its node and 0 at the others (Kronecker), they
  sum to 1 and their derivatives to 0, and the derivatives match central
  differences;
- an element whose map is a polynomial the element reproduces (tensor degree
  <= N for the (N+1)^3-node hex; the serendipity space for 20 nodes) has
  exact metrics: the oracle's adjugate and detJ at the solution points equal
  the analytic ones.
CPU only."""
import ctypes

import numpy as np
import pytest

from fr_learn_amd import cases
from oracle import Oracle
from oracle import load as load_oracle

from curved import SERENDIPITY20, curved_box, hex_lattice, node_lattice
from helpers import euler_cfg

COUNTS = [20, 27, 64, 125, 216]


def shape(nNodes, r):
    lib = load_oracle()
    P = ctypes.POINTER(ctypes.c_double)
    lib.or_shape.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P]
    sh, dsh = np.zeros(nNodes), np.zeros((nNodes, 3))
    r = np.ascontiguousarray(r, dtype=np.float64)
    lib.or_shape(3, nNodes, r.ctypes.data_as(P), sh.ctypes.data_as(P), dsh.ctypes.data_as(P))
    return sh, dsh


def node_positions(nNodes):
    nSide, lat = node_lattice(nNodes)
    return -1.0 + lat * (2.0 / (nSide - 1))


def test_27_node_order_is_gmsh():
    corners = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2]])
    edges = [(0, 1), (0, 3), (0, 4), (1, 2), (1, 5), (2, 3), (2, 6), (3, 7), (4, 5), (4, 7), (5, 6), (6, 7)]
    faces = [(0, 1, 2, 3), (0, 1, 4, 5), (0, 3, 4, 7), (1, 2, 5, 6), (2, 3, 6, 7), (4, 5, 6, 7)]
    want = [tuple(c) for c in corners]
    want += [tuple((corners[a] + corners[b]) // 2) for a, b in edges]
    want += [tuple(corners[list(f)].sum(0) // 4) for f in faces]
    want.append((1, 1, 1))
    assert [tuple(v) for v in hex_lattice(3)] == want


@pytest.mark.parametrize("nNodes", COUNTS)
def test_kronecker_partition_and_derivatives(nNodes):
    X = node_positions(nNodes)
    assert len({tuple(x) for x in X}) == nNodes  # distinct nodes
    S = np.array([shape(nNodes, x)[0] for x in X])
    assert np.abs(S - np.eye(nNodes)).max() < 1e-12
    rng = np.random.default_rng(nNodes)
    for r in rng.uniform(-1, 1, size=(5, 3)):
        sh, dsh = shape(nNodes, r)
        assert abs(sh.sum() - 1) < 1e-12
        assert np.abs(dsh.sum(0)).max() < 1e-11
        h = 1e-6
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            fd = (shape(nNodes, r + e)[0] - shape(nNodes, r - e)[0]) / (2 * h)
            assert np.abs(fd - dsh[:, d]).max() < 1e-6


def test_serendipity_node_positions():
    assert np.array_equal(node_positions(20), SERENDIPITY20)


def _poly_map(nNodes, amp=0.05):
    """A map each element reproduces exactly: tensor degree <= N for the
    Lagrange hexes, the serendipity space (x^2 y, x y z, ...) for 20 nodes."""
    if nNodes == 20:
        f = lambda x, y, z: (amp * y * y * z, amp * x * z * z, amp * x * x * y)
        df = lambda x, y, z: np.array([[0, 2 * amp * y * z, amp * y * y], [amp * z * z, 0, 2 * amp * x * z],
                                       [2 * amp * x * y, amp * x * x, 0]])
    else:
        N = int(round(nNodes ** (1 / 3))) - 1
        f = lambda x, y, z: (amp * y ** N * z, amp * x * z ** N, amp * x ** N * y ** min(N, 2))
        df = lambda x, y, z: np.array([
            [0, amp * N * y ** (N - 1) * z, amp * y ** N],
            [amp * z ** N, 0, amp * N * x * z ** (N - 1)],
            [amp * N * x ** (N - 1) * y ** min(N, 2), amp * min(N, 2) * x ** N * y ** (min(N, 2) - 1), 0]])

    def fmap(X):
        d = f(X[..., 0], X[..., 1], X[..., 2])
        return X + np.stack(d, axis=-1)

    def jac(x):  # J[dimPhys][dimX] of the map
        return np.eye(3) + df(*x)

    return fmap, jac


@pytest.mark.parametrize("nNodes", COUNTS)
def test_metrics_exact_for_reproduced_maps(nNodes):
    n, L = 2, 1.0
    fmap, jac = _poly_map(nNodes)
    m = curved_box(n, nNodes, L=L, periodic=False, bc="char", mapping=fmap)
    o = Oracle(euler_cfg(3, 3), m)
    G = o.geometry(0)       # JGinv(dimRef, spt, ele, dimPhys)
    detJ = o.geometry(1)    # (spt, ele)
    loc = o.operator(5)     # spt locations (spt, 3)
    h = L / n
    e = 0
    worst = 0.0
    for i in range(n):
        for j in range(n):
            for k in range(n):
                x0 = np.array([i, j, k]) * h
                for s in range(loc.shape[0]):
                    x = x0 + (loc[s] + 1) / 2 * h
                    J = (jac(x) * (h / 2)).T  # J[dimRef][dimPhys]
                    # JGinv(dimRef, dimPhys) = detJ * d ref / d phys
                    adj = (np.linalg.det(J) * np.linalg.inv(J)).T
                    worst = max(worst, np.abs(G[:, s, e, :] - adj).max() / np.abs(adj).max(),
                                abs(detJ[s, e] - np.linalg.det(J)) / abs(np.linalg.det(J)))
                e += 1
    assert worst < 1e-12, worst


@pytest.mark.parametrize("nNodes", [20, 27, 64])
def test_curved_residual_conserves(nNodes):
    """On a periodic curved box the residual integrates to zero
    (sum of w * divF, every field)."""
    m = curved_box(3, nNodes, L=2 * np.pi, amp=0.15)
    o = Oracle(euler_cfg(3, 2, dt=1e-3), m)
    U = o.initialize(cases.ic_tgv)
    o.set_state(U * (1 + 1e-2 * np.random.default_rng(0).standard_normal(U.shape)))
    o.calc_residual(0)
    r = o.residual(0)
    w = o.operator(8)
    tot = np.einsum("s,sek->k", w, r)
    assert (np.abs(tot) <= 1e-11 * np.einsum("s,sek->k", w, np.abs(r))).all()
