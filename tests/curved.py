"""Curved hex meshes for the tests: a box of n^3 hexes of 20 (quadratic
serendipity) or (N+1)^3 (Lagrange, order N) nodes whose node positions are a
smooth map of the box.  The node order is the reference's (shape_hex,
funcs.cpp:178-566), taken from the oracle's restatement (or_hex_lattice);
tests/test_oracle_curved.py pins that order against Gmsh's published
27-node numbering and the Kronecker property of the shape functions."""
import ctypes

import numpy as np

from fr_learn_amd import mesh
from oracle import load as load_oracle

# reference coordinates of the 20-node hex's nodes (funcs.cpp:188-212)
SERENDIPITY20 = np.array(
    [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
     [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0],
     [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1]], dtype=float)


def hex_lattice(nSide):
    """[nodes][3] lattice indices of the nSide^3 Lagrange hex's nodes, in node order."""
    lib = load_oracle()
    lib.or_hex_lattice.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    out = np.zeros((nSide ** 3, 3), dtype=np.int32)
    n = lib.or_hex_lattice(nSide, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert n == nSide ** 3
    return out


def node_lattice(nNodes):
    """(nSide, [nodes][3] lattice indices on nSide equispaced points per axis)."""
    if nNodes == 20:
        return 3, ((SERENDIPITY20 + 1).astype(np.int32))
    nSide = int(round(nNodes ** (1 / 3)))
    return nSide, hex_lattice(nSide)


def wavy(L, amp):
    """A smooth map of the box [0, L]^3 onto itself that curves every face
    and is compatible with periodicity: the displacement along x depends on
    (y, z) only and vanishes nowhere in particular, the others vanish on the
    faces normal to them."""
    k = 2 * np.pi / L

    def f(x):
        X = x.copy()
        X[..., 0] += amp * np.sin(k * x[..., 1]) * np.sin(k * x[..., 2])
        X[..., 1] += amp * np.sin(k * x[..., 0]) * np.sin(k * x[..., 2])
        X[..., 2] += amp * np.sin(k * x[..., 0]) * np.sin(k * x[..., 1])
        return X

    return f


def curved_box(n, nNodes, L=1.0, amp=0.02, periodic=True, bc=None, mapping=None):
    """n^3 hexes of nNodes nodes on [0, L]^3 mapped by `mapping` (default
    wavy(L, amp)).  The nodes sit on a global equispaced lattice (shared by
    neighbouring elements) before the map.  periodic: all three directions;
    otherwise bc (an id or name, or a dict as mesh.box takes) on the sides."""
    nSide, lat = node_lattice(nNodes)
    m = nSide - 1
    G = n * m + 1
    idx = np.arange(G)
    gx, gy, gz = np.meshgrid(idx, idx, idx, indexing="ij")
    grid = np.stack([gx, gy, gz], axis=-1).reshape(-1, 3)
    gid = lambda a, b, c: (a * G + b) * G + c
    f = mapping or wavy(L, amp)
    xv = f(grid.astype(float) * (L / (n * m)))
    c2v = np.zeros((n ** 3, nNodes), dtype=np.int32)
    e = 0
    for i in range(n):
        for j in range(n):
            for k in range(n):
                a, b, c = i * m + lat[:, 0], j * m + lat[:, 1], k * m + lat[:, 2]
                c2v[e] = gid(a, b, c)
                e += 1
    vmap, bounds = None, []
    if periodic:
        wrap = lambda v: np.where(v == G - 1, 0, v)
        vmap = gid(wrap(grid[:, 0]), wrap(grid[:, 1]), wrap(grid[:, 2])).astype(np.int32)
    else:
        sides = {"xmin": grid[:, 0] == 0, "xmax": grid[:, 0] == G - 1, "ymin": grid[:, 1] == 0,
                 "ymax": grid[:, 1] == G - 1, "zmin": grid[:, 2] == 0, "zmax": grid[:, 2] == G - 1}
        for name, sel in sides.items():
            b = bc[name] if isinstance(bc, dict) else bc
            b = mesh.BC[b] if isinstance(b, str) else int(b)
            bounds.append((b, np.nonzero(sel)[0]))
    return mesh.Mesh(3, xv, c2v, vmap, bounds)
