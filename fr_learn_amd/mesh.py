"""Hex / quad / triangle meshes for the FR engine.

A mesh is what fr_create consumes (include/fr.h `fr_mesh`): vertex
coordinates, element -> vertex lists in Gmsh corner order, an optional
canonical-vertex map used for face matching (periodicity by vertex
identification), and boundary vertex sets with a boundary-condition id each
(the reference's mesh_bound / $PhysicalNames mapping, src/geo.cpp:448-518).
"""
import itertools

import numpy as np

# boundary-condition ids (reference include/global.hpp:63-79)
BC = {
    "char": 1, "sup_in": 2, "sup_out": 3, "sub_in": 4, "sub_out": 5, "slip_wall": 8,
    "isothermal_noslip": 9, "adiabatic_noslip": 10, "symmetry": 12, "isothermal_noslip_moving": 13,
}

# Gmsh corner coordinates of the linear hex (funcs.cpp shape_hex node order)
HEX_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
QUAD_CORNERS = np.array([[0, 0], [1, 0], [1, 1], [0, 1]])


class Mesh:
    def __init__(self, nDims, xv, c2v, vmap=None, bounds=()):
        self.nDims = nDims
        self.xv = np.ascontiguousarray(xv, dtype=np.float64)
        self.c2v = np.ascontiguousarray(c2v, dtype=np.int32)
        self.vmap = None if vmap is None else np.ascontiguousarray(vmap, dtype=np.int32)
        self.bounds = [(int(bc), np.ascontiguousarray(np.unique(v), dtype=np.int32)) for bc, v in bounds]

    @property
    def nEles(self):
        return self.c2v.shape[0]

    @property
    def nNodes(self):
        return self.c2v.shape[1]

    def bnd_arrays(self):
        bcs = np.array([b for b, _ in self.bounds], dtype=np.int32)
        off = np.zeros(len(self.bounds) + 1, dtype=np.int32)
        for i, (_, v) in enumerate(self.bounds):
            off[i + 1] = off[i] + len(v)
        pts = (np.concatenate([v for _, v in self.bounds]).astype(np.int32)
               if self.bounds else np.zeros(1, dtype=np.int32))
        return bcs, off, np.ascontiguousarray(pts)

    def element_nodes(self):
        """[nEles][nNodes][3] node coordinates."""
        return self.xv[self.c2v]


def _rotations3():
    """The 24 proper rotations of the cube as signed permutation matrices."""
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3), dtype=int)
            for r in range(3):
                R[r, perm[r]] = signs[r]
            if round(np.linalg.det(R)) == 1:
                out.append(R)
    return out


def _node_perm(R, corners):
    """Permutation p with new node n = old node p[n] for a cube rotation R."""
    c = 2 * corners - 1
    rc = (c @ R.T + 1) // 2
    lookup = {tuple(v): i for i, v in enumerate(corners)}
    return np.array([lookup[tuple(v)] for v in rc])


def box(nDims, n, lo, hi, periodic=None, bc=None, perturb=0.0, seed=0, rotate=False, shuffle=False):
    """Structured box of linear hexes (nDims=3) or quads (nDims=2).

    n, lo, hi: cells and bounds per direction.  periodic: per-direction flags
    (default all periodic).  bc: boundary condition id (or name) for every
    non-periodic side, either one value or a dict keyed 'xmin', 'xmax', ...
    perturb: random displacement of vertices as a fraction of the cell size
    (the same displacement on periodic images, none normal to a wall).
    rotate / shuffle: randomly re-orient each element's corner ordering / the
    element order (exercises face rotations and left/right assignment).
    """
    D = nDims
    n = [int(v) for v in n][:D]
    lo = np.asarray(lo, dtype=np.float64)[:D]
    hi = np.asarray(hi, dtype=np.float64)[:D]
    periodic = [True] * D if periodic is None else list(periodic)[:D]
    for d in range(D):
        if periodic[d] and n[d] < 3:
            # with fewer than 3 cells a periodic face is not identified uniquely
            # by its vertex set (two faces share the same canonical vertices)
            raise ValueError("a periodic direction needs at least 3 cells")
    rng = np.random.default_rng(seed)
    dims = [v + 1 for v in n]
    grids = np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")
    ijk = np.stack([g.transpose(*reversed(range(D))).ravel() for g in grids], axis=1)  # x fastest
    nv = ijk.shape[0]
    h = (hi - lo) / np.asarray(n)
    xv = np.zeros((nv, 3))
    xv[:, :D] = lo + ijk * h
    # canonical ids
    can = ijk.copy()
    for d in range(D):
        if periodic[d]:
            can[can[:, d] == n[d], d] = 0
    strides = np.cumprod([1] + dims[:-1])
    canon = (can * strides).sum(axis=1).astype(np.int32)
    if perturb > 0:
        disp = rng.uniform(-1, 1, size=(nv, D)) * perturb * h
        disp = disp[canon]  # periodic images move together
        for d in range(D):
            if not periodic[d]:
                on = (ijk[:, d] == 0) | (ijk[:, d] == n[d])
                disp[on, d] = 0.0
        xv[:, :D] += disp
    vmap = canon if any(periodic) else None
    # elements
    corners = HEX_CORNERS if D == 3 else QUAD_CORNERS
    egrid = np.meshgrid(*[np.arange(v) for v in n], indexing="ij")
    eijk = np.stack([g.transpose(*reversed(range(D))).ravel() for g in egrid], axis=1)
    vid = lambda c: (c * strides).sum(axis=-1)
    c2v = np.stack([vid(eijk + corners[m]) for m in range(len(corners))], axis=1).astype(np.int32)
    if rotate:
        if D == 3:
            perms = [_node_perm(R, HEX_CORNERS) for R in _rotations3()]
        else:
            perms = [np.roll(np.arange(4), -r) for r in range(4)]
        choice = rng.integers(0, len(perms), size=c2v.shape[0])
        c2v = np.stack([c2v[e, perms[c]] for e, c in enumerate(choice)]).astype(np.int32)
    if shuffle:
        c2v = c2v[rng.permutation(c2v.shape[0])]
    # boundaries
    bounds = []
    names = ["x", "y", "z"]
    for d in range(D):
        if periodic[d]:
            continue
        for side, val in (("min", 0), ("max", n[d])):
            key = names[d] + side
            b = bc[key] if isinstance(bc, dict) else bc
            if b is None:
                raise ValueError(f"no boundary condition for {key}")
            b = BC[b] if isinstance(b, str) else int(b)
            bounds.append((b, np.nonzero(ijk[:, d] == val)[0]))
    return Mesh(D, xv, c2v, vmap, bounds)


def tri_box(n, lo, hi, periodic=None, bc=None, perturb=0.0, seed=0, rotate=False, shuffle=False):
    """Structured box of linear triangles: each cell of box(2, ...) split
    along its first corner's diagonal into two counter-clockwise triangles.
    Arguments as box(); rotate re-orients each triangle's corner order
    cyclically, shuffle permutes the element order."""
    q = box(2, n, lo, hi, periodic=periodic, bc=bc, perturb=perturb, seed=seed)
    c2v = np.concatenate([q.c2v[:, [0, 1, 2]], q.c2v[:, [0, 2, 3]]], axis=0)
    rng = np.random.default_rng(seed + 1)
    if rotate:
        r = rng.integers(0, 3, size=c2v.shape[0])
        c2v = np.stack([np.roll(c2v[e], -r[e]) for e in range(c2v.shape[0])])
    if shuffle:
        c2v = c2v[rng.permutation(c2v.shape[0])]
    return Mesh(2, q.xv, c2v.astype(np.int32), q.vmap, q.bounds)


def write_gmsh(mesh, path, names=None):
    """Write a Gmsh 2.2 ASCII mesh with one physical volume "fluid" and one
    physical surface per boundary (the format read by geo::readGmsh,
    src/geo.cpp:448-770).  Boundary faces are emitted as linear quads/lines;
    cells as hexes (Gmsh types 5, 17, 12, 92, 93, 94 by node count), quads or
    triangles (3, 2)."""
    D = mesh.nDims
    names = names or [f"bnd{i}" for i in range(len(mesh.bounds))]
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(mesh.bounds) + 1)]
    for i, nm in enumerate(names):
        lines.append(f"{D - 1} {i + 1} \"{nm}\"")
    lines.append(f"{D} {len(mesh.bounds) + 1} \"fluid\"")
    lines.append("$EndPhysicalNames")
    lines.append("$Nodes")
    lines.append(str(mesh.xv.shape[0]))
    for i, x in enumerate(mesh.xv):
        lines.append(f"{i + 1} {float(x[0])!r} {float(x[1])!r} {float(x[2])!r}")
    lines.append("$EndNodes")
    # boundary faces: element faces whose vertices all lie in one boundary set
    tri = D == 2 and mesh.nNodes == 3
    fv = ([[0, 1, 2, 3], [4, 5, 6, 7], [3, 0, 4, 7], [2, 1, 5, 6], [1, 0, 4, 5], [3, 2, 6, 7]] if D == 3
          else [[0, 1], [1, 2], [2, 0]] if tri else [[0, 1], [1, 2], [2, 3], [3, 0]])
    elems = []
    sets = [set(v.tolist()) for _, v in mesh.bounds]
    for e in range(mesh.nEles):
        for f in fv:
            vs = mesh.c2v[e, f]
            for b, s in enumerate(sets):
                if all(int(v) in s for v in vs):
                    elems.append((3 if D == 3 else 1, b + 1, vs))
                    break
    for e in range(mesh.nEles):
        cell = mesh.c2v[e]
        if D == 3 and mesh.nNodes == 20:
            cell = np.concatenate([cell[:8], cell[EDGE20]])  # the file order of the edge nodes
        elems.append((GMSH_HEX_TYPE[mesh.nNodes] if D == 3 else 2 if tri else 3, len(mesh.bounds) + 1, cell))
    lines.append("$Elements")
    lines.append(str(len(elems)))
    for i, (t, tag, vs) in enumerate(elems):
        lines.append(f"{i + 1} {t} 2 {tag} {tag} " + " ".join(str(int(v) + 1) for v in vs))
    lines.append("$EndElements")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


# Boundary-condition names of the mesh file / input (bcStr2Num, src/global.cpp:11-25);
# "periodic" and "overset" have no counterpart here.
BC_NAMES = dict(BC, none=-1, fluid=-1)

# cell node counts by Gmsh element type: hexahedra of 8 nodes (type 5), 20
# (quadratic serendipity, 17), 27, 64, 125, 216 (Lagrange, 12 / 92 / 93 / 94)
# (geo.cpp:586-647); in 2D linear quads (3) and triangles (2)
GMSH_HEX_NODES = {5: 8, 17: 20, 12: 27, 92: 64, 93: 125, 94: 216}
GMSH_HEX_TYPE = {v: k for k, v in GMSH_HEX_NODES.items()}
# the 20-node hex's edge nodes: the k-th edge node of the file goes to node
# EDGE20[k] (geo.cpp:603-609)
EDGE20 = [8, 11, 12, 9, 13, 10, 14, 15, 16, 19, 17, 18]

# vertices per boundary element, by Gmsh element type (geo.cpp:670-730)
_GMSH_FACE_PTS = {1: 2, 2: 3, 3: 4, 10: 4, 16: 4, 36: 4, 37: 4, 38: 4, 8: 3, 26: 4, 27: 5, 28: 6, 62: 7,
                  63: 8, 64: 9, 65: 10, 66: 11}


def read_gmsh(path, mesh_bounds=None):
    """Read a Gmsh 2.2 ASCII mesh as geo::readGmsh does (src/geo.cpp:448-770):
    $PhysicalNames gives each physical group a name (quotes removed, lower
    case) that `mesh_bounds` maps to a boundary condition (the input file's
    meshBounds; by default the name itself); the group named "fluid" holds
    the cells and fixes nDims.  Cells are linear hexes (type 5) or, in 2D,
    linear quads (type 3) or linear triangles (type 2; one type per mesh;
    triangles are turned counter-clockwise, the library's corner order --
    FlurryPP itself reads no triangles); every other element is a boundary
    face whose vertices join its group's vertex set.  Hexes may be curved:
    20-node serendipity (type 17, its edge nodes permuted as the reference
    stores them) and 27 / 64 / 125 / 216-node Lagrange (types 12, 92, 93, 94,
    in file order); faces are matched by their corners.  Vertices are read in
    file order.  Errors raise ValueError with the reference's message."""
    mesh_bounds = {k.lower(): v for k, v in (mesh_bounds or {}).items()}
    with open(path) as fh:
        lines = fh.read().split("\n")

    def section(tag):
        for i, ln in enumerate(lines):
            if tag in ln:
                return i + 1
        raise ValueError(f"{tag} tag not found in Gmsh file!")

    i = section("$PhysicalNames")
    nPhys = int(lines[i].split()[0])
    group = {}  # gmsh physical id -> bound index, -1 for the cells
    bcs, nDims = [], None
    for ln in lines[i + 1:i + 1 + nPhys]:
        tok = ln.split(None, 2)
        dim, pid, name = int(tok[0]), int(tok[1]), tok[2].strip().replace('"', "").lower()
        if mesh_bounds and name not in mesh_bounds:
            raise ValueError(f'Unrecognized mesh boundary: "{name}"\n'
                             "Boundary names in input file must match those in mesh file.")
        bc = mesh_bounds.get(name, name)
        if bc not in BC_NAMES:
            raise ValueError(f'Unrecognized boundary condition: "{bc}"')
        if bc == "fluid":
            nDims = dim
            group[pid] = -1
        else:
            group[pid] = len(bcs)
            bcs.append(BC[bc] if bc in BC else -1)
    if nDims is None:
        raise ValueError('no "fluid" physical group in the Gmsh file')
    i = section("$Nodes")
    nV = int(lines[i].split()[0])
    xv = np.zeros((nV, 3))
    for k in range(nV):
        tok = lines[i + 1 + k].split()
        xv[k] = [float(t) for t in tok[1:4]]
    i = section("$Elements")
    nEl = int(lines[i].split()[0])
    cells, bnd = [], [set() for _ in bcs]
    cell_types = GMSH_HEX_NODES if nDims == 3 else {3: 4, 2: 3}
    cell_type = None
    for k in range(nEl):
        tok = [int(t) for t in lines[i + 1 + k].split()]
        etype, ntags = tok[1], tok[2]
        g = group.get(tok[3], None)
        if g is None:
            raise ValueError(f"Gmsh element ID {k}: physical group {tok[3]} not in $PhysicalNames")
        nodes = tok[3 + ntags:]
        if g == -1:
            if etype not in cell_types:
                raise ValueError(f"Gmsh element ID {k}, Gmsh Element Type = {etype}: element type not recognized "
                                 f"({'hexahedra' if nDims == 3 else 'linear quadrilaterals or triangles'} only)")
            if cell_type is None:
                cell_type = etype
            elif etype != cell_type:
                raise ValueError(f"Gmsh element ID {k}: mixed cell types in one mesh")
            cell = [v - 1 for v in nodes[:cell_types[etype]]]
            if etype == 17:
                reordered = cell[:8] + [0] * 12
                for k2, n in enumerate(EDGE20):
                    reordered[n] = cell[8 + k2]
                cell = reordered
            cells.append(cell)
        else:
            if etype not in _GMSH_FACE_PTS:
                raise ValueError(f"Gmsh element ID {k}, Gmsh Element Type = {etype}: "
                                 "Boundary Element (Face) Type Not Recognized!")
            bnd[g].update(v - 1 for v in nodes[:_GMSH_FACE_PTS[etype]])
    bounds = [(bc, np.array(sorted(pts), dtype=np.int32)) for bc, pts in zip(bcs, bnd) if bc >= 0]
    c2v = np.array(cells, dtype=np.int32)
    if cell_type == 2:
        X = xv[c2v]
        area = ((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) -
                (X[:, 1, 1] - X[:, 0, 1]) * (X[:, 2, 0] - X[:, 0, 0]))
        cw = area < 0
        c2v[cw] = c2v[cw][:, [0, 2, 1]]
    return Mesh(nDims, xv, c2v, None, bounds)
