"""Solution files of the reference, restated in numpy -- TEST INFRASTRUCTURE ONLY.

The checker for fr_learn_amd's file output: only tests/ use it.

* paraview_text   -- writeParaview (output.cpp:103-391): primitives at the
  plot points (extrapolateUPpts, solver.cpp:449-484, with the dense
  opp_spts_to_ppts of operators.cpp:111-143 on getLocPpts' grid,
  points.cpp:130-160), plot-point positions (shape_ppts x nodes,
  solver.cpp:151-160 and 904-951), sub-hex connectivity.
* csv_text        -- writeCSV (output.cpp:26-100).
* read_restart    -- solver::readRestartFile (solver.cpp:1420-1482) and
  ele::restart (ele.cpp:1451-1880), including the inter-order interpolation.
"""
import numpy as np

# Gmsh corner order of the linear quad / hex (shape_quad / shape_hex,
# funcs.cpp:121-260)
_CORNERS = {2: [(0, 0), (1, 0), (1, 1), (0, 1)],
            3: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]}


def lagrange(x, y, i):
    """l_i(y) on the nodes x (product form, polynomials.cpp:9-19)."""
    v = 1.0
    for j, xj in enumerate(x):
        if j != i:
            v *= (y - xj) / (x[i] - xj)
    return v


def shape_linear(nDims, r):
    """Shape functions of the linear element at the reference point r."""
    out = []
    for c in _CORNERS[nDims]:
        v = 1.0
        for d in range(nDims):
            v *= (1.0 - r[d]) / 2.0 if c[d] == 0 else (1.0 + r[d]) / 2.0
        out.append(v)
    return np.array(out)


def primitives(U, cfg):
    """Primitive variables of U[..., field] (extrapolateUPpts, solver.cpp:453-476;
    ele::getPrimitives, ele.cpp:1107-1126)."""
    if cfg["equation"] == 0:
        return U[..., :1].copy()
    D = cfg["nDims"]
    rho = U[..., 0]
    V = np.empty_like(U)
    V[..., 0] = rho
    vsq = np.zeros_like(rho)
    for d in range(D):
        V[..., 1 + d] = U[..., 1 + d] / rho
        vsq = vsq + V[..., 1 + d] * V[..., 1 + d]
    V[..., D + 1] = (cfg["gamma"] - 1) * (U[..., D + 1] - 0.5 * rho * vsq)
    return V


def ppts_1d(x1d):
    return np.concatenate([[-1.0], np.asarray(x1d, dtype=np.float64), [1.0]])


def loc_ppts(x1d, nDims):
    """getLocPpts (points.cpp:130-160): i fastest."""
    p = ppts_1d(x1d)
    n = len(p)
    if nDims == 2:
        return np.array([(p[q % n], p[q // n]) for q in range(n * n)])
    return np.array([(p[q % n], p[(q // n) % n], p[q // (n * n)]) for q in range(n ** 3)])


def spts_to_ppts(x1d, nDims):
    """Dense opp_spts_to_ppts (operators.cpp:111-143)."""
    P1 = len(x1d)
    locp = loc_ppts(x1d, nDims)
    nS = P1 ** nDims
    A = np.zeros((len(locp), nS))
    for q, r in enumerate(locp):
        for s in range(nS):
            i, j, k = s % P1, (s // P1) % P1, s // (P1 * P1)
            v = lagrange(x1d, r[0], i) * lagrange(x1d, r[1], j)
            if nDims == 3:
                v *= lagrange(x1d, r[2], k)
            A[q, s] = v
    return A


def plot_data(U, cfg, mesh, x1d):
    """V_ppts[ppt, ele, field] and pos_ppts[ppt, ele, dim] of the state
    U[spt, ele, field] (reference layout, mesh element order)."""
    D = cfg["nDims"]
    V = primitives(U, cfg)
    A = spts_to_ppts(x1d, D)
    Vp = np.einsum("qs,sek->qek", A, V)
    sh = np.array([shape_linear(D, r) for r in loc_ppts(x1d, D)])  # [ppt][node]
    X = mesh.xv[mesh.c2v][:, :, :D]  # [ele][node][dim]
    pos = np.einsum("qn,end->qed", sh, X)
    return Vp, pos


def _g16(v):
    return "%.16g " % v


def paraview_text(U, cfg, mesh, x1d, iter=0, time=0.0):
    """The text writeParaview writes for the state U (3D)."""
    if cfg["nDims"] != 3:
        raise ValueError("Invalid dimensionality [nDims].")
    P = len(x1d) - 1
    NP = P + 3
    nPpts, nSub = NP ** 3, (P + 2) ** 3
    Vp, pos = plot_data(U, cfg, mesh, x1d)
    out = ['<?xml version="1.0" ?>\n',
           '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" '
           'compressor="vtkZLibDataCompressor">\n',
           "<!-- TIME %.16g -->\n" % time, "<!-- ITER %d -->\n" % iter, "\t<UnstructuredGrid>\n"]
    conn = []
    for k in range(NP - 1):
        for j in range(NP - 1):
            for i in range(NP - 1):
                q = [i + NP * (j + NP * k), i + 1 + NP * (j + NP * k), i + 1 + NP * (j + 1 + NP * k),
                     i + NP * (j + 1 + NP * k), i + NP * (j + NP * (k + 1)), i + 1 + NP * (j + NP * (k + 1)),
                     i + 1 + NP * (j + 1 + NP * (k + 1)), i + NP * (j + 1 + NP * (k + 1))]
                conn.append("".join("%d " % v for v in q) + "\n")
    conn = "".join(conn)
    offs = "".join("%d " % ((k + 1) * 8) for k in range(nSub))
    types = "12 " * nSub
    ns = cfg["equation"] != 0
    for e in range(U.shape[1]):
        out.append('\t\t<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (nPpts, nSub))
        out.append("\t\t\t<PointData>\n")
        out.append('\t\t\t\t<DataArray type="Float32" Name="Density" format="ascii">\n')
        out.append("".join(_g16(Vp[q, e, 0]) for q in range(nPpts)) + "\n\t\t\t\t</DataArray>\n")
        if ns:
            out.append('\t\t\t\t<DataArray type="Float32" NumberOfComponents="3" Name="Velocity" format="ascii">\n')
            out.append("".join(_g16(Vp[q, e, 1]) + _g16(Vp[q, e, 2]) + _g16(Vp[q, e, 3]) for q in range(nPpts)) +
                       "\n\t\t\t\t</DataArray>\n")
            out.append('\t\t\t\t<DataArray type="Float32" Name="Pressure" format="ascii">\n')
            out.append("".join(_g16(Vp[q, e, 4]) for q in range(nPpts)) + "\n\t\t\t\t</DataArray>\n")
        out.append("\t\t\t</PointData>\n\t\t\t<Points>\n")
        out.append('\t\t\t\t<DataArray type="Float32" NumberOfComponents="3" format="ascii">\n')
        out.append("".join(_g16(pos[q, e, 0]) + _g16(pos[q, e, 1]) + _g16(pos[q, e, 2]) for q in range(nPpts)) +
                   "\n\t\t\t\t</DataArray>\n")
        out.append("\t\t\t</Points>\n\t\t\t<Cells>\n")
        out.append('\t\t\t\t<DataArray type="Int32" Name="connectivity" format="ascii">\n' + conn +
                   "\t\t\t\t</DataArray>\n")
        out.append('\t\t\t\t<DataArray type="Int32" Name="offsets" format="ascii">\n' + offs +
                   "\n\t\t\t\t</DataArray>\n")
        out.append('\t\t\t\t<DataArray type="UInt8" Name="types" format="ascii">\n' + types +
                   "\n\t\t\t\t</DataArray>\n")
        out.append("\t\t\t</Cells>\n\t\t</Piece>\n")
    out.append("\t</UnstructuredGrid>\n</VTKFile>\n")
    return "".join(out)


def csv_text(U, cfg, mesh, locS, locF, spts_to_fpts):
    """The text writeCSV writes: per element, every spt then every fpt
    (U_fpts = opp_spts_to_fpts U_spts, solver.cpp:401-412)."""
    D = cfg["nDims"]
    X = mesh.xv[mesh.c2v][:, :, :D]
    shS = np.array([shape_linear(D, r) for r in locS])
    shF = np.array([shape_linear(D, r) for r in locF])
    posS = np.einsum("qn,end->qed", shS, X)
    posF = np.einsum("qn,end->qed", shF, X)
    Uf = np.einsum("fs,sek->fek", spts_to_fpts, U)
    VS, VF = primitives(U, cfg), primitives(Uf, cfg)
    out = ["x,y,z," + ("rho\n" if cfg["equation"] == 0 else "rho,u,v,p\n")]

    def row(p, v):
        s = "".join("%.15f," % x for x in p)
        if D == 2:
            s += "0.0,"
        return s + ",".join("%.15f" % x for x in v) + "\n"

    for e in range(U.shape[1]):
        out.extend(row(posS[q, e], VS[q, e]) for q in range(U.shape[0]))
        out.extend(row(posF[q, e], VF[q, e]) for q in range(Uf.shape[0]))
    return "".join(out)


def read_restart(text, cfg, x1d_target, x1d_of_order):
    """U[spt][ele][field] and the TIME of a restart .vtu (3D).  x1d_of_order(P)
    gives the 1D solution points of order P (for files of another order)."""
    lines = iter(text.splitlines())
    time = None
    for line in lines:
        w = line.split()
        if w and w[0] == "<!--" and len(w) > 2 and w[1] == "TIME":
            time = float(w[2])
        if w and w[0] == "<UnstructuredGrid>":
            break
    else:
        raise ValueError("Cannot find UnstructuredData tag in restart file.")
    D, NF = 3, cfg["nFields"] if "nFields" in cfg else (1 if cfg["equation"] == 0 else 5)
    ns = cfg["equation"] != 0
    P = len(x1d_target) - 1
    eles = []
    for line in lines:
        if "<Piece" not in line:
            if "</UnstructuredGrid>" in line:
                break
            continue
        nCells = int(line.split('NumberOfCells="')[1].split('"')[0])
        Pr = round(nCells ** (1.0 / 3.0)) - 2
        NPr = Pr + 3
        data = {}
        for line2 in lines:
            if "</Piece>" in line2:
                break
            if "<DataArray" in line2 and 'Name="' in line2:
                name = line2.split('Name="')[1].split('"')[0]
                vals = np.array([float(t) for t in next(lines).split()])
                data[name] = vals
        Pr1 = Pr + 1
        idx = [(i + 1) + NPr * ((j + 1) + NPr * (k + 1)) for k in range(Pr1) for j in range(Pr1) for i in range(Pr1)]
        rho = data["Density"][idx]
        if ns:
            vel = data["Velocity"].reshape(-1, 3)[idx]
            p = data["Pressure"][idx]
            Ue = np.empty((len(idx), 5))
            Ue[:, 0] = rho
            Ue[:, 1:4] = rho[:, None] * vel
            Ue[:, 4] = p / (cfg["gamma"] - 1) + 0.5 * rho * (vel * vel).sum(axis=1)
        else:
            Ue = rho[:, None]
        if Pr != P:
            xr = x1d_of_order(Pr)
            nSf = (P + 1) ** D
            I = np.zeros((nSf, len(idx)))
            for f in range(nSf):
                fi, fj, fk = f % (P + 1), (f // (P + 1)) % (P + 1), f // ((P + 1) ** 2)
                for r in range(len(idx)):
                    ri, rj, rk = r % Pr1, (r // Pr1) % Pr1, r // (Pr1 * Pr1)
                    I[f, r] = (lagrange(xr, x1d_target[fi], ri) * lagrange(xr, x1d_target[fj], rj) *
                               lagrange(xr, x1d_target[fk], rk))
            Ue = I @ Ue
        eles.append(Ue)
    return np.stack(eles, axis=1), time


def parse_vtu(text):
    """(header lines, pieces) of a .vtu: each piece is a dict name -> array
    of its DataArrays (the points array under "Points"), for comparisons."""
    lines = text.splitlines()
    pieces, head = [], []
    i = 0
    while i < len(lines) and "<Piece" not in lines[i]:
        head.append(lines[i])
        i += 1
    cur = None
    while i < len(lines):
        ln = lines[i]
        if "<Piece" in ln:
            cur = {"_tag": ln}
            pieces.append(cur)
        elif "<DataArray" in ln:
            name = ln.split('Name="')[1].split('"')[0] if 'Name="' in ln else "Points"
            body = []
            i += 1
            while "</DataArray>" not in lines[i]:
                body.append(lines[i])
                i += 1
            cur[name] = np.array([float(t) for t in " ".join(body).split()])
            cur["_tag_" + name] = ln
        i += 1
    return head, pieces
