"""Triangle FR operators and residual in numpy -- TEST INFRASTRUCTURE ONLY.

FlurryPP has no triangle operators (only QUAD and HEX, operators.cpp:24-26),
so this is the build's own definition, written independently of the HIP
library's host code as its checker:

* solution points: the warp-and-blend nodes of Hesthaven & Warburton
  ("Nodal Discontinuous Galerkin Methods", 2008, sec. 6.1, Nodes2D /
  Warpfactor) mapped to the reference triangle (-1,-1), (1,-1), (-1,1);
* flux points: P+1 Gauss-Legendre points on each edge, edge f running from
  vertex f to vertex f+1 (counter-clockwise), the 1D abscissae of the quad
  operators;
* the orthonormal Dubiner (PKDO) basis (sec. 6.1, Simplex2DP,
  GradSimplex2DP), Vandermonde V, differentiation Dr, Ds = V_r V^-1,
  V_s V^-1, extrapolation E = Psi(fpts) V^-1;
* the correction (lifting) operator of the DG-equivalent VCJH scheme (c = 0;
  Castonguay, Vincent & Jameson 2012): L = V Psi(fpts)^T W, W = edge
  quadrature weights, so that for every v of degree <= P,
  integral(v * sum_j L_j g_j) = sum over edge points of w_j v(x_j) g_j.

The residual follows the quad path's steps (extrapolate, common flux scaled by
dA, discontinuous normal flux, divergence plus correction) with these dense
operators; pinned in tests by free-stream preservation, polynomial exactness,
conservation and convergence.
"""
import math

import numpy as np

V0 = np.array([-1.0, -1.0])
V1 = np.array([1.0, -1.0])
V2 = np.array([-1.0, 1.0])
VERTS = [V0, V1, V2]
# outward unit normals of the reference edges (v0v1, v1v2, v2v0)
TNORM = np.array([[0.0, -1.0], [1.0 / math.sqrt(2.0), 1.0 / math.sqrt(2.0)], [-1.0, 0.0]])
EDGE_LEN = np.array([2.0, 2.0 * math.sqrt(2.0), 2.0])
# alpha_opt of Nodes2D by order P (index 0 unused)
ALPHA_OPT = [0.0, 0.0, 0.0, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832, 1.3648, 1.4773, 1.4959]


def jacobi_p(x, alpha, beta, n):
    """Orthonormal Jacobi polynomial P_n^(alpha,beta)(x) (JacobiP)."""
    x = np.asarray(x, dtype=np.float64)
    g0 = (2 ** (alpha + beta + 1) / (alpha + beta + 1) * math.gamma(alpha + 1) * math.gamma(beta + 1) /
          math.gamma(alpha + beta + 1))
    p0 = np.ones_like(x) / math.sqrt(g0)
    if n == 0:
        return p0
    g1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * g0
    p1 = ((alpha + beta + 2) * x / 2 + (alpha - beta) / 2) / math.sqrt(g1)
    if n == 1:
        return p1
    aold = 2 / (2 + alpha + beta) * math.sqrt((alpha + 1) * (beta + 1) / (alpha + beta + 3))
    pm, pc = p0, p1
    for i in range(1, n):
        h1 = 2 * i + alpha + beta
        anew = 2 / (h1 + 2) * math.sqrt((i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha) * (i + 1 + beta) /
                                        (h1 + 1) / (h1 + 3))
        bnew = -(alpha ** 2 - beta ** 2) / h1 / (h1 + 2)
        pn = 1 / anew * (-aold * pm + (x - bnew) * pc)
        aold, pm, pc = anew, pc, pn
    return pc


def grad_jacobi_p(x, alpha, beta, n):
    if n == 0:
        return np.zeros_like(np.asarray(x, dtype=np.float64))
    return math.sqrt(n * (n + alpha + beta + 1)) * jacobi_p(x, alpha + 1, beta + 1, n - 1)


def rs_to_ab(r, s):
    r, s = np.asarray(r, dtype=np.float64), np.asarray(s, dtype=np.float64)
    a = np.where(s != 1.0, 2 * (1 + r) / np.where(s != 1.0, 1 - s, 1.0) - 1, -1.0)
    return a, s


def modes(P):
    return [(i, j) for i in range(P + 1) for j in range(P + 1 - i)]


def simplex2dp(a, b, i, j):
    return math.sqrt(2.0) * jacobi_p(a, 0, 0, i) * jacobi_p(b, 2 * i + 1, 0, j) * (1 - b) ** i


def grad_simplex2dp(a, b, i, j):
    fa, dfa = jacobi_p(a, 0, 0, i), grad_jacobi_p(a, 0, 0, i)
    gb, dgb = jacobi_p(b, 2 * i + 1, 0, j), grad_jacobi_p(b, 2 * i + 1, 0, j)
    dr = dfa * gb
    if i > 0:
        dr = dr * (0.5 * (1 - b)) ** (i - 1)
    ds = dfa * (gb * (0.5 * (1 + a)))
    if i > 0:
        ds = ds * (0.5 * (1 - b)) ** (i - 1)
    tmp = dgb * (0.5 * (1 - b)) ** i
    if i > 0:
        tmp = tmp - 0.5 * i * gb * (0.5 * (1 - b)) ** (i - 1)
    ds = ds + fa * tmp
    return 2 ** (i + 0.5) * dr, 2 ** (i + 0.5) * ds


def vandermonde(P, r, s):
    a, b = rs_to_ab(r, s)
    return np.stack([simplex2dp(a, b, i, j) for i, j in modes(P)], axis=1)


def grad_vandermonde(P, r, s):
    a, b = rs_to_ab(r, s)
    g = [grad_simplex2dp(a, b, i, j) for i, j in modes(P)]
    return np.stack([x[0] for x in g], axis=1), np.stack([x[1] for x in g], axis=1)


def gauss_lobatto(P):
    """Gauss-Lobatto-Legendre points of P+1 nodes (roots of (1-x^2) P_P')."""
    if P == 1:
        return np.array([-1.0, 1.0])
    inner = np.polynomial.legendre.Legendre.basis(P).deriv().roots()
    return np.concatenate([[-1.0], np.sort(inner.real), [1.0]])


def warpfactor(P, rout):
    lgl = gauss_lobatto(P)
    req = np.linspace(-1.0, 1.0, P + 1)
    veq = np.stack([jacobi_p(req, 0, 0, i) for i in range(P + 1)], axis=1)
    pmat = np.stack([jacobi_p(rout, 0, 0, i) for i in range(P + 1)], axis=0)
    lmat = np.linalg.solve(veq.T, pmat)
    warp = lmat.T @ (lgl - req)
    zerof = np.abs(rout) < 1.0 - 1e-10
    sf = 1.0 - (zerof * rout) ** 2
    return warp / sf + warp * (zerof - 1)


def spts(P):
    """Warp-and-blend nodes on the reference triangle, (r, s) pairs."""
    alpha = ALPHA_OPT[P] if P < len(ALPHA_OPT) else 5.0 / 3.0
    L1, L3 = [], []
    for n in range(P + 1):
        for m in range(P + 1 - n):
            L1.append(n / P)
            L3.append(m / P)
    L1, L3 = np.array(L1), np.array(L3)
    L2 = 1.0 - L1 - L3
    x = -L2 + L3
    y = (-L2 - L3 + 2 * L1) / math.sqrt(3.0)
    b1, b2, b3 = 4 * L2 * L3, 4 * L1 * L3, 4 * L1 * L2
    w1 = b1 * warpfactor(P, L3 - L2) * (1 + (alpha * L1) ** 2)
    w2 = b2 * warpfactor(P, L1 - L3) * (1 + (alpha * L2) ** 2)
    w3 = b3 * warpfactor(P, L2 - L1) * (1 + (alpha * L3) ** 2)
    x = x + w1 + math.cos(2 * math.pi / 3) * w2 + math.cos(4 * math.pi / 3) * w3
    y = y + math.sin(2 * math.pi / 3) * w2 + math.sin(4 * math.pi / 3) * w3
    # equilateral -> reference right triangle
    l1 = (math.sqrt(3.0) * y + 1.0) / 3.0
    l2 = (-3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    l3 = (3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    return np.stack([-l2 + l3 - l1, -l2 - l3 + l1], axis=1)


def fpts(x1d):
    """P+1 points per edge, edge f from vertex f to vertex f+1."""
    out = []
    for f in range(3):
        va, vb = VERTS[f], VERTS[(f + 1) % 3]
        for t in x1d:
            out.append(va * (1 - t) / 2 + vb * (1 + t) / 2)
    return np.array(out)


def operators(P, x1d, w1d):
    """dict of the triangle FR operators of order P with edge points x1d."""
    S = spts(P)
    F = fpts(x1d)
    V = vandermonde(P, S[:, 0], S[:, 1])
    Vi = np.linalg.inv(V)
    Vr, Vs = grad_vandermonde(P, S[:, 0], S[:, 1])
    Psi = vandermonde(P, F[:, 0], F[:, 1])
    W = np.concatenate([np.asarray(w1d) * EDGE_LEN[f] / 2 for f in range(3)])
    L = V @ Psi.T @ np.diag(W)
    tn = np.repeat(TNORM, len(x1d), axis=0)
    return dict(spts=S, fpts=F, Dr=Vr @ Vi, Ds=Vs @ Vi, E=Psi @ Vi, L=L, W=W, tnorm=tn,
                wts=math.sqrt(2.0) * Vi[0, :])


def residual(U, cfg, ops, JG, nrm, dA, faces):
    """divF[spt][ele][field] of U[spt][ele][field] (reference layout) on a
    triangle mesh: JG[dimRef][spt][ele][dimPhys], own unit normals
    nrm[fpt][ele][dim] and dA[fpt][ele] (the library's geometry arrays),
    faces[i] = (eL, fL, eR, fR, rot, bc).  The pointwise physics are the
    oracle's restatements of the reference (flux.cpp, face.cpp, boundFace.cpp);
    the right element of a face evaluates the common flux with the left
    element's normal and scales it by -dA_R (intFace.cpp:116-120)."""
    from oracle import Point
    pt = Point(cfg)
    nS, nE, NF = U.shape
    P1 = nS and int(round((-3 + math.sqrt(1 + 8 * nS)) / 2)) + 1
    nF = 3 * P1
    F = np.zeros((2, nS, nE, NF))
    for e in range(nE):
        for s in range(nS):
            u5 = np.zeros(5)
            u5[:NF] = U[s, e]
            f = pt.inviscid_flux(u5)
            F[:, s, e, :] = f[:2, :NF]
    Ft = np.einsum("dsep,psek->dsek", JG, F)
    uo = np.einsum("qs,sek->qek", ops["E"], U)
    kind_int = 2 if cfg["equation"] == 0 else cfg["riemannType"]
    kind_bnd = 2 if cfg["equation"] == 0 else 0
    Fn = np.zeros((nF, nE, NF))

    def cf(kind, a, b, n):
        ua, ub = np.zeros(5), np.zeros(5)
        ua[:NF], ub[:NF] = a, b
        return pt.common_flux(kind, ua, ub, np.array([n[0], n[1], 0.0]))[:NF]

    for eL, fL, eR, fR, rot, bc in faces:
        for j in range(P1):
            qL = fL * P1 + j
            nL = nrm[qL, eL]
            if bc >= 0:
                u5 = np.zeros(5)
                u5[:NF] = uo[qL, eL]
                ub = pt.boundary_state(int(bc), u5, np.array([nL[0], nL[1], 0.0]))[:NF]
                Fn[qL, eL] = cf(kind_bnd, uo[qL, eL], ub, nL) * dA[qL, eL]
                continue
            qR = fR * P1 + (P1 - 1 - j)
            fc = cf(kind_int, uo[qL, eL], uo[qR, eR], nL)
            Fn[qL, eL] = fc * dA[qL, eL]
            Fn[qR, eR] = fc * (-dA[qR, eR])
    disFn = (ops["tnorm"][:, 0, None, None] * np.einsum("qs,sek->qek", ops["E"], Ft[0]) +
             ops["tnorm"][:, 1, None, None] * np.einsum("qs,sek->qek", ops["E"], Ft[1]))
    jump = Fn - disFn
    return (np.einsum("ts,sek->tek", ops["Dr"], Ft[0]) + np.einsum("ts,sek->tek", ops["Ds"], Ft[1]) +
            np.einsum("tq,qek->tek", ops["L"], jump))


def geometry(mesh, ops):
    """JG[dimRef][spt][ele][dimPhys], detJ[spt][ele], own unit normals
    nrm[fpt][ele][dim] and dA[fpt][ele] of a linear triangle mesh (affine
    elements: x = x0 + J^T (r + 1) / 2 with J[dimRef][dimPhys])."""
    X = mesh.xv[mesh.c2v][:, :, :2]
    nE = X.shape[0]
    J = np.zeros((nE, 2, 2))
    J[:, 0, :] = (X[:, 1] - X[:, 0]) / 2
    J[:, 1, :] = (X[:, 2] - X[:, 0]) / 2
    G = np.zeros_like(J)
    G[:, 0, 0], G[:, 0, 1], G[:, 1, 0], G[:, 1, 1] = J[:, 1, 1], -J[:, 1, 0], -J[:, 0, 1], J[:, 0, 0]
    det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
    nS, nF = len(ops["spts"]), len(ops["fpts"])
    JG = np.broadcast_to(G.transpose(1, 0, 2)[:, None, :, :], (2, nS, nE, 2)).copy()
    detJ = np.broadcast_to(det[None, :], (nS, nE)).copy()
    nv = np.einsum("edp,qd->qep", G, ops["tnorm"])
    dA = np.linalg.norm(nv, axis=-1)
    return JG, detJ, nv / dA[..., None], dA


def faces(mesh, P):
    """Face list (eL, fL, eR, fR, rot, bc) of a triangle mesh: interior edges
    (left = lower element index) first, then boundary edges."""
    can = (lambda v: mesh.vmap[v]) if mesh.vmap is not None else (lambda v: v)
    edges = {}
    for e in range(mesh.nEles):
        for f in range(3):
            a, b = can(mesh.c2v[e, f]), can(mesh.c2v[e, (f + 1) % 3])
            edges.setdefault(tuple(sorted((int(a), int(b)))), []).append((e, f))
    inner, outer = [], []
    bsets = [(bc, set(int(can(v)) for v in vs)) for bc, vs in mesh.bounds]
    for key in sorted(edges):
        lst = sorted(edges[key])
        if len(lst) == 2:
            inner.append((lst[0][0], lst[0][1], lst[1][0], lst[1][1], 1, -1))
        else:
            bc = next(b for b, s in bsets if key[0] in s and key[1] in s)
            outer.append((lst[0][0], lst[0][1], -1, -1, 0, bc))
    return inner + outer
