"""Independent textbook implementations used to pin the oracle.

Nothing here follows the reference's code: quadrature comes from numpy's
Gauss-Legendre routine, Lagrange bases from the barycentric formula, metrics
from the analytic derivative of the trilinear map, fluxes from their textbook
definitions, and the residual from the weak form of nodal DG.

Why the weak form pins the FR residual: with Gauss-Legendre solution points
and the DG correction function (eta = 0), flux reconstruction is the nodal DG
method (Huynh 2007): g_R'(x_i) = l_i(+1)/w_i, g_L'(x_i) = -l_i(-1)/w_i, and
Gauss quadrature integrates every product exactly, so for each element

    w_i divF_i = oint l_i Fn* dS_ref - int grad(l_i) . F~ dV_ref

where F~ = det(J) J^-1 F is the collocated contravariant flux and Fn* the
common normal flux times the face area element.  A mistake in extrapolation,
differentiation, correction sign, face pairing, normals or metrics changes
this residual at O(1).
"""
import numpy as np

# ----------------------------------------------------------------------------
# 1D bases
# ----------------------------------------------------------------------------


def gauss_legendre(P):
    return np.polynomial.legendre.leggauss(P + 1)


def bary_weights(x):
    n = len(x)
    w = np.ones(n)
    for m in range(n):
        for k in range(n):
            if k != m:
                w[m] /= (x[m] - x[k])
    return w


def lagrange_values(x, y):
    """L[j, m] = l_m(y_j) for nodes x (barycentric formula)."""
    y = np.atleast_1d(np.asarray(y, dtype=float))
    w = bary_weights(x)
    L = np.zeros((len(y), len(x)))
    for j, yy in enumerate(y):
        d = yy - x
        hit = np.nonzero(d == 0)[0]
        if len(hit):
            L[j, hit[0]] = 1.0
        else:
            t = w / d
            L[j] = t / t.sum()
    return L


def derivative_matrix(x):
    """Dm[i, m] = l_m'(x_i) (barycentric differentiation matrix)."""
    w = bary_weights(x)
    n = len(x)
    Dm = np.zeros((n, n))
    for i in range(n):
        for m in range(n):
            if i != m:
                Dm[i, m] = (w[m] / w[i]) / (x[i] - x[m])
        Dm[i, i] = -Dm[i].sum()
    return Dm


def legendre_poly(n):
    c = np.zeros(n + 1)
    c[n] = 1.0
    return c


def vcjh_right_derivative(xs, P, eta):
    """d/dx of g_R = (L_P + (eta L_{P-1} + L_{P+1}) / (1 + eta)) / 2 (Vincent et al. 2011)."""
    L = np.polynomial.legendre
    cP = np.pad(legendre_poly(P), (0, 1))
    cPm = np.pad(legendre_poly(P - 1), (0, 2))
    cPp = legendre_poly(P + 1)
    c = 0.5 * (cP + (eta * cPm + cPp) / (1.0 + eta))
    return L.legval(xs, L.legder(c))


def vcjh_eta(scheme, P):
    """eta of the VCJH family: DG 0, SD P/(P+1), HU (P+1)/P, C+ from the
    tabulated c+ values with eta = c (2P+1) (a_P P!)^2 / 2,
    a_P = (2P)! / (2^P (P!)^2) (Vincent, Castonguay, Jameson 2011)."""
    from math import factorial
    if scheme == 0:
        return 0.0
    if scheme == 1:
        return P / (P + 1)
    if scheme == 2:
        return (P + 1) / P
    c = {2: 0.206, 3: 3.80e-3, 4: 4.67e-5, 5: 4.28e-7}[P]
    a = factorial(2 * P) / (2 ** P * factorial(P) ** 2)
    return c * (2 * P + 1) / 2 * (a * factorial(P)) ** 2


# ----------------------------------------------------------------------------
# element geometry (linear hex / quad, Gmsh corner order)
# ----------------------------------------------------------------------------
HEX_S = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
QUAD_S = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=float)


def multilinear(X, xi):
    """Positions and Jacobians of linear elements.
    X: [nE, nc, 3] corners; xi: [npt, D].  Returns x [nE, npt, D] and
    J [nE, npt, D(phys), D(ref)] with J[p, r] = dx_p / dxi_r."""
    D = xi.shape[1]
    S = HEX_S if D == 3 else QUAD_S
    N = np.ones((xi.shape[0], S.shape[0]))
    for d in range(D):
        N *= 0.5 * (1 + np.outer(xi[:, d], S[:, d]))
    dN = np.zeros((D, xi.shape[0], S.shape[0]))
    for r in range(D):
        t = np.ones((xi.shape[0], S.shape[0]))
        for d in range(D):
            t *= 0.5 * S[:, d][None, :] if d == r else 0.5 * (1 + np.outer(xi[:, d], S[:, d]))
        dN[r] = t
    Xd = X[:, :, :D]
    x = np.einsum("pc,ecd->epd", N, Xd)
    J = np.einsum("rpc,ecd->epdr", dN, Xd)
    return x, J


def tensor_points(x1, D):
    """Solution points: index s = i + P1*j (+ P1^2*k), x fastest."""
    P1 = len(x1)
    idx = np.stack(np.meshgrid(*[np.arange(P1)] * D, indexing="ij"), axis=-1)
    idx = idx.transpose(*reversed(range(D)), D).reshape(-1, D)  # x fastest
    return idx, x1[idx]


# ----------------------------------------------------------------------------
# pointwise physics (textbook)
# ----------------------------------------------------------------------------


def primitives(U, D, gamma):
    rho = U[..., 0]
    vel = U[..., 1:D + 1] / rho[..., None]
    E = U[..., D + 1]
    p = (gamma - 1.0) * (E - 0.5 * rho * (vel ** 2).sum(-1))
    return rho, vel, E, p


def euler_flux(U, D, gamma):
    """F[..., dim, field] of the Euler equations."""
    rho, vel, E, p = primitives(U, D, gamma)
    F = np.zeros(U.shape[:-1] + (D, D + 2))
    for d in range(D):
        F[..., d, 0] = rho * vel[..., d]
        for c in range(D):
            F[..., d, 1 + c] = rho * vel[..., d] * vel[..., c] + (p if c == d else 0.0)
        F[..., d, D + 1] = (E + p) * vel[..., d]
    return F


def rusanov(UL, UR, n, D, gamma):
    FL = np.einsum("...dk,...d->...k", euler_flux(UL, D, gamma), n)
    FR = np.einsum("...dk,...d->...k", euler_flux(UR, D, gamma), n)
    rl, vl, _, pl = primitives(UL, D, gamma)
    rr, vr, _, pr = primitives(UR, D, gamma)
    lam = np.maximum(np.abs((vl * n).sum(-1)) + np.sqrt(gamma * pl / rl),
                     np.abs((vr * n).sum(-1)) + np.sqrt(gamma * pr / rr))
    return 0.5 * (FL + FR) - 0.5 * lam[..., None] * (UR - UL)


def roe_2d(UL, UR, n, gamma):
    """Roe flux by eigen-decomposition with Harten's entropy fix
    |lam| -> (lam^2 + delta^2) / (2 delta) when |lam| < delta,
    delta = |vnL - vnR| + |cL - cR|."""
    D = 2
    FL = np.einsum("...dk,...d->...k", euler_flux(UL, D, gamma), n)
    FR = np.einsum("...dk,...d->...k", euler_flux(UR, D, gamma), n)
    rl, vl, El, pl = primitives(UL, D, gamma)
    rr, vr, Er, pr = primitives(UR, D, gamma)
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    u = (sl[..., None] * vl + sr[..., None] * vr) / (sl + sr)[..., None]
    H = (sl * (El + pl) / rl + sr * (Er + pr) / rr) / (sl + sr)
    q2 = (u ** 2).sum(-1)
    c = np.sqrt((gamma - 1) * (H - 0.5 * q2))
    rt = sl * sr
    vn = (u * n).sum(-1)
    dp = pr - pl
    drho = rr - rl
    dv = vr - vl
    dvn = (dv * n).sum(-1)
    delta = np.abs((vl * n).sum(-1) - (vr * n).sum(-1)) + np.abs(np.sqrt(gamma * pl / rl) - np.sqrt(gamma * pr / rr))

    def fix(lam):
        a = np.abs(lam)
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(a < delta, (a * a + delta * delta) / (2 * delta), a)

    l1, l2, l3 = fix(vn - c), fix(vn), fix(vn + c)
    a1 = (dp - rt * c * dvn) / (2 * c * c)
    a2 = drho - dp / (c * c)
    a3 = (dp + rt * c * dvn) / (2 * c * c)
    ones = np.ones_like(vn)
    r1 = np.stack([ones, u[..., 0] - c * n[..., 0], u[..., 1] - c * n[..., 1], H - c * vn], -1)
    r2 = np.stack([ones, u[..., 0], u[..., 1], 0.5 * q2], -1)
    r3 = np.stack([ones, u[..., 0] + c * n[..., 0], u[..., 1] + c * n[..., 1], H + c * vn], -1)
    dvt = dv - dvn[..., None] * n
    rs = np.stack([0 * ones, dvt[..., 0], dvt[..., 1], (u * dvt).sum(-1)], -1) * rt[..., None]
    diss = (l1 * a1)[..., None] * r1 + (l2 * a2)[..., None] * r2 + (l3 * a3)[..., None] * r3 + l2[..., None] * rs
    return 0.5 * (FL + FR) - 0.5 * diss


def lax_friedrichs_adv(uL, uR, n, a, lam):
    an = (n * a).sum(-1)
    return (an * 0.5 * (uL[..., 0] + uR[..., 0]) + 0.5 * lam * np.abs(an) * (uL[..., 0] - uR[..., 0]))[..., None]


def boundary_state(bc, UL, n, cfg):
    """Ghost states of the wall / inflow / outflow conditions (static walls)."""
    D = cfg["nDims"]
    g = cfg["gamma"]
    if cfg["equation"] == 0:
        return np.zeros_like(UL)
    rho, vel, E, p = primitives(UL, D, g)
    vB = np.array([cfg["uBound"], cfg["vBound"], cfg["wBound"]])[:D]

    def cons(r, v, pp):
        return np.concatenate([r[..., None], r[..., None] * v, (pp / (g - 1) + 0.5 * r * (v ** 2).sum(-1))[..., None]],
                              -1)

    ones = np.ones_like(rho)
    if bc in (8, 12):  # slip wall / symmetry: mirror the normal velocity
        vR = vel - 2 * (vel * n).sum(-1)[..., None] * n
        return cons(rho, vR, p)
    if bc == 3:  # supersonic outflow
        return UL.copy()
    if bc == 2:  # supersonic inflow
        return cons(cfg["rhoBound"] * ones, vB * ones[..., None], cfg["pBound"] * ones)
    if bc == 4:  # subsonic inflow: rho, v fixed, p from the interior
        return cons(cfg["rhoBound"] * ones, vB * ones[..., None], p)
    if bc == 5:  # subsonic outflow: p fixed
        return cons(rho, vel, cfg["pBound"] * ones)
    if bc in (9, 13):  # isothermal no-slip wall (optionally moving)
        vw = np.array([cfg["uWall"], cfg["vWall"], cfg["wWall"]])[:D] if bc == 13 else np.zeros(D)
        return cons(p / (cfg["RGas"] * cfg["TWall"]), vw * ones[..., None], p)
    if bc == 10:  # adiabatic no-slip wall
        return cons(rho, 0 * vel, p)
    raise ValueError(f"no textbook state for bc {bc}")


def char_state_subsonic(UL, n, cfg):
    """Riemann-invariant far field, subsonic: R+ = vn + 2c/(g-1) from the
    interior, R- from the free stream; entropy and tangential velocity from
    the upwind side."""
    D = cfg["nDims"]
    g = cfg["gamma"]
    rho, vel, E, p = primitives(UL, D, g)
    vB = np.array([cfg["uBound"], cfg["vBound"], cfg["wBound"]])[:D]
    rB, pB = cfg["rhoBound"], cfg["pBound"]
    cL = np.sqrt(g * p / rho)
    cB = np.sqrt(g * pB / rB)
    vnL = (vel * n).sum(-1)
    vnB = (vB * n).sum(-1)
    Rp = vnL + 2 * cL / (g - 1)
    Rm = vnB - 2 * cB / (g - 1)
    vn = 0.5 * (Rp + Rm)
    c = 0.25 * (g - 1) * (Rp - Rm)
    inflow = vnL < 0
    s = np.where(inflow, pB / rB ** g, p / rho ** g)
    r = (c * c / (g * s)) ** (1 / (g - 1))
    pp = r * c * c / g
    vt = np.where(inflow[..., None], vB - vnB[..., None] * n, vel - vnL[..., None] * n)
    v = vt + vn[..., None] * n
    return np.concatenate([r[..., None], r[..., None] * v, (pp / (g - 1) + 0.5 * r * (v ** 2).sum(-1))[..., None]], -1)


# ----------------------------------------------------------------------------
# viscous terms (textbook Navier-Stokes / diffusion)
# ----------------------------------------------------------------------------


def _velocity_energy(U, D):
    """v_j = m_j / rho and specific internal energy e = E / rho - |v|^2 / 2
    (complex-safe, for complex-step differentiation)."""
    rho = U[..., 0]
    v = U[..., 1:D + 1] / rho[..., None]
    e = U[..., D + 1] / rho - 0.5 * (v * v).sum(-1)
    return v, e


def viscous_flux(U, q, cfg, defects=False):
    """Viscous flux with the sign convention F_total = F_inviscid + Fv, i.e.
    Fv[d] = (0, -tau_d., -(tau_d. v + kappa de/dx_d)) for Navier-Stokes with
    Stokes' hypothesis, kappa = mu gamma / Pr (heat flux in terms of e = cv T),
    mu constant (fixVis) or Sutherland's law mu_inf (T/T_ref)^1.5 (1 + S) /
    (T/T_ref + S) with T/T_ref = (gamma - 1) e / rt_inf; Fv = -diffD grad u
    for advection-diffusion.  q[..., dim, field] is the physical gradient of
    the conserved variables; the primitive gradients are taken by complex-step
    differentiation (exact to rounding), not by the quotient rule.
    defects=True reproduces three index substitutions of the reference's 3D
    flux (tau_xz = mu(du/dz + dv/dx), tau_yz = mu(du/dz + dv/dy), and the
    x-flux of z-momentum -tau_zz)."""
    D = cfg["nDims"]
    if cfg["equation"] == 0:
        return -cfg["diffD"] * q
    g = cfg["gamma"]
    v, e = _velocity_energy(U, D)
    h = 1e-30
    dv = np.zeros(U.shape[:-1] + (D, D))  # [..., dim, comp]
    de = np.zeros(U.shape[:-1] + (D,))
    for d in range(D):
        vc, ec = _velocity_energy(U + 1j * h * q[..., d, :], D)
        dv[..., d, :] = vc.imag / h
        de[..., d] = ec.imag / h
    mu = cfg["mu_inf"] * np.ones(e.shape)
    if not cfg["fixVis"]:
        T = (g - 1.0) * e / cfg["rt_inf"]
        mu = cfg["mu_inf"] * T ** 1.5 * (1.0 + cfg["c_sth"]) / (T + cfg["c_sth"])
    div = np.trace(dv, axis1=-2, axis2=-1)
    tau = mu[..., None, None] * (dv + np.swapaxes(dv, -1, -2))
    for d in range(D):
        tau[..., d, d] -= 2.0 / 3.0 * mu * div
    if defects and D == 3:
        txz = mu * (dv[..., 2, 0] + dv[..., 0, 1])
        tyz = mu * (dv[..., 2, 0] + dv[..., 1, 1])
        tau[..., 0, 2] = tau[..., 2, 0] = txz
        tau[..., 1, 2] = tau[..., 2, 1] = tyz
    kappa = mu * g / cfg["prandtl"]
    F = np.zeros(U.shape[:-1] + (D, D + 2))
    for d in range(D):
        F[..., d, 1:D + 1] = -tau[..., d, :]
        F[..., d, D + 1] = -((tau[..., d, :] * v).sum(-1) + kappa * de[..., d])
    if defects and D == 3:
        F[..., 0, 3] = -tau[..., 2, 2]
    return F


def ldg_switch(n):
    """Direction switch of the LDG one-sided terms: +1 where n . (1, 1, sqrt 2)
    >= 0 (n . (1, 1) in 2D), else -1."""
    w = np.array([1.0, 1.0, np.sqrt(2.0)])[:n.shape[-1]]
    return np.where((n * w).sum(-1) < 0, -1.0, 1.0)


def adiabatic_gradient(q, UL, Ug, cfg):
    """Ghost gradient at an adiabatic no-slip wall: the interior gradient with
    the energy gradient replaced by that of E = rho e with zero temperature
    gradient and the ghost's (zero) velocity: dE = e_g drho, e_g = p_L /
    ((gamma - 1) rho_g)."""
    D = cfg["nDims"]
    g = cfg["gamma"]
    _, _, _, pL = primitives(UL, D, g)
    eg = pL / ((g - 1.0) * Ug[..., 0])
    vg = Ug[..., 1:D + 1] / Ug[..., 0:1]
    assert np.all(vg == 0.0)
    qg = q.copy()
    qg[..., :, D + 1] = eg[..., None] * q[..., :, 0]
    return qg


# ----------------------------------------------------------------------------
# weak-form DG residual
# ----------------------------------------------------------------------------


def weak_form_residual(cfg, mesh, U, box_lo=None, box_hi=None, periodic=None, bc_of=None, defects=False):
    """Reference-space residual divF (layout [spt][ele][field]) of nodal DG in
    weak form with Gauss-Legendre collocation.  Interior face points are
    paired by physical position (modulo the box period in periodic
    directions); boundary points get the ghost state of bc_of(position).

    With cfg["viscous"], the local DG (LDG) method: the reference-space
    gradient is lifted in strong form,

        w_i q_r,i = w_i d_r u(xi_i) + oint l_i (u^ - u) nref_r dS_ref,

    with u^ = {u} - s(n) beta [u] inside and (u + u_ghost) / 2 on boundaries,
    mapped to physical space with J^-T at the solution point, and the common
    viscous normal flux is {F}.n + s(n) beta n.[F] + tau [u] (Navier-Stokes)
    or {F}.n (advection-diffusion), [a] = own minus neighbour, s(n) the
    ldg_switch of the own outward normal."""
    D, P = cfg["nDims"], cfg["order"]
    g = cfg["gamma"]
    adv = cfg["equation"] == 0
    viscous = bool(cfg.get("viscous", 0))
    NF = 1 if adv else D + 2
    x1, w1 = gauss_legendre(P)
    sidx, xis = tensor_points(x1, D)
    NS = len(xis)
    ws = np.prod(w1[sidx], axis=1)
    X = mesh.element_nodes()
    nE = X.shape[0]
    Ue = np.transpose(np.asarray(U).reshape(NS, nE, NF), (1, 0, 2))  # [e][s][k]
    _, J = multilinear(X, xis)
    det = np.linalg.det(J)
    adj = det[..., None, None] * np.linalg.inv(J)  # [e][s][ref][phys]
    Dm = derivative_matrix(x1)

    def line_matrix(r):
        """M[i, s] = d l_s / d xi_r (xi_i) on the tensor lines along r."""
        M = np.zeros((NS, NS))
        for i in range(NS):
            for t in range(NS):
                if all(sidx[i, d] == sidx[t, d] for d in range(D) if d != r):
                    M[i, t] = Dm[sidx[i, r], sidx[t, r]]
        return M

    Dr = [line_matrix(r) for r in range(D)]
    # faces: extrapolation, quadrature, positions, normals, own traces
    faces = []
    for d in range(D):
        for side in (-1.0, 1.0):
            other = [c for c in range(D) if c != d]
            fidx, fx = tensor_points(x1, D - 1)
            xi = np.zeros((fx.shape[0], D))
            xi[:, d] = side
            for c, oc in enumerate(other):
                xi[:, oc] = fx[:, c]
            wf = np.prod(w1[fidx], axis=1)
            E = np.ones((xi.shape[0], NS))
            for c in range(D):
                E *= lagrange_values(x1, xi[:, c])[:, sidx[:, c]]
            xp, Jf = multilinear(X, xi)
            adjf = np.linalg.det(Jf)[..., None, None] * np.linalg.inv(Jf)
            N = side * adjf[:, :, d, :]  # [e][fp][phys]
            dA = np.linalg.norm(N, axis=-1)
            nrm = N / dA[..., None]
            UL = np.einsum("fs,esk->efk", E, Ue)
            faces.append(dict(axis=d, side=side, E=E, wf=wf, x=xp, n=nrm, dA=dA, UL=UL))
    # pair face points by position
    allx = np.concatenate([f["x"].reshape(-1, D) for f in faces])
    key = allx.copy()
    if periodic is not None:
        for d in range(D):
            if periodic[d]:
                L = box_hi[d] - box_lo[d]
                key[:, d] = np.mod(key[:, d] - box_lo[d] + 1e-9 * L, L)
    scale = np.abs(key).max() + 1.0
    kr = np.round(key / scale * 1e6).astype(np.int64)
    order = np.lexsort(kr.T[::-1])
    ks = kr[order]
    same_next = np.all(ks[1:] == ks[:-1], axis=1)
    partner = -np.ones(len(allx), dtype=np.int64)
    i = 0
    while i < len(order):
        j = i
        while j + 1 < len(order) and same_next[j]:
            j += 1
        if j - i == 1:
            partner[order[i]] = order[j]
            partner[order[j]] = order[i]
        elif j - i > 1:
            raise RuntimeError("face point matched more than twice")
        i = j + 1
    allU = np.concatenate([f["UL"].reshape(-1, NF) for f in faces])
    # neighbour or ghost states
    off = 0
    for f in faces:
        cnt = f["UL"].shape[0] * f["UL"].shape[1]
        UL = f["UL"].reshape(-1, NF)
        n = f["n"].reshape(-1, D)
        pr = partner[off:off + cnt]
        UR = np.empty_like(UL)
        has = pr >= 0
        UR[has] = allU[pr[has]]
        bcs = np.full(cnt, -1)
        if (~has).any():
            if bc_of is None:
                raise RuntimeError("unpaired face point without a boundary condition")
            xb = f["x"].reshape(-1, D)[~has]
            bcs[~has] = np.array([bc_of(p) for p in xb])
            for b in np.unique(bcs[~has]):
                sel = np.nonzero(bcs == b)[0]
                if b == 1:
                    UR[sel] = char_state_subsonic(UL[sel], n[sel], cfg)
                else:
                    UR[sel] = boundary_state(int(b), UL[sel], n[sel], cfg)
        f.update(pr=pr, has=has, bcs=bcs, UR=UR, off=off)
        off += cnt
    # physical flux at spts
    if adv:
        a = np.array([cfg["advectVx"], cfg["advectVy"], cfg["advectVz"]])[:D]
        F = a[None, None, :, None] * Ue[:, :, None, :]
    else:
        F = euler_flux(Ue, D, g)
    if viscous:
        # LDG: common solution, lifted reference gradient, physical gradient
        qref = np.stack([np.einsum("is,esk->eik", Dr[r], Ue) for r in range(D)], axis=2)  # [e][s][r][k]
        for f in faces:
            UL, UR, has = f["UL"].reshape(-1, NF), f["UR"], f["has"]
            sw = ldg_switch(f["n"].reshape(-1, D))[:, None]
            beta = cfg["penFact"]
            UC = np.where(has[:, None], 0.5 * (UL + UR) - sw * beta * (UL - UR), 0.5 * (UL + UR))
            dU = (UC - UL).reshape(nE, -1, NF)
            qref[:, :, f["axis"], :] += f["side"] * np.einsum("f,fi,efk->eik", f["wf"], f["E"], dU) / ws[None, :, None]
        q = np.einsum("esrp,esrk->espk", adj, qref) / det[:, :, None, None]  # [e][s][phys][k]
        F = F + viscous_flux(Ue, q, cfg, defects)
        allQ = np.concatenate([np.einsum("fs,esdk->efdk", f["E"], q).reshape(-1, D, NF) for f in faces])
    Ft = np.einsum("esrp,espk->esrk", adj, F)
    # volume term: sum_s w_s d l_i / d xi_r (xi_s) F~_r(xi_s)
    vol = np.zeros((nE, NS, NF))
    for r in range(D):
        vol += np.einsum("si,s,esk->eik", Dr[r], ws, Ft[:, :, r, :])
    face_term = np.zeros((nE, NS, NF))
    for f in faces:
        nfp = f["x"].shape[1]
        UL, UR, has = f["UL"].reshape(-1, NF), f["UR"], f["has"]
        n = f["n"].reshape(-1, D)
        if adv:
            a = np.array([cfg["advectVx"], cfg["advectVy"], cfg["advectVz"]])[:D]
            Fs = lax_friedrichs_adv(UL, UR, n, a, cfg["lambda_"])
        elif D == 2 and cfg["riemannType"] == 1:
            Fs = np.where(has[:, None], roe_2d(UL, UR, n, g), rusanov(UL, UR, n, D, g))
        else:
            Fs = rusanov(UL, UR, n, D, g)
        if viscous:
            qL = allQ[f["off"]:f["off"] + len(UL)]
            qR = np.zeros_like(qL)
            qR[has] = allQ[f["pr"][has]]
            FvL = np.einsum("fdk,fd->fk", viscous_flux(UL, qL, cfg, defects), n)
            if adv:
                FvR = np.einsum("fdk,fd->fk", viscous_flux(UR, qR, cfg, defects), n)
                Fs = Fs + 0.5 * (FvL + FvR)  # boundary: zero outside gradient
            else:
                FvR = np.einsum("fdk,fd->fk", viscous_flux(UR, qR, cfg, defects), n)
                sw = ldg_switch(n)[:, None]
                inner = 0.5 * (FvL + FvR) + sw * cfg["penFact"] * (FvL - FvR) + cfg["tau"] * (UL - UR)
                bnd = FvL + cfg["tau"] * (UL - UR)
                adia = f["bcs"] == 10
                if adia.any():
                    qg = adiabatic_gradient(qL[adia], UL[adia], UR[adia], cfg)
                    Fg = np.einsum("fdk,fd->fk", viscous_flux(UR[adia], qg, cfg, defects), n[adia])
                    bnd[adia] = Fg + cfg["tau"] * (UL[adia] - UR[adia])
                Fs = Fs + np.where(has[:, None], inner, bnd)
        Fs = (Fs * f["dA"].reshape(-1)[:, None]).reshape(nE, nfp, NF)
        face_term += np.einsum("f,fi,efk->eik", f["wf"], f["E"], Fs)
    divF = (face_term - vol) / ws[None, :, None]
    return np.transpose(divF, (1, 0, 2))
