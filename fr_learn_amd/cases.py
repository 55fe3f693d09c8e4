"""Configurations and initial conditions of the benchmark workloads.

`config(**kw)` returns the parameter dict read by the engine, with the
reference's input defaults (src/input.cpp:214-481).  Initial conditions map
solution-point positions x[..., 3] to conservative states U[..., nFields].
"""
import numpy as np

DEFAULTS = dict(
    equation=1, nDims=3, order=3, riemannType=0, viscous=0, timeType=4, sptsType=0, vcjhScheme=0,
    fixVis=0, viscFix=0, dt=1e-3, gamma=1.4, advectVx=1.0, advectVy=1.0, advectVz=0.0, lambda_=1.0,
    diffD=1.0, rhoBound=1.0, uBound=0.2, vBound=0.0, wBound=0.0, pBound=0.7142857143,
    penFact=0.0, tau=1.0, mu_inf=0.0, prandtl=0.72, rt_inf=1.0, c_sth=0.0,
    TWall=300.0, RGas=286.9, uWall=0.0, vWall=0.0, wWall=0.0,
)


def config(**kw):
    c = dict(DEFAULTS)
    for k, v in kw.items():
        if k not in c:
            raise KeyError(k)
        c[k] = v
    return c


def n_fields(cfg):
    return 1 if cfg["equation"] == 0 else cfg["nDims"] + 2


def conservative(rho, vel, p, gamma):
    """(rho, [u, v(, w)], p) -> U with E = p/(gamma-1) + rho |u|^2 / 2."""
    vel = [np.asarray(v) for v in vel]
    E = p / (gamma - 1.0) + 0.5 * rho * sum(v * v for v in vel)
    return np.stack([rho] + [rho * v for v in vel] + [E], axis=-1)


def ic_uniform(x, cfg, rho=None, vel=None, p=None):
    """Uniform free stream (reference icType 0, ele.cpp:688-707)."""
    D = cfg["nDims"]
    rho = cfg["rhoBound"] if rho is None else rho
    vel = [cfg["uBound"], cfg["vBound"], cfg["wBound"]][:D] if vel is None else vel
    p = cfg["pBound"] if p is None else p
    shape = x.shape[:-1]
    return conservative(np.full(shape, rho), [np.full(shape, v) for v in vel], np.full(shape, p), cfg["gamma"])


def ic_tgv(x, cfg, mach=0.1, rho0=1.0, v0=1.0):
    """Taylor-Green vortex on [0, 2pi]^3: u = sin x cos y cos z, v = -cos x sin y
    cos z, w = 0, rho = rho0, p = p0 + rho0 v0^2/16 (cos 2x + cos 2y)(cos 2z + 2)
    with p0 = rho0 v0^2 / (gamma Ma^2)."""
    g = cfg["gamma"]
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    u = v0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -v0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(u)
    p0 = rho0 * v0 * v0 / (g * mach * mach)
    p = p0 + rho0 * v0 * v0 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    return conservative(np.full_like(u, rho0), [u, v, w], p, g)


def ic_vortex(x, cfg, beta=5.0, mach=0.5, center=(0.0, 0.0), uinf=(1.0, 0.0), rho_inf=1.0):
    """Isentropic vortex of strength beta advected by uinf:
    du = -beta/(2pi) y e^{(1-r^2)/2}, dv = beta/(2pi) x e^{(1-r^2)/2},
    T = T_inf - (gamma-1) beta^2/(8 gamma pi^2) e^{1-r^2}, T = p/rho,
    rho = rho_inf (T/T_inf)^{1/(gamma-1)}, p_inf = rho_inf |uinf|^2/(gamma Ma^2)."""
    g = cfg["gamma"]
    X = x[..., 0] - center[0]
    Y = x[..., 1] - center[1]
    r2 = X * X + Y * Y
    speed = float(np.hypot(*uinf)) or 1.0
    p_inf = rho_inf * speed * speed / (g * mach * mach)
    T_inf = p_inf / rho_inf
    f = np.exp(0.5 * (1.0 - r2))
    u = uinf[0] - beta / (2 * np.pi) * Y * f
    v = uinf[1] + beta / (2 * np.pi) * X * f
    T = T_inf - (g - 1.0) * beta * beta / (8.0 * g * np.pi * np.pi) * f * f
    rho = rho_inf * (T / T_inf) ** (1.0 / (g - 1.0))
    p = rho * T
    vel = [u, v] + ([np.zeros_like(u)] if cfg["nDims"] == 3 else [])
    return conservative(rho, vel, p, g)


def ic_gaussian(x, cfg):
    """Scalar Gaussian bump u0 = exp(-|x|^2) (reference advection icType 0,
    ele.cpp:766-772)."""
    D = cfg["nDims"]
    r2 = sum(x[..., d] ** 2 for d in range(D))
    return np.exp(-r2)[..., None]
