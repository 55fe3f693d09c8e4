"""Shared test helpers: standard small workloads and error measures."""
import numpy as np

from fr_learn_amd import cases, mesh

# FP64 parity bar: the GPU FP64 residual within 1e-12 relative of the
# reference-algorithm oracle.  Relative error is measured per field as
# max|a - b| / max|b| over the whole residual array.
TOL_FP64 = 1e-12


def rel_err(a, b, axis=(0, 1)):
    a = np.asarray(a)
    b = np.asarray(b)
    num = np.abs(a - b).max(axis=axis)
    den = np.abs(b).max(axis=axis)
    den = np.where(den == 0, 1.0, den)
    return num / den


def hex_box(n=3, periodic=True, perturb=0.15, rotate=True, shuffle=True, seed=1, bc=None, L=2 * np.pi):
    return mesh.box(3, (n, n, n), (0, 0, 0), (L, L, L), periodic=[periodic] * 3, bc=bc, perturb=perturb,
                    rotate=rotate, shuffle=shuffle, seed=seed)


def quad_box(n=4, periodic=True, perturb=0.15, rotate=True, shuffle=True, seed=1, bc=None, lo=-5.0, hi=5.0):
    return mesh.box(2, (n, n), (lo, lo), (hi, hi), periodic=[periodic] * 2, bc=bc, perturb=perturb,
                    rotate=rotate, shuffle=shuffle, seed=seed)


def perturbed_state(U, seed=0, amp=1e-2):
    """Add a smooth-ish random perturbation so element traces are discontinuous."""
    rng = np.random.default_rng(seed)
    return U * (1.0 + amp * rng.standard_normal(U.shape))


MIXED_BCS = {"xmin": "char", "xmax": "sup_out", "ymin": "slip_wall", "ymax": "sub_out", "zmin": "sub_in",
             "zmax": "isothermal_noslip"}
WALL_BCS = {"xmin": "adiabatic_noslip", "xmax": "symmetry", "ymin": "isothermal_noslip_moving",
            "ymax": "sup_in", "zmin": "slip_wall", "zmax": "char"}
QUAD_BCS = {"xmin": "char", "xmax": "sub_out", "ymin": "slip_wall", "ymax": "sup_in"}


def euler_cfg(nDims=3, order=3, **kw):
    base = dict(nDims=nDims, order=order, equation=1, uBound=0.3, vBound=0.1, wBound=-0.05, pBound=1.0 / 1.4,
                TWall=1.0, RGas=1.0 / 1.4, uWall=0.1)
    base.update(kw)
    return cases.config(**base)


def advection_cfg(nDims=3, order=3, **kw):
    base = dict(nDims=nDims, order=order, equation=0, advectVx=1.0, advectVy=0.5, advectVz=-0.25)
    base.update(kw)
    return cases.config(**base)
