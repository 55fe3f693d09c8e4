"""The workloads of the benchmark description at their full sizes:

config 1: scalar advection, periodic [-5,5]^2, 32x32 quads, P=3, velocity
(1,1), upwind Lax-Friedrichs (lambda = 1), RK4, dt = 1e-3, 1000 steps -- GPU
FP64 against the oracle after all 1000 steps;
config 2: Euler isentropic vortex, periodic [-10,10]^2, 128x128 quads, P=4,
gamma 1.4, Mach 0.5, vortex strength 5, Roe, RK4 -- GPU FP64 against the
oracle over 5 steps, and the FP32 path's relative L2 distance from FP64 after
500 steps (dt = 2e-3);
config 3: Euler Taylor-Green vortex, periodic [0,2pi]^3, 48^3 hexes, P=3,
Rusanov, RK4, dt = 2e-4 -- GPU FP64 residual and one RK4 step against the
oracle (16 threads), discrete conservation, the t = 0 kinetic energy against
its closed form, and the FP32 path's distance from FP64 after 20 steps."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases, mesh
from oracle import Oracle

from helpers import perturbed_state

pytestmark = pytest.mark.gpu


def test_config1_advection_1000_steps():
    cfg = cases.config(nDims=2, order=3, equation=0, advectVx=1.0, advectVy=1.0, lambda_=1.0, timeType=4, dt=1e-3)
    m = mesh.box(2, (32, 32), (-5.0, -5.0), (5.0, 5.0))
    print("config 1: 32x32 P3 advection, 1000 RK4 steps", flush=True)
    g = Solver(cfg, m)
    o = Oracle(cfg, m)
    U0 = g.initialize(cases.ic_gaussian)
    o.set_state(U0)
    g.update(1000)
    o.update(1000)
    a, b = g.get_state(), o.get_state()
    err = np.abs(a - b).max() / np.abs(b).max()
    print("max rel err after 1000 steps", err)
    assert err <= 1e-11
    # the bump has moved by (1, 1) and kept its mass (periodic conservation)
    w = o.operator(8)[:, None, None] * o.geometry(1)[..., None]
    assert abs((a * w).sum() - (U0 * w).sum()) <= 1e-12 * (U0 * w).sum()


def test_config2_vortex_fp64_parity_and_fp32_error():
    cfg = cases.config(nDims=2, order=4, equation=1, riemannType=1, timeType=4, dt=2e-3, gamma=1.4)
    m = mesh.box(2, (128, 128), (-10.0, -10.0), (10.0, 10.0))
    ic = lambda x, c: cases.ic_vortex(x, c, beta=5.0, mach=0.5)
    print("config 2: 128x128 P4 vortex, Roe", flush=True)
    g = Solver(cfg, m)
    o = Oracle(cfg, m)
    U0 = g.initialize(ic)
    o.set_state(U0)
    g.update(5)
    o.update(5)
    a, b = g.get_state(), o.get_state()
    err = (np.abs(a - b).max(axis=(0, 1)) / np.abs(b).max(axis=(0, 1))).max()
    print("FP64 vs oracle after 5 steps", err)
    assert err <= 1e-12
    g.set_state(U0)
    g.update(500)
    ref = g.get_state()
    g32 = Solver(cfg, m, precision="fp32")
    g32.set_state(U0)
    g32.update(500)
    l2 = np.linalg.norm(g32.get_state() - ref) / np.linalg.norm(ref)
    print("FP32 vs FP64 relative L2 after 500 steps", l2)
    assert l2 <= 1e-6


def test_config3_tgv_48_full_size():
    cfg = cases.config(nDims=3, order=3, equation=1, riemannType=0, timeType=4, dt=2e-4, gamma=1.4)
    m = mesh.box(3, (48, 48, 48), (0, 0, 0), (2 * np.pi,) * 3)
    print("config 3: 48^3 P3 TGV, Rusanov", flush=True)
    g = Solver(cfg, m)
    U0 = g.initialize(cases.ic_tgv)
    # kinetic energy of the TGV at t = 0: rho0 V0^2 / 8 per unit volume
    ke, vol = g.kinetic_energy()
    assert abs(vol - (2 * np.pi) ** 3) <= 1e-12 * vol and abs(ke / vol - 0.125) < 1e-6
    # the exact TGV state has a divergence-free velocity, so its density
    # residual is round-off small; a perturbed state gives every field a
    # residual of its own size for the per-field relative bar
    U0 = perturbed_state(U0, seed=5, amp=1e-2)
    g.set_state(U0)
    o = Oracle(cfg, m, threads=16)
    o.set_state(U0)
    g.calc_residual(0)
    o.calc_residual(0)
    a, b = g.residual(0), o.residual(0)
    err = (np.abs(a - b).max(axis=(0, 1)) / np.abs(b).max(axis=(0, 1))).max()
    print("residual rel err", err, "per-field max", np.abs(b).max(axis=(0, 1)), flush=True)
    assert err <= 1e-12
    # periodic conservation: sum over elements of the quadrature of divF
    w = o.operator(8)
    tot = np.einsum("s,sek->k", w, a)
    assert (np.abs(tot) <= 1e-12 * np.einsum("s,sek->k", w, np.abs(a))).all()
    g.update(1)
    o.update(1)
    a, b = g.get_state(), o.get_state()
    err = (np.abs(a - b).max(axis=(0, 1)) / np.abs(b).max(axis=(0, 1))).max()
    print("RK4 step rel err", err, flush=True)
    assert err <= 1e-12
    o.close()
    g.set_state(U0)
    g.update(20)
    g32 = Solver(cfg, m, precision="fp32")
    g32.set_state(U0)
    g32.update(20)
    ref = g.get_state()
    l2 = np.linalg.norm(g32.get_state() - ref) / np.linalg.norm(ref)
    print("FP32 vs FP64 relative L2 after 20 steps", l2)
    assert l2 <= 1e-6
