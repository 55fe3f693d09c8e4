"""Batched persistent stage kernel (fr_batch_*): meshes of different sizes,
orders and boundaries advanced together must reproduce each mesh advanced on
its own (same stage computation; bar 1e-14 relative), and the oracle."""
import numpy as np
import pytest

from fr_learn_amd import Batch, Solver, cases, mesh
from oracle import Oracle

from helpers import MIXED_BCS, QUAD_BCS, TOL_FP64, euler_cfg, hex_box, perturbed_state, quad_box, rel_err

pytestmark = pytest.mark.gpu


def _quad_members(rng, n, timeType=4, precision="fp64"):
    members = []
    for i in range(n):
        P = 1 + i % 5
        cells = int(rng.integers(3, 9))
        periodic = bool(i % 3)
        m = quad_box(cells, periodic=periodic, bc=None if periodic else QUAD_BCS, seed=i, lo=-5.0, hi=5.0)
        cfg = euler_cfg(2, P, riemannType=int(i % 2), timeType=timeType, dt=1e-3)
        cx, cy = rng.uniform(-1.0, 1.0, 2)
        beta, mach = rng.uniform(1.0, 5.0), rng.uniform(0.2, 0.6)
        ic = (lambda cx, cy, beta, mach: lambda x, c: cases.ic_vortex(x, c, beta=beta, mach=mach,
                                                                        center=(cx, cy)))(cx, cy, beta, mach)
        members.append((cfg, m, ic))
    print("batch of", n, "quad meshes, P 1..5", flush=True)
    return members


def _run(members, steps, precision="fp64"):
    solo, batched, U0s = [], [], []
    for cfg, m, ic in members:
        a = Solver(cfg, m, precision=precision)
        U0 = a.initialize(ic)
        b = Solver(cfg, m, precision=precision)
        b.set_state(U0)
        solo.append(a)
        batched.append(b)
        U0s.append(U0)
    for a in solo:
        a.update(steps)
    B = Batch(batched)
    grid, items, ms = B.tuning()
    print("autotuned grid", grid, "items", items, "ms", ms)
    assert grid >= 1 and items >= len(members)
    B.update(steps)
    return solo, batched, U0s, B


def _mixed_members(rng, n):
    """Quads and triangles (every second member), P 1..5."""
    members = []
    for i in range(n):
        P = 1 + (i // 2) % 5
        cells = int(rng.integers(3, 8))
        periodic = bool(i % 3)
        bc = None if periodic else QUAD_BCS
        if i % 2:
            m = mesh.tri_box((cells, cells), (-5.0, -5.0), (5.0, 5.0), periodic=[periodic] * 2, bc=bc,
                             perturb=0.15, rotate=True, shuffle=True, seed=i)
        else:
            m = quad_box(cells, periodic=periodic, bc=bc, seed=i, lo=-5.0, hi=5.0)
        cfg = euler_cfg(2, P, riemannType=int((i // 2) % 2), timeType=4, dt=1e-3)
        ic = (lambda b: lambda x, c: cases.ic_vortex(x, c, beta=b, mach=0.4))(rng.uniform(1.0, 5.0))
        members.append((cfg, m, ic))
    print("batch of", n, "quad and triangle meshes", flush=True)
    return members


def test_batch_mixes_quads_and_triangles():
    rng = np.random.default_rng(11)
    members = _mixed_members(rng, 14)
    solo, batched, _, B = _run(members, 3)
    for i, (a, b) in enumerate(zip(solo, batched)):
        err = rel_err(b.get_state(), a.get_state()).max()
        print(i, "tri" if i % 2 else "quad", "P", members[i][0]["order"], "err", err)
        assert err <= 1e-14


@pytest.mark.parametrize("layout", [0, 1])
def test_batch_tile_layouts(layout):
    """Either tile layout (threads per element a power of two, or packed to
    the element's point count, elements straddling wavefronts) gives each
    member's own update; the quads' in-block trace exchange then covers other
    faces than the per-mesh kernel's, so per-mesh steps after batched ones
    (and batched after per-mesh) must still agree."""
    rng = np.random.default_rng(21)
    members = _mixed_members(rng, 20)
    from fr_learn_amd.solver import Batch as BatchClass
    BatchClass.force_tile(layout)
    try:
        solo, batched, _, B = _run(members, 2)
    finally:
        BatchClass.force_tile(-1)
    tiles = B.group_tuning()
    print(tiles)
    for (kind, P), t in tiles.items():
        n_pts = (P + 1) ** 2 if kind == "quad" else (P + 1) * (P + 2) // 2
        n_fpts = 4 * (P + 1) if kind == "quad" else 3 * (P + 1)
        pow2 = max(8, 1 << (n_pts - 1).bit_length()) if kind == "quad" else \
            1 << (max(n_pts, n_fpts) - 1).bit_length()
        want = max(n_pts, n_fpts) if layout == 1 and max(n_pts, n_fpts) != pow2 else pow2
        assert t["threads_per_element"] == want
        assert t["elements_per_block"] == 256 // want and t["items"] >= 1 and t["grid"] >= 1
    for i, (a, b) in enumerate(zip(solo, batched)):
        err = rel_err(b.get_state(), a.get_state()).max()
        print(i, "tri" if i % 2 else "quad", "P", members[i][0]["order"], "err", err)
        assert err <= 1e-14
    for k in (0, 2, 4):
        batched[k].update(1)
        solo[k].update(1)
    B.update(1)
    for a in solo:
        a.update(1)
    for a, b in zip(solo, batched):
        assert rel_err(b.get_state(), a.get_state()).max() <= 1e-14
        assert rel_err(b.residual(0), a.residual(0)).max() <= 1e-12


@pytest.mark.parametrize("timeType", [4, 0])
def test_batch_matches_each_mesh(timeType):
    rng = np.random.default_rng(7)
    members = _quad_members(rng, 12, timeType=timeType)
    solo, batched, _, B = _run(members, 3)
    for i, (a, b) in enumerate(zip(solo, batched)):
        err = rel_err(b.get_state(), a.get_state()).max()
        print(i, "P", members[i][0]["order"], "err", err)
        assert err <= 1e-14
    # the members stay usable alone, and further batch steps continue from there
    batched[0].update(1)
    solo[0].update(1)
    B.update(2)
    for a in solo:
        a.update(2)
    for a, b in zip(solo, batched):
        assert rel_err(b.get_state(), a.get_state()).max() <= 1e-14


def test_batch_mixes_dimensions():
    """Quads, triangles and hexes (3D Euler, five fields against the 2D
    four) in one batch: each group its own kernel and field count; every
    member equals its own update, and a hex member the oracle."""
    rng = np.random.default_rng(31)
    members = _mixed_members(rng, 6)
    for P in (1, 2, 3):
        if P % 2:
            members.append((euler_cfg(3, P, dt=1e-3), hex_box(3, seed=P), cases.ic_tgv))
        else:  # boundaries: a perturbed flow near the boundary states
            ic = lambda x, c: perturbed_state(cases.ic_uniform(x, c, vel=[0.3, 0.1, -0.05], p=1.0 / 1.4),
                                              seed=4, amp=2e-2)
            members.append((euler_cfg(3, P, dt=1e-3), hex_box(3, seed=P, periodic=False, bc=MIXED_BCS, L=1.0), ic))
    solo, batched, U0s, B = _run(members, 2)
    tiles = B.group_tuning()
    print(tiles)
    assert {k for k in tiles if k[0] == "hex"} == {("hex", 1), ("hex", 2), ("hex", 3)}
    for i, (a, b) in enumerate(zip(solo, batched)):
        err = rel_err(b.get_state(), a.get_state(), axis=None)
        print(i, members[i][0]["nDims"], "P", members[i][0]["order"], "err", err)
        assert err <= 1e-14
    # the hex member with boundaries against the oracle (its fields are all of
    # order 0.1..1, unlike the Taylor-Green state's rho w)
    i = len(members) - 2
    cfg, m, _ = members[i]
    o = Oracle(cfg, m)
    o.set_state(U0s[i])
    o.update(2)
    assert rel_err(batched[i].get_state(), o.get_state()).max() <= TOL_FP64
    o.close()


def test_batch_vs_oracle_and_hex():
    rng = np.random.default_rng(9)
    members = _quad_members(rng, 5)
    _, batched, U0s, _ = _run(members, 2)
    for (cfg, m, _), b, U0 in zip(members[:2], batched[:2], U0s[:2]):
        o = Oracle(cfg, m)
        o.set_state(U0)
        o.update(2)
        assert rel_err(b.get_state(), o.get_state()).max() <= TOL_FP64
    # hexes, P1-P4, periodic
    hm = []
    for P in (1, 2, 3, 4):
        cfg = euler_cfg(3, P, dt=1e-3)
        hm.append((cfg, hex_box(3, seed=P), cases.ic_tgv))
    solo, batched, _, _ = _run(hm, 2)
    for a, b in zip(solo, batched):
        # The batched launch runs the boundary-capable instantiation of the
        # stage kernel and a lone periodic mesh the one without boundary
        # faces; the compiler may fuse their multiply-adds differently, so
        # they agree to rounding: relative to the state's largest value (the
        # TGV's rho w is 1e-4 of it, where a per-field measure would count
        # those roundings 1e4 times over).
        assert rel_err(b.get_state(), a.get_state(), axis=None) <= 1e-14


def test_batch_fp32():
    rng = np.random.default_rng(11)
    members = _quad_members(rng, 6)
    solo, batched, _, _ = _run(members, 3, precision="fp32")
    for a, b in zip(solo, batched):
        assert rel_err(b.get_state(), a.get_state()).max() <= 1e-7


def test_batch_follows_member_settings():
    """Settings changed on a member between batch calls -- a fixed dt, a CFL
    time step (global or local, recomputed every step), the geometry layout
    -- are what the next batched steps use: the batch equals the same
    members advanced alone with the same changes."""
    rng = np.random.default_rng(5)
    members = _quad_members(rng, 6)
    solo, batched, _, B = _run(members, 2)
    for i, (a, b) in enumerate(zip(solo, batched)):
        assert rel_err(b.get_state(), a.get_state()).max() <= 1e-14
    for s in (solo, batched):
        s[0].set_dt(5e-4)
        s[1].geometry_mode(0)
        s[2].calc_dt(2, 0.3)
        s[3].calc_dt(1, 0.3)
    for a in solo:
        a.update(3)
    B.update(3)
    for i, (a, b) in enumerate(zip(solo, batched)):
        err = rel_err(b.get_state(), a.get_state()).max()
        print(i, "err", err)
        assert err <= 1e-14
        assert np.abs(b.dt() - a.dt()).max() <= 1e-14 * a.dt().max()
    # back to fixed time steps
    for s in (solo, batched):
        s[2].set_dt(1e-3)
        s[3].set_dt(1e-3)
        s[1].geometry_mode(1)
    for a in solo:
        a.update(2)
    B.update(2)
    for a, b in zip(solo, batched):
        assert rel_err(b.get_state(), a.get_state()).max() <= 1e-14


def test_batch_with_destroyed_member_refuses_to_step():
    from fr_learn_amd._lib import FrError
    rng = np.random.default_rng(6)
    members = _quad_members(rng, 3)
    _, batched, _, B = _run(members, 1)
    batched[1].close()
    with pytest.raises(FrError) as ei:
        B.update(1)
    assert ei.value.code == -1
    B.close()
    batched[0].close()
    batched[2].close()
