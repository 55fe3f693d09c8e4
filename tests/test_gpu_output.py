"""Solution files through the C ABI (fr_write_data, fr_read_restart,
fr_write/read_checkpoint) vs the oracle's restatement of the reference's
writers and restart reader (oracle/output.py).

Bar: identical file structure (tags, counts, connectivity, offsets, types);
numbers within 1e-13 relative to the field's largest value (both sides print
16 significant digits of FP64 values that agree to rounding); restart
states within 1e-13; the binary checkpoint bit for bit."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases
from fr_learn_amd._lib import FrError
from oracle import Oracle
from oracle import output as oo

from helpers import advection_cfg, euler_cfg, hex_box, perturbed_state, quad_box

pytestmark = pytest.mark.gpu

def _pair(cfg, m, amp=1e-2):
    print(f"output case P={cfg['order']} eq={cfg['equation']} nEles={m.nEles}", flush=True)
    o = Oracle(cfg, m)
    g = Solver(cfg, m)
    U = perturbed_state(o.initialize(cases.ic_tgv), seed=1, amp=amp)
    o.set_state(U)
    g.set_state(U)
    return o, g, U


def _close(a, b, tol=1e-13):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * scale


def _compare_vtu(text_g, text_o):
    hg, pg = oo.parse_vtu(text_g)
    ho, po = oo.parse_vtu(text_o)
    assert hg == ho
    assert len(pg) == len(po)
    for a, b in zip(pg, po):
        assert sorted(a) == sorted(b)
        for k in a:
            if k.startswith("_tag"):
                assert a[k] == b[k], k
            elif k in ("connectivity", "offsets", "types"):
                assert np.array_equal(a[k], b[k]), k
            else:
                assert _close(a[k], b[k]), k
    assert text_g.splitlines()[-2:] == text_o.splitlines()[-2:]


@pytest.mark.parametrize("order", [1, 3])
def test_paraview_matches_oracle(order, tmp_path):
    cfg = euler_cfg(3, order)
    m = hex_box(3, perturb=0.15)
    o, g, U = _pair(cfg, m)
    g.update(2)
    o.update(2)
    name = str(tmp_path / "sim")
    g.write_data(name, iter=12, time=0.375, plot_type=1)
    text_g = open(name + "_000000012.vtu").read()
    text_o = oo.paraview_text(g.get_state(), cfg, m, o.operator(7), iter=12, time=0.375)
    _compare_vtu(text_g, text_o)


def test_paraview_advection(tmp_path):
    cfg = advection_cfg(3, 2)
    m = hex_box(3, periodic=False, bc="sup_out", L=2.0)
    o = Oracle(cfg, m)
    g = Solver(cfg, m)
    U = o.initialize(lambda x, c: cases.ic_gaussian(x - 1.0, c))
    g.set_state(U)
    name = str(tmp_path / "adv")
    g.write_data(name, iter=0, time=0.0)
    _compare_vtu(open(name + "_000000000.vtu").read(), oo.paraview_text(U, cfg, m, o.operator(7)))


def test_restart_round_trip_and_order_change(tmp_path):
    m = hex_box(3, perturb=0.15)
    cfg2, cfg3 = euler_cfg(3, 2), euler_cfg(3, 3)
    o2, g2, U2 = _pair(cfg2, m)
    name = str(tmp_path / "rst")
    g2.write_data(name, iter=40, time=1.5)
    text = open(name + "_000000040.vtu").read()
    # same order: the state back to 16 significant digits
    h = Solver(cfg2, m)
    assert h.read_restart(name, 40) == 1.5
    assert _close(h.get_state(), U2)
    # P2 file into a P3 context: the reference's interpolation of the file
    g3 = Solver(cfg3, m)
    g3.read_restart(name, 40)
    o3 = Oracle(cfg3, m)
    ref, t = oo.read_restart(text, cfg3, o3.operator(7), lambda P: Oracle(euler_cfg(3, P), m).operator(7))
    assert t == 1.5
    assert _close(g3.get_state(), ref)


def test_restart_errors(tmp_path):
    g = Solver(euler_cfg(3, 2), hex_box(3))
    with pytest.raises(FrError):
        g.read_restart(str(tmp_path / "missing"), 0)
    (tmp_path / "bad_000000000.vtu").write_text("<?xml version=\"1.0\" ?>\n<VTKFile>\n</VTKFile>\n")
    with pytest.raises(FrError):
        g.read_restart(str(tmp_path / "bad"), 0)
    q = Solver(euler_cfg(2, 2), quad_box(4))
    with pytest.raises(FrError):
        q.write_data(str(tmp_path / "q"), plot_type=1)


@pytest.mark.parametrize("D", [2, 3])
def test_csv_matches_oracle(D, tmp_path):
    cfg = euler_cfg(D, 2)
    m = hex_box(3) if D == 3 else quad_box(4)
    o = Oracle(cfg, m)
    g = Solver(cfg, m)
    ic = cases.ic_tgv if D == 3 else (lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4))
    U = perturbed_state(o.initialize(ic), seed=2)
    g.set_state(U)
    name = str(tmp_path / "c")
    g.write_data(name, iter=3, plot_type=0)
    rows_g = open(name + ".csv.000000003").read().splitlines()
    rows_o = oo.csv_text(U, cfg, m, o.operator(5), o.operator(6), o.operator(0)).splitlines()
    assert rows_g[0] == rows_o[0] and len(rows_g) == len(rows_o)
    a = np.array([[float(t) for t in r.split(",")] for r in rows_g[1:]])
    b = np.array([[float(t) for t in r.split(",")] for r in rows_o[1:]])
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_checkpoint_is_exact(tmp_path):
    cfg = euler_cfg(3, 3)
    m = hex_box(3)
    _, g, U = _pair(cfg, m)
    g.update(3)
    Ug = g.get_state()
    path = str(tmp_path / "ck.bin")
    g.write_checkpoint(path, iter=3, time=6e-4)
    h = Solver(cfg, m)
    assert h.read_checkpoint(path) == (3, 6e-4)
    assert np.array_equal(h.get_state(), Ug)
    # the restarted run continues the uninterrupted one (its first stage
    # re-extrapolates the traces, which the uninterrupted run carries over
    # from its last stage, so the two agree to rounding)
    g.update(2)
    h.update(2)
    assert _close(h.get_state(), g.get_state(), 1e-14)
    other = Solver(euler_cfg(3, 2), m)
    with pytest.raises(FrError):
        other.read_checkpoint(path)
