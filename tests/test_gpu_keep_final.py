"""fr_keep_final_residual: the final RK stage stores divF_spts[nRK-1] by
default, as solver::update leaves it (solver.cpp:212-250); switched off, the
update leaves it out and gives the same state bit for bit, the residual norms
(which read divF_spts[0]) are unchanged, and reading the dropped array fails
until a residual of that stage is computed again."""
import numpy as np
import pytest

from fr_learn_amd import Batch, Solver, cases
from fr_learn_amd._lib import FrError
from oracle import Oracle

from helpers import TOL_FP64, euler_cfg, hex_box, perturbed_state, quad_box, rel_err

pytestmark = pytest.mark.gpu


def _pair(cfg, m, ic):
    a, b = Solver(cfg, m), Solver(cfg, m)
    U = perturbed_state(a.initialize(ic), seed=5)
    a.set_state(U)
    b.set_state(U)
    return a, b, U


@pytest.mark.parametrize("order", [2, 3])
def test_kept_by_default_and_dropped_on_request(order):
    cfg = euler_cfg(3, order, dt=1e-3)
    m = hex_box(3)
    a, b, U = _pair(cfg, m, cases.ic_tgv)
    o = Oracle(cfg, m)
    o.set_state(U)
    o.update(2)
    b.keep_final_residual(False)
    a.update(2)
    b.update(2)
    assert np.array_equal(a.get_state(), b.get_state())
    assert rel_err(a.residual(3), o.residual(3)).max() <= TOL_FP64
    assert np.array_equal(a.residual(0), b.residual(0))
    assert np.array_equal(a.norm_residual(1), b.norm_residual(1))
    with pytest.raises(FrError):
        b.residual(3)
    b.calc_residual(3)
    a.calc_residual(3)
    assert np.array_equal(a.residual(3), b.residual(3))
    b.keep_final_residual(True)
    b.update(1)
    a.update(1)
    assert np.array_equal(a.residual(3), b.residual(3))
    for s in (a, b, o):
        s.close()


def test_forward_euler_keeps_its_only_residual():
    cfg = euler_cfg(2, 3, dt=1e-3, timeType=0)
    a, b, _ = _pair(cfg, quad_box(4), lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4))
    b.keep_final_residual(False)
    a.update(3)
    b.update(3)
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(a.residual(0), b.residual(0))
    a.close()
    b.close()


def test_batched_members():
    cfg = euler_cfg(2, 3, dt=1e-3)
    pairs = [_pair(cfg, quad_box(n, seed=n), lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4))
             for n in (3, 5)]
    for a, b, _ in pairs:
        b.keep_final_residual(False)
        a.update(2)
    B = Batch([b for _, b, _ in pairs])
    B.update(2)
    for a, b, _ in pairs:
        assert rel_err(b.get_state(), a.get_state(), axis=None) <= 1e-14
        with pytest.raises(FrError):
            b.residual(3)
        a.close()
        b.close()
