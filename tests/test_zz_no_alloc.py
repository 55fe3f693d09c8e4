"""fr_reduce_noalloc (the allocation-free fr_reduce) against fr_reduce through
the C ABI: every reduction kind, bit for bit, on two calls in a row -- a
larger mesh first, then a smaller one, so the second call reads partials the
first one left behind in the shared device array only if it were wrong --
in FP64 and FP32 contexts, and a residual holding a NaN."""
import numpy as np
import pytest

from fr_learn_amd import Solver, cases

from helpers import euler_cfg, hex_box, perturbed_state, quad_box

pytestmark = pytest.mark.gpu


def _reductions(g):
    """(fr_reduce, fr_reduce_noalloc) values of every kind."""
    a, b = [], []
    for kind in (1, 2, 3):
        a.append(g.norm_residual(kind))
        b.append(g.reduce_noalloc(kind))
    a.append(np.array(g.kinetic_energy()))
    b.append(g.reduce_noalloc(10))
    return a, b


def _ctx(D, n, order, precision="fp64"):
    cfg = euler_cfg(D, order, dt=1e-3)
    m = hex_box(n, perturb=0.1) if D == 3 else quad_box(n, perturb=0.1)
    g = Solver(cfg, m, precision=precision)
    U = perturbed_state(g.initialize(cases.ic_tgv if D == 3 else
                                     (lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4))), seed=1)
    g.set_state(U)
    g.calc_residual(0)
    return g


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_matches_reduce_on_two_calls(precision):
    big = _ctx(3, 12, 3, precision)      # 1728 elements: 7 blocks of partials
    small = _ctx(3, 3, 2, precision)     # 27 elements: 1 block
    print(f"no-alloc reductions {precision}: {big.nEles} then {small.nEles} elements", flush=True)
    for g in (big, small):
        a, b = _reductions(g)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (x, y)
    # and again after the small one, in the other order of sizes
    a, b = _reductions(big)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    big.close()
    small.close()


def test_noalloc_quads_and_nan():
    g = _ctx(2, 20, 4)
    a, b = _reductions(g)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    U = g.get_state()
    U[3, 5, 0] = np.nan
    g.set_state(U)
    g.calc_residual(0)
    for kind in (1, 2, 3):
        with pytest.raises(RuntimeError):
            g.reduce_noalloc(kind)
    g.close()
