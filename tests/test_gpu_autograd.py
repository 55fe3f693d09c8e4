"""The adjoint through torch.autograd (verification harness only).

The residual (fr_calc_residual with its VJP fr_vjp_residual) and the RK step
(fr_update with fr_vjp_step) are wrapped as torch.autograd.Function in a child
process (tests/autograd_client.py, CPU tensors, no libfrhip); this process
owns the library context and answers the child's forward / backward
requests.  Checked: torch.autograd.gradcheck of both functions, and the
gradient of J = KE(U_N) by torch's reverse mode through N wrapped steps
against the library's checkpointed adjoint (fr_vjp_steps) to 1e-10 and a
central difference (eps 1e-6) to 1e-6."""
import multiprocessing as mp
import os
import sys

import numpy as np
import pytest

from fr_learn_amd import Solver, cases

from helpers import euler_cfg, hex_box, perturbed_state, quad_box

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _client(conn, shape, nsteps, eps, seed):
    sys.path.insert(0, HERE)
    import autograd_client
    autograd_client.run(conn, shape, nsteps, eps, seed)


def _serve(g, U0, nsteps, seed=0):
    shape = g.state_shape()
    ctx = mp.get_context("spawn")
    parent, child = ctx.Pipe()
    proc = ctx.Process(target=_client, args=(child, shape, nsteps, 1e-6, seed))
    proc.start()
    wJ = g.geometry(1) * _weights(g)[:, None]
    result = None
    try:
        while True:
            if not parent.poll(600):
                raise TimeoutError("autograd client is silent")
            msg = parent.recv()
            op = msg[0]
            if op == "done":
                result = msg[1]
                break
            if op == "initial_state":
                parent.send(U0)
            elif op == "residual":
                g.set_state(msg[1])
                g.calc_residual(0)
                parent.send(g.residual(0))
            elif op == "vjp_residual":
                g.set_state(msg[1])
                parent.send(g.vjp_residual(msg[2]))
            elif op == "step":
                g.set_state(msg[1])
                g.update(1)
                parent.send(g.get_state())
            elif op == "vjp_step":
                g.set_state(msg[1])
                parent.send(g.vjp_step(msg[2]))
            elif op == "ke_weights":
                parent.send(wJ)
            elif op == "adjoint_ke":
                g.set_state(msg[1])
                g.update(msg[2])
                lamN = g.kinetic_energy_gradient()
                g.set_state(msg[1])
                parent.send(g.vjp_steps(msg[2], 2, lamN))
            elif op == "ke_lib":
                g.set_state(msg[1])
                g.update(msg[2])
                parent.send(g.kinetic_energy()[0])
            else:
                raise ValueError(op)
    finally:
        proc.join(60)
        if proc.is_alive():
            proc.kill()
    return result


def _weights(g):
    """The solution points' quadrature weights (tensor Gauss weights)."""
    from numpy.polynomial.legendre import leggauss
    P1 = g.cfg["order"] + 1
    w1 = leggauss(P1)[1]
    if g.nDims == 3:
        return np.einsum("k,j,i->kji", w1, w1, w1).ravel()
    return np.einsum("j,i->ji", w1, w1).ravel()


@pytest.mark.parametrize("D", [2, 3])
def test_autograd_function_matches_adjoint(D):
    if D == 2:
        cfg = euler_cfg(2, 2, riemannType=1, dt=1e-3)
        m = quad_box(3, perturb=0.1)
        ic = lambda x, c: cases.ic_vortex(x, c, beta=2.0, mach=0.4)
    else:
        cfg = euler_cfg(3, 1, dt=1e-3)
        m = hex_box(3, perturb=0.1)
        ic = cases.ic_tgv
    print(f"autograd harness D={D} P={cfg['order']} nEles={m.nEles}", flush=True)
    g = Solver(cfg, m)
    U0 = perturbed_state(g.initialize(ic), seed=3, amp=1e-2)
    r = _serve(g, U0, nsteps=3)
    print(r, flush=True)
    assert r["gradcheck_residual"] and r["gradcheck_step"]
    assert abs(r["ke"] - r["ke_lib"]) <= 1e-12 * abs(r["ke_lib"])
    assert r["grad_rel_diff_torch_vs_lib"] <= 1e-10
    assert abs(r["directional_torch"] - r["directional_fd"]) <= 1e-6 * abs(r["directional_fd"])
    g.close()
