"""torch side of the autograd verification harness (tests/test_gpu_autograd.py).

Runs in a child process that imports torch (CPU tensors only) and never loads
libfrhip: torch brings its own HIP runtime, which must not share a process
with the one the library uses.  The residual and the RK step are wrapped as
torch.autograd.Function; their forward and backward passes are requests to
the parent process, which owns the library context and answers them with the
C ABI (fr_calc_residual / fr_vjp_residual, fr_update / fr_vjp_step).

The checks (results sent back to the parent, which asserts them):
  gradcheck of the residual and of one RK4 step (torch's finite differences
  against the library's VJP through backward());
  J = kinetic energy after N steps, differentiated by torch's reverse mode
  through N wrapped steps, against the library's checkpointed adjoint
  (fr_vjp_steps) and a central difference of J.
"""
import numpy as np
import torch


def run(conn, shape, nsteps, eps, seed):
    def call(*msg):
        conn.send(msg)
        return conn.recv()

    class Residual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, U):
            ctx.save_for_backward(U)
            return torch.from_numpy(call("residual", U.detach().numpy().copy()))

        @staticmethod
        def backward(ctx, g):
            (U,) = ctx.saved_tensors
            return torch.from_numpy(call("vjp_residual", U.detach().numpy().copy(), g.detach().numpy().copy()))

    class Step(torch.autograd.Function):
        @staticmethod
        def forward(ctx, U):
            ctx.save_for_backward(U)
            return torch.from_numpy(call("step", U.detach().numpy().copy()))

        @staticmethod
        def backward(ctx, g):
            (U,) = ctx.saved_tensors
            return torch.from_numpy(call("vjp_step", U.detach().numpy().copy(), g.detach().numpy().copy()))

    out = {}
    U0 = torch.from_numpy(call("initial_state")).requires_grad_(True)
    out["gradcheck_residual"] = bool(torch.autograd.gradcheck(Residual.apply, (U0,), eps=1e-6, atol=1e-6,
                                                              rtol=1e-4, raise_exception=False))
    out["gradcheck_step"] = bool(torch.autograd.gradcheck(Step.apply, (U0,), eps=1e-6, atol=1e-6, rtol=1e-4,
                                                          raise_exception=False))
    # J = KE(U_N): reverse mode through N wrapped steps
    wJ = torch.from_numpy(call("ke_weights"))  # quadrature weight * detJ per [spt][ele]
    D = shape[-1] - 2

    def ke(U):
        rho = U[..., 0]
        m2 = (U[..., 1:1 + D] ** 2).sum(-1)
        return (wJ * 0.5 * m2 / rho).sum()

    U = U0
    for _ in range(nsteps):
        U = Step.apply(U)
    J = ke(U)
    J.backward()
    g_torch = U0.grad.detach().numpy().copy()
    g_lib = call("adjoint_ke", U0.detach().numpy().copy(), nsteps)
    dv = np.random.default_rng(seed).standard_normal(shape)

    def J_of(V):
        W = torch.from_numpy(V)
        for _ in range(nsteps):
            W = torch.from_numpy(call("step", W.numpy().copy()))
        return float(ke(W))

    base = U0.detach().numpy()
    fd = (J_of(base + eps * dv) - J_of(base - eps * dv)) / (2 * eps)
    out["ke"] = float(J)
    out["ke_lib"] = float(call("ke_lib", U0.detach().numpy().copy(), nsteps))
    out["grad_rel_diff_torch_vs_lib"] = float(np.abs(g_torch - g_lib).max() / np.abs(g_lib).max())
    out["directional_torch"] = float((g_torch * dv).sum())
    out["directional_fd"] = fd
    conn.send(("done", out))
