"""Benchmark: residual GDOF/s of the 3D Euler Taylor-Green workload on one MI355X.

Workload: periodic [0, 2pi]^3, 48^3 linear hexes, P = 3 (64 solution points /
96 flux points per element), Ma = 0.1, rho0 = 1, Rusanov flux, classic RK4
(the reference's timeType 4, "RK45"), dt = 2e-4, synthetic TGV initial state.
A step is one RK4 step = 4 residual evaluations + stage updates;
value = nEles * nSpts * 4 * steps / seconds / 1e9 (DOF counts solution points).

Prints ONE JSON line.  Besides the contract keys it carries:
  roofline      the fused residual kernel's algorithmic bytes per launch over
                its average launch time (HIP events on the launch stream),
                against MI355X's 8 TB/s;
  cpu_baseline  the CPU oracle (same algorithm, dense operators like the
                reference's cblas_dgemm calls) on about 10 s of RK steps of
                the same workload, on the host's cores;
  fp32          the FP32 residual path (FP64 state / RK / reductions) and its
                relative L2 distance from the FP64 path after the same steps
                (the FP64 path equals the oracle to 1e-12, tests/).
  copy_GBps     a device-to-device copy of 1 GiB, read + write bytes over its
                time: the streaming bandwidth this GPU reaches in practice.

--copy-calibration runs only a 256 MiB copy, three times, for scaling the
rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "residual GDOF/s, 3D Euler hex P3, FP64 vs FP32-MFMA; % HBM roofline"
HBM_PEAK = 8.0e12


def tgv_setup(n, order=3, dt=2e-4, perturb=0.0):
    """The TGV workload's config and mesh; perturb > 0 displaces the interior
    vertices randomly (a fraction of the cell size) so that no element is
    affine."""
    from fr_learn_amd import cases, mesh
    cfg = cases.config(nDims=3, order=order, equation=1, riemannType=0, timeType=4, dt=dt, gamma=1.4)
    m = mesh.box(3, (n, n, n), (0, 0, 0), (2 * np.pi,) * 3, perturb=perturb, seed=11)
    return cfg, m


GEOMETRY = {"point": 0, "compact": 1, "nodal": 2, "nodal-spt": 3}


def run_gpu(cfg, m, steps, warmup, precision, geometry="point", keep_final=False):
    """The TGV workload's RK4 steps with the stage kernels' geometry layout
    `geometry` (GEOMETRY): per-point metrics and face normals read from HBM,
    the affine mesh's per-element / per-face values, or the metrics and
    normals formed in the kernel from each element's corner nodes.
    keep_final: the final stage also stores its divergence (divF_spts[3]),
    which nothing in the step reads (fr_keep_final_residual)."""
    from fr_learn_amd import Solver, cases
    s = Solver(cfg, m, precision=precision)
    assert s.geometry_mode(GEOMETRY[geometry]) == GEOMETRY[geometry]
    s.keep_final_residual(keep_final)
    s.initialize(cases.ic_tgv)
    s.update(warmup)
    s.synchronize()
    s.set_timing(True)
    s.kernel_times(0)
    s.kernel_times(1)
    s.kernel_times(2)
    t0 = time.perf_counter()
    s.update(steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    res_ms, res_n = s.kernel_times(0)
    tr_ms, tr_n = s.kernel_times(1)
    rk_ms, rk_n = s.kernel_times(2)
    state = s.get_state()
    # residual evaluations alone (solver::calcResidual): the stage kernel
    # without the RK update and trace output
    for _ in range(3):
        s.calc_residual(0)
    s.synchronize()
    s.kernel_times(0)
    for _ in range(20):
        s.calc_residual(0)
    s.synchronize()
    r0_ms, r0_n = s.kernel_times(0)
    s.set_timing(False)
    out = dict(wall=wall, res_ms=res_ms / max(res_n, 1), res_n=res_n, trace_ms=tr_ms / max(tr_n, 1),
               rk_ms=rk_ms / max(rk_n, 1), bytes=s.residual_bytes(), nEles=s.nEles, nSpts=s.nSpts,
               state=state, res_only_ms=r0_ms / max(r0_n, 1), res_only_n=r0_n)
    s.close()
    return out


def run_cpu_baseline(cfg, m, budget_s=10.0):
    """RK4 steps of the same workload through the CPU oracle, as many as fit
    in about budget_s seconds (at least 2)."""
    from fr_learn_amd import cases
    from oracle import Oracle
    threads = min(16, os.cpu_count() or 1)
    o = Oracle(cfg, m, threads=threads)
    o.initialize(cases.ic_tgv)
    nsteps = 0
    t0 = time.perf_counter()
    while nsteps < 2 or time.perf_counter() - t0 < budget_s:
        o.update(1)
        nsteps += 1
    dt = time.perf_counter() - t0
    gdof = o.nEles * o.nSpts * 4 * nsteps / dt / 1e9
    # one step on one core, continuing from the same state
    o._lib.or_set_threads(1)
    t1 = time.perf_counter()
    o.update(1)
    dt1 = time.perf_counter() - t1
    o.close()
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=gdof, unit="GDOF/s", cores=threads, kind="port",
                sample=f"{nsteps} RK4 steps ({4 * nsteps} residuals) of the same 48^3 P3 TGV workload, "
                       f"{dt:.1f} s, dense-operator oracle with its own blocked dgemm (no OpenBLAS: the "
                       f"image has no CBLAS), {threads} OpenMP threads",
                single_core={"value": o.nEles * o.nSpts * 4 / dt1 / 1e9, "unit": "GDOF/s",
                             "sample": f"1 RK4 step of the same workload, {dt1:.1f} s, 1 thread"},
                host={"cpu": model, "logical_cpus": os.cpu_count()})


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Read + write bytes per second of torch's device-to-device copy.  Runs in
    a child process of its own (--copy-bandwidth): torch brings its own HIP
    runtime, which must not share a process with the one the library uses."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3) / 1e9


def copy_bandwidth_child():
    import subprocess
    try:
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-bandwidth"], capture_output=True,
                             text=True, timeout=300, check=True).stdout
        return json.loads(out.strip().splitlines()[-1])["copy_GBps"]
    except (subprocess.SubprocessError, ValueError, KeyError, IndexError) as exc:
        print(f"copy bandwidth not measured: {exc}", file=sys.stderr)
        return None


def copy_calibration():
    """A 256 MiB device copy, three times: the known-size kernel the PMC
    FETCH_SIZE / WRITE_SIZE readings are scaled by."""
    import torch
    n = (256 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}), flush=True)


def batch_members(n_inst=256, seed=0):
    """The instances of the batched 2D Euler workload: n_inst seeded periodic
    meshes of 8^2 to 64^2 cells, quadrilaterals or triangles (two per cell)
    with probability 1/2, P uniform in 1..5, isentropic vortices of random
    centre, strength 1..5 and Mach 0.2..0.6, Roe, RK4 (dt 1e-3).  A list of
    (config, mesh, initial-condition function, element type "quad" | "tri")."""
    from fr_learn_amd import cases, mesh
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_inst):
        cells = int(rng.integers(8, 65))
        P = int(rng.integers(1, 6))
        cx, cy = rng.uniform(-2.5, 2.5, 2)
        beta, mach = rng.uniform(1.0, 5.0), rng.uniform(0.2, 0.6)
        cfg = cases.config(nDims=2, order=P, equation=1, riemannType=1, timeType=4, dt=1e-3, gamma=1.4)
        if rng.uniform() < 0.5:
            m, kind = mesh.tri_box((cells, cells), (-5.0, -5.0), (5.0, 5.0)), "tri"
        else:
            m, kind = mesh.box(2, (cells, cells), (-5.0, -5.0), (5.0, 5.0)), "quad"
        ic = (lambda b, mc, x0, y0: lambda x, c: cases.ic_vortex(x, c, beta=b, mach=mc, center=(x0, y0)))(
            beta, mach, cx, cy)
        out.append((cfg, m, ic, kind))
    return out


def run_batch(n_inst=256, steps=20, warmup=3, seed=0):
    """The batched 2D Euler workload (batch_members): one persistent launch per
    stage, element type and order for all meshes (replayed from a hipGraph)
    against one update call per mesh."""
    from fr_learn_amd import Batch, Solver
    solvers = []
    dof = 0
    b_res = 0.0
    for cfg, m, ic, _ in batch_members(n_inst, seed):
        s = Solver(cfg, m)  # default geometry: compact (every member is affine)
        s.initialize(ic)
        solvers.append(s)
        dof += s.nEles * s.nSpts
        b_res += s.residual_bytes()
    b = Batch(solvers)
    grid, items, _ = b.tuning()
    tiles = {f"{k} P{P}": v for (k, P), v in sorted(b.group_tuning().items())}
    b.update(warmup)
    t0 = time.perf_counter()
    b.update(steps)
    wall_b = time.perf_counter() - t0
    for s in solvers:
        s.update(warmup)
    for s in solvers:
        s.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in solvers:
            s.update(1)
    for s in solvers:
        s.synchronize()
    wall_s = time.perf_counter() - t0
    b.close()
    for s in solvers:
        s.close()
    v_b = dof * 4 * steps / wall_b / 1e9
    v_s = dof * 4 * steps / wall_s / 1e9
    # batched roofline: the members' algorithmic bytes of one residual (B_res,
    # as for the main workload) per RK stage, over the measured stage time
    stage_s = wall_b / steps / 4
    achieved = b_res / stage_s / 1e9
    return {"workload": f"{n_inst} periodic meshes 8^2..64^2 cells, quads or triangles (p = 0.5), P 1..5, "
                        "2D Euler vortices, Roe, RK4",
            "instances": n_inst, "dof_per_field": dof, "steps": steps, "value": v_b, "unit": "GDOF/s",
            "ms_per_step": wall_b / steps * 1e3, "per_mesh_launches_GDOF_per_s": v_s,
            "speedup_vs_per_mesh_launches": v_b / v_s, "largest_group_grid": grid, "work_items": items,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": achieved / (HBM_PEAK / 1e9), "bytes_per_stage": b_res},
            "tiles": tiles}


def tgv_ns_adjoint(n=24, steps=200, every=20, eps=1e-6, seed=0):
    """The adjoint workload: 3D Navier-Stokes Taylor-Green (Re = 1600, Pr = 0.71,
    constant viscosity, LDG), n^3 hexes, P = 3, RK4, dt = 1e-3.  J = kinetic
    energy after `steps` steps; its gradient with respect to the initial
    velocity field by the checkpointed adjoint (a state every `every` steps),
    checked against a central finite difference (eps) along a random
    direction."""
    from fr_learn_amd import Solver, cases, mesh
    cfg = cases.config(nDims=3, order=3, equation=1, viscous=1, mu_inf=1.0 / 1600, prandtl=0.71, fixVis=1,
                       penFact=0.0, tau=1.0, timeType=4, dt=1e-3, gamma=1.4)
    m = mesh.box(3, (n, n, n), (0, 0, 0), (2 * np.pi,) * 3)
    s = Solver(cfg, m)
    U0 = s.initialize(cases.ic_tgv)
    rho, vel = U0[..., 0], U0[..., 1:4] / U0[..., :1]
    p = (cfg["gamma"] - 1) * (U0[..., 4] - 0.5 * rho * (vel ** 2).sum(-1))

    def state(v):  # U0 from the velocity field at fixed rho and p
        return cases.conservative(rho, [v[..., i] for i in range(3)], p, cfg["gamma"])

    def J(v):
        s.set_state(state(v))
        s.update(steps)
        return s.kinetic_energy()[0]

    s.set_state(U0)
    s.set_timing(True)
    s.kernel_times(0)
    s.kernel_times(3)
    t0 = time.perf_counter()
    s.update(steps)
    s.synchronize()
    t_fwd = time.perf_counter() - t0
    st_ms, st_n = s.kernel_times(0)
    gr_ms, gr_n = s.kernel_times(3)
    s.set_timing(False)
    # viscous roofline: B_res plus the gradient traffic E*Ns*2*D*Nf*s of a
    # viscous residual (the gradient written and read once per point)
    b_vis = s.residual_bytes() + s.nEles * s.nSpts * 2 * 3 * 5 * 8
    stage_ms = (st_ms + gr_ms) / max(st_n, 1)
    lamN = s.kinetic_energy_gradient()
    s.set_state(U0)
    t0 = time.perf_counter()
    lam0 = s.vjp_steps(steps, every, lamN)
    t_adj = time.perf_counter() - t0
    grad_v = rho[..., None] * lam0[..., 1:4] + (rho[..., None] * vel) * lam0[..., 4:5]
    dv = np.random.default_rng(seed).standard_normal(vel.shape)
    fd = (J(vel + eps * dv) - J(vel - eps * dv)) / (2 * eps)
    adj = float(np.sum(grad_v * dv))
    s.close()
    return {"metric": "adjoint gradient of KE(T) w.r.t. the initial velocity, 3D NS TGV",
            "workload": f"{n}^3 hexes P3, Navier-Stokes Re 1600 Pr 0.71 (LDG), RK4 dt 1e-3, {steps} steps, "
                        f"checkpoint every {every}",
            "forward_s": t_fwd, "adjoint_s": t_adj, "adjoint_over_forward": t_adj / t_fwd,
            "viscous_stage_roofline": {
                "bound": "hbm", "algorithmic_bytes_per_stage": b_vis, "stage_ms": stage_ms,
                "gradient_and_face_flux_ms": gr_ms / max(gr_n, 1), "stage_kernel_ms": st_ms / max(st_n, 1),
                "achieved": b_vis / (stage_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": b_vis / (stage_ms * 1e-3) / HBM_PEAK,
                "note": "k_grad + k_vflux + k_stage<...,VIS> per RK stage (HIP events); bytes = B_res + "
                        "E*Ns*2*D*Nf*s"},
            "fd_directional": fd, "adjoint_directional": adj, "rel_diff_vs_fd": abs(fd - adj) / abs(fd),
            "eps": eps}


def csrc_digest():
    """sha256 over the library's kernel and host sources (name and bytes of every
    .hip / .hpp / .cpp under fr_learn_amd/csrc, in name order)."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "fr_learn_amd", "csrc")
    for fn in sorted(os.listdir(csrc)):
        if fn.endswith((".hip", ".hpp", ".cpp")):
            h.update(fn.encode())
            with open(os.path.join(csrc, fn), "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()


def traffic_from_profile(precision, geometry):
    """HBM bytes per RK stage launch of the stage kernel from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    factor), if there is one and it was measured on the current sources (its
    csrc_sha256 equals csrc_digest()); None otherwise, so that a kernel change
    never reports the traffic of the kernel before it."""
    path = os.path.join(ROOT, "profiles", f"k_stage_tgv48_p3_{precision}_{geometry}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        js = json.load(fh)
    if js.get("csrc_sha256") != csrc_digest():
        return None
    return js.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=48, help="cells per direction")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 path")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-mesh workload")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--geometry", choices=list(GEOMETRY), default="nodal",
                    help="stage-kernel geometry of the headline run (default: metrics and normals formed in the "
                         "kernel from each hex's corner nodes, valid for any linear hex)")
    ap.add_argument("--lib", default=None,
                    help="load this build of libfrhip.so instead of the package's (e.g. one built by "
                         "`python -m fr_learn_amd.build --lib PATH -D FR_MFMA=0`)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the extra FP64 runs with the other geometry layouts and the perturbed mesh")
    ap.add_argument("--copy-calibration", action="store_true", help="only the 256 MiB copy for PMC scaling")
    ap.add_argument("--copy-bandwidth", action="store_true", help="only measure the 1 GiB copy bandwidth")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched 256-mesh workload (--steps / --warmup steps), its own JSON line")
    ap.add_argument("--adjoint", action="store_true",
                    help="only the Navier-Stokes TGV adjoint workload (24^3 P3, 200 steps), its own JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the library has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    if args.lib:
        import fr_learn_amd._lib as frlib
        frlib.LIB_PATH = os.path.abspath(args.lib)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.copy_bandwidth:
        print(json.dumps({"copy_GBps": copy_bandwidth()}), flush=True)
        return
    if args.adjoint:
        print(json.dumps(tgv_ns_adjoint()), flush=True)
        return
    if args.batch_only:
        print(json.dumps(run_batch(steps=args.steps, warmup=args.warmup)), flush=True)
        return

    copy_gbps = copy_bandwidth_child()  # before this process touches the GPU
    cfg, m = tgv_setup(args.n)
    main_prec = args.precision
    r = run_gpu(cfg, m, args.steps, args.warmup, main_prec, geometry=args.geometry)
    dof = r["nEles"] * r["nSpts"]
    value = dof * 4 * args.steps / r["wall"] / 1e9
    achieved = r["bytes"] / (r["res_ms"] * 1e-3)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32",
        "data": "synthetic (Taylor-Green initial state, generated mesh)",
        "config": {
            "workload": "3D Euler Taylor-Green vortex, periodic [0,2pi]^3, 48^3 linear hexes, P=3, Ma=0.1, "
                        "Rusanov, RK4 (timeType 4), dt=2e-4",
            "cells": [args.n] * 3, "order": 3, "nEles": r["nEles"], "dof_per_field": dof, "fields": 5,
            "precision": main_prec, "geometry": {
                "nodal": "metrics at the solution points and normals at the flux points formed in the kernel "
                         "from each element's 8 corner nodes (any linear hex; no affine assumption)",
                "point": "per-point metrics and face normals read from HBM",
                "compact": "one metric set per element and one normal per face (affine meshes only)",
                "nodal-spt": "solution-point metrics formed in the kernel, face normals read from HBM"}[args.geometry],
            "final_stage_divF_stored": False,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic_from_profile(main_prec, args.geometry),
            "kernel": "k_stage<3,3,5,...> (residual + RK stage update + next-stage face traces; "
                      "average over the RK4 stages, the final stage's variant included)",
            "algorithmic_bytes_per_launch": r["bytes"],
            "avg_launch_ms": r["res_ms"], "launches": r["res_n"],
            "copy_GBps": copy_gbps,
        },
        "kernels_ms": {"stage": r["res_ms"], "trace_init": r["trace_ms"]},
        "residual_only": {
            "note": "calcResidual launches alone (no RK update, no trace output), same kernel, mode 0",
            "avg_launch_ms": r["res_only_ms"], "launches": r["res_only_n"],
            "achieved": r["bytes"] / (r["res_only_ms"] * 1e-3) / 1e9, "unit": "GB/s",
            "frac": r["bytes"] / (r["res_only_ms"] * 1e-3) / HBM_PEAK,
            "gdof_per_s": dof / (r["res_only_ms"] * 1e-3) / 1e9,
        },
    }
    if not args.no_fp32 and main_prec == "fp64":
        r32 = run_gpu(cfg, m, args.steps, args.warmup, "fp32", geometry=args.geometry)
        v32 = dof * 4 * args.steps / r32["wall"] / 1e9
        a, b = r32["state"], r["state"]
        l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        l2f = [float(np.linalg.norm(a[..., k] - b[..., k]) / np.linalg.norm(b[..., k])) for k in range(5)]
        # the TGV's rho w is small (w = 0 initially): its error measured
        # against the momentum scale ||rho u||, as the tests do
        l2w = float(np.linalg.norm(a[..., 3] - b[..., 3]) / np.linalg.norm(b[..., 1]))
        a32 = r32["bytes"] / (r32["res_ms"] * 1e-3)
        line["fp32"] = {
            "value": v32, "unit": "GDOF/s", "ms_per_step": r32["wall"] / args.steps * 1e3,
            "speedup_vs_fp64": v32 / value, "rel_l2_vs_fp64": l2, "rel_l2_per_field_vs_fp64": l2f,
            "rel_l2_rho_w_vs_norm_rho_u": l2w,
            "roofline": {"bound": "hbm", "achieved": a32 / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": a32 / HBM_PEAK, "avg_launch_ms": r32["res_ms"],
                         "algorithmic_bytes_per_launch": r32["bytes"],
                         "traffic": traffic_from_profile("fp32", args.geometry)},
            "kernels_ms": {"stage": r32["res_ms"]},
        }
    if not args.no_variants and main_prec == "fp64":
        # the same FP64 workload with the other geometry layouts, and the
        # headline layout on a mesh whose interior vertices are displaced by
        # up to 10 % of the cell size (no element affine)
        var = {}
        for geo in [g for g in ("point", "compact") if g != args.geometry]:
            rv = run_gpu(cfg, m, args.steps, args.warmup, "fp64", geometry=geo)
            var[geo] = {"value": dof * 4 * args.steps / rv["wall"] / 1e9, "unit": "GDOF/s",
                        "avg_launch_ms": rv["res_ms"], "frac": rv["bytes"] / (rv["res_ms"] * 1e-3) / HBM_PEAK,
                        "max_abs_state_diff_vs_headline": float(np.abs(rv["state"] - r["state"]).max())}
        rk = run_gpu(cfg, m, args.steps, args.warmup, "fp64", geometry=args.geometry, keep_final=True)
        var["final_stage_divF_stored"] = {
            "value": dof * 4 * args.steps / rk["wall"] / 1e9, "unit": "GDOF/s", "avg_launch_ms": rk["res_ms"],
            "frac": rk["bytes"] / (rk["res_ms"] * 1e-3) / HBM_PEAK,
            "state_equal_to_headline": bool(np.array_equal(rk["state"], r["state"])),
            "note": "the headline run with the final stage's divF_spts[3] written, as solver::update keeps it"}
        cfgp, mp = tgv_setup(args.n, perturb=0.1)
        rp = run_gpu(cfgp, mp, args.steps, args.warmup, "fp64", geometry=args.geometry)
        var[f"{args.geometry}_perturbed_mesh"] = {
            "value": dof * 4 * args.steps / rp["wall"] / 1e9, "unit": "GDOF/s", "avg_launch_ms": rp["res_ms"],
            "frac": rp["bytes"] / (rp["res_ms"] * 1e-3) / HBM_PEAK,
            "mesh": "interior vertices displaced by up to 0.1 cell sizes (seed 11)"}
        line["geometry_variants"] = var
    if not args.no_batch:
        line["batched"] = run_batch()
    if not args.no_cpu:
        line["cpu_baseline"] = run_cpu_baseline(cfg, m)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
