"""Python mirror of the reference `solver` verbs over the libfrhip C ABI.

Method names follow reference include/solver.hpp (setup, initializeSolution,
calcResidual, update, timeStepA/B, ...); arrays use the reference layout
U(spt, ele, field) (src/matrix.cpp:260-272).  Every call goes to the HIP
library; there is no CPU path.
"""
import ctypes

import numpy as np

from . import _lib
from .cases import n_fields


class Solver:
    def __init__(self, cfg, mesh, precision="fp64"):
        lib = _lib.load()
        self.cfg = dict(cfg)
        self.mesh = mesh
        c = _lib.FrConfig()
        for name, _ in _lib.FrConfig._fields_:
            setattr(c, name, self.cfg[name])
        self._c = c
        bcs, off, pts = mesh.bnd_arrays()
        self._keep = (mesh.xv, mesh.c2v, mesh.vmap, bcs, off, pts)
        m = _lib.FrMesh()
        m.nVerts = mesh.xv.shape[0]
        m.xv = _lib.dptr(mesh.xv)
        m.nEles = mesh.nEles
        m.nNodes = mesh.nNodes
        m.c2v = _lib.iptr(mesh.c2v)
        m.vmap = _lib.iptr(mesh.vmap) if mesh.vmap is not None else ctypes.POINTER(ctypes.c_int)()
        m.nBounds = len(mesh.bounds)
        m.bcType = _lib.iptr(bcs)
        m.bndOff = _lib.iptr(off)
        m.bndPts = _lib.iptr(pts)
        self._m = m
        prec = {"fp64": _lib.FR_PREC_FP64, "fp32": _lib.FR_PREC_FP32}[precision]
        self.precision = precision
        h = ctypes.c_void_p()
        _lib.check(lib.fr_create(ctypes.byref(c), ctypes.byref(m), prec, ctypes.byref(h)))
        self._h = h
        self._lib = lib
        sz = np.zeros(8, dtype=np.int32)
        _lib.check(lib.fr_get_sizes(h, _lib.iptr(sz)))
        (self.nSpts, self.nFpts, self.nEles, self.nFields, self.nDims, self.nFaces, self.nIntFaces,
         self.nRKSteps) = [int(v) for v in sz]
        assert self.nFields == n_fields(self.cfg)

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.fr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ----------------------------------------------------------------
    def state_shape(self):
        return (self.nSpts, self.nEles, self.nFields)

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.state_shape())
        _lib.check(self._lib.fr_set_state(self._h, _lib.dptr(U), _lib.FR_LAYOUT_REF))

    def get_state(self):
        out = np.zeros(self.state_shape())
        _lib.check(self._lib.fr_get_state(self._h, _lib.dptr(out), _lib.FR_LAYOUT_REF))
        return out

    def positions(self):
        out = np.zeros((self.nSpts, self.nEles, self.nDims))
        _lib.check(self._lib.fr_get_geometry(self._h, 4, _lib.dptr(out)))
        return out

    def initialize(self, ic, **kw):
        """solver::initializeSolution with an initial-condition function of the
        solution-point positions (cases.ic_*)."""
        x = self.positions()
        if self.nDims == 2:
            x = np.concatenate([x, np.zeros(x.shape[:-1] + (1,))], axis=-1)
        U = ic(x, self.cfg, **kw)
        self.set_state(U)
        return U

    # -- verbs ----------------------------------------------------------------
    def calc_residual(self, step=0):
        _lib.check(self._lib.fr_calc_residual(self._h, step))

    def residual(self, step=0):
        out = np.zeros(self.state_shape())
        _lib.check(self._lib.fr_get_residual(self._h, step, _lib.dptr(out), _lib.FR_LAYOUT_REF))
        return out

    def traces(self):
        out = np.zeros((self.nFpts, self.nEles, self.nFields))
        _lib.check(self._lib.fr_get_traces(self._h, _lib.dptr(out)))
        return out

    # the residual's intermediate arrays (fr_get_array), with their shapes
    ARRAYS = {"F_spts": 2, "Fn_fpts": 3, "disFn_fpts": 4, "dU_spts": 5, "dU_fpts": 6}

    def array(self, which):
        """F_spts (2), Fn_fpts (3), disFn_fpts (4), dU_spts (5), dU_fpts (6)
        of a residual evaluation of the current state, in the reference
        layout (a debug path: see fr_get_array); `which` is the number or
        the name."""
        which = self.ARRAYS.get(which, which)
        D, nS, nF, nE, K = self.nDims, self.nSpts, self.nFpts, self.nEles, self.nFields
        shape = {2: (D, nS, nE, K), 3: (nF, nE, K), 4: (nF, nE, K), 5: (D, nS, nE, K), 6: (D, nF, nE, K)}
        if which not in shape:
            raise ValueError(f"no array {which}")
        out = np.zeros(shape[which])
        _lib.check(self._lib.fr_get_array(self._h, int(which), _lib.dptr(out)))
        return out

    def keep_final_residual(self, keep=True):
        """Whether update() stores the final stage's divF_spts (fr_keep_final_residual)."""
        _lib.check(self._lib.fr_keep_final_residual(self._h, int(bool(keep))))

    def update(self, n=1):
        _lib.check(self._lib.fr_update(self._h, n))

    def time_step_a(self, step, rkval):
        _lib.check(self._lib.fr_time_step_a(self._h, step, rkval))

    def time_step_b(self, step):
        _lib.check(self._lib.fr_time_step_b(self._h, step))

    def copy_u_u0(self):
        _lib.check(self._lib.fr_copy_u_u0(self._h))

    def copy_u0_u(self):
        _lib.check(self._lib.fr_copy_u0_u(self._h))

    def synchronize(self):
        _lib.check(self._lib.fr_synchronize(self._h))

    # -- diagnostics ----------------------------------------------------------
    def norm_residual(self, kind=2):
        out = np.zeros(self.nFields)
        _lib.check(self._lib.fr_reduce(self._h, kind, _lib.dptr(out)))
        return out

    def integrate_error(self, testCase=1, errorNorm=1, quadOrder=10, icType=0, time=0.0, rhoIC=1.0, pIC=1.0,
                        vortexAngle=0.4636476090008061, box=(-5.0, 5.0, -5.0, 5.0)):
        """Error norms of the current state (fr_integrate_error; the reference's
        solver::integrateError).  Defaults are the reference's (errorNorm 1,
        quadOrder 10, vortexAngle atan(1/2), the standard box [-5, 5]^2)."""
        p = _lib.FrErrorParams(int(testCase), int(errorNorm), int(quadOrder), int(icType), float(time),
                               float(rhoIC), float(pIC), float(vortexAngle), *[float(b) for b in box])
        out = np.zeros(self.nFields)
        _lib.check(self._lib.fr_integrate_error(self._h, ctypes.byref(p), _lib.dptr(out)))
        return out

    def reduce_noalloc(self, kind):
        """fr_reduce_noalloc: fr_reduce's values (kind 1-3 residual norms of
        divF_spts[0], 10 kinetic energy and volume) with no device allocation."""
        out = np.zeros(max(self.nFields, 2))
        _lib.check(self._lib.fr_reduce_noalloc(self._h, int(kind), _lib.dptr(out)))
        return out[:2] if kind == 10 else out[:self.nFields]

    def kinetic_energy(self):
        out = np.zeros(2)
        _lib.check(self._lib.fr_reduce(self._h, 10, _lib.dptr(out)))
        return out[0], out[1]

    def operator(self, which):
        D, nS, nF = self.nDims, self.nSpts, self.nFpts
        shape = {0: (nF, nS), 1: (D, nS, nS), 2: (D, nF, nS), 3: (nS, nF), 4: (D, nS, nF)}[which]
        out = np.zeros(shape)
        _lib.check(self._lib.fr_get_operator(self._h, which, _lib.dptr(out)))
        return out

    def geometry(self, which):
        D, nS, nF, nE = self.nDims, self.nSpts, self.nFpts, self.nEles
        shape = {0: (D, nS, nE, D), 1: (nS, nE), 2: (nF, nE, D), 3: (nF, nE), 4: (nS, nE, D)}[which]
        out = np.zeros(shape)
        _lib.check(self._lib.fr_get_geometry(self._h, which, _lib.dptr(out)))
        return out

    def faces(self):
        out = np.zeros((self.nFaces, 6), dtype=np.int32)
        _lib.check(self._lib.fr_get_faces(self._h, _lib.iptr(out)))
        return out

    def set_dt(self, dt):
        """Fixed time step params->dt (dtType 0)."""
        _lib.check(self._lib.fr_set_dt(self._h, float(dt)))
        self.cfg["dt"] = float(dt)

    def calc_dt(self, dtType, CFL):
        """CFL time step from the current state (solver::calcDt); dtType 1
        global, 2 local per element.  Returns the minimum over elements."""
        out = ctypes.c_double()
        _lib.check(self._lib.fr_calc_dt(self._h, int(dtType), float(CFL), ctypes.byref(out)))
        self.cfg["dt"] = out.value
        return out.value

    def dt(self):
        """The time step each element advances with (mesh element order)."""
        out = np.zeros(self.nEles)
        _lib.check(self._lib.fr_get_dt(self._h, _lib.dptr(out)))
        return out

    def vjp_residual(self, lam):
        """(d divF / dU)^T lam at the current state (layout [spt][ele][field])."""
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(self.state_shape())
        out = np.zeros(self.state_shape())
        _lib.check(self._lib.fr_vjp_residual(self._h, _lib.dptr(lam), _lib.dptr(out)))
        return out

    def vjp_step(self, lam1):
        """(d U_{n+1} / d U_n)^T lam1 for one update from the current state."""
        lam1 = np.ascontiguousarray(lam1, dtype=np.float64).reshape(self.state_shape())
        out = np.zeros(self.state_shape())
        _lib.check(self._lib.fr_vjp_step(self._h, _lib.dptr(lam1), _lib.dptr(out)))
        return out

    def vjp_steps(self, n, every, lamN):
        """Advance n updates (the state ends at U_n) and return
        (d U_n / d U_0)^T lamN, with a checkpoint every `every` steps."""
        lamN = np.ascontiguousarray(lamN, dtype=np.float64).reshape(self.state_shape())
        out = np.zeros(self.state_shape())
        _lib.check(self._lib.fr_vjp_steps(self._h, int(n), int(every), _lib.dptr(lamN), _lib.dptr(out)))
        return out

    def kinetic_energy_gradient(self):
        """d(kinetic energy)/dU at the current state."""
        out = np.zeros(self.state_shape())
        _lib.check(self._lib.fr_reduce_gradient(self._h, 10, _lib.dptr(out)))
        return out

    def element_order(self):
        """order[i] = mesh element at device position i (FR_LAYOUT_DEVICE order)."""
        out = np.zeros(self.nEles, dtype=np.int32)
        _lib.check(self._lib.fr_get_element_order(self._h, _lib.iptr(out)))
        return out

    # -- files ----------------------------------------------------------------
    def write_data(self, data_file_name, iter=0, time=0.0, plot_type=1):
        """writeData (output.cpp:11-24): plot_type 1 writes the ParaView file
        <data_file_name>_<iter %09d>.vtu (3D), 0 the CSV file
        <data_file_name>.csv.<iter %09d>."""
        _lib.check(self._lib.fr_write_data(self._h, str(data_file_name).encode(), int(plot_type), int(iter),
                                           float(time)))

    def read_restart(self, data_file_name, restart_iter, time=0.0):
        """solver::readRestartFile (solver.cpp:1420-1482): sets the state from
        <data_file_name>_<restart_iter %09d>.vtu; returns the file's TIME
        (`time` when the file has none)."""
        t = ctypes.c_double(float(time))
        _lib.check(self._lib.fr_read_restart(self._h, str(data_file_name).encode(), int(restart_iter),
                                             ctypes.byref(t)))
        return t.value

    def write_checkpoint(self, path, iter=0, time=0.0):
        """Exact binary copy of the FP64 state with iter and time."""
        _lib.check(self._lib.fr_write_checkpoint(self._h, str(path).encode(), int(iter), float(time)))

    def read_checkpoint(self, path):
        """Sets the state from a checkpoint; returns (iter, time)."""
        it, t = ctypes.c_int(), ctypes.c_double()
        _lib.check(self._lib.fr_read_checkpoint(self._h, str(path).encode(), ctypes.byref(it), ctypes.byref(t)))
        return it.value, t.value

    def geometry_mode(self, request=-1):
        """Stage-kernel geometry layout (fr_geometry_mode): request 0 per-point
        arrays, 1 compact per element / face when the mesh is affine (the
        default), 2 nodal (metrics and normals formed from each linear
        element's corner nodes), 3 nodal metrics with the stored face
        geometry, -1 query.  Returns the layout in use (0..3)."""
        act = ctypes.c_int()
        _lib.check(self._lib.fr_geometry_mode(self._h, int(request), ctypes.byref(act)))
        return int(act.value)

    def residual_bytes(self):
        b = ctypes.c_double()
        _lib.check(self._lib.fr_residual_bytes(self._h, ctypes.byref(b)))
        return b.value

    def set_timing(self, on):
        _lib.check(self._lib.fr_set_timing(self._h, int(on)))

    def kernel_times(self, kernel=0):
        ms = ctypes.c_double()
        cnt = ctypes.c_long()
        _lib.check(self._lib.fr_kernel_times(self._h, kernel, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value


class Batch:
    """Several Solvers advanced together (fr_batch_*): one persistent launch
    per RK stage, element type and order over all their element blocks,
    replayed from a hipGraph.  The solvers must share equation, precision and
    timeType (quads, triangles and hexes may mix)."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        self._lib = _lib.load()
        arr = (ctypes.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        h = ctypes.c_void_p()
        _lib.check(self._lib.fr_batch_create(arr, len(self.solvers), ctypes.byref(h)))
        self._h = h

    def update(self, n=1):
        _lib.check(self._lib.fr_batch_update(self._h, int(n)))

    @staticmethod
    def force_tile(layout=-1):
        """Tile layout of the batches created afterwards (fr_batch_force_tile):
        -1 tuned, 0 per-mesh tile, 1 packed 2D tile."""
        _lib.check(_lib.load().fr_batch_force_tile(int(layout)))

    def tuning(self):
        """(largest persistent grid, work items, sum over the groups of their
        residual-only launch time in ms)."""
        g, n, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _lib.check(self._lib.fr_batch_tuning(self._h, ctypes.byref(g), ctypes.byref(n), ctypes.byref(ms)))
        return g.value, n.value, ms.value

    def group_tuning(self):
        """The tuner's choice per kernel group present: {("quad" | "hex" |
        "tri", P): {"threads_per_element", "elements_per_block", "grid",
        "items", "ms"}}."""
        out = {}
        for grp in range(1, 16):
            v = [ctypes.c_int() for _ in range(4)]
            ms = ctypes.c_double()
            _lib.check(self._lib.fr_batch_group_tuning(self._h, grp, *[ctypes.byref(x) for x in v],
                                                       ctypes.byref(ms)))
            if v[3].value:
                key = ("hex", grp - 10) if grp > 10 else ("tri", grp - 5) if grp > 5 else ("quad", grp)
                out[key] = {"threads_per_element": v[0].value, "elements_per_block": v[1].value,
                            "grid": v[2].value, "items": v[3].value, "ms": ms.value}
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.fr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
