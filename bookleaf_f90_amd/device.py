"""Host-side mirror of the reference's hot-path drivers over the C-ABI.

``BookLeaf(problem)`` owns a device-resident dataAPI registry on the current
HIP device (one MI355X) and exposes one method per reference driver with the
reference's name (hydro_getdt, hydro_getforce, ..., ale_advect), plus the
solver sequence of src/infrastructure/solver (getdt / lagstep / alestep) and
the device-resident multi-step runner.  There is no CPU fallback: the HIP
library must be built (``python -m bookleaf_f90_amd.build``) or construction
fails.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_LIB = None
LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbookleaf_hip.so")


class BookLeafError(RuntimeError):
    def __init__(self, msg, err=None):
        super().__init__(msg)
        self.err = err
        self.ierr = err.ierr if err is not None else None


def load_library(path=LIBPATH):
    """Load libbookleaf_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise ImportError("libbookleaf_hip.so not built: run python -m bookleaf_f90_amd.build")
        lib = C.CDLL(path)
        abi.declare(lib, "bl")
        P, vp = C.POINTER, C.c_void_p
        lib.bl_create.argtypes = [P(abi.Sizes), P(vp)]
        lib.bl_create.restype = C.c_int
        lib.bl_destroy.argtypes = [vp]
        lib.bl_set_config.argtypes = [vp, P(abi.Config)]
        lib.bl_set_regions.argtypes = [vp, C.c_int32, P(C.c_double), P(C.c_double),
                                       P(C.c_int32), P(C.c_int32)]
        lib.bl_set_eos.argtypes = [vp, C.c_int32, P(C.c_int32), P(C.c_double)]
        lib.bl_field_info.argtypes = [vp, C.c_int32, P(C.c_int32), P(C.c_int64)]
        lib.bl_field_device_ptr.argtypes = [vp, C.c_int32]
        lib.bl_field_device_ptr.restype = vp
        lib.bl_upload.argtypes = [vp, C.c_int32, vp, C.c_int64]
        lib.bl_download.argtypes = [vp, C.c_int32, vp, C.c_int64]
        lib.bl_build_maps.argtypes = [vp]
        lib.bl_init_mesh.argtypes = [vp]
        lib.bl_synchronize.argtypes = [vp]
        lib.bl_run_graph.argtypes = lib.bl_run.argtypes
        lib.bl_ensemble_create.argtypes = [P(abi.Instance), C.c_int32, P(abi.Config), P(vp)]
        lib.bl_ensemble_run.argtypes = [vp, C.c_int32, P(C.c_double), P(C.c_int32),
                                        P(C.c_int32), P(C.c_int32), P(abi.Error)]
        lib.bl_ensemble_offsets.argtypes = [vp, P(C.c_int32), P(C.c_int32)]
        lib.bl_ensemble_times.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_int32)]
        fp = P(C.c_float)
        lib.bl_mlp_create.argtypes = [C.c_uint64, C.c_int32, P(vp)]
        lib.bl_mlp_destroy.argtypes = [vp]
        lib.bl_mlp_weights.argtypes = [vp, fp, fp]
        lib.bl_mlp_forward.argtypes = [vp, fp, C.c_int32, fp]
        lib.bl_mlp_normalize.argtypes = [vp]
        lib.bl_mlp_norm.argtypes = [vp, fp, fp]
        lib.bl_mlp_add_pairs.argtypes = [vp, fp, fp, C.c_int32]
        lib.bl_mlp_pairs.argtypes = [vp, P(C.c_int32), fp, fp]
        lib.bl_mlp_gradient.argtypes = [vp, C.c_int32, fp, fp]
        lib.bl_mlp_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, fp]
        lib.bl_relax_collect.argtypes = [vp, vp, C.c_int32]
        lib.bl_relax_learned.argtypes = [vp, vp]
        lib.bl_set_learned_check.argtypes = [vp, C.c_double, C.c_int32]
        lib.bl_relax_learned_stats.argtypes = [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32),
                                               P(C.c_int32)]
        lib.bl_relax_features.argtypes = [vp, fp]
        lib.bl_autotune.argtypes = [vp, C.c_int32, P(abi.Tuning)]
        lib.bl_run_stats.argtypes = [vp, P(C.c_int64), P(C.c_int64)]
        lib.bl_set_tuning.argtypes = [vp, P(abi.Tuning)]
        lib.bl_get_tuning.argtypes = [vp, P(abi.Tuning)]
        lib.bl_set_tiling.argtypes = [vp, C.c_int32]
        lib.bl_get_tiling.argtypes = [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32),
                                      P(C.c_int32)]
        lib.bl_timing_enable.argtypes = [vp, C.c_int32]
        lib.bl_timing_reset.argtypes = [vp]
        lib.bl_timing_query.argtypes = [vp, C.c_char_p, P(C.c_double), P(C.c_int64)]
        lib.bl_timing_names.argtypes = [vp, C.c_char_p, C.c_int32]
        lib.bl_last_error.restype = C.c_char_p
        _LIB = lib
    return _LIB


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


class DriverState:
    """Common driver-level interface of the HIP library and the CPU oracle.

    Subclasses set ``self.lib``, ``self.prefix`` and ``self.h`` and implement
    ``get``/``set`` for dataAPI fields.
    """

    def _f(self, name):
        return getattr(self.lib, self.prefix + "_" + name)

    def _check(self, rc, what, err=None):
        if rc != abi.SUCCESS:
            msg = what
            if err is not None and err.ierr != 0:
                msg += ": " + err.message + " (ierr=%d)" % err.ierr
            raise BookLeafError(msg, err)

    def _setup(self, problem):
        """inf_init order (inf_init.F90:53-101) with the problem's mesh and ICs."""
        p = problem
        self.problem = p
        self.sizes = p.sizes()
        for name in ("ielnd", "ielreg", "ielmat", "indtype", "ndx", "ndy"):
            self.set(name, getattr(p, name))
        self._check(self._f("init_mesh")(self.h), "init_mesh")
        for name in ("ndu", "ndv"):
            self.set(name, getattr(p, name))
        self.set("eldensity", p.density)
        self.set("elenergy", p.energy)
        if p.ncp:  # mixed cells: the flattened component lists and states
            for name in ("imxel", "imxfcp", "imxncp", "icpmat", "icpnext", "icpprev",
                         "frvolume", "cpdensity", "cpenergy"):
                self.set(name, getattr(p, name))
        err = abi.Error()
        self._check(self._f("geometry_getgeometry")(self.h, C.byref(err)), "getgeometry", err)
        self._check(self._f("init_state")(self.h), "init_state")
        self._check(self._f("hydro_init")(self.h), "hydro_init")
        if p.relax:
            self.set_relaxation(*p.relax)
        self.ts = abi.Timestep()
        self.ts.time = p.config["time_start"]
        self.ts.dt = p.config["dt_initial"]
        self.ts.dts = self.ts.dt
        self.ts.nstep = 0
        self.ts.idtel = 0
        self.ts.zcorrector = 0
        self.ts.sdt = abi.DT_INITIAL

    # ---- drivers (src/packages/*/driver) -------------------------------
    def hydro_getdt(self):
        cfl, div, err = abi.Dt(), abi.Dt(), abi.Error()
        self._check(self._f("hydro_getdt")(self.h, C.byref(cfl), C.byref(div), C.byref(err)),
                    "hydro_getdt", err)
        return cfl, div

    def hydro_getdt_noalloc(self):
        """hydro_getdt through bl_hydro_getdt_noalloc: the same kernels and
        results, with no device allocation in the call (meshes up to
        abi.NOALLOC_MAX_EL elements)."""
        cfl, div, err = abi.Dt(), abi.Dt(), abi.Error()
        self._check(self._f("hydro_getdt_noalloc")(self.h, C.byref(cfl), C.byref(div),
                                                   C.byref(err)),
                    "hydro_getdt_noalloc", err)
        return cfl, div

    def hydro_setpredictor(self):
        self._check(self._f("hydro_setpredictor")(self.h), "hydro_setpredictor")

    def hydro_setcorrector(self):
        self._check(self._f("hydro_setcorrector")(self.h), "hydro_setcorrector")

    def hydro_getforce(self):
        self._check(self._f("hydro_getforce")(self.h, C.byref(self.ts)), "hydro_getforce")

    def hydro_getstate(self):
        err = abi.Error()
        self._check(self._f("hydro_getstate")(self.h, C.byref(self.ts), C.byref(err)),
                    "hydro_getstate", err)

    def hydro_getenergy(self):
        self._check(self._f("hydro_getenergy")(self.h, C.byref(self.ts)), "hydro_getenergy")

    def hydro_geteos(self):
        self._check(self._f("hydro_geteos")(self.h), "hydro_geteos")

    def hydro_initacceleration(self):
        self._check(self._f("hydro_initacceleration")(self.h), "hydro_initacceleration")

    def hydro_scatteracceleration(self):
        self._check(self._f("hydro_scatteracceleration")(self.h), "hydro_scatteracceleration")

    def hydro_getacceleration(self):
        self._check(self._f("hydro_getacceleration")(self.h), "hydro_getacceleration")

    def hydro_applyacceleration(self):
        self._check(self._f("hydro_applyacceleration")(self.h, C.byref(self.ts)),
                    "hydro_applyacceleration")

    def geometry_getgeometry(self):
        err = abi.Error()
        self._check(self._f("geometry_getgeometry")(self.h, C.byref(err)), "getgeometry", err)

    def ale_active(self):
        return bool(self._f("ale_active")(self.h, C.byref(self.ts)))

    def ale_getmeshstatus(self):
        self._check(self._f("ale_getmeshstatus")(self.h), "ale_getmeshstatus")

    def ale_getfluxvolume(self):
        self._check(self._f("ale_getfluxvolume")(self.h, C.byref(self.ts)), "ale_getfluxvolume")

    def ale_advect(self):
        err = abi.Error()
        self._check(self._f("ale_advect")(self.h, C.byref(self.ts), C.byref(err)),
                    "ale_advect", err)

    def ale_hydroupdate(self):
        err = abi.Error()
        self._check(self._f("ale_hydroupdate")(self.h, C.byref(err)), "ale_hydroupdate", err)

    # ---- time package (time_dr_advance.f90:38-132) -----------------------
    def time_select(self, cfl, div):
        c = self.problem.config
        cand = [(c["dt_max"], -1, abi.DT_MAXIMUM), (c["dt_g"] * self.ts.dt, -1, abi.DT_GROWTH)]
        if self.ts.nstep == 0:
            cand.append((c["dt_initial"], -1, abi.DT_INITIAL))
        cand += [(cfl.rdt, cfl.idt, cfl.sdt), (div.rdt, div.idt, div.sdt)]
        best = cand[0]
        for cnd in cand[1:]:
            if cnd[0] < best[0]:
                best = cnd
        self.ts.nstep += 1
        self.ts.dt, self.ts.idtel, self.ts.sdt = best
        self.ts.time = self.ts.time + self.ts.dt
        if self.ts.dt < c["dt_min"]:
            raise BookLeafError("ERROR: dt < dt_min")

    # ---- solver sequence (src/infrastructure/solver) ----------------------
    def getdt(self):
        cfl, div = self.hydro_getdt()
        self.time_select(cfl, div)

    def lagstep(self):
        self.ts.zcorrector = 0
        self.ts.dts = 0.5 * self.ts.dt
        self.hydro_setpredictor()
        self.hydro_getforce()
        self.hydro_getstate()
        self.hydro_getenergy()
        self.hydro_geteos()
        self.ts.zcorrector = 1
        self.ts.dts = self.ts.dt
        self.hydro_setcorrector()
        self.hydro_getforce()
        self.hydro_initacceleration()
        self.hydro_scatteracceleration()
        self.hydro_getacceleration()
        self.hydro_applyacceleration()
        self.hydro_getstate()
        self.hydro_getenergy()
        self.hydro_geteos()

    def alestep(self):
        if not self.ale_active():
            return
        self.ale_getmeshstatus()
        self.ale_getfluxvolume()
        self.ale_advect()
        self.ale_hydroupdate()

    def step(self):
        self.getdt()
        self.lagstep()
        self.alestep()

    def run(self, nsteps, stop_at_time_end=False, runner="run"):
        """Run up to nsteps full steps in the library; returns the per-step
        (dt, idtel, control) history."""
        dt = np.zeros(max(nsteps, 1))
        idt = np.zeros(max(nsteps, 1), np.int32)
        ctl = np.zeros(max(nsteps, 1), np.int32)
        taken = C.c_int32(0)
        err = abi.Error()
        rc = self._f(runner)(self.h, nsteps, int(bool(stop_at_time_end)), C.byref(self.ts),
                             _dptr(dt), _iptr(idt), _iptr(ctl), C.byref(taken), C.byref(err))
        n = taken.value
        self.last_history = (dt[:n].copy(), idt[:n].copy(), ctl[:n].copy())
        self._check(rc, runner, err)
        return self.last_history

    def totals(self):
        t = abi.Totals()
        self._check(self._f("totals")(self.h, C.byref(t)), "totals")
        return t

    def shortprint(self):
        """hydro_dr_shortprint's tables: (region, material) as float arrays of
        shape (nreg, 9) and (nmat, 9), columns in abi.FLAG_TOTALS order."""
        nreg, nmat = self.sizes.nreg, self.sizes.nmat
        reg = (abi.FlagTotals * max(nreg, 1))()
        mat = (abi.FlagTotals * max(nmat, 1))()
        self._check(self._f("shortprint")(self.h, reg, mat), "shortprint")
        as_np = lambda t, n: np.array([[getattr(t[k], q) for q in abi.FLAG_TOTALS]
                                       for k in range(n)]).reshape(n, len(abi.FLAG_TOTALS))
        return as_np(reg, nreg), as_np(mat, nmat)

    def sod_l1(self, time=None, x0=50.0, scale=1.0):
        """check_kn_sod's (L1 density, L1 energy) at time (default: now)."""
        out = (C.c_double * 2)()
        t = self.ts.time if time is None else time
        self._check(self._f("sod_l1")(self.h, t, x0, scale, out), "sod_l1")
        return out[0], out[1]

    def set_relaxation(self, rtype=abi.RELAX_EQUI, nsweep=1, om=1.0, tol=0.0):
        """The ALE step's node-relaxation mesh motion (zexist, not zeul)."""
        self._check(self._f("set_relaxation")(self.h, rtype, nsweep, om, tol),
                    "set_relaxation")

    def relax_mesh(self, rtype=abi.RELAX_EQUI, nsweep=1, om=1.0, tol=0.0):
        """nsweep relaxation sweeps on ndx, ndy in place."""
        self._check(self._f("relax_mesh")(self.h, rtype, nsweep, om, tol), "relax_mesh")

    def state(self, names=None):
        names = names or abi.FIELDS
        return {n: self.get(n) for n in names}


class BookLeaf(DriverState):
    """Device-resident BookLeaf state on the current HIP device."""

    def __init__(self, problem):
        self.lib = load_library()
        self.prefix = "bl"
        sz = problem.sizes()
        h = C.c_void_p()
        if self.lib.bl_create(C.byref(sz), C.byref(h)) != abi.SUCCESS:
            raise BookLeafError("bl_create: " + self.lib.bl_last_error().decode())
        self.h = h
        cfg = problem.cconfig()
        self._check(self.lib.bl_set_config(self.h, C.byref(cfg)), "set_config")
        self._check(self.lib.bl_set_regions(
            self.h, problem.nreg, _dptr(np.ascontiguousarray(problem.kappareg, np.float64)),
            _dptr(np.ascontiguousarray(problem.pmeritreg, np.float64)),
            _iptr(np.ascontiguousarray(problem.zdtnotreg, np.int32)),
            _iptr(np.ascontiguousarray(problem.zmidlength, np.int32))), "set_regions")
        self._check(self.lib.bl_set_eos(
            self.h, problem.nmat, _iptr(np.ascontiguousarray(problem.eos_itype, np.int32)),
            _dptr(np.ascontiguousarray(problem.eos_param, np.float64))), "set_eos")
        self._setup(problem)

    def _check(self, rc, what, err=None):
        if rc != abi.SUCCESS:
            msg = what + ": " + self.lib.bl_last_error().decode()
            if err is not None and err.ierr != 0:
                msg += " [" + err.message + ", ierr=%d]" % err.ierr
            raise BookLeafError(msg, err)

    def _info(self, name):
        kind, count = C.c_int32(), C.c_int64()
        self.lib.bl_field_info(self.h, abi.ID[name], C.byref(kind), C.byref(count))
        return kind.value, count.value

    def get(self, name):
        kind, count = self._info(name)
        out = np.empty(count, abi.DTYPE[kind])
        self._check(self.lib.bl_download(self.h, abi.ID[name], out.ctypes.data, count),
                    "download " + name)
        return out

    def set(self, name, values):
        kind, count = self._info(name)
        a = np.ascontiguousarray(values, abi.DTYPE[kind])
        self._check(self.lib.bl_upload(self.h, abi.ID[name], a.ctypes.data, len(a)),
                    "upload " + name)

    def device_ptr(self, name):
        return self.lib.bl_field_device_ptr(self.h, abi.ID[name])

    def synchronize(self):
        self._check(self.lib.bl_synchronize(self.h), "synchronize")

    def set_diagnostics(self, on=True, x0=50.0, scale=1.0, maxrec=4096):
        """Record, every step of run/run_graph, nstep, time, the region table
        and the Sod L1 pair on the device."""
        self._check(self._f("set_diagnostics")(self.h, int(bool(on)), x0, scale, maxrec),
                    "set_diagnostics")
        self._diag_max = maxrec

    def diagnostics(self):
        """The records since set_diagnostics: dict of nstep (n,), time (n,),
        region (n, nreg, 9) and l1 (n, 2)."""
        nreg = self.sizes.nreg
        rec = 2 + len(abi.FLAG_TOTALS) * nreg + 2
        cap = getattr(self, "_diag_max", 0)
        buf = np.zeros(max(cap, 1) * rec)
        n = C.c_int32(0)
        self._check(self._f("diagnostics")(self.h, cap, _dptr(buf), C.byref(n)), "diagnostics")
        r = buf[:n.value * rec].reshape(n.value, rec)
        return {"nstep": r[:, 0].astype(np.int64), "time": r[:, 1],
                "region": r[:, 2:2 + 9 * nreg].reshape(n.value, nreg, 9), "l1": r[:, -2:]}

    def relax_collect(self, net, stride=1):
        """Append (stencil, displacement) pairs of the plain relaxation to net
        on every ALE step (every stride-th relaxed node); net None stops."""
        self._check(self.lib.bl_relax_collect(self.h, net.m if net else None, stride),
                    "relax_collect")
        self._collect_net = net

    def relax_learned(self, net):
        """The network's node motion as the ALE mesh motion (the plain
        relaxation as its fallback); net None goes back to the plain one."""
        self._check(self.lib.bl_relax_learned(self.h, net.m if net else None), "relax_learned")
        self._learned_net = net

    def set_learned_check(self, max_aspect=10.0, relative=True):
        """the check of the network's mesh (relative: only elements it makes
        worse than the Lagrangian mesh count)"""
        self._check(self.lib.bl_set_learned_check(self.h, max_aspect, int(bool(relative))),
                    "set_learned_check")

    def learned_stats(self, full=False):
        """(ALE steps run with the network, steps that fell back); with full
        (steps, guarded steps, fallbacks, reason bits: 1 a corner Jacobian
        <= 0, 2 an aspect ratio > max_aspect)"""
        a, g, b, c = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.bl_relax_learned_stats(self.h, C.byref(a), C.byref(g), C.byref(b),
                                                    C.byref(c)), "relax_learned_stats")
        return (a.value, g.value, b.value, c.value) if full else (a.value, b.value)

    def relax_features(self):
        """The network's input features of every node (nnd, 26)."""
        out = np.zeros((self.sizes.nnd, abi.MLP_NIN), np.float32)
        self._check(self.lib.bl_relax_features(self.h, _fptr(out)), "relax_features")
        return out

    def run_stats(self):
        """(moving, zeroed) of the last graph or ensemble run call: elements
        in motion and elements whose zero viscosity the step wrote, summed
        over its steps (bl_run_stats)."""
        m, z = C.c_int64(0), C.c_int64(0)
        self._check(self.lib.bl_run_stats(self.h, C.byref(m), C.byref(z)), "run_stats")
        return m.value, z.value

    def autotune(self, reps=5):
        """Pick the limiter pass's launch shape by timing (bl_autotune);
        returns the choice as a dict.  Results do not depend on it."""
        t = abi.Tuning()
        self._check(self.lib.bl_autotune(self.h, int(reps), C.byref(t)), "autotune")
        return {k: getattr(t, k) for k, _ in abi.Tuning._fields_}

    def set_tuning(self, limiter_block=256, limiter_grid=512, ens_chunk=4096):
        """Apply a launch shape without timing (bl_set_tuning)."""
        t = abi.Tuning(limiter_block=limiter_block, limiter_grid=limiter_grid,
                       ens_chunk=ens_chunk)
        self._check(self.lib.bl_set_tuning(self.h, C.byref(t)), "set_tuning")

    def get_tuning(self):
        """The launch shape in use (bl_get_tuning)."""
        t = abi.Tuning()
        self._check(self.lib.bl_get_tuning(self.h, C.byref(t)), "get_tuning")
        return {k: getattr(t, k) for k, _ in abi.Tuning._fields_[:3]}

    def set_tiling(self, on=True):
        """The tile step on or off (bl_set_tiling)."""
        self._check(self.lib.bl_set_tiling(self.h, int(bool(on))), "set_tiling")

    def get_tiling(self):
        """(tile step in use, tiles, 0 (no boundary elements: every element
        completes in its own tile), element entries) (bl_get_tiling)."""
        v = [C.c_int32(0) for _ in range(4)]
        self._check(self.lib.bl_get_tiling(self.h, *(C.byref(x) for x in v)), "get_tiling")
        return bool(v[0].value), v[1].value, v[2].value, v[3].value

    def timing(self, on=True):
        self._check(self.lib.bl_timing_enable(self.h, int(on)), "timing_enable")

    def timing_reset(self):
        self._check(self.lib.bl_timing_reset(self.h), "timing_reset")

    def timing_query(self, name):
        ms, n = C.c_double(), C.c_int64()
        self._check(self.lib.bl_timing_query(self.h, name.encode(), C.byref(ms), C.byref(n)),
                    "timing_query")
        return ms.value, n.value

    def timing_names(self):
        buf = C.create_string_buffer(4096)
        self.lib.bl_timing_names(self.h, buf, 4096)
        return [s for s in buf.value.decode().split(",") if s]

    def close(self):
        if getattr(self, "h", None):
            self.lib.bl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RelaxNet:
    """The node-relaxation network (bl_mlp_*) on the current HIP device."""

    def __init__(self, seed=0, capacity=0):
        self.lib = load_library()
        m = C.c_void_p()
        if self.lib.bl_mlp_create(seed, capacity, C.byref(m)) != abi.SUCCESS:
            raise BookLeafError("bl_mlp_create: " + self.lib.bl_last_error().decode())
        self.m = m
        self.capacity = capacity

    def _check(self, rc, what):
        if rc != abi.SUCCESS:
            raise BookLeafError(what + ": " + self.lib.bl_last_error().decode())

    def weights(self):
        w = np.empty(abi.MLP_NPARAM, np.float32)
        self._check(self.lib.bl_mlp_weights(self.m, _fptr(w), None), "mlp_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.shape == (abi.MLP_NPARAM,)
        self._check(self.lib.bl_mlp_weights(self.m, None, _fptr(w)), "mlp_weights")

    def normalize(self):
        """standardisation from the stored pairs"""
        self._check(self.lib.bl_mlp_normalize(self.m), "mlp_normalize")

    def norm(self):
        """(mu, inverse sigma, output mu, output sigma)"""
        v = np.empty(abi.MLP_NNORM, np.float32)
        self._check(self.lib.bl_mlp_norm(self.m, _fptr(v), None), "mlp_norm")
        return v[:26], v[26:52], v[52:54], v[54:56]

    def set_norm(self, mu, isig, mu_t, sig_t):
        v = np.concatenate([mu, isig, mu_t, sig_t]).astype(np.float32)
        assert v.shape == (abi.MLP_NNORM,)
        self._check(self.lib.bl_mlp_norm(self.m, None, _fptr(v)), "mlp_norm")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, abi.MLP_NIN)
        y = np.empty((len(x), abi.MLP_NOUT), np.float32)
        self._check(self.lib.bl_mlp_forward(self.m, _fptr(x), len(x), _fptr(y)), "mlp_forward")
        return y

    def add_pairs(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, abi.MLP_NIN)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, abi.MLP_NOUT)
        assert len(x) == len(t)
        self._check(self.lib.bl_mlp_add_pairs(self.m, _fptr(x), _fptr(t), len(x)),
                    "mlp_add_pairs")

    def npairs(self):
        n = C.c_int32()
        self._check(self.lib.bl_mlp_pairs(self.m, C.byref(n), None, None), "mlp_pairs")
        return n.value

    def pairs(self):
        n = self.npairs()
        x = np.empty((n, abi.MLP_NIN), np.float32)
        t = np.empty((n, abi.MLP_NOUT), np.float32)
        c = C.c_int32()
        self._check(self.lib.bl_mlp_pairs(self.m, C.byref(c), _fptr(x), _fptr(t)), "mlp_pairs")
        return x, t

    def gradient(self, n):
        """(gradient, loss) of the loss over pairs 0..n-1"""
        g = np.empty(abi.MLP_NPARAM, np.float32)
        loss = C.c_float()
        self._check(self.lib.bl_mlp_gradient(self.m, n, _fptr(g),
                                             C.cast(C.byref(loss), C.POINTER(C.c_float))),
                    "mlp_gradient")
        return g, loss.value

    def train(self, iters, batch=4096, lr=1e-3, seed=0):
        """iters Adam steps; the loss of each batch"""
        hist = np.empty(max(iters, 1), np.float32)
        self._check(self.lib.bl_mlp_train(self.m, iters, batch, lr, seed, _fptr(hist)),
                    "mlp_train")
        return hist[:iters]

    def close(self):
        if getattr(self, "m", None):
            self.lib.bl_mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble(BookLeaf):
    """Independent instances generated and stepped together on the device
    (bl_ensemble_create / bl_ensemble_run); params as problem.ensemble_params."""

    def __init__(self, params, config=None):
        from . import problem as P
        self.lib = load_library()
        self.prefix = "bl"
        self.params = list(params)
        cfg = abi.Config()
        for k, v in (config or P.ENSEMBLE_CONFIG).items():
            setattr(cfg, k, v)
        arr = P.instances_array(self.params)
        h = C.c_void_p()
        if self.lib.bl_ensemble_create(arr, len(arr), C.byref(cfg), C.byref(h)) != abi.SUCCESS:
            raise BookLeafError("bl_ensemble_create: " + self.lib.bl_last_error().decode())
        self.h = h
        n = len(arr)
        self.eoff = np.zeros(n + 1, np.int32)
        self.noff = np.zeros(n + 1, np.int32)
        self._check(self.lib.bl_ensemble_offsets(self.h, _iptr(self.eoff), _iptr(self.noff)),
                    "ensemble_offsets")
        self.nel, self.nnd = int(self.eoff[-1]), int(self.noff[-1])

    def run(self, nsteps):
        """nsteps steps of every instance; returns (dt, idtel, control), each
        (steps, ninst)."""
        n = len(self.params)
        dt = np.zeros(max(nsteps, 1) * n)
        idt = np.zeros(max(nsteps, 1) * n, np.int32)
        ctl = np.zeros(max(nsteps, 1) * n, np.int32)
        taken = C.c_int32(0)
        err = abi.Error()
        rc = self.lib.bl_ensemble_run(self.h, nsteps, _dptr(dt), _iptr(idt), _iptr(ctl),
                                      C.byref(taken), C.byref(err))
        m = taken.value
        self.last_history = tuple(a[:m * n].reshape(m, n).copy() for a in (dt, idt, ctl))
        self._check(rc, "ensemble_run", err)
        return self.last_history

    def times(self):
        n = len(self.params)
        t, dt, ns = np.zeros(n), np.zeros(n), np.zeros(n, np.int32)
        self._check(self.lib.bl_ensemble_times(self.h, _dptr(t), _dptr(dt), _iptr(ns)),
                    "ensemble_times")
        return t, dt, ns

    def instance_field(self, name, k, data=None):
        """Instance k's part of a field (element, corner or node layout), with
        node and element index contents made instance-local."""
        a = self.get(name) if data is None else data
        kind, count = self._info(name)
        e0, e1 = int(self.eoff[k]), int(self.eoff[k + 1])
        n0, n1 = int(self.noff[k]), int(self.noff[k + 1])
        if name in ("ndx", "ndy", "ndu", "ndv", "indtype"):
            return a[n0:n1].copy()
        if count >= 4 * self.nel and count != self.nnd:
            out = a[4 * e0:4 * e1].copy()
        else:
            out = a[e0:e1].copy()
        if name == "ielnd":
            out = out - n0
        elif name in ("ielel", "ielsort1", "ielsort2"):
            out = out - e0
        elif name == "ielreg":
            out = out - 2 * k
        elif name == "ielmat":
            out = out - k
        return out
