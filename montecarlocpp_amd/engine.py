"""ctypes binding of the engine library (montecarlocpp_amd/libmcg.so).

The library is built in-tree by :mod:`montecarlocpp_amd.build`; this module
never builds it and never falls back to another implementation: if the
library is missing, :func:`lib` raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmcg.so")

# every function include/mcg.h declares
EXPORTS = ("mcg_abi_version", "mcg_last_error", "mcg_material_compute", "mcg_problem_plan",
           "mcg_cell_volumes",
           "mcg_alias_build", "mcg_philox4x32_10", "mcg_scene_create", "mcg_scene_destroy",
           "mcg_scene_plan", "mcg_scene_tune", "mcg_autotune", "mcg_autotune_batch",
           "mcg_server_create", "mcg_server_submit", "mcg_server_step", "mcg_server_completed",
           "mcg_server_poll",
           "mcg_server_cancel", "mcg_server_info", "mcg_server_destroy",
           "mcg_solve", "mcg_trace", "mcg_trace_from", "mcg_batch_solve", "mcg_batch_create", "mcg_batch_run",
           "mcg_batch_destroy", "mcg_solve_static")

TRACE_WIDTH = abi.TRACE_WIDTH

_lib = None


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mcg status {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError(abi.MCG_EDEVICE,
                              f"{LIB_PATH} is not built; run python -m montecarlocpp_amd.build")
        L = C.CDLL(LIB_PATH)
        L.mcg_last_error.restype = C.c_char_p
        L.mcg_abi_version.restype = C.c_int
        MS, DS, PS = (C.POINTER(abi.MaterialSpec), C.POINTER(abi.DomainSpec),
                      C.POINTER(abi.ProblemSpec))
        L.mcg_material_compute.argtypes = [MS, C.POINTER(abi.MaterialTables)]
        L.mcg_problem_plan.argtypes = [MS, DS, PS, C.POINTER(abi.ProblemInfo),
                                       C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int64]
        L.mcg_cell_volumes.argtypes = [DS, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64)]
        L.mcg_alias_build.argtypes = [C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32)]
        L.mcg_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.mcg_philox4x32_10.restype = None
        L.mcg_scene_create.argtypes = [MS, DS, C.POINTER(C.c_void_p)]
        L.mcg_scene_destroy.argtypes = [C.c_void_p]
        L.mcg_scene_plan.argtypes = [C.c_void_p, PS, C.POINTER(abi.ProblemInfo)]
        L.mcg_scene_tune.argtypes = [C.c_void_p, C.POINTER(abi.Tune)]
        L.mcg_server_create.argtypes = [C.c_uint64, C.c_int64, C.POINTER(abi.Tune),
                                        C.POINTER(C.c_void_p)]
        L.mcg_server_submit.argtypes = [C.c_void_p, MS, DS, PS, C.POINTER(C.c_int64)]
        L.mcg_server_step.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]
        L.mcg_server_completed.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int64,
                                           C.POINTER(C.c_int64)]
        L.mcg_server_poll.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.c_int64,
                                      C.POINTER(abi.Stats)]
        L.mcg_server_cancel.argtypes = [C.c_void_p, C.c_int64]
        L.mcg_server_info.argtypes = [C.c_void_p, C.POINTER(abi.ServerStats)]
        L.mcg_server_destroy.argtypes = [C.c_void_p]
        L.mcg_autotune.argtypes = [C.c_void_p, PS, C.c_int64, C.POINTER(abi.Tune),
                                   C.POINTER(C.c_double)]
        L.mcg_autotune_batch.argtypes = [C.c_int32, C.POINTER(C.c_void_p), PS, C.c_double,
                                         C.POINTER(abi.Tune), C.POINTER(C.c_double)]
        L.mcg_solve.argtypes = [C.c_void_p, PS, C.POINTER(C.c_double), C.c_int64,
                                C.POINTER(abi.Counts), C.POINTER(abi.Stats)]
        L.mcg_solve_static.argtypes = [C.c_void_p, PS, C.POINTER(C.c_double), C.c_int64,
                                       C.POINTER(abi.Stats)]
        L.mcg_trace.argtypes = [C.c_void_p, PS, C.c_int64, C.POINTER(C.c_double), C.c_int64,
                                C.POINTER(C.c_int64)]
        L.mcg_trace_from.argtypes = [C.c_void_p, PS, C.c_int64, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64,
                                     C.POINTER(C.c_int64)]
        L.mcg_batch_solve.argtypes = [C.c_int32, C.POINTER(C.c_void_p), PS,
                                      C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_int64),
                                      C.POINTER(abi.Stats), C.POINTER(abi.Tune)]
        L.mcg_batch_create.argtypes = [C.c_int32, C.POINTER(C.c_void_p), PS, C.POINTER(abi.Tune),
                                       C.POINTER(C.c_void_p)]
        L.mcg_batch_run.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.POINTER(C.c_double)),
                                    C.POINTER(C.c_int64), C.POINTER(abi.Stats)]
        L.mcg_batch_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise EngineError(rc, lib().mcg_last_error().decode())


def plan(mat, dom, prob):
    info = abi.ProblemInfo()
    w = np.zeros(256)
    n = np.zeros(256, np.int64)
    check(lib().mcg_problem_plan(C.byref(mat.spec), C.byref(dom.spec()), C.byref(prob),
                                 C.byref(info), abi.dptr(w),
                                 n.ctypes.data_as(C.POINTER(C.c_int64)), 256))
    k = info.nemitter
    return info, w[:k].copy(), n[:k].copy()


def cell_volumes(dom):
    """Per-column tally cell volumes (host only)."""
    n = C.c_int64()
    check(lib().mcg_cell_volumes(C.byref(dom.spec()), None, 0, C.byref(n)))
    out = np.zeros(n.value)
    check(lib().mcg_cell_volumes(C.byref(dom.spec()), abi.dptr(out), out.size, C.byref(n)))
    return out


def material_tables(mat):
    n = mat.nw * mat.np
    names = ("tau", "vel", "energy_pdf", "flux_pdf", "scat_pdf")
    arrs = {k: np.zeros(n) for k in names}
    t = abi.MaterialTables(*(abi.dptr(arrs[k]) for k in names))
    check(lib().mcg_material_compute(C.byref(mat.spec), C.byref(t)))
    arrs.update(energy_sum=t.energy_sum, flux_sum=t.flux_sum, scat_sum=t.scat_sum, cond=t.cond,
                nmodes=t.nmodes)
    return arrs


def alias(pdf):
    pdf = np.ascontiguousarray(pdf, np.float64)
    thr = np.zeros(pdf.size, np.uint32)
    al = np.zeros(pdf.size, np.uint32)
    check(lib().mcg_alias_build(abi.dptr(pdf), pdf.size, abi.u32ptr(thr), abi.u32ptr(al)))
    return thr, al


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().mcg_philox4x32_10(c, k, o)
    return list(o)


class Scene:
    """Material tables + geometry resident on the current GPU
    (mcg_scene_create).  Keeps the material and domain alive."""

    def __init__(self, mat, dom):
        self.mat, self.dom = mat, dom
        h = C.c_void_p()
        check(lib().mcg_scene_create(C.byref(mat.spec), C.byref(dom.spec()), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().mcg_scene_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_tune(self, tune):
        """Launch shape of this scene's single-solve kernel (None: defaults)."""
        check(lib().mcg_scene_tune(self.handle, C.byref(tune) if tune is not None else None))

    def autotune(self, prob, sample_bundles=0):
        """mcg_autotune: pick and install the fastest launch shape."""
        t = abi.Tune()
        r = C.c_double()
        check(lib().mcg_autotune(self.handle, C.byref(prob), int(sample_bundles), C.byref(t),
                                 C.byref(r)))
        return t, r.value

    def plan(self, prob):
        info = abi.ProblemInfo()
        check(lib().mcg_scene_plan(self.handle, C.byref(prob), C.byref(info)))
        return info

    def solve(self, prob, counts=False):
        """FieldProblem::solve on the GPU.  Returns (field rows x cols, stats
        dict, CountsBuffers or None)."""
        info = self.plan(prob)
        out = np.zeros(info.rows * info.cols)
        nb = prob.bundle_count if prob.bundle_count else info.nemit - prob.first_bundle
        cb = abi.CountsBuffers(nb) if counts else None
        st = abi.Stats()
        check(lib().mcg_solve(self.handle, C.byref(prob), abi.dptr(out), out.size,
                              C.byref(cb.struct) if cb else None, C.byref(st)))
        return out.reshape(info.cols, info.rows).T.copy(), st.as_dict(), cb

    def trace(self, prob, bundle, cap=4096):
        """Step records of one bundle (mcg_trace): array (nsteps, TRACE_WIDTH)."""
        rec = np.zeros(cap * TRACE_WIDTH)
        n = C.c_int64()
        check(lib().mcg_trace(self.handle, C.byref(prob), int(bundle), abi.dptr(rec), cap,
                              C.byref(n)))
        return rec.reshape(cap, TRACE_WIDTH)[:n.value].copy()

    def trace_from(self, prob, pos, dir=None, prop_cell=-1, cap=4096):
        """Step records of one trajectory from a given start (mcg_trace_from,
        TrajProblem with pos): array (nsteps, TRACE_WIDTH)."""
        rec = np.zeros(cap * TRACE_WIDTH)
        n = C.c_int64()
        p = np.ascontiguousarray(pos, dtype=np.float64)
        d = None if dir is None else np.ascontiguousarray(dir, dtype=np.float64)
        check(lib().mcg_trace_from(self.handle, C.byref(prob), int(prop_cell), abi.dptr(p),
                                   None if d is None else abi.dptr(d), abi.dptr(rec), cap, C.byref(n)))
        return rec.reshape(cap, TRACE_WIDTH)[:n.value].copy()


def batch_solve(scenes_, probs, tune=None):
    """mcg_batch_solve: all solves in one persistent launch (per precision).
    Returns (list of fields, list of stats dicts)."""
    n = len(scenes_)
    infos = [sc.plan(p) for sc, p in zip(scenes_, probs)]
    outs = [np.zeros(i.rows * i.cols) for i in infos]
    handles = (C.c_void_p * n)(*[sc.handle.value for sc in scenes_])
    parr = (abi.ProblemSpec * n)(*probs)
    optr = (C.POINTER(C.c_double) * n)(*[abi.dptr(o) for o in outs])
    lens = (C.c_int64 * n)(*[o.size for o in outs])
    stats = (abi.Stats * n)()
    t = tune if tune is not None else abi.Tune(0, 0, 0, 0, 0)
    check(lib().mcg_batch_solve(n, handles, parr, optr, lens, stats, C.byref(t)))
    fields = [o.reshape(i.cols, i.rows).T.copy() for o, i in zip(outs, infos)]
    return fields, [s.as_dict() for s in stats]


class Batch:
    """mcg_batch_*: a batch captured once into a HIP graph, replayed by run()."""

    def __init__(self, scenes_, probs, tune=None):
        n = len(scenes_)
        self._scenes = list(scenes_)           # the scenes must outlive the batch
        self.infos = [sc.plan(p) for sc, p in zip(scenes_, probs)]
        self.outs = [np.zeros(i.rows * i.cols) for i in self.infos]
        self._optr = (C.POINTER(C.c_double) * n)(*[abi.dptr(o) for o in self.outs])
        self._lens = (C.c_int64 * n)(*[o.size for o in self.outs])
        self.stats = (abi.Stats * n)()
        handles = (C.c_void_p * n)(*[sc.handle.value for sc in scenes_])
        parr = (abi.ProblemSpec * n)(*probs)
        h = C.c_void_p()
        check(lib().mcg_batch_create(n, handles, parr, C.byref(tune) if tune is not None else None,
                                     C.byref(h)))
        self.handle = h
        self.n = n

    def run(self, seeds=None):
        """Replay the graph (optionally with new per-solve seeds); returns
        (list of fields, list of stats dicts) as batch_solve does."""
        sp = (C.c_uint64 * self.n)(*[int(x) for x in seeds]) if seeds is not None else None
        check(lib().mcg_batch_run(self.handle, sp, self._optr, self._lens, self.stats))
        fields = [o.reshape(i.cols, i.rows).T.copy() for o, i in zip(self.outs, self.infos)]
        return fields, [s.as_dict() for s in self.stats]

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().mcg_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def autotune_batch(scenes_, probs, sample_fraction=0.1):
    """mcg_autotune_batch: the fastest batch launch shape on a sample."""
    n = len(scenes_)
    handles = (C.c_void_p * n)(*[sc.handle.value for sc in scenes_])
    parr = (abi.ProblemSpec * n)(*probs)
    t = abi.Tune()
    r = C.c_double()
    check(lib().mcg_autotune_batch(n, handles, parr, sample_fraction, C.byref(t), C.byref(r)))
    return t, r.value


def tally_rows(tally, size):
    """Rows of a problem's output (TempProblem 1, FluxProblem 3, MultiProblem
    4, CumTempProblem size + 1, CumFluxProblem 3 (size + 1))."""
    return {abi.TEMP: 1, abi.FLUX: 3, abi.MULTI: 4, abi.CUMTEMP: size + 1}.get(tally, 3 * (size + 1))


class Server:
    """Continuous-batching request server (mcg_server_*)."""

    def __init__(self, budget_bytes=200 << 30, step_bundles=1 << 24, tune=None):
        h = C.c_void_p()
        check(lib().mcg_server_create(int(budget_bytes), int(step_bundles),
                                      C.byref(tune) if tune is not None else None, C.byref(h)))
        self.handle = h
        self._keep = {}          # request id -> (mat, dom, rows, cols) kept alive until polled

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().mcg_server_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, mat, dom, prob):
        # the output shape without building the material: rows from the
        # tally kind, columns from the domain alone
        ds = dom.spec()
        cols = C.c_int64()
        check(lib().mcg_cell_volumes(C.byref(ds), None, 0, C.byref(cols)))
        rows = tally_rows(prob.tally, prob.size)
        rid = C.c_int64()
        check(lib().mcg_server_submit(self.handle, C.byref(mat.spec), C.byref(ds), C.byref(prob), C.byref(rid)))
        self._keep[rid.value] = (mat, dom, rows, cols.value)
        return rid.value

    def step(self):
        a, q, d = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().mcg_server_step(self.handle, C.byref(a), C.byref(q), C.byref(d)))
        return a.value, q.value, d.value

    def completed(self):
        """Ids the last step retired."""
        ids = (C.c_int64 * 65536)()
        n = C.c_int64()
        check(lib().mcg_server_completed(self.handle, ids, 65536, C.byref(n)))
        return list(ids[:n.value])

    def poll(self, rid):
        """(field, stats) when done, None while running."""
        if rid not in self._keep:      # unknown or already collected: let the server say so
            check(lib().mcg_server_poll(self.handle, rid, None, 0, None))
        _, _, rows, cols = self._keep[rid]
        out = np.zeros(rows * cols)
        st = abi.Stats()
        rc = lib().mcg_server_poll(self.handle, rid, abi.dptr(out), out.size, C.byref(st))
        if rc == abi.MCG_EAGAIN:
            return None
        del self._keep[rid]    # done or failed: the server has forgotten it
        check(rc)
        return out.reshape(cols, rows).T.copy(), st.as_dict()

    def cancel(self, rid):
        check(lib().mcg_server_cancel(self.handle, rid))
        self._keep.pop(rid, None)

    def info(self):
        s = abi.ServerStats()
        check(lib().mcg_server_info(self.handle, C.byref(s)))
        return {n: getattr(s, n) for n, _ in s._fields_}

