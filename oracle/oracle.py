"""Python binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, never by the engine.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from montecarlocpp_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_solve.argtypes = [C.POINTER(abi.MaterialSpec), C.POINTER(abi.DomainSpec),
                                   C.POINTER(abi.ProblemSpec), C.c_int32,
                                   C.POINTER(C.c_double), C.c_int64, C.POINTER(abi.Counts),
                                   C.POINTER(abi.Stats)]
        L.oracle_plan.argtypes = [C.POINTER(abi.MaterialSpec), C.POINTER(abi.DomainSpec),
                                  C.POINTER(abi.ProblemSpec), C.POINTER(abi.ProblemInfo),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int64]
        L.oracle_material.argtypes = [C.POINTER(abi.MaterialSpec), C.POINTER(abi.MaterialTables)]
        L.oracle_alias.argtypes = [C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32)]
        L.oracle_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32)]
        L.oracle_philox.restype = None
        L.oracle_cell_volumes.argtypes = [C.POINTER(abi.DomainSpec), C.POINTER(C.c_double),
                                          C.c_int64, C.POINTER(C.c_int64)]
        L.oracle_accumulate.argtypes = [C.POINTER(abi.DomainSpec), C.c_int32,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64]
        L.oracle_sample_dirs.argtypes = [C.c_int32, C.c_uint64, C.c_int64, C.POINTER(C.c_double)]
        L.oracle_sample_dirs.restype = None
        L.oracle_sample_emission.argtypes = [C.POINTER(abi.MaterialSpec),
                                             C.POINTER(abi.DomainSpec),
                                             C.POINTER(abi.ProblemSpec), C.c_int64,
                                             C.POINTER(C.c_double)]
        L.oracle_trace.argtypes = [C.POINTER(abi.MaterialSpec), C.POINTER(abi.DomainSpec),
                                   C.POINTER(abi.ProblemSpec), C.c_int64, C.POINTER(C.c_double),
                                   C.c_int64, C.POINTER(C.c_int64)]
        L.oracle_trace_from.argtypes = [C.POINTER(abi.MaterialSpec), C.POINTER(abi.DomainSpec),
                                        C.POINTER(abi.ProblemSpec), C.c_int64, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64,
                                        C.POINTER(C.c_int64)]
        L.oracle_trajectory_from.argtypes = [C.POINTER(abi.MaterialSpec), C.POINTER(abi.DomainSpec),
                                             C.POINTER(abi.ProblemSpec), C.c_int64, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64,
                                             C.POINTER(C.c_int64)]
        L.oracle_math.argtypes = [C.c_int32, C.POINTER(C.c_double), C.c_int64,
                                  C.POINTER(C.c_double)]
        L.oracle_math.restype = None
        L.oracle_disk_area.argtypes = [C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_double)]
        L.oracle_disk_area.restype = None
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(f"oracle status {rc}: {lib().oracle_last_error().decode()}")


def plan(mat, dom, prob):
    info = abi.ProblemInfo()
    w = np.zeros(256)
    n = np.zeros(256, np.int64)
    _check(lib().oracle_plan(C.byref(mat.spec), C.byref(dom.spec()), C.byref(prob),
                             C.byref(info), abi.dptr(w), n.ctypes.data_as(C.POINTER(C.c_int64)),
                             256))
    k = info.nemitter
    return info, w[:k].copy(), n[:k].copy()


def solve(mat, dom, prob, threads=0, counts=False):
    """Run the oracle; returns (field rows x cols, stats dict, CountsBuffers|None)."""
    info, _, _ = plan(mat, dom, prob)
    out = np.zeros(info.rows * info.cols)
    nb = prob.bundle_count if prob.bundle_count else info.nemit - prob.first_bundle
    cb = abi.CountsBuffers(nb) if counts else None
    st = abi.Stats()
    _check(lib().oracle_solve(C.byref(mat.spec), C.byref(dom.spec()), C.byref(prob), threads,
                              abi.dptr(out), out.size,
                              C.byref(cb.struct) if cb else None, C.byref(st)))
    return out.reshape(info.cols, info.rows).T.copy(), st.as_dict(), cb


def material(mat):
    n = mat.nw * mat.np
    arrs = {k: np.zeros(n) for k in ("tau", "vel", "energy_pdf", "flux_pdf", "scat_pdf")}
    t = abi.MaterialTables(*(abi.dptr(arrs[k]) for k in
                             ("tau", "vel", "energy_pdf", "flux_pdf", "scat_pdf")))
    _check(lib().oracle_material(C.byref(mat.spec), C.byref(t)))
    arrs.update(energy_sum=t.energy_sum, flux_sum=t.flux_sum, scat_sum=t.scat_sum,
                cond=t.cond, nmodes=t.nmodes)
    return arrs


def alias(pdf):
    pdf = np.ascontiguousarray(pdf, np.float64)
    thr = np.zeros(pdf.size, np.uint32)
    al = np.zeros(pdf.size, np.uint32)
    _check(lib().oracle_alias(abi.dptr(pdf), pdf.size, abi.u32ptr(thr), abi.u32ptr(al)))
    return thr, al


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox(c, k, o)
    return list(o)


def cell_volumes(dom):
    out = np.zeros(1 << 16)
    n = C.c_int64()
    _check(lib().oracle_cell_volumes(C.byref(dom.spec()), abi.dptr(out), out.size, C.byref(n)))
    return out[:n.value].copy()


def accumulate(dom, sdom, bpos, epos, amount, ncols):
    amount = np.ascontiguousarray(amount, np.float64)
    rows = amount.size
    field = np.zeros(rows * ncols)
    b = np.ascontiguousarray(bpos, np.float64)
    e = np.ascontiguousarray(epos, np.float64)
    _check(lib().oracle_accumulate(C.byref(dom.spec()), sdom, abi.dptr(b), abi.dptr(e), rows,
                                   abi.dptr(amount), abi.dptr(field), field.size))
    return field.reshape(ncols, rows).T.copy()


def sample_dirs(kind, seed, n):
    out = np.zeros(3 * n)
    lib().oracle_sample_dirs(kind, seed, n, abi.dptr(out))
    return out.reshape(n, 3)


def sample_emission(mat, dom, prob, n):
    out = np.zeros(10 * n)
    _check(lib().oracle_sample_emission(C.byref(mat.spec), C.byref(dom.spec()), C.byref(prob), n,
                                        abi.dptr(out)))
    return out.reshape(n, 10)


def trace(mat, dom, prob, bundle, cap=4096):
    rec = np.zeros(cap * abi.TRACE_WIDTH)
    n = C.c_int64()
    _check(lib().oracle_trace(C.byref(mat.spec), C.byref(dom.spec()), C.byref(prob), int(bundle),
                              abi.dptr(rec), cap, C.byref(n)))
    return rec.reshape(cap, abi.TRACE_WIDTH)[:n.value].copy()


def math_eval(fn, x):
    """fn 0: log, 1: sin(pi x), 2: cos(pi x) with the portable FP64 functions."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    lib().oracle_math(fn, abi.dptr(x), x.size, abi.dptr(out))
    return out



def disk_area(q):
    """Area of rectangle [x0,x1]x[y0,y1] inside disk (cx,cy,r); q is (n, 7)."""
    q = np.ascontiguousarray(q, np.float64).reshape(-1, 7)
    out = np.zeros(len(q))
    lib().oracle_disk_area(abi.dptr(q), len(q), abi.dptr(out))
    return out


def trace_from(mat, dom, prob, pos, dir=None, prop_cell=-1, cap=4096):
    """TrajProblem with a given start: the checker's records, as engine.Scene.trace_from."""
    rec = np.zeros(cap * abi.TRACE_WIDTH)
    n = C.c_int64()
    p = np.ascontiguousarray(pos, dtype=np.float64)
    d = None if dir is None else np.ascontiguousarray(dir, dtype=np.float64)
    _check(lib().oracle_trace_from(C.byref(mat.spec), C.byref(dom.spec()), C.byref(prob), int(prop_cell),
                                   abi.dptr(p), None if d is None else abi.dptr(d), abi.dptr(rec), cap,
                                   C.byref(n)))
    return rec.reshape(cap, abi.TRACE_WIDTH)[:n.value].copy()


def trajectory_from(mat, dom, prob, pos, dir=None, prop_cell=-1, cap=1 << 16):
    """The points TrajProblem::solve returns for a given start, collected as
    TrkPhonon does (the start, then every new position): array (3, n)."""
    pts = np.zeros(3 * cap)
    n = C.c_int64()
    p = np.ascontiguousarray(pos, dtype=np.float64)
    d = None if dir is None else np.ascontiguousarray(dir, dtype=np.float64)
    _check(lib().oracle_trajectory_from(C.byref(mat.spec), C.byref(dom.spec()), C.byref(prob), int(prop_cell),
                                        abi.dptr(p), None if d is None else abi.dptr(d), abi.dptr(pts), cap,
                                        C.byref(n)))
    assert n.value <= cap, "trajectory longer than the point buffer"
    return pts[:3 * n.value].reshape(n.value, 3).T.copy()

