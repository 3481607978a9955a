"""ctypes mirror of the C ABI declared in include/mcg.h.

Only data layouts and constants live here; loading the engine library is
done by :mod:`montecarlocpp_amd.engine`.
"""
import ctypes as C

import numpy as np

MCG_OK, MCG_EINVAL, MCG_ENOMEM, MCG_EDEVICE, MCG_ERANGE, MCG_EAGAIN, MCG_ENOTFOUND = range(7)

# boundary kinds (boundary.h:137-281 plus the partial-specularity extension)
SPEC, DIFF, ISOT, PERI, INTER, PARTIAL = range(6)
# cell kinds (subdomain.h:128-630); PRISM / PYRAMID have an N-gon base
PARALLELEPIPED, TRIPRISM, TETRAHEDRON, PRISM, PYRAMID = range(5)
MAX_SIDES = 9                  # N of a prism's or pyramid's base: 4..MAX_SIDES
MAX_FACES = MAX_SIDES + 2
# tally kinds (problem.h:157-239)
TEMP, FLUX, MULTI, CUMTEMP, CUMFLUX = range(5)
# precision
FP64, FP32_FLIGHT = range(2)
# bundle fates
ABSORBED, ESCAPED, MAXSCAT, MAXLOOP = range(4)
# doubles per trajectory record (mcg_trace / mcg_trace_from)
TRACE_WIDTH = 17

NFACE = {PARALLELEPIPED: 6, TRIPRISM: 5, TETRAHEDRON: 4}


def nface(cell, nside=None):
    """Faces of a cell kind; a prism or pyramid needs its base's sides."""
    if cell == PRISM:
        return nside + 2
    if cell == PYRAMID:
        return nside + 1
    return NFACE[cell]

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)


class MaterialSpec(C.Structure):
    _fields_ = [("nw", C.c_int64), ("np", C.c_int64), ("disp", _dp), ("relax", _dp),
                ("temp", C.c_double)]


class MaterialTables(C.Structure):
    _fields_ = [("tau", _dp), ("vel", _dp), ("energy_pdf", _dp), ("flux_pdf", _dp),
                ("scat_pdf", _dp), ("energy_sum", C.c_double), ("flux_sum", C.c_double),
                ("scat_sum", C.c_double), ("cond", C.c_double), ("nmodes", C.c_int64)]


class FaceSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("temp", C.c_double),
                ("specularity", C.c_double)]


class SubdomainSpec(C.Structure):
    _fields_ = [("cell", C.c_int32), ("nface", C.c_int32), ("origin", C.c_double * 3),
                ("mat", C.c_double * (3 * MAX_SIDES)), ("div", C.c_int64 * 3),
                ("grad_t", C.c_double * 3), ("face", FaceSpec * MAX_FACES)]


class PairSpec(C.Structure):
    _fields_ = [("sdom_a", C.c_int32), ("face_a", C.c_int32), ("sdom_b", C.c_int32),
                ("face_b", C.c_int32), ("transl", C.c_double * 3), ("rot", C.c_double * 9)]


class PoreSpec(C.Structure):
    _fields_ = [("sdom", C.c_int32), ("kind", C.c_int32), ("center", C.c_double * 2),
                ("radius", C.c_double), ("specularity", C.c_double)]


class DomainSpec(C.Structure):
    _fields_ = [("nsdom", C.c_int32), ("npair", C.c_int32), ("npore", C.c_int32),
                ("reserved", C.c_int32), ("sdom", C.POINTER(SubdomainSpec)),
                ("pair", C.POINTER(PairSpec)), ("pore", C.POINTER(PoreSpec))]


class ProblemSpec(C.Structure):
    _fields_ = [("tally", C.c_int32), ("precision", C.c_int32), ("nemit", C.c_int64),
                ("maxscat", C.c_int64), ("maxloop", C.c_int64), ("size", C.c_int64),
                ("seed", C.c_uint64), ("first_bundle", C.c_int64), ("bundle_count", C.c_int64)]


class ProblemInfo(C.Structure):
    _fields_ = [("nemit", C.c_int64), ("rows", C.c_int64), ("cols", C.c_int64),
                ("maxloop", C.c_int64), ("nemitter", C.c_int64), ("power", C.c_double),
                ("weight_sum", C.c_double)]


class Counts(C.Structure):
    _fields_ = [("nscat", _u32p), ("nbdry", _u32p), ("nstep", _u32p), ("fate", _u32p)]


class Stats(C.Structure):
    _fields_ = [("bundles", C.c_int64), ("steps", C.c_int64), ("escapes", C.c_int64),
                ("absorbed", C.c_int64), ("truncated", C.c_int64), ("kernel_ms", C.c_double),
                ("wall_ms", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# engines of a single solve
ENGINE_HISTORY, ENGINE_EVENT = range(2)


class Tune(C.Structure):
    _fields_ = [("block", C.c_int32), ("per_thread", C.c_int32), ("chunk", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("lds_tables", C.c_int32), ("engine", C.c_int32),
                ("compact_every", C.c_int32), ("defer_walks", C.c_int32), ("live", C.c_int64)]


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def u32ptr(a):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_u32p)


class CountsBuffers:
    """Per-bundle parity outputs (nscat, nbdry, nstep, fate) for n bundles."""

    def __init__(self, n):
        self.nscat = np.zeros(n, np.uint32)
        self.nbdry = np.zeros(n, np.uint32)
        self.nstep = np.zeros(n, np.uint32)
        self.fate = np.zeros(n, np.uint32)
        self.struct = Counts(u32ptr(self.nscat), u32ptr(self.nbdry), u32ptr(self.nstep),
                             u32ptr(self.fate))


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_int64), ("active", C.c_int64), ("bytes_used", C.c_uint64),
                ("bytes_budget", C.c_uint64), ("kernel_ms", C.c_double), ("steps", C.c_int64)]

