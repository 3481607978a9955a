"""Pins for the random-number generator and the portable FP64 math.

* Philox4x32-10: the known-answer vectors published with the Random123
  library (Salmon et al., SC'11), for the checker's generator and the engine's
  host copy of the device rounds; plus a cross-check against rocRAND's
  open-source implementation of the same rounds when hipcc is available.
* log / sin(pi x) / cos(pi x): within a few ulp of glibc / long double.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from montecarlocpp_amd import engine
from oracle import oracle

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert tuple(oracle.philox(ctr, key)) == expect
    assert tuple(engine.philox(ctr, key)) == expect


def test_philox_matches_rocrand(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc) or not os.path.exists("/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"):
        pytest.skip("rocRAND headers / hipcc not available")
    src = tmp_path / "r.cpp"
    src.write_text(r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
struct P : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
    P p; unsigned s = 12345;
    for (int i = 0; i < 64; ++i) {
        s = s * 1664525u + 1013904223u; unsigned a = s; s = s * 1664525u + 1013904223u; unsigned b = s;
        s = s * 1664525u + 1013904223u; unsigned c = s; s = s * 1664525u + 1013904223u; unsigned d = s;
        s = s * 1664525u + 1013904223u; unsigned e = s; s = s * 1664525u + 1013904223u; unsigned f = s;
        uint4 o = p.run(make_uint4(a, b, c, d), make_uint2(e, f));
        printf("%u %u %u %u %u %u %u %u %u %u\n", a, b, c, d, e, f, o.x, o.y, o.z, o.w);
    }
}
""")
    exe = tmp_path / "r"
    r = subprocess.run([hipcc, "-O1", "-std=c++17", str(src), "-o", str(exe)], capture_output=True,
                       text=True)
    if r.returncode != 0:
        pytest.skip("could not build the rocRAND cross-check: " + r.stderr[-300:])
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    n = 0
    for line in out:
        if not line.strip():
            continue
        v = [int(t) for t in line.split()]
        assert oracle.philox(v[0:4], v[4:6]) == v[6:10]
        assert engine.philox(v[0:4], v[4:6]) == v[6:10]
        n += 1
    assert n == 64


def ulps(a, b):
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    return np.abs(ia - ib)


def test_log_accuracy():
    rng = np.random.default_rng(1)
    u = rng.random(2_000_000)
    x = np.concatenate([1.0 - u, rng.random(1000) * 1e-300 + 1e-300, [1.0, 0.5, 2.0, 1e300]])
    got = oracle.math_eval(0, x)
    assert ulps(got, np.log(x)).max() <= 2
    assert oracle.math_eval(0, np.array([1.0]))[0] == 0.0


def test_sincospi_accuracy():
    rng = np.random.default_rng(2)
    r = rng.random(2_000_000) * 2.0 - 1.0
    r = np.concatenate([r, [-1.0, -0.75, -0.5, -0.25, 0.0, 0.25, 0.5, 0.75]])
    s = oracle.math_eval(1, r)
    c = oracle.math_eval(2, r)
    pi_l = np.longdouble("3.14159265358979323846264338327950288")
    ref_s = np.sin(pi_l * r.astype(np.longdouble)).astype(np.float64)
    ref_c = np.cos(pi_l * r.astype(np.longdouble)).astype(np.float64)
    # the long-double reference itself is only accurate away from the zeros
    ms = np.abs(ref_s) > 1e-3
    mc = np.abs(ref_c) > 1e-3
    assert ulps(s[ms], ref_s[ms]).max() <= 3
    assert ulps(c[mc], ref_c[mc]).max() <= 3
    assert np.abs(s ** 2 + c ** 2 - 1.0).max() <= 1e-15
    # exact at the quarter turns
    assert oracle.math_eval(1, np.array([0.5]))[0] == 1.0
    assert oracle.math_eval(2, np.array([1.0]))[0] == -1.0
