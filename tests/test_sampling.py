"""Direction, position and free-path draws (random.cpp:16-44,
boundary.cpp:147-152, 418-431, subdomain.cpp:255-281, material.cpp:215-224),
checked by Kolmogorov-Smirnov and moment tests against their closed forms."""
import numpy as np
import pytest
from scipy import stats

from montecarlocpp_amd import abi, materials, scenes
from oracle import oracle

N = 200_000
P_MIN = 1e-4


def test_isotropic_directions():
    d = oracle.sample_dirs(0, 11, N)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-15)
    assert stats.kstest(d[:, 2], stats.uniform(-1, 2).cdf).pvalue > P_MIN
    phi = np.arctan2(d[:, 1], d[:, 0])
    assert stats.kstest(phi, stats.uniform(-np.pi, 2 * np.pi).cdf).pvalue > P_MIN


def test_lambertian_directions():
    d = oracle.sample_dirs(1, 12, N)
    assert np.all(d[:, 2] >= 0.0)
    sin2 = 1.0 - d[:, 2] ** 2
    assert stats.kstest(sin2, stats.uniform(0, 1).cdf).pvalue > P_MIN
    assert abs(d[:, 2].mean() - 2.0 / 3.0) < 5 * np.sqrt(1 / 18 / N)   # var cos = 1/2 - 4/9


def test_bidirectional_directions():
    d = oracle.sample_dirs(2, 13, N)
    frac = np.mean(d[:, 2] > 0)
    assert abs(frac - 0.5) < 5 * 0.5 / np.sqrt(N)
    assert stats.kstest(1.0 - d[:, 2] ** 2, stats.uniform(0, 1).cdf).pvalue > P_MIN


def test_wall_emission():
    """Slab hot wall at x = 0 (Bac face): positions uniform on the face,
    Lambertian into +x, sign positive; cold wall negative."""
    mat = materials.si_like(300.0)
    L, W = 1e-7, 2e-7
    dom = scenes.slab(L, 10, lateral=W)
    e = oracle.sample_emission(mat, dom, scenes.problem(N, seed=3), N)
    hot = e[:, 9] == 0
    assert abs(hot.mean() - 0.5) < 5 * 0.5 / np.sqrt(N)
    h, c = e[hot], e[~hot]
    assert np.all(h[:, 0] == 0.0) and np.all(np.abs(c[:, 0] - L) < 1e-22)
    for col in (1, 2):
        assert stats.kstest(h[:, col] / W, stats.uniform(0, 1).cdf).pvalue > P_MIN
    assert np.all(h[:, 3] >= 0) and np.all(c[:, 3] <= 0)
    assert stats.kstest(1 - h[:, 3] ** 2, stats.uniform(0, 1).cdf).pvalue > P_MIN
    assert np.all(h[:, 6] == 1) and np.all(c[:, 6] == -1)


def test_volume_emission_and_free_paths():
    """BulkDomain gradient source: positions uniform in the box, direction
    bidirectional along the gradient, sign = -(dir . gradT); free paths are
    exponential with mean v tau of the drawn mode."""
    mat = materials.si_like(300.0)
    dim = (2e-7, 1e-7, 1e-7)
    dom = scenes.bulk(dim, (5, 0, 0), 0.4)
    e = oracle.sample_emission(mat, dom, scenes.problem(N, seed=4, maxscat=10), N)
    for k in range(3):
        assert stats.kstest(e[:, k] / dim[k], stats.uniform(0, 1).cdf).pvalue > P_MIN
    # gradient is -x: sign +1 exactly when the direction has dx > 0
    assert np.all((e[:, 3] > 0) == (e[:, 6] > 0))
    t = oracle.material(mat)
    mfp = (t["vel"] * t["tau"])[e[:, 7].astype(np.int64)]
    assert stats.kstest(e[:, 8] / mfp, stats.expon().cdf).pvalue > P_MIN
