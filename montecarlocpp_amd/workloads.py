"""The benchmark workloads, as scene + problem builders.

* ``slab``: 1-D cross-plane slab, 100 nm, isothermal walls at +/-1 K about
  T_ref = 300 K, 1e5 bundles, 50 bins (the CPU-runnable case).
* ``bar``: 1 um x 100 nm x 100 nm bar, ends at +/-10 K, diffuse side walls,
  1e7 bundles, 100 x 1 x 1 grid.
* ``slab_batch``: 256 independent slabs.  Per solve i, four words of the
  Philox block (counter (i, 0, 0, 0), key (0x5EED0002, 0)) give, as
  uniforms u = w / 2^32: thickness L = 10^(-8 + 3 u0) m, T_ref = 100 + 400 u1
  K, lateral-wall specularity p = u2 on an L x L cross-section, bundle count
  N = 10^(4 + 2 u3); walls at +/-1 K (dT = 2 K); bins
  round(20 + 180 log10(L / 10 nm) / 3); seed = i.
* ``porous_film``: 300 x 300 x 100 nm unit cell with a centred cylindrical
  pore of radius 100 nm (diffuse wall), periodic in x (gradient 1 K / 300 nm)
  and in y, diffuse top and bottom walls, 1e8 bundles, 64 x 64 map, FP32
  flight, maxscat 100 (no absorber: scatter count is what ends a bundle; a
  scatter-order tally shows the cumulative flux flat within its noise past
  about 40 scatters).
"""
import math

import numpy as np

from . import abi, engine, materials, scenes

BATCH_KEY = (0x5EED0002, 0)


def slab(nemit=100_000, precision=abi.FP64, seed=1):
    return materials.si_like(300.0), scenes.slab(100e-9, 50), scenes.problem(
        nemit, seed=seed, precision=precision)


def bar(nemit=10_000_000, precision=abi.FP32_FLIGHT, seed=1):
    return (materials.si_like(300.0), scenes.bar(1e-6, 1e-7, 1e-7, 100, 10.0, -10.0),
            scenes.problem(nemit, seed=seed, precision=precision))


def slab_batch_params(n=256):
    out = []
    for i in range(n):
        w = engine.philox([i, 0, 0, 0], list(BATCH_KEY))
        u = [x / 2.0 ** 32 for x in w]
        L = 10.0 ** (-8.0 + 3.0 * u[0])
        T = 100.0 + 400.0 * u[1]
        p = u[2]
        N = int(round(10.0 ** (4.0 + 2.0 * u[3])))
        bins = int(round(20 + 180 * math.log10(L / 1e-8) / 3))
        out.append(dict(index=i, thickness=L, temp=T, specularity=p, nemit=N, bins=bins))
    return out


def slab_batch(n=256, precision=abi.FP32_FLIGHT, base=None):
    """(material, domain, problem) for each solve of the batch."""
    base = base or materials.si_like(300.0)
    cases = []
    for prm in slab_batch_params(n):
        mat = base.at_temperature(prm["temp"])
        dom = scenes.slab(prm["thickness"], prm["bins"], side=abi.PARTIAL,
                          specularity=prm["specularity"])
        prob = scenes.problem(prm["nemit"], seed=prm["index"], precision=precision)
        cases.append((mat, dom, prob))
    return cases


PORE_DIM = (300e-9, 300e-9, 100e-9)
PORE_RADIUS = 100e-9
PORE_MAXSCAT = 100


def porous_film(nemit=100_000_000, precision=abi.FP32_FLIGHT, seed=1, maxscat=PORE_MAXSCAT,
                div=(64, 64, 0)):
    return (materials.si_like(300.0), scenes.porous_film(PORE_DIM, PORE_RADIUS, div, dT=1.0),
            scenes.problem(nemit, seed=seed, precision=precision, maxscat=maxscat))


def k_eff_porous(field, cell_vol, dim=PORE_DIM, dT=1.0):
    """Effective in-plane conductivity of the porous film [W/m/K]: the
    x-flux (Multi row 1) averaged over the whole unit cell, pore included,
    divided by the imposed gradient dT / dim[0]."""
    vol = dim[0] * dim[1] * dim[2]
    qx = float(np.dot(np.asarray(field)[1], np.asarray(cell_vol))) / vol
    return qx / (dT / dim[0])
