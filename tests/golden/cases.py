"""The seeded deterministic-mode cases of the regression vectors in
oracle_vectors.npz: one small solve per feature family (1-D slab with
isothermal and specular walls, diffuse bar, periodic film with a volume
source, Inter-joined tee, porous cell on a 2-D map, triangular prism, a
cumulative tally kind, a pentagonal prism with a volume source, the octet
truss).  The reference holds no golden vectors of its own,
so these record the FP64 oracle's per-bundle counts and tallies; they catch
any change to the bundle streams or the arithmetic of either side."""
import numpy as np

from montecarlocpp_amd import abi, materials, octet, scenes
from tests import ngon_scenes

NBUNDLES = 2000


def cases():
    f = scenes.face
    si = materials.si_like(300.0)
    prism = scenes.Domain("prism")
    prism.add_cell(abi.TRIPRISM, (0.0, 0.0, 0.0), np.array([[2e-7, 0, 0], [0, 1.5e-7, 0], [0, 0, 1e-7]]),
                   (5, 4, 3), [f(abi.ISOT, 1.0), f(abi.DIFF), f(abi.SPEC), f(abi.PARTIAL, 0.0, 0.4), f(abi.DIFF)])
    return {
        "slab": (si, scenes.slab(1e-7, 50), scenes.problem(NBUNDLES, seed=1)),
        "bar": (si, scenes.bar(1e-6, 1e-7, 1e-7, 20), scenes.problem(NBUNDLES, seed=2)),
        "film": (si, scenes.film((2e-7, 5e-8, 2e-7), (0, 6, 0), 0.2), scenes.problem(NBUNDLES, maxscat=40, seed=3)),
        "tee": (si, scenes.tee((1e-7, 5e-8, 5e-8, 6e-8, 4e-8), (4, 2, 2, 3, 1), 1.0),
                scenes.problem(NBUNDLES, maxscat=40, seed=4)),
        "porous": (si, scenes.porous_film(div=(8, 8, 0)), scenes.problem(NBUNDLES, maxscat=30, seed=5)),
        "prism": (si, prism, scenes.problem(NBUNDLES, maxscat=60, seed=6)),
        "cumflux": (materials.si_like(200.0), scenes.slab(2e-7, 10), scenes.problem(NBUNDLES, tally=abi.CUMFLUX,
                                                                                      size=4, seed=7)),
        "ngon": (si, ngon_scenes.prism(), scenes.problem(NBUNDLES, maxscat=60, seed=8)),
        "octet": (si, octet.octet((6e-7, 1e-7, 1e-7, 1e-8), (4, 2, 2, 1), 1.0),
                  scenes.problem(NBUNDLES, maxscat=40, seed=9)),
    }
