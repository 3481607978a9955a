"""Material inputs in the reference's two text formats.

The reference reads a dispersion file and a relaxation-time file
(Material::Material, material.cpp:82-114):

* dispersion: a header ``nw np`` then ``nw`` rows of ``2 + 2*np`` numbers:
  omega, domega, then (vel_p, dos_p) for each branch p;
* relaxation: ``np`` rows of 8 numbers, two mechanisms of (c0, c1, c2, c3)
  with 1/tau = sum c0 * omega**c1 * T**c2 * exp(-c3 / T).

The reference ships none of its material files (main.cpp:269-273 names
``Si_disp.txt`` etc.), so the workloads use an analytic model written into
that same format: isotropic branches with the quadratic dispersion
omega = v*k + c*k**2, sampled on uniform omega bins at their midpoints, with
1/tau = A*omega**4 + B*omega**2*T*exp(-C/T).  The branches are stacked row
blocks (LA rows, then TA rows, ...), each row populated only in its own
branch's (vel, dos) columns, which encodes per-branch omega grids in the
reference's single-omega-column format.
"""
import dataclasses

import numpy as np

from . import abi


@dataclasses.dataclass(frozen=True)
class Branch:
    v: float          # sound velocity (m/s)
    c: float          # quadratic coefficient (m^2/s)


# Si-like isotropic model used by every workload (LA + two TA branches)
SI_LIKE_BRANCHES = (Branch(9.01e3, -2.0e-7), Branch(5.23e3, -2.26e-7), Branch(5.23e3, -2.26e-7))
SI_LIKE_KMAX = 1.157e10
SI_LIKE_RELAX = (1.32e-45, 2e-24, 140.0)    # A (s^3), B (s/K), C (K)


class MaterialInput:
    """Parsed dispersion + relaxation data and a temperature; owns the arrays
    an ``mcg_material_spec`` points at."""

    def __init__(self, disp, relax, temp):
        self.disp = np.ascontiguousarray(disp, dtype=np.float64)
        self.relax = np.ascontiguousarray(relax, dtype=np.float64)
        nw, ncol = self.disp.shape
        np_ = (ncol - 2) // 2
        if ncol != 2 + 2 * np_ or self.relax.shape != (np_, 8):
            raise ValueError("dispersion/relaxation arrays do not match the reference format")
        self.nw, self.np, self.temp = nw, np_, float(temp)
        self.spec = abi.MaterialSpec(nw, np_, abi.dptr(self.disp), abi.dptr(self.relax),
                                     self.temp)

    def at_temperature(self, temp):
        return MaterialInput(self.disp, self.relax, temp)


def quadratic_dispersion(branches=SI_LIKE_BRANCHES, kmax=SI_LIKE_KMAX, nbins=1000):
    """Dispersion rows (nw = nbins * len(branches), np = len(branches))."""
    nb = len(branches)
    rows = np.zeros((nbins * nb, 2 + 2 * nb))
    for b, br in enumerate(branches):
        wmax = br.v * kmax + br.c * kmax ** 2
        dw = wmax / nbins
        w = (np.arange(nbins) + 0.5) * dw
        k = (-br.v + np.sqrt(br.v ** 2 + 4.0 * br.c * w)) / (2.0 * br.c)
        vg = br.v + 2.0 * br.c * k
        dos = k ** 2 / (2.0 * np.pi ** 2 * vg)
        blk = rows[b * nbins:(b + 1) * nbins]
        blk[:, 0] = w
        blk[:, 1] = dw
        blk[:, 2 + 2 * b] = vg
        blk[:, 3 + 2 * b] = dos
    return rows


def power_law_relaxation(nbranch=3, A=SI_LIKE_RELAX[0], B=SI_LIKE_RELAX[1], Cc=SI_LIKE_RELAX[2]):
    """Relaxation rows ``A 4 0 0  B 2 1 C`` for every branch."""
    return np.tile(np.array([A, 4.0, 0.0, 0.0, B, 2.0, 1.0, Cc]), (nbranch, 1))


def si_like(temp=300.0, nbins=1000):
    """The Si-like three-branch model at temperature ``temp``."""
    return MaterialInput(quadratic_dispersion(nbins=nbins), power_law_relaxation(), temp)


def grey(vel, mfp, temp=300.0, nbins=10, omega_max=1e13):
    """A one-branch test material with constant velocity and constant
    relaxation time tau = mfp / vel (the relax law with c1 = 0)."""
    w = (np.arange(nbins) + 0.5) * omega_max / nbins
    rows = np.zeros((nbins, 4))
    rows[:, 0] = w
    rows[:, 1] = omega_max / nbins
    rows[:, 2] = vel
    rows[:, 3] = w ** 2 / (2 * np.pi ** 2 * vel ** 3)
    relax = np.array([[vel / mfp, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]])
    return MaterialInput(rows, relax, temp)


def write_reference_files(mat, disp_path, relax_path):
    """Write ``mat`` in the reference's text formats (material.cpp:87-114)."""
    with open(disp_path, "w") as f:
        f.write(f"{mat.nw} {mat.np}\n")
        for row in mat.disp:
            f.write(" ".join(f"{v:.17g}" for v in row) + "\n")
    with open(relax_path, "w") as f:
        for row in mat.relax:
            f.write(" ".join(f"{v:.17g}" for v in row) + "\n")


def read_reference_files(disp_path, relax_path, temp):
    """Parse the reference's text formats (material.cpp:87-114)."""
    with open(disp_path) as f:
        tok = f.read().split()
    nw, np_ = int(tok[0]), int(tok[1])
    disp = np.array([float(t) for t in tok[2:2 + nw * (2 + 2 * np_)]]).reshape(nw, 2 + 2 * np_)
    with open(relax_path) as f:
        tok = f.read().split()
    relax = np.array([float(t) for t in tok[:8 * np_]]).reshape(np_, 8)
    return MaterialInput(disp, relax, temp)
