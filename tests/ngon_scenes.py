"""Prism and pyramid cells with N-gon bases (subdomain.h:364-630), shared by
the CPU and GPU tests.  Lengths in units of 100 nm."""
import numpy as np

from montecarlocpp_amd import abi, scenes

f = scenes.face
U = 1e-7

# convex pentagon in the x-y plane, counter-clockwise from the origin vertex
PENTAGON = np.array([[0.0, 0.0], [2.0, 0.0], [2.6, 1.2], [1.0, 2.0], [-0.4, 1.1]]) * U
PRISM_AXIS = np.array([0.3, 0.2, 1.5]) * U
QUAD = np.array([[0.0, 0.0], [2.0, 0.0], [2.2, 1.8], [0.0, 1.6]]) * U
APEX = np.array([0.9, 0.8, 1.5]) * U


def ngon_mat(first, base):
    """mat columns: the axis (prism) or apex (pyramid), then vertices 1..N-1."""
    cols = [first] + [np.array([x, y, 0.0]) for x, y in base[1:]]
    return np.array(cols).T


def regular(n, r=1.0):
    """Regular n-gon of circumradius r (units U) with vertex 0 at the origin."""
    ang = 2 * np.pi * np.arange(n) / n - np.pi / 2
    p = np.stack([np.cos(ang), np.sin(ang)], axis=1) * r * U
    return p - p[0]


def tri_area(a, b, c):
    return 0.5 * abs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))


def shoelace(p):
    x, y = p[:, 0], p[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def side_kind(i):
    return f((abi.DIFF, abi.SPEC, abi.PARTIAL)[i % 3], 0.0, 0.4 if i % 3 == 2 else 0.0)


def prism(base=PENTAGON, axis=PRISM_AXIS, div=0, grad=(0.0, 0.0, 2e6), t_bot=1.0, t_top=-1.0, sides=None):
    """Prism<Isot, Isot, (Diff, Spec, Partial, ...)>; sides may be given."""
    n = len(base)
    d = scenes.Domain("prism")
    faces = [f(abi.ISOT, t_bot) if t_bot else f(abi.SPEC), f(abi.ISOT, t_top) if t_top else f(abi.SPEC)]
    faces += sides if sides is not None else [side_kind(i) for i in range(n)]
    d.add_cell(abi.PRISM, (0.0, 0.0, 0.0), ngon_mat(axis, base), div, faces, grad_t=grad)
    return d


def pyramid(base=QUAD, apex=APEX, div=0, grad=(0.0, 0.0, 0.0), t_bot=1.0, t_side=-1.0):
    """Pyramid<Isot, (Diff, Spec, Partial, Isot)>."""
    n = len(base)
    d = scenes.Domain("pyramid")
    faces = [f(abi.ISOT, t_bot)] + [side_kind(i) for i in range(n - 1)] + [f(abi.ISOT, t_side)]
    d.add_cell(abi.PYRAMID, (0.0, 0.0, 0.0), ngon_mat(apex, base), div, faces, grad_t=grad)
    return d


def box_pair(L=2.0 * U, W=1.5 * U, H=1.0 * U):
    """The same box as a 4-gon prism and as a parallelepiped: isothermal at
    x = 0 (T = 1) and x = L (T = -1), diffuse elsewhere, one tally cell."""
    base = np.array([[0.0, 0.0], [L, 0.0], [L, W], [0.0, W]])
    # prism sides run y = 0, x = L, y = W, x = 0
    sides = [f(abi.DIFF), f(abi.ISOT, -1.0), f(abi.DIFF), f(abi.ISOT, 1.0)]
    p = prism(base, np.array([0.0, 0.0, H]), 0, (0.0, 0.0, 0.0), 0.0, 0.0, sides)
    p.cells[0].face[0].kind = p.cells[0].face[1].kind = abi.DIFF
    b = scenes.Domain("box")
    b.add_box((0.0, 0.0, 0.0), (L, W, H), (0, 0, 0),
              [f(abi.ISOT, 1.0), f(abi.DIFF), f(abi.DIFF), f(abi.ISOT, -1.0), f(abi.DIFF), f(abi.DIFF)])
    return p, b
