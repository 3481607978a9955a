"""OctetDomain (domain.h:299-350, domain.cpp:591-1225): one periodic unit of
an octet nanotruss, 42 convex cells -- 4-gon prisms and pyramids, triangular
prisms and parallelepipeds -- joined by 71 interface pairs and closed by 5
periodic pairs that map x to -x (rotation diag(-1, 1, 1), translation
(s/2, 0, -s/2)).  The ten struts between the nodes (cells 16-25) carry the
volumetric gradient source and the tallies; the node cells carry neither.

The cell table below restates the reference's constructor in this package's
terms: each cell's kind, origin, edge columns (the reference's mat, column by
column), tally divisions and face kinds in the reference's container order.
The numbers are the reference's geometry (a table that has to be alike for
the same truss to come out).
"""
import math

import numpy as np

from . import abi, scenes

_F = {"I": abi.INTER, "D": abi.DIFF, "S": abi.SPEC, "P": abi.PERI}

# face kinds per cell, in container order (domain.h:302-348):
# prism Bot Top Sid0-3, pyramid Bot Sid0-3, triangular prism Bac Lef Bot Dia
# Top, parallelepiped Bac Lef Bot Fro Rig Top
_FACES = [
    "IPIDSD", "DIIDSI", "DDIISS", "IPDIID", "IIDIII", "ISDDII",
    "SIPDID", "IDPDII", "DDPIII", "DDIIII", "IDIIDI", "SIDIDI", "DIISI", "IDISI", "IIDSI", "DDIISS",
    "DSIDII", "IIIDDI", "IDIIDI", "DIIIDI", "DSIDII",
    "DSIDII", "DIIIDI", "IDIIDI", "IIIDDI", "DSIDII",
    "DDIISS", "IIDSI", "IDISI", "DIISI", "SIDIDI", "IDIIDI", "DDIIII", "DDPIII", "IDPDII", "SIPDID",
    "ISDDII", "IIDIII", "IPDIID", "DDIISS", "DIIDSI", "IPIDSD",
]

# interface pairs (sdom, face, sdom, face), domain.cpp:595-675
_INTER = [
    (0, 0, 1, 1), (1, 5, 2, 3), (0, 2, 3, 4), (1, 2, 4, 4), (2, 2, 5, 5), (3, 0, 4, 1), (4, 5, 5, 4),
    (3, 3, 8, 5), (4, 3, 8, 5), (4, 0, 9, 2), (5, 0, 9, 2), (5, 0, 12, 4),
    (6, 1, 7, 0), (7, 5, 8, 3), (9, 4, 10, 2), (10, 0, 11, 1), (9, 5, 12, 1), (9, 5, 13, 0),
    (10, 5, 14, 1), (11, 5, 15, 2), (12, 2, 13, 2), (13, 4, 14, 4), (14, 0, 15, 3),
    (6, 4, 16, 5), (7, 4, 17, 5), (8, 4, 18, 5), (9, 3, 18, 5), (10, 3, 19, 5), (11, 3, 20, 5),
    (16, 4, 17, 1), (17, 0, 18, 3), (18, 0, 19, 3), (19, 1, 20, 4),
    (16, 2, 21, 2), (17, 2, 22, 2), (18, 2, 23, 2), (19, 2, 24, 2), (20, 2, 25, 2),
    (21, 4, 22, 1), (22, 3, 23, 0), (23, 3, 24, 0), (24, 1, 25, 4),
    (21, 5, 30, 3), (22, 5, 31, 3), (23, 5, 32, 3), (23, 5, 33, 4), (24, 5, 34, 4), (25, 5, 35, 4),
    (26, 3, 27, 0), (27, 4, 28, 4), (28, 2, 29, 2), (26, 2, 30, 5), (27, 1, 31, 5), (28, 0, 32, 5),
    (29, 1, 32, 5), (30, 1, 31, 0), (31, 2, 32, 4), (33, 3, 34, 5), (34, 0, 35, 1),
    (29, 4, 36, 0), (32, 2, 36, 0), (32, 2, 37, 0), (33, 5, 37, 3), (33, 5, 38, 3),
    (36, 4, 37, 5), (37, 1, 38, 0), (36, 5, 39, 2), (37, 4, 40, 2), (38, 4, 41, 2), (39, 3, 40, 5),
    (40, 1, 41, 0),
]

# periodic pairs (domain.cpp:678-687)
_PERI = [(0, 1, 41, 1), (3, 1, 38, 1), (6, 2, 35, 2), (7, 2, 34, 2), (8, 2, 33, 2)]


def cells(dim, div):
    """The 42 cells: (cell kind, origin, edge columns, divisions, gradient
    flag) and the 52 check points (domain.cpp:693-1225)."""
    s, t = float(dim[0]), float(dim[3])
    r2 = math.sqrt(2.0)
    a = math.pi / 5.0 * (dim[1] + dim[2]) - (4.0 - math.pi) / 5.0 * t
    b = a / 4.0
    if not b > t:
        raise ValueError("Invalid dimensions")
    if not a > (r2 + 1.0) * (b + t):
        raise ValueError("Invalid dimensions")
    if not s / 2.0 > a + (r2 + 1.0) * (b + t) + t:
        raise ValueError("Invalid dimensions")
    d0, d1, d2, d3 = (int(v) for v in div)
    P, T, R, Y = abi.PRISM, abi.TRIPRISM, abi.PARALLELEPIPED, abi.PYRAMID
    out, pts = [], []

    def v(x, y, z):
        return np.array([x, y, z], dtype=np.float64)

    def add(kind, o, cols, dv=None, grad=False):
        out.append((kind, o, np.array(cols).T, dv, grad))
        return o, cols

    def mid(o, cols):                      # o + (mat.col(0) + mat.col(2)) / 2
        pts.append(o + (cols[0] + cols[2]) / 2.0)

    # horizontal top: 00-05
    c = [v((r2 * t), 0, 0),
         v(s / 4 - (2 + r2) / 2 * b - (2 - r2) / 2 * t, s / 4 - (2 + r2) / 2 * (b + t), 0),
         v(s / 4 - (2 + r2) / 2 * b - t, s / 4 - (2 + r2) / 2 * b - t, 0)]
    mid(*add(P, v((r2 + 1) * b + t, b + t, -a), [v(0, 0, a)] + c))
    mid(*add(P, v((r2 + 1) * b + t, b + t, -a - t), [v(0, 0, t)] + c))
    mid(*add(P, v(b + t, b + t, -a - t),
             [v(0, 0, t), v(r2 * b, 0, 0), v(s / 4 - (2 - r2) / 2 * b - t, s / 4 - (2 + r2) / 2 * b - t, 0),
              v(s / 4 - b - t, s / 4 - b - t, 0)]))
    c = [v((r2 + 1) * t, 0, 0), v((r2 + 1) * t, t, 0), v(t, t, 0)]
    mid(*add(P, v((r2 + 1) * b, b, -a), [v(0, 0, a)] + c))
    mid(*add(P, v((r2 + 1) * b, b, -a - t), [v(0, 0, t)] + c))
    o, m = add(R, v(b, b, -a - t), [v(0, 0, t), v(r2 * b, 0, 0), v(t, t, 0)])
    u = v(t, 0, 0)
    pts.append(o + m[0] / 2 + (m[2] + u) / 3)
    pts.append(o + (m[0] + m[1] + m[2] + u) / 2)

    # node top: 06-15
    c = [v(r2 * t, 0, 0), v((r2 + 1) / 2 * b + (r2 + 1) * t, 0, -(r2 + 1) / 2 * b - t),
         v((r2 + 1) / 2 * b + (2 + r2) / 2 * t, 0, -(r2 + 1) / 2 * b - (2 + r2) / 2 * t)]
    mid(*add(P, v(a, 0, 0), [v(0, b, 0)] + c))
    mid(*add(P, v(a, b, 0), [v(0, t, 0)] + c))
    o, m = add(P, v((r2 + 1) * (b + t), b, 0),
               [v(0, t, 0), v(a - (r2 + 1) * (b + t), 0, 0),
                v(a - (r2 + 1) / 2 * b - t / r2, 0, -(r2 + 1) / 2 * b - (2 + r2) / 2 * t), v(0, 0, -a - t)])
    w = v(0, 0, -a)
    pts.append(o + (m[0] + m[1] + w) / 2)
    pts.append(o + (m[0] + m[2] + w) / 2)
    o, m = add(P, v(b + t, b, -a - t),
               [v(0, t, 0), v(r2 * (b + t), 0, 0), v((r2 - 1) / 2 * b + t / r2, 0, -(r2 + 1) / 2 * b - t / r2),
                v(0, 0, -b)])
    u, w = v(r2 * b, 0, 0), v(0, 0, -b + t)
    pts.append(o + (m[0] + m[2] + u) / 2)
    pts.append(o + (m[0] + m[2] + w) / 2)
    pts.append(o + (m[0] + u + w) / 2)
    c = [v((r2 - 1) / 2 * b + t / r2, 0, -(r2 - 1) / 2 * b - t / r2),
         v((r2 - 1) / 2 * b, 0, -(r2 - 1) / 2 * b - r2 * t), v(0, 0, -r2 * t)]
    mid(*add(P, v(b + t, b, -a - b - t), [v(0, t, 0)] + c))
    mid(*add(P, v(b + t, 0, -a - b - t), [v(0, b, 0)] + c))
    o, m = add(T, v(b + t, b, -a - b), [v(0, t, 0), v(-t, 0, 0), v(0, 0, b - t)])
    pts.append(o + (m[0] + m[1]) / 3 + m[2] / 2)
    o, m = add(Y, v(b + t, b, -a - b - t), [v(-t, 0, t), v(0, 0, t), v(0, t, t), v(0, t, 0)])
    pts.append(o + sum(m) / 5)
    o, m = add(T, v(b + t, b, -a - b - (r2 + 1) * t), [v(0, t, 0), v(-t, 0, t), v(0, 0, r2 * t)])
    pts.append(o + (m[0] + m[1]) / 3 + m[2] / 2)
    mid(*add(P, v(b + t, 0, -a - b - (r2 + 1) * t),
             [v(0, 0, r2 * t), v(0, b, 0), v(-t, b, t), v(-b - t, 0, b + t)]))

    # struts: 16-20 (top) and 21-25 (bottom), with the gradient source
    W = (r2 + 1) / 2 * b + (2 + r2) / 2 * t
    K = s / 4 - a / 2 - W

    def strut(o, c0, c1, dv, extra=None):
        o, m = add(R, o, [c0, c1, v(-K, 0, K) if c0[0] > 0 else v(K, 0, -K)], dv, True)
        if extra is None:
            pts.append(o + sum(m) / 2)
        else:
            pts.append(o + (m[1] + m[2] + extra) / 2)
            pts.append(o + (m[0] + m[1] + m[2] + extra) / 2)

    h = t / r2
    strut(v(s / 4 + a / 2, 0, -s / 4 + a / 2), v(h, 0, h), v(0, b, 0), (d3, d2, d0))
    strut(v(s / 4 + a / 2, b, -s / 4 + a / 2), v(h, 0, h), v(0, t, 0), (d3, d3, d0))
    strut(v(s / 4 - a / 2, b, -s / 4 - a / 2), v(a, 0, a), v(0, t, 0), (d1, d3, d0),
          extra=v((r2 + 1) / 2 * b + h, 0, (r2 + 1) / 2 * b + h))
    strut(v(s / 4 - a / 2 - h, b, -s / 4 - a / 2 - h), v(h, 0, h), v(0, t, 0), (d3, d3, d0))
    strut(v(s / 4 - a / 2 - h, 0, -s / 4 - a / 2 - h), v(h, 0, h), v(0, b, 0), (d3, d2, d0))
    strut(v(s / 4 + a / 2 + h, 0, -s / 4 + a / 2 + h), v(-h, 0, -h), v(0, b, 0), (d3, d2, d0))
    strut(v(s / 4 + a / 2 + h, b, -s / 4 + a / 2 + h), v(-h, 0, -h), v(0, t, 0), (d3, d3, d0))
    strut(v(s / 4 + a / 2, b, -s / 4 + a / 2), v(-a, 0, -a), v(0, t, 0), (d1, d3, d0),
          extra=v(-(r2 + 1) / 2 * b - h, 0, -(r2 + 1) / 2 * b - h))
    strut(v(s / 4 - a / 2, b, -s / 4 - a / 2), v(-h, 0, -h), v(0, t, 0), (d3, d3, d0))
    strut(v(s / 4 - a / 2, 0, -s / 4 - a / 2), v(-h, 0, -h), v(0, b, 0), (d3, d2, d0))

    # node bottom: 26-35
    z0 = -s / 2 + a
    mid(*add(P, v(s / 2 - b - t, 0, z0 + b + (r2 + 1) * t),
             [v(0, 0, -r2 * t), v(0, b, 0), v(t, b, -t), v(b + t, 0, -b - t)]))
    o, m = add(T, v(s / 2 - b - t, b, z0 + b + (r2 + 1) * t), [v(0, t, 0), v(t, 0, -t), v(0, 0, -r2 * t)])
    pts.append(o + (m[0] + m[1]) / 3 + m[2] / 2)
    o, m = add(Y, v(s / 2 - b - t, b, z0 + b + t), [v(t, 0, -t), v(0, 0, -t), v(0, t, -t), v(0, t, 0)])
    pts.append(o + sum(m) / 5)
    o, m = add(T, v(s / 2 - b - t, b, z0 + b), [v(0, t, 0), v(t, 0, 0), v(0, 0, -b + t)])
    pts.append(o + (m[0] + m[1]) / 3 + m[2] / 2)
    c = [v(-(r2 - 1) / 2 * b - t / r2, 0, (r2 - 1) / 2 * b + t / r2),
         v(-(r2 - 1) / 2 * b, 0, (r2 - 1) / 2 * b + r2 * t), v(0, 0, r2 * t)]
    mid(*add(P, v(s / 2 - b - t, 0, z0 + b + t), [v(0, b, 0)] + c))
    mid(*add(P, v(s / 2 - b - t, b, z0 + b + t), [v(0, t, 0)] + c))
    o, m = add(P, v(s / 2 - b - t, b, z0 + t),
               [v(0, t, 0), v(-r2 * (b + t), 0, 0), v(-(r2 - 1) / 2 * b - t / r2, 0, (r2 + 1) / 2 * b + t / r2),
                v(0, 0, b)])
    u, w = v(-r2 * b, 0, 0), v(0, 0, b - t)
    pts.append(o + (m[0] + m[2] + u) / 2)
    pts.append(o + (m[0] + m[2] + w) / 2)
    pts.append(o + (m[0] + u + w) / 2)
    o, m = add(P, v(s / 2 - (r2 + 1) * (b + t), b, -s / 2),
               [v(0, t, 0), v(-a + (r2 + 1) * (b + t), 0, 0),
                v(-a + (r2 + 1) / 2 * b + t / r2, 0, (r2 + 1) / 2 * b + (2 + r2) / 2 * t), v(0, 0, a + t)])
    w = v(0, 0, a)
    pts.append(o + (m[0] + m[1] + w) / 2)
    pts.append(o + (m[0] + m[2] + w) / 2)
    c = [v(-r2 * t, 0, 0), v(-(r2 + 1) / 2 * b - (r2 + 1) * t, 0, (r2 + 1) / 2 * b + t),
         v(-(r2 + 1) / 2 * b - (2 + r2) / 2 * t, 0, (r2 + 1) / 2 * b + (2 + r2) / 2 * t)]
    mid(*add(P, v(s / 2 - a, b, -s / 2), [v(0, t, 0)] + c))
    mid(*add(P, v(s / 2 - a, 0, -s / 2), [v(0, b, 0)] + c))

    # horizontal bottom: 36-41
    o, m = add(R, v(s / 2 - b, b, z0 + t), [v(0, 0, -t), v(-r2 * b, 0, 0), v(-t, t, 0)])
    u = v(-t, 0, 0)
    pts.append(o + m[0] / 2 + (m[2] + u) / 3)
    pts.append(o + (m[0] + m[1] + m[2] + u) / 2)
    c = [v(-(r2 + 1) * t, 0, 0), v(-(r2 + 1) * t, t, 0), v(-t, t, 0)]
    mid(*add(P, v(s / 2 - (r2 + 1) * b, b, z0 + t), [v(0, 0, -t)] + c))
    mid(*add(P, v(s / 2 - (r2 + 1) * b, b, z0), [v(0, 0, -a)] + c))
    mid(*add(P, v(s / 2 - b - t, b + t, z0 + t),
             [v(0, 0, -t), v(-r2 * b, 0, 0), v(-s / 4 + (2 - r2) / 2 * b + t, s / 4 - (2 + r2) / 2 * b - t, 0),
              v(-s / 4 + b + t, s / 4 - b - t, 0)]))
    c = [v(-r2 * t, 0, 0),
         v(-s / 4 + (2 + r2) / 2 * b + (2 - r2) / 2 * t, s / 4 - (2 + r2) / 2 * (b + t), 0),
         v(-s / 4 + (2 + r2) / 2 * b + t, s / 4 - (2 + r2) / 2 * b - t, 0)]
    mid(*add(P, v(s / 2 - (r2 + 1) * b - t, b + t, z0 + t), [v(0, 0, -t)] + c))
    mid(*add(P, v(s / 2 - (r2 + 1) * b - t, b + t, z0), [v(0, 0, -a)] + c))
    assert len(out) == 42 and len(pts) == 52
    return out, np.array(pts)


def octet(dim, div, dT):
    """OctetDomain(dim, div, dT) (domain.cpp:693-1225): dim = (s, d1, d2, t)
    (the unit's period s, the strut's size from d1 + d2 and its wall t), div
    = (n0, n1, n2, n3) (strut tally divisions), dT the temperature difference
    across the struts; the command line's octet deck (main.cpp:376-386)."""
    table, _ = cells(dim, div)
    s, t = float(dim[0]), float(dim[3])
    r2 = math.sqrt(2.0)
    a = math.pi / 5.0 * (dim[1] + dim[2]) - (4.0 - math.pi) / 5.0 * t
    b = a / 4.0
    g = dT / (s - 2.0 * a - 2.0 * (r2 + 1.0) * b - 2.0 * (2.0 + r2) * t)
    d = scenes.Domain("octet")
    for (kind, o, m, dv, grad), fk in zip(table, _FACES):
        faces = [scenes.face(_F[ch]) for ch in fk]
        if kind in (abi.PRISM, abi.PYRAMID):
            d.add_cell(kind, o, m, -1, faces)            # noDiv(0), noGrad
        elif grad:
            d.add_cell(kind, o, m, dv, faces, grad_t=(g, 0.0, -g))
        else:
            d.add_cell(kind, o, m, (-1, -1, -1), faces)  # noDiv, noGrad
    for sa, fa, sb, fb in _INTER:
        d.pair(sa, fa, sb, fb)
    for sa, fa, sb, fb in _PERI:
        d.pair(sa, fa, sb, fb, transl=(s / 2.0, 0.0, -s / 2.0), rot=np.diag([-1.0, 1.0, 1.0]))
    return d


def checkpoints(dim, div=(0, 0, 0, 0)):
    """OctetDomain::checkpoints (domain.cpp:1247-1250): 52 points inside the
    truss, two to three per cell, as 52 x 3."""
    return cells(dim, div)[1]


def average_weights(dom, cell_volumes):
    """OctetDomain::WeightF (domain.cpp:1260-1280) over the field's columns:
    a strut's (cells 16-25) cell volume on its first layer along its third
    axis, zero elsewhere.  ``cell_volumes`` are the per-column volumes (for
    example ``engine.cell_volumes(dom)``)."""
    w = np.zeros(len(cell_volumes))
    col = 0
    for i, c in enumerate(dom.cells):
        if c.cell in (abi.PRISM, abi.PYRAMID):
            sh = (0, 0, 0) if c.div[0] < 0 else (1, 1, 1)
        elif min(c.div[:3]) < 0:
            sh = (0, 0, 0)
        else:
            sh = tuple(max(int(v), 1) for v in c.div[:3])
        layer = sh[0] * sh[1]
        if 16 <= i < 26:
            w[col:col + layer] = cell_volumes[col:col + layer]
        col += layer * sh[2]
    return w


def average(data, dom, cell_volumes):
    """OctetDomain::average (domain.cpp:1252-1258): the rows x cols field
    reduced to one column, weighted by :func:`average_weights`."""
    w = average_weights(dom, cell_volumes)
    return (np.asarray(data) * w).sum(axis=1, keepdims=True) / w.sum()

