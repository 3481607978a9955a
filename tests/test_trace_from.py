"""TrajProblem with a given start (problem.cpp:226-299, the pos_ branch) in
the checker: the reference's geometry check (checkDomain, main.cpp:104-143:
two-scatter probes from every check point of the domain along six axis
directions) finds no escape in the octet truss; a first flight moves the start
by distance x direction; the start cell is the first containing the point;
and the construction errors."""
import numpy as np
import pytest

from montecarlocpp_amd import abi, materials, octet, scenes
from oracle import oracle

DIM, DIV = (6e-7, 1e-7, 1e-7, 1e-8), (4, 2, 2, 1)


@pytest.fixture(scope="module")
def mat():
    return materials.si_like(300.0)


def test_check_probes_stay_in_the_truss(mat):
    d = octet.octet(DIM, DIV, 1.0)
    prob = scenes.problem(1, maxscat=2, maxloop=2, seed=3)
    n = 0
    for p in octet.checkpoints(DIM, DIV):
        for i in range(6):
            u = np.zeros(3)
            u[i % 3] = 1.0 if i < 3 else -1.0
            rec = oracle.trace_from(mat, d, prob, p, u)
            assert len(rec) >= 1 and not np.any(rec[:, 13] == abi.ESCAPED)
            n += 1
    assert n == 52 * 6


def test_first_flight_geometry(mat):
    d = scenes.slab(1e-7, 10)
    start = np.array([3e-8, 4e-8, 5e-8])
    u = np.array([0.3, -0.2, 0.9])
    rec = oracle.trace_from(mat, d, scenes.problem(1, maxscat=5, seed=4), start, u)
    first = rec[0]
    np.testing.assert_allclose(first[3:6], start + first[2] * u / np.linalg.norm(u), rtol=0, atol=1e-20)
    assert first[12] == 0


def test_given_mode_and_drawn_direction(mat):
    """A given (w, p) cell is the mode of the whole flight (its index among
    the populated cells in the records); without a direction one is drawn
    and the first flight has the recorded length."""
    d = scenes.slab(1e-7, 10)
    pdf = oracle.material(mat)["energy_pdf"]
    cells = np.flatnonzero(pdf > 0)
    k = 7
    start = np.array([5e-8, 5e-8, 5e-8])
    rec = oracle.trace_from(mat, d, scenes.problem(1, maxscat=1, seed=4), start, prop_cell=int(cells[k]))
    assert rec[0, 1] != -1 and rec[0, 10] == k   # a wall first: the mode is kept
    assert np.linalg.norm(rec[0, 3:6] - start) == pytest.approx(rec[0, 2], rel=1e-12)
    with pytest.raises(Exception, match="prop has no energy pdf"):
        oracle.trace_from(mat, d, scenes.problem(1, seed=4), start, prop_cell=int(np.flatnonzero(pdf == 0)[0]))


@pytest.mark.parametrize("case,msg", [
    ("outside", "Position not inside domain"),
    ("zero_dir", "direction must be nonzero"),
])
def test_trace_from_errors(mat, case, msg):
    d = scenes.slab(1e-7, 10)
    prob = scenes.problem(1, maxscat=2, seed=1)
    with pytest.raises(Exception, match=msg):
        if case == "outside":
            oracle.trace_from(mat, d, prob, (2e-7, 5e-8, 5e-8), (1.0, 0.0, 0.0))
        else:
            oracle.trace_from(mat, d, prob, (5e-8, 5e-8, 5e-8), (0.0, 0.0, 0.0))


def test_trajectory_points_cross_periodic_faces(mat):
    """TrajProblem::solve returns TrkPhonon's points (phonon.cpp:133-169): the
    start, the end of every flight and, after a periodic face, the mapped
    position (boundary.cpp:517-522).  In a film periodic along x (makePair
    with translation L, domain.cpp:188-199) each crossing is a point on one
    x face followed by the point with the same y, z on the other."""
    L = 2e-7
    d = scenes.film((L, 5e-8, L), (0, 6, 0), 1e6 * L)
    start = np.array([5e-8, 2e-8, 1e-7])
    prob = scenes.problem(1, maxscat=40, seed=2)
    pts = oracle.trajectory_from(mat, d, prob, start, (0.9, 0.1, 0.3))
    rec = oracle.trace_from(mat, d, prob, start, (0.9, 0.1, 0.3))
    np.testing.assert_array_equal(pts[:, 0], start)
    x = pts[0]
    on_face = (np.abs(x) < 1e-20) | (np.abs(x - L) < 1e-20)
    cross = [i for i in range(1, pts.shape[1] - 1) if on_face[i] and on_face[i + 1]
             and abs(x[i] - x[i + 1]) > 0.5 * L]
    assert len(cross) >= 2
    for i in cross:
        np.testing.assert_allclose(pts[1:, i + 1], pts[1:, i], rtol=0, atol=1e-22)
    # one point per flight and one per crossing: nothing else moves a bundle
    assert pts.shape[1] == 1 + len(rec) + len(cross)


def test_trace_from_needs_no_source(mat):
    """A geometry check traces from a given start, so a domain without an
    emitting face or volume source is accepted (the reference's TrajProblem
    with pos_ never touches the emitters); a field solve of it is not."""
    d = scenes.Domain("closed")
    sp = scenes.face(abi.SPEC)
    d.add_box((0, 0, 0), (1e-7, 1e-7, 1e-7), (0, 0, 0), [sp, sp, sp, sp, sp, sp])
    prob = scenes.problem(1, maxscat=5, seed=3)
    rec = oracle.trace_from(mat, d, prob, (5e-8, 5e-8, 5e-8), (1.0, 0.2, 0.1))
    assert len(rec) >= 1 and not np.any(rec[:, 13] == abi.ESCAPED)
    with pytest.raises(Exception, match="No phonons to emit"):
        oracle.solve(mat, d, prob)


def test_default_maxloop_overflow_rejected(mat):
    """maxloop_ = 100 maxscat_ (problem.h:124) must fit an int64: a larger
    maxscat with maxloop 0 is an error in the checker and in the engine's
    host planner, not a wrapped step cap."""
    from montecarlocpp_amd import engine
    d = scenes.slab(1e-7, 10)
    big = scenes.problem(10, maxscat=2 ** 62)
    with pytest.raises(Exception, match="maxscat too large"):
        oracle.plan(mat, d, big)
    with pytest.raises(Exception, match="maxscat too large"):
        engine.plan(mat, d, big)
    ok = scenes.problem(10, maxscat=2 ** 62, maxloop=1000)
    assert oracle.plan(mat, d, ok)[0].maxloop == 1000
    assert engine.plan(mat, d, ok)[0].maxloop == 1000
