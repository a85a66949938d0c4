"""The brute close-point search (src/line.f90:135-160) against the oracle,
bit for bit, on inputs built to stress its pruning: exact distance ties
between many candidates (points on a binary grid), coincident points
(distance 0), loops of three points (no valid seed candidate), and a
tangle large enough that every target's sources are split into several
chunks."""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams
from nuvort_amd.state import FilamentState
from oracle import port

pytestmark = pytest.mark.gpu


def params():
    return RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                     overide_timestep_check=True)


def loops_state(loops):
    return tangles._assemble([np.asarray(p, np.float64) for p in loops], (0.1, 0.1, 0.1))


def check(st):
    from nuvort_amd.engine import FilamentSolver
    p = params()
    ps = port.PortState(st)
    port.lib().nvo_pclose(ps.ref(), port.C.byref(port.params_struct(p, 0)))
    ref = ps.to_state()
    s = FilamentSolver(p)
    s.set_state(st)
    s.pclose()
    got = s.get_state()
    s.close()
    assert np.array_equal(got.closest, ref.closest)
    assert np.array_equal(got.closestd, ref.closestd)
    return ref


def grid_loops(seed, nloops, h=2.0 ** -9):
    """Rectangular loops along grid lines of spacing h (exact in binary),
    so that many candidates lie at exactly the same distance."""
    g = np.random.default_rng(seed)
    loops = []
    for _ in range(nloops):
        a, b = g.integers(2, 9, size=2)
        o = g.integers(-15, 15, size=3) * h
        ax = g.permutation(3)[:2]
        pts = []
        for k in range(a):
            pts.append((k, 0))
        for k in range(b):
            pts.append((a, k))
        for k in range(a, 0, -1):
            pts.append((k, b))
        for k in range(b, 0, -1):
            pts.append((0, k))
        xyz = np.zeros((len(pts), 3))
        for q, (u, v) in enumerate(pts):
            xyz[q] = o
            xyz[q, ax[0]] += u * h
            xyz[q, ax[1]] += v * h
        loops.append(xyz)
    return loops


def test_close_search_exact_ties():
    ref = check(loops_state(grid_loops(3, 60)))
    # the fixture holds ties: several points at the minimum distance
    assert np.any(np.isin(ref.closestd, [2.0 ** -9]))


def test_close_search_coincident_points():
    loops = grid_loops(4, 20)
    loops.append(loops[0][::-1].copy())        # a copy of a loop: distance 0 to every point
    loops.append(loops[3] + 0.0)
    ref = check(loops_state(loops))
    assert np.any(ref.closestd == 0.0)


def test_close_search_triangles():
    g = np.random.default_rng(5)
    loops = [g.uniform(-0.01, 0.01, 3) + g.uniform(-2e-4, 2e-4, (3, 3)) for _ in range(50)]
    loops += grid_loops(6, 10)
    check(loops_state(loops))


def test_close_search_many_chunks():
    st = tangles.random_rings(1236, 64, 0.01, 0.03, 0.1, 5e-4)
    check(st)
