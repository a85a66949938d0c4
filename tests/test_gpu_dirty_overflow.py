"""pinsert and premove when more rows are queued for a second visit than the
staged kernels' queue holds (8,192 per tangle). Then the kernel's serial
thread continues with the reference's plain loop from the next row
(src/line.f90:18-96, :102-130). The results must match the oracle bit for
bit: links, slots, positions, pcount and counts. For pinsert the check also
covers total_length and avg_sep, to 1e-12 since they are summed in parallel.

* pinsert: 9,000 points on rings whose chords are 2.5 delta, followed by
  12,000 empty rows. Each point inserts into an empty slot the loop has still
  to visit, and the new point's segment (about 1.25 delta) inserts again when
  the loop reaches it. That is 9,000 queued revisits; later insertions reuse
  the last empty slots and then grow the array.
* premove: 9,000 pentagons of circumradius 0.2 delta. The first point of
  each sits in the low rows and the other four in the high rows. Each first
  point removes its neighbour, which queues its own behind, a high row, for a
  second visit; those visits then remove again and clear the loops.
"""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams
from nuvort_amd.state import FilamentState
from golden_util import compare_states
from oracle import port

pytestmark = pytest.mark.gpu

DELTA = 5e-4
BOX = 0.1
QUEUE = 8192


def params():
    return RunParams(dt=5e-6, delta=DELTA, corea=1.3e-8, box_size=(BOX, BOX, BOX), velocity="BS",
                     recon_shots=1, overide_timestep_check=True)


def place(loops, rows_of, nrows):
    """A state of nrows rows holding the closed loops (point k of a loop has
    infront k+1, cyclic); point k of loop i sits in row rows_of[i][k]."""
    st = FilamentState.empty(nrows)
    for pts, rows in zip(loops, rows_of):
        m = len(pts)
        for k in range(m):
            r = rows[k]
            st.x[r] = pts[k]
            st.infront[r] = rows[(k + 1) % m] + 1
            st.behind[r] = rows[(k - 1) % m] + 1
    return st


def gpu(st, op):
    from nuvort_amd.engine import FilamentSolver
    s = FilamentSolver(params(), check=False, capacity=3 * st.pcount)
    try:
        s.set_state(st)
        getattr(s, op)()
        got = s.get_state()
        d = s.diagnostics() if op == "pinsert" else None
        return got, d
    finally:
        s.close()


def test_pinsert_past_the_queue():
    g = np.random.default_rng(91)
    loops, n = [], 0
    while n < 9000:
        loops.append(tangles.ring(g.uniform(-0.035, 0.035, 3), 0.01, g.normal(size=3), DELTA, chord=2.5))
        n += len(loops[-1])
    rows_of, off = [], 0
    for pts in loops:
        rows_of.append(list(range(off, off + len(pts))))
        off += len(pts)
    st = port.ghostp(place(loops, rows_of, n + 12000), params())
    rc, ref, tl, sep = port.pinsert(st, params())
    assert rc == 0
    # the case is the one described: more insertions above the cursor than the queue holds
    assert int((ref.infront[n:n + 12000] != 0).sum()) == 12000 and ref.pcount > st.pcount
    assert ref.pcount - st.pcount + 12000 > 2 * QUEUE
    got, d = gpu(st, "pinsert")
    assert compare_states(ref, got, st.pcount, searched=False) == []
    assert d["total_length"] == pytest.approx(tl, rel=1e-12, abs=0)   # a parallel sum
    assert d["avg_sep"] == pytest.approx(sep, rel=1e-12, abs=0)


def test_premove_past_the_queue():
    L = 9000
    g = np.random.default_rng(92)
    rho = 0.2 * DELTA
    ang = 2 * np.pi * np.arange(5) / 5
    loops, rows_of = [], []
    for i in range(L):
        c = g.uniform(-0.04, 0.04, 3)
        nrm = g.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        e1 = np.cross(nrm, [1.0, 0.0, 0.0] if abs(nrm[0]) < 0.9 else [0.0, 1.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(nrm, e1)
        loops.append(c + rho * (np.cos(ang)[:, None] * e1 + np.sin(ang)[:, None] * e2))
        rows_of.append([i] + [L + 4 * i + k for k in range(4)])
    st = port.ghostp(place(loops, rows_of, 5 * L), params())
    ref = port.premove(st, params())
    # every first point removed its neighbour (a queued revisit each), and the
    # revisits cleared the loops
    assert ref.extra["remove_count"] >= L + QUEUE
    assert int((ref.infront != 0).sum()) < L
    got, _ = gpu(st, "premove")
    assert compare_states(ref, got, st.pcount) == []
    assert got.extra["remove_count"] == ref.extra["remove_count"]
