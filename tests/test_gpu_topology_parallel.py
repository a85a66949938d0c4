"""Chunks of independent insertions and removals, which the staged kernels
apply together, against the oracle bit for bit (links, slots, positions,
counts; total_length and avg_sep to 1e-12, being summed in parallel).

Rings are built from prescribed spacings. Most segments are 0.7 delta for
removals and 0.9 delta for insertions; neither inserts or removes. Isolated
spots differ:
  * removals: two segments of 0.35 delta. The point before them removes its
    infront. The neighbouring points' tests stay false (1.05 delta).
  * insertions: one segment of 1.3 delta, whose insertion changes no other
    candidate, or runs of two to four such segments, where each insertion
    changes the next candidate's inputs (planned again in rounds).
Rings made of spots only are mixed in (every segment short or long), so that
some chunks are still made one visit at a time.
"""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams
from golden_util import compare_states
from oracle import port

pytestmark = pytest.mark.gpu

DELTA = 5e-4
BOX = 0.1


def params():
    return RunParams(dt=5e-6, delta=DELTA, corea=1.3e-8, box_size=(BOX, BOX, BOX), velocity="BS",
                     recon_shots=1, overide_timestep_check=True)


def spaced_ring(center, radius, normal, spacing):
    """Points on a circle whose successive arcs are proportional to spacing."""
    normal = np.asarray(normal, np.float64) / np.linalg.norm(normal)
    e1 = np.cross(normal, [1.0, 0.0, 0.0] if abs(normal[0]) < 0.9 else [0.0, 1.0, 0.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(normal, e1)
    arc = np.concatenate([[0.0], np.cumsum(spacing)[:-1]]) * (2 * np.pi / np.sum(spacing))
    return center + radius * (np.cos(arc)[:, None] * e1 + np.sin(arc)[:, None] * e2)


def rings(seed, nrings, base, spots, every, runs):
    """nrings rings of radius 5e-3 cm: segments of `base` delta with a spot
    (a list of segment lengths, taken in turn from `spots`) after every
    `every` segments; every `runs`-th ring is made of its spot only."""
    g = np.random.default_rng(seed)
    loops = []
    R = 5e-3
    for k in range(nrings):
        sp = []
        spot = spots[k % len(spots)]
        while sum(sp) * DELTA < 2 * np.pi * R:
            if k % runs == runs - 1:
                sp += spot
            else:
                sp += [base] * every + spot
        loops.append(spaced_ring(g.uniform(-0.04, 0.04, 3), R, g.normal(size=3), np.array(sp)))
    return port.ghostp(tangles._assemble(loops, (BOX, BOX, BOX)), params())


def solver():
    from nuvort_amd.engine import FilamentSolver
    return FilamentSolver(params(), check=False)


def test_independent_removals():
    st = rings(71, 120, 0.7, [[0.35, 0.35]], 5, 10)
    ref = port.premove(st, params())
    assert ref.extra["remove_count"] > 1500
    s = solver()
    try:
        s.set_state(st)
        s.premove()
        got = s.get_state()
    finally:
        s.close()
    assert compare_states(ref, got, st.pcount) == []
    assert got.extra["remove_count"] == ref.extra["remove_count"]


def test_independent_insertions():
    st = rings(72, 120, 0.9, [[1.3], [1.3] * 2, [1.3], [1.3] * 4], 6, 10)
    rc, ref, tl, sep = port.pinsert(st, params())
    assert rc == 0 and ref.pcount > st.pcount + 1500
    s = solver()
    try:
        s.set_state(st)
        s.pinsert()
        got = s.get_state()
        d = s.diagnostics()
    finally:
        s.close()
    assert compare_states(ref, got, st.pcount, searched=False) == []
    assert d["total_length"] == pytest.approx(tl, rel=1e-12, abs=0)   # a parallel sum
    assert d["avg_sep"] == pytest.approx(sep, rel=1e-12, abs=0)
