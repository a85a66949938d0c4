"""Octree rebuilds in one context. The first build runs the level loop with a
host round trip per level; later builds launch every level at once, on launch
bounds taken from the previous build's level sizes, and fall back to the
synchronous loop when the new tree outgrows them (more nodes in a level, or
more levels). Whichever loop ran, the tree (depth-first nodes: boxes, counts,
moments, leaf links) and the FP64 walk's velocities must equal a fresh
context's first build bit for bit.  A tree deeper than 16 levels is built
again with longer paths (32, then 48 levels), and the context keeps the
longer paths for its later builds, which must still equal fresh builds."""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams

pytestmark = pytest.mark.gpu

BOX = 0.1


def params():
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(BOX, BOX, BOX), velocity="Tree", tree_theta=0.5,
                  overide_timestep_check=True)
    p.hip_precision = "fp64"
    return p


def build(s, st):
    s.set_state(st)
    s.ghostp()
    s.construct_tree()
    nodes = s.tree_nodes()
    s.velocity()
    return nodes, s.get_state().u_sup, s.stats()


def fresh(st):
    from nuvort_amd.engine import FilamentSolver
    s = FilamentSolver(params())
    try:
        nodes, u, stats = build(s, st)
    finally:
        s.close()
    assert stats["tree_sync_builds"] >= 1
    return nodes, u, stats


def with_tiny_ring(st):
    """st plus a ring of radius 1e-5 cm with points about 7e-6 cm apart: the
    tree needs levels down to widths of a few 1e-6 cm (14 levels against 10
    for the 64 rings alone)."""
    loops = split_loops(st)
    loops.append(tangles.ring(np.array([0.011, -0.023, 0.031]), 1e-5, np.array([0.3, 0.5, 0.8]), 8e-6))
    return tangles._assemble(loops, (BOX, BOX, BOX))


def split_loops(st):
    seen = np.zeros(st.pcount, bool)
    out = []
    for r in range(st.pcount):
        if seen[r] or st.infront[r] == 0:
            continue
        loop, q = [], r
        while not seen[q]:
            seen[q] = True
            loop.append(st.x[q])
            q = st.infront[q] - 1
        out.append(np.array(loop))
    return out


def test_rebuilds_match_fresh_builds():
    from nuvort_amd.engine import FilamentSolver
    base = tangles.random_rings(1236, 64, 0.01, 0.03, BOX, 5e-4)
    moved = base.copy()
    moved.x[:moved.pcount] += np.random.default_rng(5).normal(scale=1e-5, size=(moved.pcount, 3))
    bigger = tangles.random_rings(1237, 160, 0.01, 0.03, BOX, 5e-4)
    deeper = with_tiny_ring(base)
    smaller = tangles.random_rings(1238, 16, 0.01, 0.03, BOX, 5e-4)
    # (name, state, whether the synchronous loop must run again)
    from golden_util import Fixture
    deep = Fixture("tree_deep").state("in")   # pairs 4e-7 and 5e-12 cm apart: 34 levels
    seq = [("base", base, True), ("moved", moved, False), ("bigger", bigger, True), ("deeper", deeper, None),
           ("smaller", smaller, False), ("smaller again", smaller, False), ("deep", deep, True),
           ("base after deep", base, True), ("moved after deep", moved, False)]
    s = FilamentSolver(params())
    try:
        prev_sync, prev_levels = 0, 0
        for name, st, sync in seq:
            nodes, u, stats = build(s, st)
            ref_nodes, ref_u, ref_stats = fresh(st)
            assert len(nodes) == len(ref_nodes), name
            for k in ref_nodes.dtype.names:
                assert np.array_equal(nodes[k], ref_nodes[k]), (name, k)
            assert stats["tree_levels"] == ref_stats["tree_levels"], name
            if name == "deep":
                assert stats["tree_levels"] > 33
            assert stats["tree_nodes"] == ref_stats["tree_nodes"], name
            assert np.array_equal(u[:st.pcount], ref_u[:st.pcount]), name
            ran_sync = stats["tree_sync_builds"] > prev_sync
            if sync is None:
                # one level more than the previous tree is launched; a tree
                # deeper than that cannot fit
                sync = stats["tree_levels"] > prev_levels + 1
            assert ran_sync == sync, (name, stats["tree_sync_builds"], stats["tree_levels"], prev_levels)
            prev_sync, prev_levels = stats["tree_sync_builds"], stats["tree_levels"]
        assert stats["tree_builds"] == len(seq)
    finally:
        s.close()
