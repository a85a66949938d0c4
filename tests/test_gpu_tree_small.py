"""The octree path on small tangles against the oracle: fewer targets than
one walk item (128 targets, two per lane), a three-point loop, and a few
rings over several steps, so that later builds of tiny trees go through the
level loop launched on the previous build's sizes. Per phase, from the
oracle's own state before it: tree nodes bit for bit (moments to the tree
tolerance), FP64 walk velocities to 1e-12, the FP32 walk against the FP64
walk, and the stale-tree close search bit for bit."""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams
from oracle import port

pytestmark = pytest.mark.gpu

BOX = 0.1


def params(theta, prec="fp64"):
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(BOX, BOX, BOX), velocity="Tree", tree_theta=theta,
                  recon_shots=1, overide_timestep_check=True)
    p.hip_precision = prec
    return p


def rel(a, b):
    s = max(np.abs(a).max(), 1e-300)
    return float(np.abs(a - b).max() / s)


def small_cases():
    g = np.random.default_rng(7)
    one = tangles._assemble([tangles.ring(np.zeros(3), 0.003, np.array([0.2, 0.3, 1.0]), 5e-4)], (BOX,) * 3)
    tri = tangles._assemble([np.array([[0.0, 0.0, 0.0], [4e-4, 0.0, 0.0], [2e-4, 3.5e-4, 0.0]])], (BOX,) * 3)
    few = tangles._assemble([tangles.ring(g.uniform(-0.01, 0.01, 3), g.uniform(0.002, 0.004), g.normal(size=3), 5e-4)
                             for _ in range(5)], (BOX,) * 3)
    return {"one_ring": one, "triangle": tri, "five_rings": few}


CASES = small_cases()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("theta", [0.5, 0.9])
def test_small_tree_phases(name, theta):
    from nuvort_amd.engine import FilamentSolver
    st = CASES[name]
    p = params(theta)
    ref = port.run_phases(st, p, threads=4)
    assert ref["rc_tree"] == 0
    s = FilamentSolver(p, check=False)
    try:
        # build twice: the second build runs the level loop on the first's sizes
        for _ in range(2):
            s.set_state(ref["ghost"])
            s.construct_tree()
            nodes = s.tree_nodes()
            want = ref["tree"]
            assert len(nodes) == len(want["pcount"])
            for k in ("pos", "width", "pcount", "leaf_infront"):
                assert np.array_equal(nodes[k], want[k]), k
            assert rel(want["cent"], nodes["cent"]) <= 1e-14
        s.pmotion()
        got = s.get_state()
        mot = ref["motion"]
        act = mot.infront[:mot.pcount] != 0
        assert rel(mot.u[:mot.pcount][act], got.u[:mot.pcount][act]) <= 1e-12
        # the stale-tree close search from the oracle's state after pinsert
        s.set_state(ref["insert"])
        s.pclose_tree()
        got = s.get_state()
        cl = ref["close"]
        n = cl.pcount
        assert np.array_equal(got.closestd[:n], cl.closestd[:n])
        found = cl.closestd[:n] < 100.0
        assert np.array_equal(got.closest[:n][found], cl.closest[:n][found])
    finally:
        s.close()


@pytest.mark.parametrize("name", list(CASES))
def test_small_tree_fp32_walk(name):
    from nuvort_amd.engine import FilamentSolver
    st = CASES[name]
    out = {}
    for prec in ("fp64", "fp32c"):
        s = FilamentSolver(params(0.5, prec), check=False)
        try:
            s.set_state(st)
            s.ghostp()
            s.construct_tree()
            s.velocity()
            out[prec] = s.get_state()
        finally:
            s.close()
    a, b = out["fp64"], out["fp32c"]
    act = a.infront[:a.pcount] != 0
    assert rel(a.u_sup[:a.pcount][act], b.u_sup[:a.pcount][act]) <= 1e-4


def test_small_tree_steps():
    """Five rings, 4 steps with the octree rebuilt every step (the later
    builds on the previous build's sizes): links exact, positions to 1e-12."""
    from nuvort_amd.engine import FilamentSolver
    st = CASES["five_rings"]
    p = params(0.5)
    rc, ref, log = port.step(st, p, 4, threads=4)
    assert rc == 0
    s = FilamentSolver(p, check=False)
    try:
        s.set_state(st)
        s.step(4)
        got = s.get_state()
        stats = s.stats()
    finally:
        s.close()
    n = ref.pcount
    assert got.pcount == n
    assert np.array_equal(got.infront[:n], ref.infront[:n])
    assert np.array_equal(got.behind[:n], ref.behind[:n])
    act = ref.infront[:n] != 0
    assert rel(ref.x[:n][act], got.x[:n][act]) <= 1e-12
    assert stats["tree_builds"] >= 4 and stats["tree_sync_builds"] < stats["tree_builds"]
