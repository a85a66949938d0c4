"""The stale-tree close search (src/tree.f90:190-244) at scale against the
oracle: the tree is built from the step's start state and searched with the
positions after pmotion and pinsert, and the walk skips every subtree that
lies farther than the lane's best distance plus the subtree's radius. closest
and closestd must equal the oracle's bit for bit on every row."""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams
from oracle import port

pytestmark = pytest.mark.gpu


def params(box, theta):
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(box, box, box), velocity="Tree", tree_theta=theta,
                  normal_velocity="xflow", norm_vel_xflow=0.5, alpha=(0.206, 0.0083), recon_shots=1,
                  overide_timestep_check=True)
    p.hip_precision = "fp64"
    return p


def check(st, p):
    from nuvort_amd.engine import FilamentSolver
    ref = port.run_phases(st, p, threads=16)
    assert ref["rc_tree"] == 0
    s = FilamentSolver(p)
    try:
        s.set_state(ref["ghost"])
        s.construct_tree()
        s.set_state(ref["insert"])
        s.pclose_tree()
        got = s.get_state()
    finally:
        s.close()
    want = ref["close"]
    n = want.pcount
    assert got.pcount == n
    bad = np.flatnonzero(got.closest[:n] != want.closest[:n])
    assert bad.size == 0, (bad[:10], got.closest[bad[:10]], want.closest[bad[:10]])
    bad = np.flatnonzero(got.closestd[:n] != want.closestd[:n])
    assert bad.size == 0, (bad[:10], got.closestd[bad[:10]], want.closestd[bad[:10]])
    return want


@pytest.mark.parametrize("theta", [0.4, 0.8])
def test_close_tree_rings(theta):
    st = tangles.random_rings(1236, 64, 0.01, 0.03, 0.1, 5e-4)
    want = check(st, params(0.1, theta))
    act = want.infront[:want.pcount] != 0
    # the search found real neighbours, not the initial 100
    assert np.median(want.closestd[:want.pcount][act]) < 2e-3


def test_close_tree_dense_tangle():
    # 150 small rings in a 0.015 cm box: hundreds of close approaches,
    # so best distances fall far below delta and most subtrees are pruned
    st = tangles.random_rings(33, 150, 0.002, 0.005, 0.015, 5e-4)
    want = check(st, params(0.015, 0.4))
    act = want.infront[:want.pcount] != 0
    assert (want.closestd[:want.pcount][act] < 2.5e-4).sum() > 100
