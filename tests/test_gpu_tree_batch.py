"""Batches of tangles with the octree: every tangle gets its own octree (the
reference's tree of its own points, src/tree.f90:38-157) and its steps run
concurrently with the others'.  Each tangle's result must equal the same
tangle run alone, bit for bit -- positions, velocities, links, the close
search and the reconnection log -- recorded and eager, FP64 and FP32 walks,
and phase by phase; the single-tangle octree path is itself checked against
the oracle and the reference's fixtures elsewhere."""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams

pytestmark = pytest.mark.gpu

FIELDS = ("x", "u", "u1", "u2", "u_sup", "u_mf", "infront", "behind", "closest", "closestd")


def params(**kw):
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.03, 0.03, 0.03), velocity="Tree", tree_theta=0.5,
                  overide_timestep_check=True)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def ensemble():
    # uneven sizes (974 to 2,300 points); crossing rings that reconnect within
    # five steps (9 reconnections in the oracle)
    return [tangles.random_rings(400 + k, 8 + 4 * k, 0.006, 0.012, 0.03, 5e-4) for k in range(4)]


def same_state(a, b, what):
    assert a.pcount == b.pcount, what
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), (what, f)


def events_of(evs, tangle=None):
    return [e[:7] for e in evs if tangle is None or e[7] == tangle]


@pytest.mark.parametrize("prec,launch", [("fp64", "graph"), ("fp32c", "graph"), ("fp64", "eager")])
def test_tree_batch_equals_tangles_alone(prec, launch):
    from nuvort_amd.engine import FilamentSolver
    sts = ensemble()
    p = params(hip_precision=prec, hip_step_launch=launch, normal_velocity="xflow", norm_vel_xflow=0.5,
               alpha=(0.206, 0.0083))
    b = FilamentSolver(p)
    b.set_state_batch(sts)
    b.step(5)
    got = b.get_state_batch()
    ev = b.recon_events()
    st_b = b.stats()
    b.close()
    assert len(got) == len(sts)
    nrec = 0
    for k, st in enumerate(sts):
        s = FilamentSolver(p)
        s.set_state(st)
        s.step(5)
        ref = s.get_state()
        same_state(got[k], ref, "tangle %d" % k)
        assert got[k].t == ref.t and got[k].itime == ref.itime
        ev_k = s.recon_events()
        assert events_of(ev, k) == events_of(ev_k)
        nrec += len(ev_k)
        s.close()
    assert nrec > 0, "the ensemble should reconnect within five steps"
    if launch == "graph":
        assert st_b["tree_replays"] > 0


def test_tree_batch_phase_calls():
    """The reference's phases called one by one on a batch (the Fortran
    host's sequence) forward to every tangle."""
    from nuvort_amd.engine import FilamentSolver
    sts = ensemble()
    p = params(hip_precision="fp32c")
    b = FilamentSolver(p)
    b.set_state_batch(sts)
    alone = []
    for st in sts:
        s = FilamentSolver(p)
        s.set_state(st)
        alone.append(s)
    for phase in ("ghostp", "construct_tree", "pmotion", "pinsert", "pclose_tree", "precon", "premove",
                  "enforce_boundary", "empty_tree", "nan_finder"):
        getattr(b, phase)()
        for s in alone:
            getattr(s, phase)()
        got = b.get_state_batch()
        for k, s in enumerate(alone):
            same_state(got[k], s.get_state(), "%s, tangle %d" % (phase, k))
    for s in alone:
        s.close()
    b.close()


def test_tree_batch_then_single_tangle():
    """A context that held a batch takes a single tangle again, and the
    single-tangle calls a batch refuses say so."""
    from nuvort_amd import lib as nvlib
    from nuvort_amd.engine import FilamentSolver
    sts = ensemble()
    p = params()
    b = FilamentSolver(p)
    b.set_state_batch(sts)
    with pytest.raises(nvlib.NuvortError):
        b.get_state()
    with pytest.raises(nvlib.NuvortError):
        b.tree_nodes()
    b.set_state(sts[1])
    b.step(2)
    s = FilamentSolver(p)
    s.set_state(sts[1])
    s.step(2)
    same_state(b.get_state(), s.get_state(), "single tangle after a batch")
    s.close()
    b.close()
