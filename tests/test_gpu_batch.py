"""Batches of independent tangles advance in one launch per phase and give
the same result as each tangle run on its own (reference summation order:
bit for bit), and as the oracle (to the FP64 tolerances)."""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams

pytestmark = pytest.mark.gpu


def params(**kw):
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.03, 0.03, 0.03), velocity="BS",
                  overide_timestep_check=True)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def ensemble():
    # uneven sizes, dense enough to reconnect and re-mesh within a few steps
    return [tangles.random_rings(100 + k, 4 + 3 * k, 0.003, 0.008, 0.03, 5e-4) for k in range(5)]


def test_batch_equals_tangles_run_alone():
    from nuvort_amd.engine import FilamentSolver
    sts = ensemble()
    p = params()
    b = FilamentSolver(p, exact_order=True)
    b.set_state_batch(sts)
    b.step(6)
    got = b.get_state_batch()
    ev = b.recon_events()
    for k, st in enumerate(sts):
        s = FilamentSolver(p, exact_order=True)
        s.set_state(st)
        s.step(6)
        ref = s.get_state()
        assert got[k].pcount == ref.pcount
        for f in ("x", "u", "u1", "u2", "infront", "behind", "closest", "closestd"):
            assert np.array_equal(getattr(got[k], f), getattr(ref, f)), (k, f)
        mine = [e[:7] for e in ev if e[7] == k]
        theirs = [e[:7] for e in s.recon_events()]
        assert mine == theirs


@pytest.mark.parametrize("tpt,kb,ts", [(1, 8, 128), (2, 32, 128), (1, 4, 256), (2, 16, 256)])
def test_fp32c_variants_match_fp64(tpt, kb, ts):
    """Every points-per-lane / Kahan-block / source-tile variant the
    autotuner picks from gives FP32C within 1e-5 of FP64, on a batch and
    over recorded steps."""
    from nuvort_amd.engine import FilamentSolver
    sts = ensemble()
    s64 = FilamentSolver(params(hip_precision="fp64"))
    s64.set_state_batch(sts)
    s64.ghostp()
    s64.velocity()
    ref = s64.get_state_batch()
    s64.close()
    v = {"targets_per_lane": tpt, "kahan_block": kb, "source_tile": ts}
    s = FilamentSolver(params(hip_precision="fp32c"), variant=v)
    s.set_state_batch(sts)
    s.ghostp()
    s.velocity()
    t = s.stats()
    assert (t["targets_per_lane"], t["kahan_block"], t["source_tile"]) == (tpt, kb, ts)
    for a, b in zip(ref, s.get_state_batch()):
        act = a.infront != 0
        mag = np.linalg.norm(a.u_sup[act], axis=1)
        assert (np.linalg.norm(a.u_sup[act] - b.u_sup[act], axis=1) / mag).max() <= 1e-5
    # two recorded steps and two eager ones from the same start agree bit for bit
    s.set_state_batch(sts)
    s.step(2)
    rec = s.get_state_batch()
    s.close()
    e = FilamentSolver(params(hip_precision="fp32c", hip_step_launch="eager"), variant=v)
    e.set_state_batch(sts)
    e.step(2)
    for a, b in zip(rec, e.get_state_batch()):
        assert np.array_equal(a.x, b.x) and np.array_equal(a.infront, b.infront)
    e.close()


def test_batch_fp32c_matches_fp64():
    from nuvort_amd.engine import FilamentSolver
    sts = ensemble()
    out = {}
    for prec in ("fp64", "fp32c"):
        s = FilamentSolver(params(hip_precision=prec))
        s.set_state_batch(sts)
        s.ghostp()
        s.velocity()
        out[prec] = s.get_state_batch()
    for a, b in zip(out["fp64"], out["fp32c"]):
        act = a.infront != 0
        mag = np.linalg.norm(a.u_sup[act], axis=1)
        e = np.linalg.norm(a.u_sup[act] - b.u_sup[act], axis=1) / mag
        assert e.max() <= 1e-5


def test_batch_against_oracle():
    from nuvort_amd.engine import FilamentSolver
    from oracle import port
    sts = ensemble()
    p = params()
    s = FilamentSolver(p)
    s.set_state_batch(sts)
    s.step(3)
    got = s.get_state_batch()
    for k, st in enumerate(sts):
        rc, ref, _ = port.step(st, p, 3)
        assert rc == 0
        assert np.array_equal(got[k].infront, ref.infront)
        act = ref.infront != 0
        assert np.abs(got[k].x[act] - ref.x[act]).max() <= 1e-12
