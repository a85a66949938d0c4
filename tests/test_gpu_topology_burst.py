"""Bursts of topology changes in one step against the oracle: the FP64 direct
sum in the reference's summation order, so links, slots, counts and the
reconnection log must be identical, and velocities agree to the last bits
(the local term's log comes from the GPU's math library, which rounds a few
arguments differently from libm), positions to 1e-15 cm.

* Reconnections: 150 rings of radii 2-5e-3 cm in a 0.015 cm box (6,620
  points; the oracle reconnects about 300 pairs in the first step), so
  that candidate chunks hold many reconnections one after another, each
  changing the rows later visits read.
* Insertions and removals: rings whose chords are 1.9 delta (every point
  inserts, and each insertion changes the next candidate's inputs) beside
  rings whose chords are 0.3 delta (runs of removals, each changing the
  next candidate's links), plus a dense tangle that reconnects.
"""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams
from oracle import port

pytestmark = pytest.mark.gpu

EXACT = ("infront", "behind")
POSITIONS = ("x", "ghosti", "ghostb")
VELOCITIES = ("u", "u1", "u2", "u_sup")


def params(box):
    return RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(box, box, box), velocity="BS",
                     recon_shots=1, overide_timestep_check=True)


def run_gpu(p, st, steps, launch):
    from nuvort_amd.engine import FilamentSolver
    q = p.__class__(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    q.hip_step_launch = launch
    s = FilamentSolver(q, exact_order=True)
    s.set_state(st)
    s.step(steps)
    got = s.get_state()
    ev = s.recon_events()
    s.close()
    return got, ev


def compare(st, p, steps):
    rc, ref, log = port.step(st, p, steps, threads=16)
    assert rc == 0
    for launch in ("eager", "graph"):
        got, ev = run_gpu(p, st, steps, launch)
        n = ref.pcount
        assert got.pcount == n, launch
        assert got.recon_count == ref.recon_count, launch
        for f in EXACT:
            a, b = getattr(got, f)[:n], getattr(ref, f)[:n]
            assert np.array_equal(a, b), (launch, f, int(np.sum(a != b)))
        for f in POSITIONS:
            d = np.abs(getattr(got, f)[:n] - getattr(ref, f)[:n]).max()
            assert d <= 1e-15, (launch, f, d)
        umax = np.abs(ref.u[:n]).max()
        for f in VELOCITIES:
            d = np.abs(getattr(got, f)[:n] - getattr(ref, f)[:n]).max()
            assert d <= 1e-14 * umax, (launch, f, d / umax)
        assert [e[:3] for e in ev] == [e[:3] for e in log], launch
    return ref


def test_reconnection_burst():
    st = tangles.random_rings(33, 150, 0.002, 0.005, 0.015, 5e-4)
    ref = compare(st, params(0.015), 2)
    assert ref.recon_count - st.recon_count > 200


def test_insertion_and_removal_burst():
    g = np.random.default_rng(21)
    box = 0.03
    loops = []
    for k in range(24):
        chord = 1.9 if k % 2 == 0 else 0.3
        loops.append(tangles.ring(g.uniform(-0.012, 0.012, 3), g.uniform(0.002, 0.004), g.normal(size=3), 5e-4,
                                  chord=chord))
    # a reconnecting tangle after the stretched and compressed rings
    dense = tangles.random_rings(34, 40, 0.002, 0.004, box, 5e-4)
    both = tangles._assemble(loops + split_loops(dense), (box, box, box))
    ref = compare(both, params(box), 2)
    # rows were added (insertions) and points cleared (removals)
    assert ref.pcount > both.pcount + 200
    assert int((ref.infront[:ref.pcount] != 0).sum()) < both.pcount - 300


def split_loops(st):
    """The closed loops of a state as (n, 3) position arrays, in row order."""
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
