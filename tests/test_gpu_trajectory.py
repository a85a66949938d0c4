"""Short-horizon trajectories against the oracle at full size.

* Kelvin-wave lattice (8 periodic lines, 1,752 points, dt 1e-5 s, direct
  FP64, 100 steps): positions at steps 10, 50 and 100, FP64 in the
  reference's summation order and FP32C, links bit for bit.
* 64 random rings (16,677 points, dt 5e-6 s, reconnection and re-meshing
  every step), 3 steps: links, slots and the reconnection log bit for bit,
  positions to a tolerance, for the chunked FP64 kernel and for FP32C; and
  FP32C against FP64 on the GPU over 100 steps.

Position differences are taken modulo the periodic box; the tolerances are
written relative to max|u| times the elapsed time (the displacement scale).
"""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams
from oracle import port

pytestmark = pytest.mark.gpu


def wrapped_diff(a, b, box):
    d = a - b
    return d - box * np.round(d / box)


def run_gpu(p, st, checkpoints, exact=False):
    from nuvort_amd.engine import FilamentSolver
    s = FilamentSolver(p, exact_order=exact)
    s.set_state(st)
    out, done = [], 0
    for c in checkpoints:
        s.step(c - done)
        done = c
        out.append(s.get_state())
    ev = s.recon_events()
    s.close()
    return out, ev


def run_oracle(p, st, checkpoints):
    out, done, cur, ev = [], 0, st, []
    for c in checkpoints:
        rc, cur, log = port.step(cur, p, c - done, threads=16)
        assert rc == 0
        done = c
        out.append(cur)
        ev += log
    return out, ev


def kelvin():
    p = RunParams(dt=1e-5, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                  quant_circ=9.97e-4, overide_timestep_check=True)
    return p, tangles.kelvin_lattice(1235)


@pytest.fixture(scope="module")
def kelvin_oracle():
    p, st = kelvin()
    return run_oracle(p, st, [10, 50, 100])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_kelvin_lattice_trajectory(kelvin_oracle, prec, tol):
    p, st = kelvin()
    p.hip_precision = prec
    ref, _ = kelvin_oracle
    got, _ = run_gpu(p, st, [10, 50, 100], exact=(prec == "fp64"))
    for k, (r, g) in enumerate(zip(ref, got)):
        assert g.pcount == r.pcount
        assert np.array_equal(g.infront, r.infront) and np.array_equal(g.behind, r.behind)
        act = r.infront != 0
        disp = np.abs(r.u[act]).max() * p.dt * [10, 50, 100][k]
        err = np.abs(wrapped_diff(g.x[act], r.x[act], 0.1)).max()
        print("kelvin %s step %d: max |dx| %.3e cm, max|u| t %.3e cm, ratio %.3e"
              % (prec, [10, 50, 100][k], err, disp, err / disp))
        assert err <= tol * disp


def rings():
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                  recon_shots=1)
    return p, tangles.random_rings(1236, 64, 0.01, 0.03, 0.1, 5e-4)


@pytest.fixture(scope="module")
def rings_oracle():
    p, st = rings()
    return run_oracle(p, st, [1, 3])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32c", 1e-5)])
def test_ring_tangle_trajectory(rings_oracle, prec, tol):
    p, st = rings()
    p.hip_precision = prec
    ref, ref_ev = rings_oracle
    got, ev = run_gpu(p, st, [1, 3])
    for k, (r, g) in enumerate(zip(ref, got)):
        assert g.pcount == r.pcount
        assert np.array_equal(g.infront, r.infront) and np.array_equal(g.behind, r.behind)
        assert g.recon_count == r.recon_count
        act = r.infront != 0
        disp = np.abs(r.u[act]).max() * p.dt * [1, 3][k]
        err = np.abs(wrapped_diff(g.x[act], r.x[act], 0.1)).max()
        print("rings %s step %d: max |dx| %.3e cm, max|u| t %.3e cm, ratio %.3e, reconnections %d"
              % (prec, [1, 3][k], err, disp, err / disp, r.recon_count))
        assert err <= tol * disp
    assert [e[:3] for e in ev] == [e[:3] for e in ref_ev]


def test_ring_tangle_fp32c_follows_fp64():
    """The 64-ring tangle over 100 steps (21 reconnections): the FP32C run
    keeps the FP64 run's links and reconnection log at steps 10, 50 and 100,
    and its positions within 1e-6 of max|u| times the elapsed time (measured
    8e-9).  The FP64 run is pinned to the oracle by the test above."""
    p, st = rings()
    got = {}
    for prec in ("fp64", "fp32c"):
        p.hip_precision = prec
        got[prec] = run_gpu(p, st, [10, 50, 100])
    (s64, ev64), (s32, ev32) = got["fp64"], got["fp32c"]
    assert len(ev64) > 10
    assert [e[:3] for e in ev32] == [e[:3] for e in ev64]
    for k, (a, b) in enumerate(zip(s64, s32)):
        assert b.pcount == a.pcount
        assert np.array_equal(b.infront, a.infront) and np.array_equal(b.behind, a.behind)
        act = a.infront != 0
        disp = np.linalg.norm(a.u[act], axis=1).max() * p.dt * [10, 50, 100][k]
        err = np.linalg.norm(wrapped_diff(b.x[act], a.x[act], 0.1), axis=1).max()
        print("rings fp32c vs fp64 step %d: ratio %.3e" % ([10, 50, 100][k], err / disp))
        assert err <= 1e-6 * disp
