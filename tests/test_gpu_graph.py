"""nv_step records one step's launches in a HIP graph and replays it; the
replayed steps must give exactly what eager launches give: the same
positions, links, slots and reconnection log, bit for bit, across both x /
xn parities, steps with and without the reconnection phase
(recon_shots > 1), a layout change in the middle of a run (insertions
outgrowing the device rows), eager calls between recorded steps, and
batches of tangles in FP32C.  The octree step likewise, including a step
whose build outgrows the recorded plan (taken again eagerly)."""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams
from nuvort_amd.state import FilamentState

pytestmark = pytest.mark.gpu

FIELDS = ("x", "u", "u1", "u2", "u_sup", "u_mf", "ghosti", "ghostb", "infront", "behind", "closest", "closestd")


def params(**kw):
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.03, 0.03, 0.03), velocity="BS",
                  overide_timestep_check=True)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def run(p, st, launch, steps, batch=False, between=None, capacity=0):
    from nuvort_amd.engine import FilamentSolver
    q = p.__class__(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    q.hip_step_launch = launch
    s = FilamentSolver(q, capacity=capacity)
    if batch:
        s.set_state_batch(st)
    else:
        s.set_state(st)
    for k in steps:
        s.step(k)
        if between:
            between(s)
    out = s.get_state_batch() if batch else [s.get_state()]
    ev = s.recon_events()
    if not batch:
        # every reconnection of the run is in the log
        assert len(ev) == out[0].recon_count - st.recon_count
    stats = s.stats()
    s.close()
    return out, ev, stats


def assert_same(a, b):
    (sa, ea, ta), (sb, eb, tb) = a, b
    assert len(sa) == len(sb)
    for k, (x, y) in enumerate(zip(sa, sb)):
        assert x.pcount == y.pcount, k
        assert (x.t, x.itime, x.recon_count) == (y.t, y.itime, y.recon_count), k
        for f in FIELDS:
            assert np.array_equal(getattr(x, f), getattr(y, f)), (k, f)
    assert ea == eb
    assert ta["point_updates"] == tb["point_updates"]


def dense():
    # 1408 points; the oracle logs 4 reconnections in its first 7 steps
    return tangles.random_rings(10, 20, 0.003, 0.008, 0.03, 5e-4)


@pytest.mark.parametrize("shots", [1, 3])
@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_recorded_step_equals_eager(shots, prec):
    p = params(recon_shots=shots, hip_precision=prec)
    st = dense()
    ref = run(p, st, "eager", [7])
    got = run(p, st, "graph", [7])
    if shots == 1:
        assert len(ref[1]) > 0      # the run reconnects
    assert_same(ref, got)


def test_recorded_step_split_calls():
    # several nv_step calls continue the device clock and parity
    p = params(recon_shots=2)
    st = dense()
    assert_same(run(p, st, "eager", [7]), run(p, st, "graph", [2, 1, 3, 1]))


def concat(a, b):
    st = FilamentState.empty(a.pcount + b.pcount)
    st.x[:a.pcount], st.x[a.pcount:] = a.x, b.x
    st.infront[:a.pcount], st.behind[:a.pcount] = a.infront, a.behind
    st.infront[a.pcount:] = np.where(b.infront > 0, b.infront + a.pcount, 0)
    st.behind[a.pcount:] = np.where(b.behind > 0, b.behind + a.pcount, 0)
    return st


def test_recorded_step_after_layout_change():
    # the reconnecting tangle plus rings with segments of 1.9 delta: those
    # insert a point on every segment in step 1, so step 2 outgrows the
    # initial rows (2 x pcount + 1024, below the default) and the state moves
    # to a wider layout between two replays; the reconnections of step 1 stay
    # in the log
    g = np.random.default_rng(11)
    loops = [tangles.ring(g.uniform(-0.015, 0.015, 3), 0.01, g.normal(size=3), 5e-4, chord=1.9)
             for _ in range(10)]
    st = concat(dense(), tangles._assemble(loops, (0.03, 0.03, 0.03)))
    p = params(recon_shots=1, alpha=(0.206, 0.0083), normal_velocity="xflow", norm_vel_xflow=0.5)
    cap = 2 * st.pcount + 1024
    ref = run(p, st, "eager", [5], capacity=cap)
    got = run(p, st, "graph", [5], capacity=cap)
    assert ref[0][0].pcount > st.pcount + 600
    assert len(ref[1]) > 0
    assert_same(ref, got)


def test_eager_calls_between_recorded_steps():
    def velocity(s):
        s.ghostp()
        s.velocity()
    p = params(hip_precision="fp32c")
    st = dense()
    assert_same(run(p, st, "eager", [2, 2, 2], between=velocity), run(p, st, "graph", [2, 2, 2], between=velocity))


def test_recorded_step_batch():
    sts = [tangles.random_rings(200 + k, 3 + 2 * k, 0.003, 0.008, 0.03, 5e-4) for k in range(4)]
    p = params(hip_precision="fp32c")
    assert_same(run(p, sts, "eager", [5], batch=True), run(p, sts, "graph", [5], batch=True))


def test_recorded_lia_step():
    p = params(velocity="LIA")
    st = dense()
    assert_same(run(p, st, "eager", [5]), run(p, st, "graph", [5]))


def tree_params(**kw):
    return params(velocity="Tree", tree_theta=0.5, **kw)


@pytest.mark.parametrize("shots", [1, 3])
@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_recorded_tree_step_equals_eager(shots, prec):
    p = tree_params(recon_shots=shots, hip_precision=prec)
    st = dense()
    ref = run(p, st, "eager", [7])
    got = run(p, st, "graph", [7])
    if shots == 1:
        assert len(ref[1]) > 0
    assert ref[2]["tree_replays"] == 0
    # the first step has no build to plan from; the others replay
    assert got[2]["tree_replays"] == 6 and got[2]["tree_redone"] == 0
    assert got[2]["tree_builds"] == ref[2]["tree_builds"] == 7
    assert_same(ref, got)


def test_recorded_tree_step_split_calls_and_eager_calls():
    def tree_velocity(s):
        s.ghostp()
        s.construct_tree()
        s.velocity()
    p = tree_params(recon_shots=2, hip_precision="fp32c")
    st = dense()
    assert_same(run(p, st, "eager", [7]), run(p, st, "graph", [2, 1, 3, 1]))
    assert_same(run(p, st, "eager", [2, 2, 2], between=tree_velocity),
                run(p, st, "graph", [2, 2, 2], between=tree_velocity))


def test_recorded_tree_step_outgrowing_its_plan():
    # rings with segments of 1.9 delta: step 1 inserts a point on every
    # segment, so step 2's build (planned on step 1's tree) outgrows its
    # level bounds; that replay stops after the build and the step is taken
    # again eagerly, after which the plan is recorded anew
    g = np.random.default_rng(11)
    loops = [tangles.ring(g.uniform(-0.015, 0.015, 3), 0.01, g.normal(size=3), 5e-4, chord=1.9)
             for _ in range(60)]
    st = concat(dense(), tangles._assemble(loops, (0.03, 0.03, 0.03)))
    p = tree_params(recon_shots=1, alpha=(0.206, 0.0083), normal_velocity="xflow", norm_vel_xflow=0.5)
    cap = 4 * st.pcount + 4096
    ref = run(p, st, "eager", [5], capacity=cap)
    got = run(p, st, "graph", [5], capacity=cap)
    assert ref[0][0].pcount > st.pcount + 3000
    assert got[2]["tree_redone"] >= 1 and got[2]["tree_replays"] >= 2
    assert_same(ref, got)


@pytest.mark.parametrize("velocity", ["BS", "Tree"])
@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_recorded_step_with_walls(velocity, prec):
    """Solid x and z walls (mirror images, pinned points, wall_recon) in the
    recorded step, against eager launches."""
    from nuvort_amd import tangles
    p = params(box_size=(0.1, 0.1, 0.1), boundary_x="solid", boundary_z="solid", velocity=velocity,
               hip_precision=prec, recon_shots=1)
    if velocity == "Tree":
        p.tree_theta = 0.5
    st = tangles.wall_tangle(41)
    ref = run(p, st, "eager", [4])
    got = run(p, st, "graph", [4])
    assert ref[0][0].extra["wall_recon_count"] > 0
    assert_same(ref, got)
    assert np.array_equal(ref[0][0].pinnedi, got[0][0].pinnedi)
    assert np.array_equal(ref[0][0].wpinned, got[0][0].wpinned)
