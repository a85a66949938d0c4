"""The autotuner picks a direct-sum variant by timing it on the uploaded
state, leaves the state untouched, keeps the FP32C error bound, and reads
its choice back from the cache for the same GPU and size class."""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams

pytestmark = pytest.mark.gpu

FIELDS = ("x", "u", "u1", "u2", "u_sup", "u_mf", "ghosti", "ghostb", "infront", "behind", "closest", "closestd")


def params(**kw):
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                  hip_precision="fp32c", overide_timestep_check=True)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_autotune_keeps_state_and_accuracy(tmp_path):
    from nuvort_amd.engine import FilamentSolver
    st = tangles.random_rings(21, 24, 0.01, 0.03, 0.1, 5e-4)
    s = FilamentSolver(params())
    s.set_state(st)
    s.ghostp()
    before = s.get_state()
    cache = str(tmp_path / "tune.txt")
    t = s.autotune(cache)
    assert t["tuned"] == 1 and t["tune_ms"] > 0
    assert t["targets_per_lane"] in (1, 2) and t["kahan_block"] in (4, 8, 16, 32) and t["items_per_cu"] in (8, 16, 32)
    assert t["source_tile"] in (128, 256)
    after = s.get_state()
    for f in FIELDS:
        assert np.array_equal(getattr(before, f), getattr(after, f)), f
    s.velocity()
    u32 = s.get_state().u_sup
    s.close()
    r = FilamentSolver(params(hip_precision="fp64"))
    r.set_state(st)
    r.ghostp()
    r.velocity()
    u64 = r.get_state().u_sup
    r.close()
    act = st.infront != 0
    mag = np.linalg.norm(u64[act], axis=1)
    assert (np.linalg.norm(u32[act] - u64[act], axis=1) / mag).max() <= 1e-5
    # the same size class on the same GPU reads the cached choice
    s2 = FilamentSolver(params())
    s2.set_state(tangles.random_rings(22, 24, 0.01, 0.03, 0.1, 5e-4))
    t2 = s2.autotune(cache)
    if t2["tuned"] == 2:
        assert (t2["targets_per_lane"], t2["kahan_block"], t2["items_per_cu"], t2["source_tile"]) == \
               (t["targets_per_lane"], t["kahan_block"], t["items_per_cu"], t["source_tile"])
    else:   # a different size class (log2 of the point count changed)
        assert t2["tuned"] == 1
    s2.close()
    assert len(open(cache).read().strip().splitlines()) in (1, 2)


def test_autotune_then_recorded_steps_match_eager(tmp_path):
    # the eager run reads the graph run's choice from the cache, so both
    # run the same variant and must agree bit for bit
    from nuvort_amd.engine import FilamentSolver
    st = tangles.random_rings(10, 20, 0.003, 0.008, 0.03, 5e-4)
    cache = str(tmp_path / "tune.txt")
    out = []
    for launch in ("graph", "eager"):
        p = params(box_size=(0.03, 0.03, 0.03), hip_step_launch=launch)
        s = FilamentSolver(p)
        s.set_state(st)
        s.step(2)
        t = s.autotune(cache)
        assert t["tuned"] == (1 if launch == "graph" else 2)
        s.step(3)
        out.append(s.get_state())
        s.close()
    for f in FIELDS:
        assert np.array_equal(getattr(out[0], f), getattr(out[1], f)), f
