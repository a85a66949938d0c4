"""The FP32 tree walk (FP32 node terms, FP64 leaves within r_near) against
the FP64 walk, which follows the reference's arithmetic (src/tree.f90:250-316),
on the same tree.  The FP32 walk decides every opening test the reference's
way (in FP64 where the FP32 squared distance lies within rounding of the
opening radius), so both walks visit and accept the same nodes for every
target and image, take the same interactions, and the velocities differ only
by the FP32 rounding of the node terms (amplified by their cancellation)."""
import numpy as np
import pytest

from nuvort_amd import tangles
from nuvort_amd.runfile import RunParams

pytestmark = pytest.mark.gpu


def tree_velocity(st, p, prec):
    from nuvort_amd.engine import FilamentSolver
    q = p.__class__(**{k: getattr(p, k) for k in p.__dataclass_fields__})
    q.hip_precision = prec
    s = FilamentSolver(q)
    s.set_state(st)
    s.ghostp()
    s.construct_tree()
    s.velocity()
    g = s.get_state()
    st_ = s.stats()
    s.close()
    return g, st_["tree_evals"], st_["tree_nodes_visited"]


@pytest.mark.parametrize("theta,xflow", [(0.5, False), (0.4, True)])
def test_fp32_walk_matches_fp64_walk(theta, xflow):
    st = tangles.random_rings(1236, 64, 0.01, 0.03, 0.1, 5e-4)
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="Tree",
                  tree_theta=theta, overide_timestep_check=True)
    if xflow:
        p.normal_velocity, p.norm_vel_xflow, p.alpha = "xflow", 0.5, (0.206, 0.0083)
    g64, ev64, vis64 = tree_velocity(st, p, "fp64")
    g32, ev32, vis32 = tree_velocity(st, p, "fp32c")
    act = g64.infront != 0
    mag = np.linalg.norm(g64.u_sup[act], axis=1)
    err = np.linalg.norm(g32.u_sup[act] - g64.u_sup[act], axis=1)
    sel = mag >= 1e-3 * mag.max()
    rel = err[sel] / mag[sel]
    # targets within r_near of a periodic face, whose shifted images hold
    # FP64 near-field terms
    rn = p.hip_near_radius * p.delta
    face = (np.abs(np.abs(g64.x[act]) - p.box_size[0] / 2) < rn).any(axis=1)[sel]
    print("theta %.1f: max rel %.2e (near a face %.2e), 99%% %.2e, 99.9%% %.2e, rms %.2e, interactions %d / %d, "
          "visits %d / %d" % (theta, rel.max(), rel[face].max(), np.quantile(rel, 0.99), np.quantile(rel, 0.999),
                              np.sqrt(np.mean(rel ** 2)), ev32, ev64, vis32, vis64))
    assert vis32 == vis64            # the same nodes visited: every opening test decided alike
    assert ev32 == ev64              # and the same interactions taken
    # A node term in FP32 carries ~1e-6 relative error; the node terms of a
    # walk cancel (their magnitudes sum to 10-100x |u|)
    assert rel.max() <= 1e-4
    assert rel[face].max() <= 1e-4
    assert np.sqrt(np.mean(rel ** 2)) <= 1e-5
