"""The fatal conditions through the C-ABI: on each fatal_* fixture (a state
on which the reference build stopped in fatal_error, with the message it
printed) the library returns the matching status and message instead of
stopping, from the eager phases and from nv_step.  Every kernel exits
normally: the conditions are status bits the kernels set.

  * a NaN position: NAN_finder, src/general.f90:172-191 -> NV_ERR_NAN
  * every slot empty: pinsert, src/line.f90:25-27 -> NV_ERR_ZERO_LENGTH
  * a target on a leaf of a shifted image: tree_walk, src/tree.f90:289-295
    -> NV_ERR_SINGULAR (FP64 and FP32 walks)
"""
import dataclasses

import pytest

from golden_util import Fixture, fatal_names

pytestmark = pytest.mark.gpu

CODE = {"there is a NAN": 6, "vortex line length is 0": 4, "singularity in BS (tree)": 5}


def expected(fx):
    hits = [(text, code) for text, code in CODE.items() if text in fx.fatal_text]
    assert len(hits) == 1, fx.fatal_text
    return hits[0]


def run(fx, launch, precision="fp64"):
    from nuvort_amd import lib
    from nuvort_amd.engine import FilamentSolver
    p = dataclasses.replace(fx.params, hip_step_launch=launch, hip_precision=precision)
    s = FilamentSolver(p, check=False)
    s.set_state(fx.state("in"))
    print("fatal case %s, %s launch, %s" % (fx.name, launch, precision), flush=True)
    try:
        s.step(fx.nsteps)
    except lib.NuvortError as e:
        return e.code, str(e)
    finally:
        s.close()
    return 0, ""


@pytest.mark.parametrize("launch", ["graph", "eager"])
@pytest.mark.parametrize("name", fatal_names())
def test_step_returns_the_fatal_status(name, launch):
    fx = Fixture(name)
    text, code = expected(fx)
    precisions = ["fp64", "fp32c"] if fx.params.velocity == "Tree" else ["fp64"]
    for prec in precisions:
        got, msg = run(fx, launch, prec)
        assert got == code, (prec, msg)
        assert text in msg


def test_nan_check_phase():
    from nuvort_amd import lib
    from nuvort_amd.engine import FilamentSolver
    fx = Fixture("fatal_nan")
    s = FilamentSolver(fx.params, check=False)
    s.set_state(fx.state("in"))
    with pytest.raises(lib.NuvortError) as e:
        s.nan_finder()
    assert e.value.code == 6
    s.close()


def test_zero_length_phase():
    from nuvort_amd import lib
    from nuvort_amd.engine import FilamentSolver
    fx = Fixture("fatal_zero_length")
    s = FilamentSolver(fx.params, check=False)
    s.set_state(fx.state("in"))
    s.ghostp()
    s.pmotion()
    with pytest.raises(lib.NuvortError) as e:
        s.pinsert()
    assert e.value.code == 4
    s.close()


@pytest.mark.parametrize("launch", ["graph", "eager"])
def test_noisey_xflow_without_its_table(launch):
    """noisey_xflow reads its noise table at set-up (src/normal_fluid.f90:
    25-60); a context given none refuses to step (NV_ERR_STATE), and steps
    once the table is set."""
    import numpy as np
    from nuvort_amd import lib, tangles
    from nuvort_amd.engine import FilamentSolver
    from nuvort_amd.runfile import RunParams
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                  normal_velocity="noisey_xflow", norm_vel_xflow=0.5, alpha=(0.206, 0.0083),
                  hip_step_launch=launch)
    s = FilamentSolver(p, check=False)
    try:
        s.set_state(tangles.random_rings(5, 3, 0.01, 0.02, 0.1, 5e-4))
        with pytest.raises(lib.NuvortError) as e:
            s.step(1)
        assert e.value.code == 10 and "noise" in str(e.value)
        s.set_noise(tangles.noise_table(3))
        s.step(1)
        assert s.get_state().itime == 1
    finally:
        s.close()
