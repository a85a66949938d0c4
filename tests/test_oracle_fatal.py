"""The fatal conditions of the path on the oracle: each fatal_* fixture is a
state on which the reference build stopped in fatal_error (the fixture keeps
the message it printed); the oracle must stop with the matching status.

  * a NaN position: NAN_finder, src/general.f90:172-191
  * every slot empty: pinsert, src/line.f90:25-27
  * a target on a leaf of a shifted image: tree_walk, src/tree.f90:289-295
"""
import pytest

from golden_util import Fixture, fatal_names
from oracle import port

# oracle status (oracle/nv_oracle.h) for each reference message
ORACLE_CODE = {"there is a NAN": 4, "vortex line length is 0": 2, "singularity in BS (tree)": 3}


def expected(fx):
    hits = [code for text, code in ORACLE_CODE.items() if text in fx.fatal_text]
    assert len(hits) == 1, fx.fatal_text
    return hits[0]


def test_every_fatal_condition_has_a_fixture():
    texts = " ".join(Fixture(n).fatal_text for n in fatal_names())
    for text in ORACLE_CODE:
        assert text in texts, text


@pytest.mark.parametrize("name", fatal_names())
def test_oracle_stops_where_the_reference_does(name):
    fx = Fixture(name)
    rc, _, _ = port.step(fx.state("in"), fx.params, fx.nsteps, threads=2)
    assert rc == expected(fx)
