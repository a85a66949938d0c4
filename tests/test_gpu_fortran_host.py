"""The Fortran host (nuvort_amd/fortran/run_hip) drives the library through
the iso_c_binding layer in the reference's call order; its result must match
the reference on the golden fixtures exactly as the Python host does."""
import os
import subprocess

import numpy as np
import pytest

from golden_util import Fixture, names
from oracle import ref_io

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "nuvort_amd", "fortran", "run_hip")


@pytest.mark.parametrize("name", [n for n in names() if n in ("recon_bs", "recon_tree", "remesh_bs",
                                                               "tree04_xflow", "kelvin_traj20", "tree_mesh")])
def test_fortran_host_matches_reference(name, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("Fortran host not built (make -C nuvort_amd/fortran)")
    fx = Fixture(name)
    (tmp_path / "run.in").write_text(str(fx.z["run_text"]))
    ref_io.write_state(str(tmp_path / "in.bin"), fx.state("in"))
    (tmp_path / "data").mkdir()
    r = subprocess.run([EXE, "in.bin", "out.bin", str(fx.nsteps)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = ref_io.read_state(str(tmp_path / "out.bin"))
    ref = fx.state("final")
    assert got.pcount == ref.pcount
    assert np.array_equal(got.infront, ref.infront)
    assert np.array_equal(got.behind, ref.behind)
    assert got.recon_count == ref.recon_count
    act = ref.infront != 0
    assert np.abs(got.x[act] - ref.x[act]).max() <= 1e-10 * fx.params.box_size[0]
    for fname in fx.files():
        # files the reference wrote that the host writes too (tree_print's leaf list)
        a = np.array([float(v) for v in (tmp_path / "data" / fname).read_text().split()])
        b = np.array([float(v) for v in fx.file(fname).decode().split()])
        assert np.array_equal(a, b), fname
