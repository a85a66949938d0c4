"""Loading and comparing the golden fixtures of tests/golden/ (written by
tests/golden/make_golden.py from the reference build)."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from nuvort_amd.runfile import parse_run_text
from nuvort_amd.state import FilamentState

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STATE_FIELDS = ("x", "u", "u1", "u2", "u_sup", "u_mf", "ghosti", "ghostb", "infront", "behind", "closest",
                "closestd", "pinnedi", "pinnedb", "wpinned")


def names():
    """Fixtures of normal runs (the fatal_* ones end in fatal_error, the
    restart_* ones hold the output phase's files and a restarted run)."""
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(p).startswith(("fatal_", "restart_")))


def restart_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "restart_*.npz")))


def fatal_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "fatal_*.npz")))


class Fixture:
    def __init__(self, name: str):
        self.name = name
        self.z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        self.params = parse_run_text(str(self.z["run_text"]))
        self.nsteps = int(self.z["nsteps"])
        self.timing = json.loads(str(self.z["timing"])) if "timing" in self.z.files else {}

    def has(self, phase: str) -> bool:
        return ("%s__x" % phase) in self.z.files

    def state(self, phase: str) -> FilamentState:
        x = self.z["%s__x" % phase]
        st = FilamentState.empty(x.shape[0])
        for f in STATE_FIELDS:
            key = "%s__%s" % (phase, f)
            if key in self.z.files:   # (fixtures without walls hold no wpinned)
                getattr(st, f)[:] = self.z[key]
        meta = self.z["%s__meta" % phase]
        st.t, st.itime, st.recon_count = float(meta[0]), int(meta[1]), int(meta[2])
        return st

    def tree(self):
        if "tree__pcount" not in self.z.files:
            return None
        return {k[len("tree__"):]: self.z[k] for k in self.z.files if k.startswith("tree__")}

    @property
    def recon_log(self) -> np.ndarray:
        return self.z["recon_log"]

    @property
    def fatal_text(self) -> str:
        """What the reference printed when it stopped (fatal_* fixtures)."""
        return str(self.z["fatal_text"]) if "fatal_text" in self.z.files else ""

    def file(self, name: str) -> bytes:
        """A file the reference's output phase wrote (restart_* fixtures)."""
        return self.z["file__" + name].tobytes()

    def files(self):
        return sorted(k[len("file__"):] for k in self.z.files if k.startswith("file__"))

    def noise(self):
        """The noisey_xflow table the reference read, or None."""
        if "noise_seed" not in self.z.files:
            return None
        from nuvort_amd.tangles import noise_table
        return noise_table(int(self.z["noise_seed"]))


def compare_states(ref: FilamentState, got: FilamentState, n_defined: int, fields=None, rtol=0.0,
                   skip_u_aux=False, searched=True):
    """Return a list of mismatch descriptions (empty when equal).

    Rows at or beyond ``n_defined`` were grown by point insertion; the
    reference leaves u_sup, u_mf and closest undefined there (uninitialised
    memory, src/line.f90:51), so those fields are compared only below it.
    ``closest`` is compared only where a candidate was found (closestd < 100):
    otherwise the reference keeps its stale value (src/line.f90:142-155).
    ``searched`` says whether a close-point search ran since the rows grew;
    before it, closest and closestd of grown rows are undefined too.
    With ``rtol`` > 0 floating fields are compared to that relative tolerance
    (normalised by the field's max magnitude) instead of bit for bit.
    """
    bad = []
    if ref.pcount != got.pcount:
        return ["pcount %d != %d" % (ref.pcount, got.pcount)]
    fields = fields or STATE_FIELDS
    act = ref.infront != 0
    for f in fields:
        a, b = getattr(ref, f), getattr(got, f)
        rows = np.ones(a.shape[0], bool)
        if f in ("u_sup", "u_mf") or (not searched and f in ("closest", "closestd")):
            rows[n_defined:] = False
            if skip_u_aux and f in ("u_sup", "u_mf"):
                continue
        if f == "closest":
            rows &= act & (ref.closestd < 100.0)
        if f == "closestd":
            rows &= act
        a, b = a[rows], b[rows]
        if a.dtype.kind == "f" and rtol > 0:
            scale = max(np.abs(a).max() if a.size else 0.0, 1e-300)
            err = np.abs(a - b).max() / scale if a.size else 0.0
            if not err <= rtol:
                bad.append("%s: rel err %.3e > %.1e" % (f, err, rtol))
        else:
            neq = a != b
            if a.dtype.kind == "f":
                neq &= ~(np.isnan(a) & np.isnan(b))
            if np.any(neq):
                idx = np.nonzero(neq.reshape(neq.shape[0], -1).any(1))[0]
                bad.append("%s: %d rows differ (first rows %s)" % (f, idx.size, idx[:5].tolist()))
    return bad
