"""Test infrastructure only: the speed of the oracle's C restatement against
the reference build (oracle/_ref/nuvort_ref) on the same input, so that the
CPU baseline bench.py times with the restatement (the reference build cannot
travel) can be related to the reference itself.

    make -C oracle ref port && python oracle/calibrate.py [threads]

Both run one direct-sum velocity pass (pmotion: every point against every
segment in the 27 images) of the same 16-ring tangle with the same OpenMP
thread count; the ratio is written to oracle/calibration.json.  Run in the
development container (it needs the reference tree's build).
"""
from __future__ import annotations

import ctypes as C
import json
import os
import platform
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from nuvort_amd import tangles  # noqa: E402
from nuvort_amd.runfile import RunParams, format_run_text  # noqa: E402
from oracle import port, ref_io  # noqa: E402


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main(threads: int) -> None:
    st = tangles.random_rings(1236, 16, 0.01, 0.03, 0.1, 5e-4)
    p = RunParams(nsteps=1, shots=1, dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                  overide_timestep_check=True, recon_shots=1000)
    ref_s = []
    for _ in range(2):
        with tempfile.TemporaryDirectory(dir="/tmp") as wd:
            out = ref_io.run_reference(wd, st, format_run_text(p), 1, dump_phases=False, threads=threads)
        ref_s.append(out["timing"]["velocity_s"])
    L = port.lib()
    port_s = []
    for _ in range(2):
        ps = port.PortState(st)
        pr = port.params_struct(p, threads)
        L.nvo_ghostp(ps.ref(), C.byref(pr))
        t0 = time.perf_counter()
        rc = L.nvo_pmotion(ps.ref(), C.byref(pr), None)
        port_s.append(time.perf_counter() - t0)
        assert rc == 0
    res = {"points": st.active, "threads": threads, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "reference_velocity_s": min(ref_s), "restatement_velocity_s": min(port_s),
           "reference_over_restatement_time": min(ref_s) / min(port_s),
           "what": "one pmotion velocity pass (direct sum, 27 images) of a 16-ring tangle, seed 1236; reference "
                   "build amdflang -O2 -fdefault-real-8 -fopenmp vs the C restatement gcc -O2 -fopenmp "
                   "-ffp-contract=off, same thread count, best of two"}
    with open(os.path.join(HERE, "calibration.json"), "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1))
