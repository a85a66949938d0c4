"""Write the golden fixtures under tests/golden/ from the reference build.

Each fixture is an input state made by this project's seeded generators,
run through oracle/_ref/nuvort_ref (the reference's own Fortran modules,
compiled unmodified by ``make -C oracle ref``) with OMP_NUM_THREADS=1, and the
phase-by-phase outputs it dumps.  Run in the development container only (it
needs the reference tree and amdflang):

    make -C oracle ref && python tests/golden/make_golden.py

Fixture file layout (npz): ``run_text`` (the run.in used), ``nsteps``,
``in__<field>`` (input state), ``<phase>__<field>`` for the phases ghost,
motion, insert, close, recon, remove (first step) and final (after nsteps),
``tree__<field>`` (depth-first node list of the first step's tree when the
tree is on) and ``recon_log`` (rows of src/reconnection.f90:48).
"""
from __future__ import annotations

import json
import math
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nuvort_amd import tangles  # noqa: E402
from nuvort_amd.runfile import RunParams, format_run_text  # noqa: E402
from nuvort_amd.state import FilamentState  # noqa: E402
from oracle import ref_io  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
FIELDS = ("x", "u", "u1", "u2", "u_sup", "u_mf", "ghosti", "ghostb", "infront", "behind", "closest",
          "closestd", "pinnedi", "pinnedb", "wpinned")


def base_params(**kw) -> RunParams:
    p = RunParams(nsteps=1, shots=1, dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1),
                  velocity="BS", overide_timestep_check=True)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def pack_state(prefix: str, st: FilamentState, d: dict) -> None:
    for name in FIELDS:
        d["%s__%s" % (prefix, name)] = getattr(st, name)
    d["%s__meta" % prefix] = np.array([st.t, st.itime, st.recon_count], np.float64)


ALL_PHASES = ("ghost", "motion", "insert", "close", "recon", "wall", "remove", "final")


def make(name: str, st: FilamentState, p: RunParams, nsteps: int, note: str, phases=ALL_PHASES,
         noise_seed: int | None = None, keep_files=()) -> None:
    """Run the reference build and store the fixture.  The ghost phase (the
    input of every per-phase test) is always kept.  With ``noise_seed`` the
    run reads the noisey_xflow table tangles.noise_table(noise_seed) from
    ``noise.dat``; the fixture stores the seed, not the table.  ``keep_files``
    names files of data/ the run writes (tree_print's leaf lists) to store as
    file__<name>."""
    files = None
    if noise_seed is not None:
        p.noise_file = "noise.dat"
        files = {"noise.dat": "\n".join("%.17g" % v for v in tangles.noise_table(noise_seed)) + "\n"}
    txt = format_run_text(p)
    with tempfile.TemporaryDirectory(dir="/tmp") as wd:
        out = ref_io.run_reference(wd, st, txt, nsteps, dump_phases=True, threads=1, files=files)
    d = {"run_text": np.array(txt), "nsteps": np.array(nsteps), "note": np.array(note)}
    for fname in keep_files:
        d["file__" + fname] = np.frombuffer(out["files"][fname], np.uint8)
    if noise_seed is not None:
        d["noise_seed"] = np.array(noise_seed)
    pack_state("in", st, d)
    for ph in ("ghost",) + tuple(phases) + ("final",):
        if ph in out and "%s__x" % ph not in d:
            pack_state(ph, out[ph], d)
    if "tree" in out:
        for k in out["tree"].dtype.names:
            d["tree__" + k] = out["tree"][k]
    d["recon_log"] = out["recon_log"]
    d["timing"] = np.array(json.dumps(out["timing"]))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
    print("%-22s N=%5d steps=%3d recon=%d final_pcount=%d" % (
        name, st.pcount, nsteps, len(out["recon_log"]), out["final"].pcount))


def make_restart(name: str, st: FilamentState, p: RunParams, nsteps: int, cont: int, note: str) -> None:
    """The reference's output phase over nsteps steps (shots 1: diagnostics,
    var.dat, var%04d.log and the ts.log line every step), then a run of
    `cont` steps restarted from the var.dat it wrote.  Stores the files'
    bytes (file__<name>), the state after nsteps, and the restarted run's
    final state (cont__*)."""
    p.shots = 1
    txt = format_run_text(p)
    with tempfile.TemporaryDirectory(dir="/tmp") as wd:
        out = ref_io.run_reference(wd, st, txt, nsteps, dump_phases=False, threads=1, output=True)
    with tempfile.TemporaryDirectory(dir="/tmp") as wd:
        var = os.path.join(wd, "var_saved.dat")
        with open(var, "wb") as fh:
            fh.write(out["files"]["var.dat"])
        c = ref_io.run_reference(os.path.join(wd, "run"), None, txt, cont, dump_phases=False, threads=1,
                                 state_path=var)
    d = {"run_text": np.array(txt), "nsteps": np.array(nsteps), "cont_steps": np.array(cont), "note": np.array(note)}
    pack_state("in", st, d)
    pack_state("final", out["final"], d)
    pack_state("cont", c["final"], d)
    for fname, data in out["files"].items():
        d["file__" + fname] = np.frombuffer(data, np.uint8)
    d["recon_log"] = out["recon_log"]
    d["timing"] = np.array(json.dumps(out["timing"]))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
    print("%-22s N=%5d steps=%3d+%d files=%s" % (name, st.pcount, nsteps, cont, sorted(out["files"])))


def crossing_rings(y0: float, delta: float) -> FilamentState:
    """Ring A in the xy-plane (centre 0, R=0.01); ring B in the plane y=y0,
    centred at (0.004, y0, 0) with R=0.006, so B passes y0 away from A's
    point (0.01, 0, 0) with a perpendicular tangent: a reconnection pair."""
    def loop(center, r, e1, e2, n):
        phi = 2 * math.pi * np.arange(n) / n
        return center[None, :] + r * np.cos(phi)[:, None] * e1[None, :] + r * np.sin(phi)[:, None] * e2[None, :]
    nA = math.ceil(2 * math.pi * 0.01 / (0.98 * delta))
    nB = math.ceil(2 * math.pi * 0.006 / (0.98 * delta))
    A = loop(np.zeros(3), 0.01, np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), nA)
    B = loop(np.array([0.004, y0, 0.0]), 0.006, np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), nB)
    return tangles._assemble([A, B], (0.1, 0.1, 0.1))


def remesh_case(delta: float) -> FilamentState:
    """Three rings: one stretched (chord ~1.3 delta, insertions), one
    compressed (chord ~0.45 delta, removals), one regular; two empty slots
    interleaved so insertions reuse them (first empty slot, src/line.f90:44-47)."""
    A = tangles.ring([0.01, 0.0, 0.0], 0.006, [0, 0, 1], delta, chord=1.3)
    B = tangles.ring([-0.02, 0.01, 0.0], 0.004, [0, 1, 0], delta, chord=0.45)
    Cc = tangles.ring([0.0, -0.025, 0.01], 0.005, [1, 1, 0], delta)
    st = tangles._assemble([A, B, Cc], (0.1, 0.1, 0.1))
    n = st.pcount
    holes = {len(A), len(A) + len(B) + len(Cc) // 2 + 1}   # rows left empty
    rows = []
    for old in range(n):
        while len(rows) in holes:
            rows.append(-1)
        rows.append(old)
    mapping = np.zeros(n + 1, np.int64)
    for r, old in enumerate(rows):
        if old >= 0:
            mapping[old + 1] = r + 1
    out = FilamentState.empty(len(rows))
    for r, old in enumerate(rows):
        if old < 0:
            continue
        for name, arr in st.arrays():
            getattr(out, name)[r] = arr[old]
        out.infront[r] = mapping[st.infront[old]]
        out.behind[r] = mapping[st.behind[old]]
    return out


def make_fatal(name: str, st: FilamentState, p: RunParams, nsteps: int, note: str) -> None:
    """A state on which the reference stops in fatal_error: the fixture keeps
    the input and the message the reference printed."""
    txt = format_run_text(p)
    with tempfile.TemporaryDirectory(dir="/tmp") as wd:
        out = ref_io.run_reference(wd, st, txt, nsteps, dump_phases=False, threads=1, expect_fatal=True)
    text = out["stdout"] + out["stderr"]
    lines = [ln.strip() for ln in text.splitlines() if ln.strip()]
    fatal = [ln for ln in lines if ".mod" in ln or "run.x" in ln or "FATAL" in ln.upper()]
    assert out["returncode"] != 0 or fatal, "the reference did not stop: %r" % text[-500:]
    d = {"run_text": np.array(txt), "nsteps": np.array(nsteps), "note": np.array(note),
         "fatal_text": np.array("\n".join(fatal[-6:])), "returncode": np.array(out["returncode"])}
    pack_state("in", st, d)
    np.savez_compressed(os.path.join(OUT, "fatal_" + name + ".npz"), **d)
    print("%-22s N=%5d fatal: %s" % ("fatal_" + name, st.pcount, " | ".join(fatal[-3:])))


def singular_state(delta: float, L: float) -> FilamentState:
    """Two rings that touch through the periodic x faces: a point of ring A
    sits exactly on the +x face and a point of ring B exactly on the -x
    face, with the same y and z, so that in the image shifted by +L the tree
    leaf of B's point coincides with A's point (src/tree.f90:289-295).  The
    +x face point lies in no child of the root (half-open boxes), so only
    that one target meets the singularity."""
    A = tangles.ring([L / 2 - 0.006, 0.0, 0.0], 0.006, [0, 0, 1], delta)
    B = tangles.ring([-L / 2 - 0.006, 0.0, 0.0], 0.006, [0, 1, 0], delta, phase=math.pi)
    A[int(np.argmax(A[:, 0]))] = (L / 2, 0.0, 0.0)
    B[int(np.argmin(np.abs(B[:, 0] + L / 2)))] = (-L / 2, 0.0, 0.0)
    return tangles._assemble([A, B], (L, L, L))


def deep_tree_state(delta: float) -> FilamentState:
    """Three rings; a point of ring B is moved to 4e-7 cm from a point of ring
    A and a point of ring C to 5e-12 cm from another point of ring A, so the
    octree needs more than 16 (box / 2^16 = 1.5e-6 cm) and more than 32
    levels (box / 2^32 = 2.3e-11 cm) to separate them (src/tree.f90:64-104
    recurses while a node holds more than one point)."""
    A = tangles.ring([0.0, 0.0, 0.0], 0.01, [0, 0, 1], delta)
    B = tangles.ring([0.0, 0.0105, 0.0], 0.006, [1, 0, 0], delta)
    Cc = tangles.ring([0.0, -0.0105, 0.0], 0.006, [1, 0, 0], delta)
    ia = int(np.argmax(A[:, 1]))          # A's point nearest +y
    ib = int(np.argmin(B[:, 1]))          # B's point nearest A
    B[ib] = A[ia] + np.array([4.0e-7, 1.0e-7, -2.0e-7])
    ja = int(np.argmin(A[:, 1]))
    jc = int(np.argmax(Cc[:, 1]))
    Cc[jc] = A[ja] + np.array([-5.0e-12, 3.0e-12, 0.0])
    return tangles._assemble([A, B, Cc], (0.1, 0.1, 0.1))


def fatal_cases(d5: float) -> None:
    rings = tangles.random_rings(7, 4, 0.004, 0.012, 0.1, d5)
    nan = rings.copy()
    nan.x[5, 1] = float("nan")
    make_fatal("nan", nan, base_params(), 1, "a NaN position: NAN_finder stops the run (src/general.f90:172-191)")
    empty = rings.copy()
    empty.infront[:] = 0
    empty.behind[:] = 0
    make_fatal("zero_length", empty, base_params(), 1,
               "every slot empty: pinsert stops the run (src/line.f90:25-27)")
    make_fatal("singular", singular_state(d5, 0.1), base_params(velocity="Tree", tree_theta=0.5), 1,
               "a target on a shifted image's leaf: tree_walk stops the run (src/tree.f90:289-295)")


def main() -> None:
    d5 = 5e-4
    rings = tangles.random_rings(7, 4, 0.004, 0.012, 0.1, d5)
    make("bs_rings_1step", rings, base_params(), 1, "BS, 4 random rings, one step with phase dumps")
    make("bs_rings_5step", rings, base_params(), 5, "BS, Euler->AB2->AB3 start-up over 5 steps", ("motion", "final"))
    make("tree05_rings", rings, base_params(velocity="Tree", tree_theta=0.5), 2,
         "Tree theta=0.5, tree node list + pclose_tree")
    mf = base_params(velocity="Tree", tree_theta=0.4, normal_velocity="xflow", norm_vel_xflow=0.5,
                     alpha=(0.206, 0.0083))
    make("tree04_xflow", rings, mf, 2, "Tree theta=0.4 with xflow mutual friction",
         ("motion", "final"))
    make("bs_xflow", rings, base_params(normal_velocity="xflow", norm_vel_xflow=0.5, alpha=(0.206, 0.0083),
                                        super_velocity=(0.0, 0.01, 0.0)), 1,
         "BS with xflow mutual friction and an imposed super_velocity", ("motion",))
    ab = rings.copy()
    g = np.random.Generator(np.random.PCG64(11))
    ab.u1[:] = g.normal(scale=0.05, size=ab.u1.shape)
    ab.u2[:] = g.normal(scale=0.05, size=ab.u2.shape)
    sel = g.uniform(size=ab.pcount)
    ab.u1[sel < 0.3] = 0.0          # Euler
    ab.u2[(sel >= 0.3) & (sel < 0.6)] = 0.0   # AB2
    ab.u1[sel > 0.95, 0] = -0.0     # a negative zero among the history
    make("bs_ab_startup", ab, base_params(), 1, "per-point Euler/AB2/AB3 selection from u1/u2 history",
         ("motion",))
    kel = tangles.kelvin_lattice(3, box=0.1, nx=2, ny=1, amp=0.002, waves=3, delta=2e-3)
    make("kelvin_traj20", kel, base_params(delta=2e-3, dt=1e-5), 20, "2 Kelvin-wave lines, 20-step trajectory",
         ("motion", "final"))
    cr = crossing_rings(1.0e-4, d5)
    make("recon_bs", cr, base_params(), 3, "crossing rings: brute close search + reconnection")
    make("recon_tree", cr, base_params(velocity="Tree", tree_theta=0.5), 3,
         "crossing rings: stale-tree close search + reconnection")
    rm = remesh_case(d5)
    make("remesh_bs", rm, base_params(), 2, "stretched/compressed rings with empty slots: pinsert/premove")
    make("remesh_tree", rm, base_params(velocity="Tree", tree_theta=0.5), 2,
         "remesh case under Tree: stale tree with inserted points")
    # LIA (src/timestep.f90:60-73) with every topology phase
    make("lia_rings", rings, base_params(velocity="LIA"), 3, "LIA, 4 random rings, 3 steps")
    # the closed-form normal-fluid models and the noise table
    # (src/normal_fluid.f90:94-115), with friction, started at t > 0 so the
    # time-dependent shear has a phase
    late = rings.copy()
    late.t = 2.5e-5
    mfk = dict(norm_vel_xflow=0.5, alpha=(0.206, 0.0083), super_velocity=(0.0, 0.003, 0.0))
    make("nf_shear_xflow", late, base_params(normal_velocity="shear_xflow", norm_shear_omega=4.0e4, **mfk), 2,
         "BS, shear_xflow normal flow with friction", ("motion",))
    make("nf_ychannel", late, base_params(normal_velocity="ychannel", **mfk), 2,
         "BS, ychannel normal flow with friction", ("motion",))
    make("nf_channel", late, base_params(normal_velocity="channel", **mfk), 2,
         "BS, channel normal flow with friction", ("motion",))
    make("nf_noisey_xflow", late, base_params(normal_velocity="noisey_xflow", **mfk), 3,
         "BS, noisey_xflow normal flow (seeded noise table) with friction", ("motion",), noise_seed=77)
    make("nf_channel_tree", late, base_params(velocity="Tree", tree_theta=0.4, normal_velocity="channel", **mfk), 2,
         "Tree theta 0.4, channel normal flow with friction", ("motion",))
    make_restart("restart_bs", rings, base_params(normal_velocity="xflow", norm_vel_xflow=0.5,
                                                 alpha=(0.206, 0.0083)), 3, 2,
                 "BS with friction: 3 steps with the output phase (var.dat, var%04d.log, ts.log), 2 restarted")
    make_restart("restart_tree", crossing_rings(1.0e-4, d5), base_params(velocity="Tree", tree_theta=0.5), 2, 2,
                 "Tree with a reconnection: 2 steps with the output phase, 2 restarted from var.dat")
    make_restart("restart_remesh", remesh_case(d5), base_params(), 2, 1,
                 "stretched/compressed rings with empty slots: pinsert's total length over reused slots, "
                 "the output phase over 2 steps, 1 restarted")
    make_restart("restart_walls", tangles.wall_tangle(41), base_params(boundary_x="solid", boundary_z="solid"), 2, 1,
                 "solid x and z walls: pinned points in var.dat, wall_recon_count in ts.log, 1 step restarted")
    # tree_print (src/tree.f90:96-102): the leaf list of the step's octree
    make("tree_mesh", rings, base_params(velocity="Tree", tree_theta=0.5, tree_print=True), 1,
         "Tree theta 0.5 with tree_print: the leaf list of the step's octree", keep_files=("tree_mesh001.log",))
    make("tree_deep", deep_tree_state(d5), base_params(velocity="Tree", tree_theta=0.5), 1,
         "Tree theta 0.5 with point pairs 4e-7 and 5e-12 cm apart (octree deeper than 16 and 32 levels)")
    # solid walls at x = +-L/2 and z = +-L/2, y periodic: pinned lines and
    # half loops, a ring cut by wall_recon, a pinned pair cleared
    # (src/timestep.f90:201-234, src/tree.f90:321-392, src/boundary.f90:31-38,
    # 124-135, src/reconnection.f90:66-150)
    walls = dict(boundary_x="solid", boundary_z="solid")
    wst = tangles.wall_tangle(41)
    make("walls_bs", wst, base_params(**walls), 3, "BS with solid x and z walls: mirror images, pinned points, "
         "wall reconnection, 3 steps")
    make("walls_tree", wst, base_params(velocity="Tree", tree_theta=0.5, **walls), 3,
         "Tree theta 0.5 with solid x and z walls: mirrored tree walks, pinned points, wall reconnection")
    make("walls_xflow", wst, base_params(normal_velocity="xflow", norm_vel_xflow=0.5, alpha=(0.206, 0.0083),
                                         boundary_x="solid", boundary_y="solid", boundary_z="solid"), 2,
         "BS with friction and walls in all three directions (26 mirror images)", ("motion", "wall", "final"))
    dense = tangles.random_rings(21, 24, 0.003, 0.008, 0.025, d5)
    few = ("motion", "close", "final")
    make("dense_bs_10step", dense, base_params(box_size=(0.025, 0.025, 0.025)), 10,
         "dense random rings in a 0.025 cm box: 10 steps with reconnections", few)
    make("dense_tree_10step", dense, base_params(box_size=(0.025, 0.025, 0.025), velocity="Tree",
                                                 tree_theta=0.5), 10,
         "dense random rings, Tree theta=0.5, 10 steps", few)


if __name__ == "__main__":
    if sys.argv[1:] == ["fatal"]:
        fatal_cases(5e-4)
    elif sys.argv[1:2] == ["only"]:
        # one fixture by name (the others are left as they are)
        want = set(sys.argv[2:])
        real_make, real_restart = make, make_restart

        def make(name, *a, **kw):   # noqa: F811
            if name in want:
                real_make(name, *a, **kw)

        def make_restart(name, *a, **kw):   # noqa: F811
            if name in want:
                real_restart(name, *a, **kw)
        main()
    else:
        main()
        fatal_cases(5e-4)
