"""run.in parameters and parser with the reference's semantics.

``read_run_file`` restates src/cdata.f90:131-241:

* each line (at most 100 characters) splits at its first blank into a label
  and the rest;
* the rest is read list-directed into the parameter the label names;
* unknown labels (and comment lines) are skipped silently;
* the FIRST read error ends parsing: that parameter and every later line keep
  their defaults (e.g. a one-value ``box_size`` line, which is how the
  reference's shipped run.in fails).

Defaults are those of src/cdata.f90:76-127.  Keys of this library's own
(prefixed ``hip_``) are unknown labels to the reference parser, so a run.in
carrying them still runs the reference unchanged.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field


class RunFileError(ValueError):
    pass


@dataclass
class RunParams:
    nsteps: int | None = None
    shots: int | None = None
    recon_shots: int = 1
    pcount: int | None = None
    dt: float | None = None
    delta: float | None = None
    quant_circ: float = 9.97e-4
    corea: float = 8.244023e-9
    box_size: tuple = (0.0, 0.0, 0.0)
    velocity: str | None = None
    boundary_x: str = "periodic"
    boundary_y: str = "periodic"
    boundary_z: str = "periodic"
    tree_theta: float = 0.0
    tree_print: bool = False
    tree_extra_correction: bool = True
    normal_velocity: str = "zero"
    normal_fluid_cutoff: float = 1e8
    t_zero_normal_fluid: float = 1e8
    norm_vel_xflow: float = 0.5
    alpha: tuple = (0.0, 0.0)
    super_velocity: tuple = (0.0, 0.0, 0.0)
    initf: str | None = None
    scaling_factor: float = 1.0
    loop_translate: tuple = (1.0, 1.0, 1.0)
    rotation_factor: float = 1.0
    line_count: int = 1
    NAN_test: bool = True
    max_warn_count: int = 5
    overide_timestep_check: bool = False
    serial_run: bool = False
    qvort_nproc: int = 0
    wave_amp: float = 1.0
    noise_file: str = "set_me!"
    norm_shear_omega: float = 0.0
    lattice_ratio: float = 1.0
    # this library's own keys (ignored by the reference parser)
    hip_precision: str = "fp64"          # fp64 | fp32c
    hip_near_radius: float = 8.0         # FP64 near field, in units of delta
    hip_close_search: str = "auto"       # auto | brute | tree
    hip_step_launch: str = "graph"       # graph | eager (how nv_step issues a step's kernels)
    extra: dict = field(default_factory=dict)


# label -> (field, kind, count); kinds: i integer, r real, l logical, c<len> char
_KEYS = {
    "nsteps": ("nsteps", "i", 1), "shots": ("shots", "i", 1), "recon_shots": ("recon_shots", "i", 1),
    "pcount": ("pcount", "i", 1), "dt": ("dt", "r", 1), "delta": ("delta", "r", 1),
    "quant_circ": ("quant_circ", "r", 1), "corea": ("corea", "r", 1),
    "box_size": ("box_size", "r", 3), "velocity": ("velocity", "c30", 1),
    "boundary_x": ("boundary_x", "c40", 1), "boundary_y": ("boundary_y", "c40", 1),
    "boundary_z": ("boundary_z", "c40", 1), "tree_theta": ("tree_theta", "r", 1),
    "tree_print": ("tree_print", "l", 1), "tree_extra_correction": ("tree_extra_correction", "l", 1),
    "normal_velocity": ("normal_velocity", "c30", 1),
    "normal_fluid_cutoff": ("normal_fluid_cutoff", "r", 1),
    "t_zero_normal_fluid": ("t_zero_normal_fluid", "r", 1),
    "norm_vel_xflow": ("norm_vel_xflow", "r", 1), "alpha": ("alpha", "r", 2),
    "super_velocity": ("super_velocity", "r", 3), "initf": ("initf", "c30", 1),
    "scaling_factor": ("scaling_factor", "r", 1), "loop_translate": ("loop_translate", "r", 3),
    "rotation_factor": ("rotation_factor", "r", 1), "line_count": ("line_count", "i", 1),
    "NAN_test": ("NAN_test", "l", 1), "max_warn_count": ("max_warn_count", "i", 1),
    "overide_timestep_check": ("overide_timestep_check", "l", 1),
    "serial_run": ("serial_run", "l", 1), "qvort_nproc": ("qvort_nproc", "i", 1),
    "wave_amp": ("wave_amp", "r", 1), "noise_file": ("noise_file", "c30", 1),
    "norm_shear_omega": ("norm_shear_omega", "r", 1), "lattice_ratio": ("lattice_ratio", "r", 1),
    "hip_precision": ("hip_precision", "c30", 1), "hip_near_radius": ("hip_near_radius", "r", 1),
    "hip_close_search": ("hip_close_search", "c30", 1),
    "hip_step_launch": ("hip_step_launch", "c30", 1),
}


def _tokens(s: str):
    """List-directed value separators: blanks and commas; '/' ends input."""
    out, cur, quote = [], "", None
    i = 0
    while i < len(s):
        ch = s[i]
        if quote:
            if ch == quote:
                quote = None
                out.append(cur)
                cur = ""
                i += 1
                continue
            cur += ch
        elif ch in "'\"":
            quote = ch
        elif ch == "/":
            break
        elif ch in " ,\t":
            if cur:
                out.append(cur)
                cur = ""
        else:
            cur += ch
        i += 1
    if cur and not quote:
        out.append(cur)
    return out


def _conv(tok: str, kind: str):
    if kind == "i":
        return int(tok)  # '0.4' into an integer is a read error, as in Fortran
    if kind == "r":
        return float(tok.replace("d", "e").replace("D", "e"))
    if kind == "l":
        t = tok[1:] if tok.startswith(".") else tok
        if t[:1] in ("T", "t"):
            return True
        if t[:1] in ("F", "f"):
            return False
        raise ValueError(tok)
    return tok[: int(kind[1:])]


def parse_run_text(text: str) -> RunParams:
    p = RunParams()
    for raw in text.splitlines():
        buf = raw[:100]
        pos = buf.find(" ")
        label = buf[:pos] if pos > 0 else ("" if pos < 0 else "")
        rest = buf[pos + 1:] if pos >= 0 else ""
        if label not in _KEYS:
            continue
        name, kind, cnt = _KEYS[label]
        toks = _tokens(rest)
        try:
            if len(toks) < cnt:
                raise ValueError("end of record")
            vals = [_conv(toks[k], kind) for k in range(cnt)]
        except ValueError:
            break  # src/cdata.f90:145: any read error ends parsing
        setattr(p, name, vals[0] if cnt == 1 else tuple(vals))
    return p


def read_run_file(path: str = "run.in") -> RunParams:
    with open(path, "r", errors="replace") as fh:
        return parse_run_text(fh.read())


def dt_max(p: RunParams) -> float:
    """src/initial.f90:169-180 timestep_check bound."""
    dmin = p.delta / 2.0
    return abs((dmin ** 2) / (p.quant_circ * math.log(dmin / (math.pi * p.corea))))


def validate(p: RunParams) -> None:
    """The fatal checks of src/initial.f90:14-146 that concern this path."""
    if p.delta is None or p.delta < 2.220446049250313e-16:
        raise RunFileError("init.mod: delta must be set in run.in")
    if p.dt is None:
        raise RunFileError("dt must be set in run.in")
    if not p.overide_timestep_check and not (p.dt < dt_max(p)):
        raise RunFileError("initial.mod:timestep_check: dt is too large")
    if not (min(p.box_size) > 0.0):
        raise RunFileError("init_setup: box size is less than or equal zero")
    for b in (p.boundary_x, p.boundary_y, p.boundary_z):
        if b not in ("periodic", "open", "solid"):
            raise RunFileError("init_setup: incorrect boundary parameter")
    if p.velocity not in ("LIA", "BS", "Tree"):
        raise RunFileError('init.mod:init_setup: correct value for "velocity" in run.in has not been set')
    if p.tree_theta > 0 and "open" in (p.boundary_x, p.boundary_y, p.boundary_z):
        raise RunFileError("init.mod:init_setup: tree approximation cannot run with open boundaries")
    if p.normal_velocity not in ("zero", "xflow", "noisey_xflow", "shear_xflow", "ychannel", "channel"):
        if p.normal_velocity in ("laizetDNS", "turbDNS"):
            raise RunFileError("normal_velocity %r needs a DNS data file and is not supported by this library"
                               % p.normal_velocity)
        raise RunFileError("normal_fluid.mod:get_normal_fluid: correct parameter for normal_veloctity not set")
    if p.hip_precision not in ("fp64", "fp32c"):
        raise RunFileError("hip_precision must be fp64 or fp32c")
    if p.hip_step_launch not in ("graph", "eager"):
        raise RunFileError("hip_step_launch must be graph or eager")


def format_run_text(p: RunParams, keys=None) -> str:
    """Write a run.in the reference parser reads back to the same values."""
    def fmt(v):
        if isinstance(v, bool):
            return "T" if v else "F"
        if isinstance(v, float):
            return repr(v).replace("e", "E")
        if isinstance(v, tuple):
            return " ".join(fmt(x) for x in v)
        return str(v)
    lines = []
    for label, (name, _, _) in _KEYS.items():
        if keys is not None and label not in keys:
            continue
        v = getattr(p, name)
        if v is None:
            continue
        lines.append("%s %s" % (label, fmt(v)))
    return "\n".join(lines) + "\n"
