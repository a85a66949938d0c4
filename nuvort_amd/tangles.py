"""Seeded synthetic vortex tangles (initial conditions).

The reference's own initial conditions draw from Fortran ``random_number``,
whose stream differs between compilers (src/initial_cond.f90:145-148), so the
project generates its tangles itself and hands them to every implementation
(the HIP library, the oracle and the reference build) as the same state.

Random numbers come from ``numpy.random.Generator(PCG64(seed))``.  Filaments
are discretised with chords of at most ``0.98 * delta`` so that neither point
insertion (``|x - x_infront| > delta``, src/line.f90:40) nor removal
(``|x_i - x_{i+2}| < 0.99 delta``, src/line.f90:115) fires at t = 0.
Positions are wrapped into the periodic box the way src/boundary.f90:91-138
wraps them (a single +-L shift per component).
"""
from __future__ import annotations

import math

import numpy as np

from .state import FilamentState


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _wrap(x: np.ndarray, box) -> np.ndarray:
    box = np.asarray(box, np.float64)
    x = x.copy()
    for k in range(3):
        L = box[k]
        hi = x[:, k] > L / 2.0
        x[hi, k] -= L
        lo = x[:, k] < -L / 2.0
        x[lo, k] += L
    return x


def _assemble(loops, box) -> FilamentState:
    """Concatenate closed loops (each an (n,3) array, consecutive points along
    the line) into one state; point k of a loop has infront k+1 (cyclic)."""
    n = sum(len(p) for p in loops)
    st = FilamentState.empty(n)
    off = 0
    for pts in loops:
        m = len(pts)
        idx = np.arange(m)
        st.x[off:off + m] = pts
        st.infront[off:off + m] = off + (idx + 1) % m + 1
        st.behind[off:off + m] = off + (idx - 1) % m + 1
        off += m
    st.x = _wrap(st.x, box)
    return st


def ring(center, radius, normal, delta, phase=0.0, chord=0.98):
    """Points of one circular loop (chord <= chord*delta)."""
    normal = np.asarray(normal, np.float64)
    normal = normal / np.linalg.norm(normal)
    a = np.array([1.0, 0.0, 0.0]) if abs(normal[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(normal, a)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(normal, e1)
    n = max(5, math.ceil(2.0 * math.pi * radius / (chord * delta)))
    phi = phase + 2.0 * math.pi * np.arange(n) / n
    return (np.asarray(center, np.float64)[None, :] + radius * np.cos(phi)[:, None] * e1[None, :]
            + radius * np.sin(phi)[:, None] * e2[None, :])


def random_rings(seed: int, nrings: int, rmin: float, rmax: float, box: float, delta: float,
                 target_points: int | None = None, sides: tuple | None = None) -> FilamentState:
    """Rings with radii ~U[rmin, rmax], centres ~U(box), normals uniform on
    the sphere (normalised Gaussian), phase ~U[0, 2pi).  If ``target_points``
    is given, rings are added until the point count reaches it instead.
    ``sides`` gives a box of unequal sides instead of the cube of side box:
    the centres are drawn in it and the points wrapped into it."""
    g = _rng(seed)
    lo = -np.asarray(sides, np.float64) / 2.0 if sides is not None else -box / 2.0
    loops, total = [], 0
    k = 0
    while (target_points is None and k < nrings) or (target_points is not None and total < target_points):
        r = g.uniform(rmin, rmax)
        c = g.uniform(lo, -lo, size=3)
        nv = g.normal(size=3)
        ph = g.uniform(0.0, 2.0 * math.pi)
        pts = ring(c, r, nv, delta, ph)
        loops.append(pts)
        total += len(pts)
        k += 1
    return _assemble(loops, tuple(sides) if sides is not None else (box, box, box))


def kelvin_lattice(seed: int, box: float = 0.1, nx: int = 4, ny: int = 2, amp: float = 0.002,
                   waves: int = 3, delta: float = 5e-4) -> FilamentState:
    """Straight periodic lines along z on an nx-by-ny lattice (the lattice
    formula of src/initial_cond.f90:225-228), each carrying a helical Kelvin
    wave x += A cos(kz + phi), y += A sin(kz + phi) (the form of
    src/initial_cond.f90:367-375) with k = 2 pi waves / box and a random phase
    per line.  infront points towards -z as in src/initial_cond.f90:233."""
    g = _rng(seed)
    k = 2.0 * math.pi * waves / box
    n = math.ceil(box * math.sqrt(1.0 + (amp * k) ** 2) / (0.98 * delta))
    loops = []
    for i in range(nx):
        for j in range(ny):
            x0 = -box / 2.0 + box * (2 * i + 1) / (2 * nx)
            y0 = -box / 2.0 + box * (2 * j + 1) / (2 * ny)
            phi = 2.0 * math.pi * g.uniform()
            z = -box / 2.0 + box * np.arange(n) / n
            pts = np.stack([x0 + amp * np.cos(k * z + phi), y0 + amp * np.sin(k * z + phi), z], 1)
            loops.append(pts[::-1].copy())  # consecutive points run towards -z
    return _assemble(loops, (box, box, box))


def ensemble_sizes(seed: int, ntangles: int, nmin: int, nmax: int) -> np.ndarray:
    """Target point counts log-uniform in [nmin, nmax]."""
    g = _rng(seed)
    return np.floor(np.exp(g.uniform(math.log(nmin), math.log(nmax), size=ntangles))).astype(np.int64)


def ensemble(seed: int, ntangles: int, nmin: int, nmax: int, box: float = 0.1, delta: float = 5e-4,
             rmin: float = 0.01, rmax: float = 0.03):
    """Independent random-ring tangles, one per target size; returns a list of
    states (tangle k seeded with seed*1000003 + k)."""
    sizes = ensemble_sizes(seed, ntangles, nmin, nmax)
    return [random_rings(seed * 1000003 + k, 0, rmin, rmax, box, delta, target_points=int(s))
            for k, s in enumerate(sizes)]


def noise_table(seed: int, n: int = 100000, spread: float = 0.2) -> np.ndarray:
    """A noisey_xflow multiplier table (the role of the reference's
    noise/wn_snr_*.log files, which hold 100000 values): 1 + spread * N(0, 1),
    seeded."""
    return 1.0 + spread * _rng(seed).normal(size=n)


WALL = 0.499999999   # where src/boundary.f90:124-135 puts a pinned point: this fraction of the box


def wall_tangle(seed: int, box: float = 0.1, delta: float = 5e-4, nlines: int = 3, nhalf: int = 2,
                nrings: int = 2) -> FilamentState:
    """A tangle for solid walls at x = +-box/2 and z = +-box/2 (y periodic).

    * ``nlines`` wavy lines from the +z wall to the -z wall (the shape of the
      reference's lattice lines, src/initial_cond.f90:296-318): the first
      point pinned behind on the +z wall (behind = itself, wpinned (0,0,1)),
      the last pinned in front on the -z wall (infront = itself, wpinned
      (0,0,-1));
    * ``nhalf`` half loops standing on the x walls, alternately +x and -x
      (the shape of src/initial_cond.f90:41-75, in a random plane through the
      wall normal), both ends pinned to that wall;
    * ``nrings`` rings, the first one reaching to within delta/5 of the +x
      wall, which wall_recon (src/reconnection.f90:84-93) cuts there;
    * one two-point segment lying on the -z wall (pinned behind, then pinned
      in front), which wall_recon and premove clear.

    Pinned points sit at +-0.499999999 box, where enforce_boundary puts them.
    Chords are at most 0.98 delta."""
    g = _rng(seed)
    L = box
    segs = []   # (points, pinned ends: None for a closed loop, else (wpinned of the first, of the last))
    for _ in range(nlines):
        x0, y0 = g.uniform(-0.3 * L, 0.3 * L, size=2)
        amp, waves, phi = g.uniform(0.001, 0.003), int(g.integers(1, 4)), g.uniform(0, 2 * math.pi)
        k = 2.0 * math.pi * waves / L
        n = math.ceil(2 * WALL * L * math.sqrt(1.0 + (amp * k) ** 2) / (0.98 * delta)) + 1
        z = WALL * L - 2 * WALL * L * np.arange(n) / (n - 1)
        pts = np.stack([x0 + amp * np.cos(k * z + phi), y0 + amp * np.sin(k * z + phi), z], 1)
        segs.append((pts, ((0, 0, 1), (0, 0, -1))))
    for h in range(nhalf):
        side = 1 if h % 2 == 0 else -1
        r = g.uniform(0.008, 0.02)
        ang = g.uniform(0, 2 * math.pi)
        d = np.array([0.0, math.cos(ang), math.sin(ang) * 0.3])   # a direction in the wall, mostly along y
        d /= np.linalg.norm(d)
        c = np.array([side * WALL * L, g.uniform(-0.3 * L, 0.3 * L), g.uniform(-0.2 * L, 0.2 * L)])
        n = max(5, math.ceil(math.pi * r / (0.98 * delta)) + 1)
        th = math.pi * np.arange(n) / (n - 1)
        pts = c[None, :] + r * np.cos(th)[:, None] * d[None, :]
        pts[:, 0] = side * WALL * L - side * r * np.sin(th)
        segs.append((pts, ((side, 0, 0), (side, 0, 0))))
    for q in range(nrings):
        r = g.uniform(0.006, 0.012)
        if q == 0:
            # in a plane holding the x axis, its +x point delta/5 inside the wall
            c = np.array([0.5 * L - 0.2 * delta - r, g.uniform(-0.3 * L, 0.3 * L), g.uniform(-0.2 * L, 0.2 * L)])
            nv = np.array([0.0, math.cos(0.7), math.sin(0.7)])
        else:
            c = g.uniform(-0.25 * L, 0.25 * L, size=3)
            nv = g.normal(size=3)
        segs.append((ring(c, r, nv, delta, g.uniform(0, 2 * math.pi)), None))
    zf = -WALL * L
    segs.append((np.array([[0.01, 0.02, zf], [0.01 + 0.5 * delta, 0.02, zf]]), ((0, 0, -1), (0, 0, -1))))
    n = sum(len(p) for p, _ in segs)
    st = FilamentState.empty(n)
    off = 0
    for pts, pins in segs:
        m = len(pts)
        idx = np.arange(m)
        st.x[off:off + m] = pts
        if pins is None:
            st.infront[off:off + m] = off + (idx + 1) % m + 1
            st.behind[off:off + m] = off + (idx - 1) % m + 1
        else:
            st.infront[off:off + m] = off + idx + 2
            st.behind[off:off + m] = off + idx
            st.behind[off] = off + 1                  # pinned behind: behind = itself
            st.infront[off + m - 1] = off + m         # pinned in front: infront = itself
            st.pinnedb[off] = 1
            st.pinnedi[off + m - 1] = 1
            st.wpinned[off] = pins[0]
            st.wpinned[off + m - 1] = pins[1]
        off += m
    # y is periodic: wrap it alone
    y = st.x[:, 1]
    y[y > L / 2] -= L
    y[y < -L / 2] += L
    return st
