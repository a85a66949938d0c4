"""Filament point state held as structure-of-arrays.

This is the host-side mirror of the reference's ``type qvort`` array ``f(:)``
(src/cdata.f90:19-33).  Each field of the Fortran struct becomes one numpy
array; vector fields are ``(n, 3)`` float64, links are int32.  Point ``p``
(1-based, as the reference numbers them) lives in row ``p - 1``; ``infront``,
``behind`` and ``closest`` hold 1-based indices and ``infront == 0`` marks an
empty slot, exactly as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass, field, fields

import numpy as np

VEC_FIELDS = ("x", "u", "u1", "u2", "u_sup", "u_mf", "ghosti", "ghostb")
INT_FIELDS = ("infront", "behind", "closest", "pinnedi", "pinnedb")


@dataclass
class FilamentState:
    x: np.ndarray
    u: np.ndarray
    u1: np.ndarray
    u2: np.ndarray
    u_sup: np.ndarray
    u_mf: np.ndarray
    ghosti: np.ndarray
    ghostb: np.ndarray
    infront: np.ndarray
    behind: np.ndarray
    closest: np.ndarray
    closestd: np.ndarray
    pinnedi: np.ndarray
    pinnedb: np.ndarray
    wpinned: np.ndarray
    t: float = 0.0
    itime: int = 0
    recon_count: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def pcount(self) -> int:
        return int(self.x.shape[0])

    @property
    def active(self) -> int:
        return int(np.count_nonzero(self.infront))

    @classmethod
    def empty(cls, n: int) -> "FilamentState":
        z3 = lambda: np.zeros((n, 3), np.float64)  # noqa: E731
        zi = lambda: np.zeros(n, np.int32)  # noqa: E731
        return cls(x=z3(), u=z3(), u1=z3(), u2=z3(), u_sup=z3(), u_mf=z3(), ghosti=z3(),
                   ghostb=z3(), infront=zi(), behind=zi(), closest=zi(),
                   closestd=np.zeros(n, np.float64), pinnedi=zi(), pinnedb=zi(),
                   wpinned=np.zeros((n, 3), np.int32))

    def copy(self) -> "FilamentState":
        kw = {}
        for f in fields(self):
            v = getattr(self, f.name)
            kw[f.name] = v.copy() if isinstance(v, (np.ndarray, dict)) else v
        return FilamentState(**kw)

    def arrays(self):
        """(name, array) for every per-point field, in file order."""
        for name in VEC_FIELDS:
            yield name, getattr(self, name)
        yield "infront", self.infront
        yield "behind", self.behind
        yield "closest", self.closest
        yield "closestd", self.closestd
        yield "pinnedi", self.pinnedi
        yield "pinnedb", self.pinnedb
        yield "wpinned", self.wpinned

    def resized(self, n: int) -> "FilamentState":
        """A copy with ``n`` rows: truncated, or padded with empty slots."""
        out = FilamentState.empty(n)
        m = min(n, self.pcount)
        for name, arr in self.arrays():
            getattr(out, name)[:m] = arr[:m]
        out.t, out.itime, out.recon_count = self.t, self.itime, self.recon_count
        return out

    def check(self) -> None:
        """Link sanity: every non-empty point's neighbours are non-empty and
        point back (the invariant the reference's topology routines keep)."""
        n = self.pcount
        act = self.infront != 0
        inf = self.infront[act]
        beh = self.behind[act]
        if np.any((inf < 1) | (inf > n) | (beh < 1) | (beh > n)):
            raise ValueError("link out of range")
        idx = np.nonzero(act)[0] + 1
        if np.any(self.behind[inf - 1] != idx) or np.any(self.infront[beh - 1] != idx):
            raise ValueError("infront/behind links are not mutual")
