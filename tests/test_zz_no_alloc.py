"""nv_pclose_noalloc (the close-point search that allocates nothing: static
device working memory, caller-held device inputs and outputs) against the
context entry point nv_pclose on the same states, bit for bit: two calls in
a row on one GPU, the second on a smaller input, so a stale partial of the
first call would show in the second."""
import ctypes as C

import numpy as np
import pytest

from golden_util import Fixture

pytestmark = pytest.mark.gpu


def context_pclose(st, p):
    from nuvort_amd.engine import FilamentSolver
    s = FilamentSolver(p, check=False)
    s.set_state(st)
    s.pclose()
    got = s.get_state()
    s.close()
    return got


class Hip:
    """The HIP runtime the library links (found by its soname once the library
    is loaded), for the caller's side of the test: device buffers and a
    stream, made here so that the entry point under test allocates nothing."""

    def __init__(self):
        from nuvort_amd import lib
        lib.load()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.bufs = []

    def ok(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), max(1, a.nbytes)))
        self.bufs.append(p)
        self.ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self.ok(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2))
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def noalloc_pclose(hip, stream, st, pad):
    from nuvort_amd import lib
    L = lib.load()
    n = st.pcount
    ld = n + pad
    xs = np.full((3, ld), 7.0)   # the padding rows hold values the search must not read
    xs[:, :n] = st.x.T
    x = hip.upload(xs)
    inf, beh = hip.upload(st.infront), hip.upload(st.behind)
    clo, cld = hip.upload(st.closest), hip.upload(st.closestd)
    print("nv_pclose_noalloc n=%d ld=%d" % (n, ld), flush=True)
    rc = L.nv_pclose_noalloc(n, ld, x, inf, beh, clo, cld, stream)
    assert rc == 0, rc
    hip.ok(hip.h.hipStreamSynchronize(stream))
    out = hip.download(clo, st.closest), hip.download(cld, st.closestd)
    hip.free()
    return out


def test_noalloc_close_search_matches_nv_pclose():
    from nuvort_amd import tangles
    from nuvort_amd.runfile import RunParams
    p = RunParams(dt=5e-6, delta=5e-4, corea=1.3e-8, box_size=(0.1, 0.1, 0.1), velocity="BS",
                  overide_timestep_check=True)
    big = tangles.random_rings(31, 0, 0.004, 0.012, 0.1, 5e-4, target_points=9000)   # many chunks per target
    big.closestd[:] = 3.0
    small = Fixture("remesh_bs").state("insert")   # empty slots among the rows
    hip = Hip()
    stream = C.c_void_p()
    hip.ok(hip.h.hipStreamCreate(C.byref(stream)))
    for st, pad in ((big, 100), (small, 3)):
        want = context_pclose(st, p)
        clo, cld = noalloc_pclose(hip, stream, st, pad)
        act = st.infront != 0
        assert np.array_equal(clo, want.closest), st.pcount
        assert np.array_equal(cld[act], want.closestd[act]), st.pcount
        assert np.array_equal(cld[~act], st.closestd[~act])
    hip.ok(hip.h.hipStreamDestroy(stream))


def test_noalloc_rejects_sizes_beyond_its_buffer():
    from nuvort_amd import lib
    L = lib.load()
    n = 5341953   # one point above NV_NOALLOC_MAX_POINTS: refused before any launch
    assert L.nv_pclose_noalloc(n, n, 1, 1, 1, 1, 1, None) == 3
    assert L.nv_pclose_noalloc(10, 5, 1, 1, 1, 1, 1, None) == 1
