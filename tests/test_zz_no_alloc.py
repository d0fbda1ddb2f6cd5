"""emst_weight_noalloc: the FP64 forest weight of a device edge list with no
device allocation (partial sums in the library's static device array),
against emst_get_weight and the oracle, on two calls in a row, the second
with a smaller input."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(pkg, xyz_dev, dim, eu_dev, ev_dev, m, stream=None):
    w = ctypes.c_double(-1.0)
    rc = pkg.lib().emst_weight_noalloc(ctypes.c_void_p(xyz_dev), dim, ctypes.c_void_p(eu_dev),
                                       ctypes.c_void_p(ev_dev), m, 0, stream, ctypes.byref(w))
    return rc, w.value


class _Dev:
    """A device copy of a host array, made with the HIP runtime that
    libemst_amd.so itself links (looked up through the library's handle, so
    no second runtime is loaded into the process)."""

    def __init__(self, pkg, arr):
        L = pkg.lib()
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipFree.argtypes = [ctypes.c_void_p]
        self._L = L
        self.ptr = ctypes.c_void_p()
        arr = np.ascontiguousarray(arr)
        assert L.hipMalloc(ctypes.byref(self.ptr), max(arr.nbytes, 4)) == 0
        assert L.hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, 1) == 0      # host to device

    def __enter__(self):
        return self.ptr.value

    def __exit__(self, *a):
        self._L.hipFree(self.ptr)


def _tree(pkg, pts):
    with pkg.Calculateur(dim=pts.shape[1]) as c:
        c.initialize(pts)
        c.run()
        u, v, _ = c.getLinks()
        w, m = c.evaluate()
    return u, v, w


@pytest.mark.parametrize("first,second", [((3, 1, 20000), (2, 0, 3000))])
def test_weight_noalloc_matches_context_weight(pkg, oracle_mod, first, second):
    for dim, kind, n in (first, second):
        pts = pkg.synth(kind, n, dim, 90 + n)
        print("case noalloc", dim, kind, n, flush=True)
        u, v, w_ctx = _tree(pkg, pts)
        # a slot of the context's form (eu < 0: empty) is skipped
        eu_h = np.concatenate([u, [-1]]).astype(np.int32)
        ev_h = np.concatenate([v, [0]]).astype(np.int32)
        with _Dev(pkg, pts) as xyz, _Dev(pkg, eu_h) as eu, _Dev(pkg, ev_h) as ev:
            rc, w = _noalloc(pkg, xyz, dim, eu, ev, len(u) + 1)
            assert rc == 0, pkg.lib().emst_last_error(None)
            w_or = oracle_mod.weight(pts, u, v)
            assert abs(w - w_ctx) <= 1e-12 * w_ctx, (w, w_ctx)
            assert abs(w - w_or) <= 1e-12 * w_or, (w, w_or)
            # a prefix of the edges: the scratch of the previous call is rewritten
            k = len(u) // 3
            rc, wk = _noalloc(pkg, xyz, dim, eu, ev, k)
            assert rc == 0
            wk_or = oracle_mod.weight(pts, u[:k], v[:k])
            assert abs(wk - wk_or) <= 1e-12 * wk_or, (wk, wk_or)


def test_weight_noalloc_refusals(pkg):
    L = pkg.lib()
    w = ctypes.c_double(0)
    assert L.emst_weight_noalloc(None, 4, None, None, 0, 0, None, ctypes.byref(w)) == 1
    assert L.emst_weight_noalloc(None, 3, None, None, 1 << 31, 0, None, ctypes.byref(w)) == 7
    assert L.emst_weight_noalloc(None, 3, None, None, 5, 0, None, ctypes.byref(w)) == 1
    assert L.emst_weight_noalloc(None, 3, None, None, 0, 0, None, None) == 1
    assert L.emst_weight_noalloc(None, 3, None, None, 0, 0, None, ctypes.byref(w)) == 0 and w.value == 0.0
