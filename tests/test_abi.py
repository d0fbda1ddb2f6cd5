"""The C-ABI library loads and exports every symbol include/emst.h declares
(no compute call needs a GPU here)."""
import ctypes
import os
import re

import numpy as np

from conftest import ROOT


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "emst.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(emst_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol(pkg):
    lib = ctypes.CDLL(pkg.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(pkg.EXPORTS) == names


def test_version_and_defaults(pkg):
    L = pkg.lib()
    assert L.emst_version().startswith(b"emst_amd")
    o = pkg.Options()
    L.emst_default_options(ctypes.byref(o))
    assert (o.dim, o.level_radius, o.max_rounds) == (3, 1, 100)


def test_synth_is_deterministic(pkg):
    for kind in (0, 1, 2):
        for dim in (2, 3):
            a = pkg.synth(kind, 2000, dim, 7)
            b = pkg.synth(kind, 2000, dim, 7)
            c = pkg.synth(kind, 2000, dim, 8)
            assert a.dtype == np.float32 and a.shape == (2000, dim)
            assert np.array_equal(a, b) and not np.array_equal(a, c)
            assert a.min() >= 0 and a.max() < 1


def test_synth_uniform_first_values(pkg):
    """SplitMix64 -> xoshiro128+, u = (x >> 8) * 2^-24, restated in Python."""
    M64 = (1 << 64) - 1
    M32 = (1 << 32) - 1

    def splitmix(s):
        s = (s + 0x9E3779B97F4A7C15) & M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return s, z ^ (z >> 31)

    s, a = splitmix(123)
    s, b = splitmix(s)
    st = [a & M32, a >> 32, b & M32, b >> 32]
    out = []
    for _ in range(12):
        r = (st[0] + st[3]) & M32
        t = (st[1] << 9) & M32
        st[2] ^= st[0]; st[3] ^= st[1]; st[1] ^= st[2]; st[0] ^= st[3]
        st[2] ^= t
        st[3] = ((st[3] << 11) | (st[3] >> 21)) & M32
        out.append(np.float32((r >> 8) * 2.0 ** -24))
    got = pkg.synth(0, 4, 3, 123).ravel()
    assert np.array_equal(got, np.array(out, np.float32))


def test_create_without_device_fails_cleanly(pkg):
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        return
    L = pkg.lib()
    o = pkg.Options()
    L.emst_default_options(ctypes.byref(o))
    assert not L.emst_create(ctypes.byref(o))
    assert b"device" in L.emst_last_error(None)
