import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__ as g
    return g.load_package()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


def read_points(path):
    """.2d/.3d: N then N rows of coordinates (InputRW.h:516-623); .tsp: TSPLIB
    NODE_COORD_SECTION rows 'id x y' (InputRW.h:411-513)."""
    if path.endswith(".tsp"):
        rows = []
        started = False
        for line in open(path):
            t = line.split()
            if not t:
                continue
            if t[0].startswith("NODE_COORD_SECTION"):
                started = True
                continue
            if started:
                if t[0] == "EOF":
                    break
                rows.append((float(t[1]), float(t[2])))
        return np.array(rows, dtype=np.float32)
    toks = open(path).read().split()
    n = int(toks[0])
    dim = 3 if path.endswith(".3d") else 2
    return np.array(toks[1:1 + n * dim], dtype=np.float32).reshape(n, dim)


def golden_path(name):
    return os.path.join(GOLDEN, "data", name)


def g6(x):
    """Text of a float as the reference's logs print it: 6 significant digits,
    %g layout, halves rounded away from zero (the MSVC C runtime the reference
    was built with; glibc rounds exact halves to even)."""
    from decimal import ROUND_HALF_UP, Decimal

    d = Decimal(float(x))
    if d == 0:
        return "0"
    exp = d.adjusted()
    q = d.scaleb(-exp).quantize(Decimal("1.00000"), rounding=ROUND_HALF_UP)
    if q >= 10:
        q = q / 10
        exp += 1
    digits = format(q, "f").rstrip("0").rstrip(".")
    if -5 <= exp < 6:
        v = (q.scaleb(exp)).quantize(Decimal(1).scaleb(min(0, exp - 5)))
        s = format(v, "f")
        if "." in s:
            s = s.rstrip("0").rstrip(".")
        return s
    return f"{digits}e{'+' if exp >= 0 else '-'}{abs(exp):02d}"
