"""Solution files and the CLI against the reference's own output of its
3000randomCity.2d run (outypu.*): the maps byte for byte, the link lists as
sets per node (the reference writes each node's links in the order its
atomics ran), and the stats file's layout and per-round sizes."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, golden_path, read_points

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(GOLDEN, "fixtures.json")))


def _links(path):
    txt = open(path).read()
    head = txt.split("\n", 1)[0]
    nodes = {}
    for m in re.finditer(r"NodeP = (\d+) 0 NumLinks = (\d+) ((?:\d+ 0 )*)\n", txt):
        ids = [int(x) for x in m.group(3).split()[0::2]]
        assert len(ids) == int(m.group(2))
        nodes[int(m.group(1))] = sorted(ids)
    return head, nodes, txt.count("\n")


def _check_maps(base):
    for sfx in ("disjointSetMap", "correspondenceMap", "points"):
        got = open(f"{base}.{sfx}", "rb").read()
        want = open(golden_path(f"outypu.{sfx}"), "rb").read()
        assert got == want, sfx
    gh, gn, gl = _links(f"{base}.links")
    wh, wn, wl = _links(golden_path("outypu.links"))
    assert gh == wh and gl == wl and gn == wn


def test_write_solution_matches_reference_files(pkg, tmp_path):
    pts = read_points(golden_path("3000randomCity.2d"))
    with pkg.Calculateur(dim=2) as c:
        c.initialize(pts)
        c.run()
        c.writeSolution(str(tmp_path / "outypu"))
    _check_maps(tmp_path / "outypu")


def test_cli_run_matches_reference_outputs(pkg, tmp_path):
    assert os.path.exists(pkg.CLI_PATH), "emst_cli is not built"
    stats = tmp_path / "outypu.stats"
    print("cli 3000randomCity.2d", flush=True)
    r = subprocess.run([pkg.CLI_PATH, golden_path("3000randomCity.2d"), str(tmp_path / "outypu"), str(stats)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _check_maps(tmp_path / "outypu")
    lines = open(stats).read().split("\n")
    assert lines[0] == FIX["stats_header"]
    rows = [l.split() for l in lines[1:] if l.strip()]
    assert all(len(t) == 18 for t in rows)
    trace = FIX["traces"]["3000randomCity.2d"]
    # one row per round (traceActive defaults on), then the final row
    assert [int(t[6]) for t in rows[:-1]] == [t["size"] for t in trace]
    assert [int(t[0]) for t in rows[:-1]] == list(range(1, len(trace) + 1))
    assert rows[-1][0] == rows[-2][0] and rows[-1][6] == rows[-2][6]
    assert abs(float(rows[-1][7]) - float(trace[-1]["length"])) <= 1.0
    # a second run appends rows without a second header
    r = subprocess.run([pkg.CLI_PATH, golden_path("3000randomCity.2d"), str(tmp_path / "outypu"), str(stats)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert open(stats).read().count("iteration") == 1


def test_cli_3d_mode_from_config(pkg, tmp_path, oracle_mod):
    cfg = tmp_path / "config.cfg"
    cfg.write_text("[coalition_global_param]\nfunctionModeChoice = 5\n[global_param]\ntraceActive = 0\n")
    r = subprocess.run([pkg.CLI_PATH, golden_path("1000randomCity.3d"), str(tmp_path / "o3"),
                        str(tmp_path / "o3.stats"), str(cfg)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l.split() for l in open(tmp_path / "o3.stats").read().split("\n")[1:] if l.strip()]
    assert len(rows) == 1                       # no per-round rows without tracing
    _, nodes, _ = _links(tmp_path / "o3.links")
    edges = sorted({(min(i, j), max(i, j)) for i, js in nodes.items() for j in js})
    ref = oracle_mod.emst(read_points(golden_path("1000randomCity.3d")), mode=0)
    assert np.array_equal(np.array(edges, np.int64), oracle_mod.canon(ref["u"], ref["v"]))


def test_cli_2p5d_mode_from_config(pkg, tmp_path, oracle_mod):
    """functionModeChoice = 6 (RUN_2DPOINT5): 3D points with cones over the
    xy plane; the forest equals the oracle's 2.5D forest, per round."""
    cfg = tmp_path / "config.cfg"
    cfg.write_text("[coalition_global_param]\nfunctionModeChoice = 6\n[global_param]\ntraceActive = 1\n")
    r = subprocess.run([pkg.CLI_PATH, golden_path("2000randomCity.3d"), str(tmp_path / "o25"),
                        str(tmp_path / "o25.stats"), str(cfg)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, nodes, _ = _links(tmp_path / "o25.links")
    edges = sorted({(min(i, j), max(i, j)) for i, js in nodes.items() for j in js})
    ref = oracle_mod.emst(read_points(golden_path("2000randomCity.3d")), mode=0, cdim=2)
    assert np.array_equal(np.array(edges, np.int64), oracle_mod.canon(ref["u"], ref["v"]))
    rows = [l.split() for l in open(tmp_path / "o25.stats").read().split("\n")[1:] if l.strip()]
    assert [int(r[6]) for r in rows[:-1]] == ref["cum_edges"]      # size column, one row per round
