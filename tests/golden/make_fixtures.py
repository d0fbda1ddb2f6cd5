"""Extract the reference's own golden data into small fixtures.

Run once in a container that has the reference checkout mounted; the outputs
(tests/golden/fixtures.json and tests/golden/data/*) are committed and the
tests read only those.  Nothing here imports or executes reference code: it
copies data files and parses the reference's committed result logs.

Sources (relative to the reference root):
  emst_coalition/test_emst/test_2D_2/output, test_3D_2/output
      per-round rows "iteration ... benchMark size length ..." written by
      SolutionEMSTRW.cu:133-165 (size = cumulative forest edges, length =
      forest length as float, 6 significant digits)
  emst_coalition/test_emst/test_2D_2/outypu.links / .points / .disjointSetMap /
      .correspondenceMap
      full output of a 3000randomCity.2d run (NeuralNetEMST.h:275-326)
  emst_coalition/test_emst/test_*_2/*randomCity.{2d,3d}, *.tsp  input clouds
  emst_coalition/test_emst/test_2D_2/*.tsp  all 79 TSPLIB instances, gzipped
      into tests/golden/tsplib/ (the sweep's inputs)
"""
import collections
import gzip
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
T = os.path.join(REF, "emst_coalition", "test_emst")
HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "data")

INPUTS = [
    "test_2D_2/1000randomCity.2d", "test_2D_2/2000randomCity.2d", "test_2D_2/3000randomCity.2d",
    "test_2D_2/4000randomCity.2d", "test_2D_2/5000randomCity.2d", "test_2D_2/6000randomCity.2d",
    "test_3D_2/1000randomCity.3d", "test_3D_2/2000randomCity.3d", "test_3D_2/3000randomCity.3d",
    "test_3D_2/10000randomCity.3d", "test_3D_2/20000randomCity.3d",
    "test_2D_2/test20.tsp", "test_2D_2/wi29.tsp", "test_2D_2/dj38.tsp", "test_2D_2/uy734.tsp",
    "test_2D_2/lu980.tsp", "test_2D_2/rw1621.tsp", "test_2D_2/dka1376.tsp",
]


def traces(path):
    """benchmark -> iteration -> Counter((size, length string))."""
    out = collections.defaultdict(lambda: collections.defaultdict(collections.Counter))
    for line in open(path):
        tok = line.split()
        if len(tok) < 8 or not tok[0].isdigit() or not tok[6].isdigit():
            continue
        it, bench, size, length = int(tok[0]), tok[5], int(tok[6]), tok[7]
        if size == 0:
            continue    # failed runs in the logs (size 0 / 8.96831e-44 rows)
        out[bench][it][(size, length)] += 1
    return out


def modal_rounds(per_it):
    rounds = []
    for it in sorted(per_it):
        (size, length), count = per_it[it].most_common(1)[0]
        if rounds and size == rounds[-1]["size"]:
            break       # the reference logs the final state twice
        rounds.append({"iteration": it, "size": size, "length": length, "runs": count})
    return rounds


def main():
    os.makedirs(DATA, exist_ok=True)
    for rel in INPUTS:
        shutil.copyfile(os.path.join(T, rel), os.path.join(DATA, os.path.basename(rel)))
    fx = {"traces": {}}
    for sub in ("test_2D_2", "test_3D_2"):
        for bench, per_it in traces(os.path.join(T, sub, "output")).items():
            if os.path.exists(os.path.join(DATA, bench)):
                fx["traces"][bench] = modal_rounds(per_it)
    # full output of the 3000randomCity.2d run
    txt = open(os.path.join(T, "test_2D_2", "outypu.links")).read()
    edges = set()
    for m in re.finditer(r"NodeP = (\d+) 0 NumLinks = (\d+)((?: \d+ 0)*)", txt):
        i = int(m.group(1))
        for j in (int(x) for x in m.group(3).split()[0::2]):
            edges.add((min(i, j), max(i, j)))
    dsm = open(os.path.join(T, "test_2D_2", "outypu.disjointSetMap")).read().split()[6:]
    fx["outypu"] = {
        "input": "3000randomCity.2d",
        "edges": sorted(edges),
        "disjoint_set_values": sorted(set(int(x) for x in dsm)),
    }
    # the stats header line as the reference writes it (SolutionEMSTRW.cu:104-128)
    fx["stats_header"] = open(os.path.join(T, "test_2D_2", "output")).readline().rstrip("\n")
    # the run's written maps, for byte comparisons of the solution writers
    for sfx in ("points", "links", "disjointSetMap", "correspondenceMap"):
        shutil.copyfile(os.path.join(T, "test_2D_2", "outypu." + sfx), os.path.join(DATA, "outypu." + sfx))
    # the TSPLIB instances of the reference's 2D tests, gzip-compressed
    tsp_dir = os.path.join(HERE, "tsplib")
    os.makedirs(tsp_dir, exist_ok=True)
    for name in sorted(os.listdir(os.path.join(T, "test_2D_2"))):
        if name.endswith(".tsp"):
            with open(os.path.join(T, "test_2D_2", name), "rb") as fi, \
                    gzip.GzipFile(os.path.join(tsp_dir, name + ".gz"), "wb", mtime=0) as fo:
                fo.write(fi.read())
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(fx, f, indent=0)
    print("traces:", sorted(fx["traces"]))


if __name__ == "__main__":
    main()
