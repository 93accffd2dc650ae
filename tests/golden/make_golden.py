"""Writes the golden fixtures under tests/golden/ (run in the build container).

* toy.json      -- the reference's only shipped fixture (graph_converter/
                   toy_edge_list.txt:1-18 and its fw_begin.bin/fw_adjacent.bin),
                   with the canonical labels and the raw scc_id the compiled
                   reference produced for it.
* patterns.json -- hand graphs, one per trim pattern of src/trim_2_3.h and one
                   per reference hazard (H0 livelock, H1 label collision,
                   H2 stale visit stamps), with Tarjan's canonical labels and
                   what the compiled reference did on each.
* hashes.json   -- larger seeded graphs: sha256 of the four CSR arrays and of
                   the canonical labels (Tarjan), SCC count, giant size, and
                   whether the compiled reference agreed.

Usage: python tests/golden/make_golden.py [--big] [--case KIND,SIZE,PARAM,SEED]
(--big adds R-MAT s24 and the 4096^2 grid, which take a few minutes and
~10 GB of host memory; --case hashes that one graph only, e.g.
--case rmat,26,16,1 for R-MAT s26, ~30 GB of host memory.)
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

REF_TOY = "/root/reference/graph_converter"


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    return h.hexdigest()


def csr_sha(g) -> str:
    return sha(g.fw_beg, g.fw_adj, g.bw_beg, g.bw_adj)


def ref_result(g, timeout=60.0):
    """Canonical labels from the compiled reference, or a reason it has none."""
    if not O.have_ref():
        return None, "reference not built"
    try:
        raw, _out = O.run_ref(g, threads=4, timeout=timeout)
    except subprocess.TimeoutExpired:
        return None, "timeout (did not terminate)"
    except RuntimeError as e:
        return None, f"failed: {e}"
    try:
        return O.canon_ref_raw(g, raw), "ok"
    except ValueError:
        return None, "left vertices unassigned"


def edges_of(g):
    src = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.fw_beg))
    return [[int(a), int(b)] for a, b in zip(src, g.fw_adj)]


def pattern_graphs():
    """Hand graphs: name -> (V, edges).  Each trim pattern sits next to a
    larger SCC with one-directional edges attached, so only the pattern
    rule (not trim-1) can peel it."""
    G = {}
    # giant SCC on 10..19 used as a neighbour: a ring plus a star around 10,
    # so 10 has the largest out*in and is the pivot
    ring = [(10 + i, 10 + (i + 1) % 10) for i in range(10)] + [(10, 10 + k) for k in range(2, 10)] + \
        [(10 + k, 10) for k in range(1, 9)]
    G["pair_sink"] = (20, ring + [(0, 1), (1, 0), (11, 0), (13, 1)])
    G["pair_source"] = (20, ring + [(0, 1), (1, 0), (0, 11), (1, 13)])
    G["cycle3_sink"] = (20, ring + [(0, 1), (1, 2), (2, 0), (1, 0), (14, 0), (16, 2)])
    G["cycle3_source"] = (20, ring + [(0, 1), (1, 2), (2, 0), (2, 1), (0, 14), (2, 16)])
    G["hub3_sink"] = (20, ring + [(0, 1), (1, 0), (0, 2), (2, 0), (11, 1), (12, 2)])
    G["hub3_source"] = (20, ring + [(0, 1), (1, 0), (0, 2), (2, 0), (1, 11), (2, 12)])
    # several patterns chained so later phases must finish what trims leave
    G["mixed_chain"] = (30, ring + [(0, 1), (1, 0), (1, 2), (2, 3), (3, 4), (4, 2), (4, 5), (5, 6), (6, 5),
                                    (6, 7), (7, 8), (8, 6), (20, 21), (21, 22), (22, 23), (23, 20), (23, 10),
                                    (9, 20), (24, 25), (25, 24), (25, 26), (26, 27), (27, 25), (27, 28),
                                    (28, 29), (29, 28)])
    # Hazards of the reference, each next to a star SCC around vertex 200
    # whose degree product makes 200 the giant pivot.
    star = [(200, 200 + k) for k in range(1, 12)] + [(200 + k, 200) for k in range(1, 12)]
    # H0: a chord-free 4-cycle {0,3,4,5} downstream of the star; 0 is the
    # first mice pivot, its label 0 reads as "unassigned" and the backward
    # search re-enqueues the cycle forever (src/fw_bw.h:1658-1683).
    G["H0_pivot0"] = (220, star + [(0, 3), (3, 4), (4, 5), (5, 0), (201, 0), (202, 4)])
    # H1: mice SCC {1,3,4,5} downstream of the star; its pivot label 1 is the
    # giant's label (src/fw_bw.h:1658 vs :974).
    G["H1_pivot1"] = (220, star + [(1, 3), (3, 4), (4, 5), (5, 1), (4, 1), (5, 3), (203, 1), (207, 4)])
    # H2: SCC {100..103} reached by the giant traversal at level 2
    # (200 -> 201 -> 100) and upstream of SCC {2,3,4,5}.  Ids >= the queue
    # length keep their stale level (src/fw_bw.h:1560-1563 resets queue
    # positions, not ids), so the search from pivot 2 absorbs vertex 100.
    G["H2_stale_stamp"] = (220, star + [(2, 3), (3, 4), (4, 5), (5, 2), (4, 2), (100, 101), (101, 102),
                                        (102, 103), (103, 100), (102, 100), (201, 100), (100, 2), (101, 3)])
    # small edge cases
    G["single_vertex"] = (1, [])
    G["self_loop_only"] = (3, [(0, 0), (1, 1), (1, 2)])
    G["two_cycle"] = (2, [(0, 1), (1, 0)])
    G["dag_chain"] = (8, [(i, i + 1) for i in range(7)])
    G["long_cycle"] = (50, [(i, (i + 1) % 50) for i in range(50)])
    G["complete_6"] = (6, [(i, j) for i in range(6) for j in range(6) if i != j])
    G["two_cycles_bridge"] = (8, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 6), (6, 7), (7, 4)])
    return G


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true")
    ap.add_argument("--toy-only", action="store_true")
    ap.add_argument("--case", help="KIND,SIZE,PARAM,SEED: hash only this graph")
    a = ap.parse_args(argv)
    O.build()
    if a.case:
        kind, size, param, seed = a.case.split(",")
        return hash_cases([(kind, int(size), int(param) if kind == "rmat" else float(param), int(seed))])

    # ---- toy fixture (reference data)
    fb = np.fromfile(os.path.join(REF_TOY, "fw_begin.bin"), np.int32)
    fa = np.fromfile(os.path.join(REF_TOY, "fw_adjacent.bin"), np.int32)
    V = fb.shape[0] - 1
    src = np.repeat(np.arange(V, dtype=np.int32), np.diff(fb))
    g = O.build_csr(V, src, fa)
    assert np.array_equal(g.fw_beg, fb) and np.array_equal(g.fw_adj, fa)
    raw, out = O.run_ref(g, threads=2)
    toy = {"source": "graph_converter/toy_edge_list.txt:1-18, graph_converter/fw_begin.bin, fw_adjacent.bin",
           "V": int(V), "fw_begin": fb.tolist(), "fw_adjacent": fa.tolist(),
           "bw_begin": g.bw_beg.tolist(), "bw_adjacent": g.bw_adj.tolist(),
           "canonical": O.tarjan(g).tolist(), "ref_raw_scc_id": raw.tolist(),
           "ref_result_block": [l for l in out.splitlines() if "," in l and l.split(",")[0] in
                                ("largest", "trimmed size_1", "trimmed size_2", "trimmed size_3", "others",
                                 "total")],
           "edge_list_text": open(os.path.join(REF_TOY, "toy_edge_list.txt")).read(),
           "head": np.fromfile(os.path.join(REF_TOY, "head.bin"), np.int32).tolist(),
           "out_degree": np.fromfile(os.path.join(REF_TOY, "out_degree.bin"), np.int32).tolist()}
    with open(os.path.join(HERE, "toy.json"), "w") as f:
        json.dump(toy, f, indent=1)
    print("toy", toy["canonical"], toy["ref_raw_scc_id"])
    if a.toy_only:
        return 0

    # ---- hand graphs
    pats = {}
    for name, (V, edges) in pattern_graphs().items():
        g = O.graph_from_edges(V, edges)
        tj = O.tarjan(g)
        rl, why = ref_result(g, timeout=20.0)
        pats[name] = {"V": V, "edges": [list(e) for e in edges], "canonical": tj.tolist(),
                      "ref": why if rl is None else ("agrees" if np.array_equal(rl, tj) else "wrong partition")}
        print(name, pats[name]["ref"])
    with open(os.path.join(HERE, "patterns.json"), "w") as f:
        json.dump(pats, f, indent=1)

    # ---- hashed seeded graphs
    cases = [("rmat", 16, 16, 1), ("rmat", 16, 16, 2), ("rmat", 18, 8, 3), ("rmat", 20, 16, 1),
             ("grid", 512, 0.6, 1), ("grid", 1024, 0.6, 1)]
    if a.big:
        cases += [("rmat", 24, 16, 1), ("grid", 4096, 0.6, 1)]
    return hash_cases(cases)


def hash_cases(cases) -> int:
    path = os.path.join(HERE, "hashes.json")
    hashes = json.load(open(path)) if os.path.exists(path) else {}
    for kind, size, param, seed in cases:
        key = f"{kind}-{size}-{param}-seed{seed}"
        g = O.rmat(size, param, seed) if kind == "rmat" else O.grid(size, param, seed)
        print(key, "generated", g.V, g.E, flush=True)
        tj = O.tarjan(g)
        print(key, "tarjan done", flush=True)
        n_scc = int(np.count_nonzero(tj == np.arange(g.V)))
        giant = int(np.bincount(tj).max())
        rl, why = ref_result(g, timeout=1800.0)
        hashes[key] = {"kind": kind, "size": size, "param": param, "seed": seed, "V": g.V, "E": g.E,
                       "csr_sha256": csr_sha(g), "labels_sha256": sha(tj), "n_scc": n_scc, "giant": giant,
                       "ref": why if rl is None else ("agrees" if np.array_equal(rl, tj) else "wrong partition")}
        print(key, hashes[key])
        with open(path, "w") as f:
            json.dump(hashes, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
