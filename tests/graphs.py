"""Fixture loaders and seeded graph builders shared by the tests."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_json(name: str):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def toy_graph():
    d = load_json("toy.json")
    g = O.Graph(*(np.asarray(d[k], np.int32) for k in ("fw_begin", "fw_adjacent", "bw_begin", "bw_adjacent")))
    return g, d


def pattern_cases():
    return sorted(load_json("patterns.json").items())


def pattern_graph(d):
    return O.graph_from_edges(d["V"], d["edges"])


def hash_cases(max_v: int | None = None):
    cases = sorted(load_json("hashes.json").items())
    return [(k, d) for k, d in cases if max_v is None or d["V"] <= max_v]


def build_hashed(d):
    if d["kind"] == "rmat":
        return O.rmat(d["size"], d["param"], d["seed"])
    return O.grid(d["size"], d["param"], d["seed"])


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    return h.hexdigest()


def random_graph(seed: int, vmax: int = 300, density: float = 4.0, self_loops: bool = True):
    rng = np.random.default_rng(seed)
    V = int(rng.integers(1, vmax))
    m = int(rng.integers(0, int(density * V) + 1))
    e = rng.integers(0, V, size=(m, 2)).astype(np.int32)
    if not self_loops:
        e = e[e[:, 0] != e[:, 1]]
    return O.build_csr(V, e[:, 0].copy(), e[:, 1].copy())


def cycles_graph(seed: int, n_cycles: int = 40, max_len: int = 6):
    """Many small cycles (sizes 1..max_len) wired by one-directional edges:
    exercises the size-2/3 patterns, WCC and mice phases."""
    rng = np.random.default_rng(seed)
    edges = []
    base = 0
    groups = []
    for _ in range(n_cycles):
        k = int(rng.integers(1, max_len + 1))
        vs = list(range(base, base + k))
        groups.append(vs)
        base += k
        if k > 1:
            edges += [(vs[i], vs[(i + 1) % k]) for i in range(k)]
            if k > 2 and rng.random() < 0.5:
                edges.append((vs[1], vs[0]))
    for _ in range(n_cycles):  # DAG edges between groups (lower -> higher index)
        a, b = sorted(rng.choice(len(groups), 2, replace=False))
        edges.append((int(rng.choice(groups[a])), int(rng.choice(groups[b]))))
    perm = rng.permutation(base)
    e = np.asarray(edges, np.int64)
    e = perm[e].astype(np.int32)
    return O.build_csr(base, e[:, 0].copy(), e[:, 1].copy())


def brute_force_labels(g) -> np.ndarray:
    """SCCs from the transitive closure (tiny graphs only)."""
    V = g.V
    R = np.eye(V, dtype=bool)
    for u in range(V):
        R[u, g.fw_adj[g.fw_beg[u]:g.fw_beg[u + 1]]] = True
    for k in range(V):
        R |= np.outer(R[:, k], R[k, :])
    mutual = R & R.T
    return np.argmax(mutual, axis=1).astype(np.int32)  # first (smallest) mutually reachable id


def cycle_chain(groups: int, seed: int):
    """Groups of 1-6 vertices on a cycle each, group g pointing forward to
    groups g+1..g+3: one weakly connected piece of small SCCs, so the WCC's
    label propagation runs many rounds with path compression in every one."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 7, size=groups)
    first = np.concatenate([[0], np.cumsum(sizes)])
    edges = []
    for i, k in enumerate(sizes):
        b = int(first[i])
        edges += [(b + j, b + (j + 1) % int(k)) for j in range(int(k))]
        for t in (1, 2, 3):
            if i + t < groups:
                edges.append((b + int(rng.integers(0, k)), int(first[i + t])))
    return O.graph_from_edges(int(first[-1]), edges)
