"""Writes the golden fixtures of this directory from the oracle.

The reference ships no tests, fixtures or golden vectors for this path and
cannot be built here, so these fixtures pin the oracle itself (a later change
to its restatement that moves any value fails the CPU tests) and give the GPU
tests committed expected outputs:

* philox4x32_10_kat.json -- the Random123 known-answer vectors of
  Philox4x32-10 (Salmon et al., SC'11, kat_vectors), which the oracle's and
  the kernels' generator must reproduce;
* trace_2k.npz -- a seeded 2,000-triangle scene (conftest.random_triangles,
  seed 7, max leaf 8): SHA-256 and sizes of its Compact2 buffers, the tree
  statistics, and for 2,048 seeded rays (conftest.random_rays, seed 9) the
  closest-hit id / t, the any-hit id / t and the per-ray visit counts
  (inner nodes, triangles, leaves) of the reference traversal order;
* trace_1m_tree.json -- the 1,000,000-triangle trace scene the headline
  benchmark runs (scenes.trace_1m(), seed 0): SHA-256 of its vertices and
  triangles, and SHA-256, sizes and statistics of the Compact2 buffers of
  the single-threaded oracle's SBVH over it (default build parameters).
  The product's multi-threaded build must give the same bytes.

Run from the repository root: python tests/golden/make_fixtures.py
(trace_1m_tree.json only: python tests/golden/make_fixtures.py 1m; the
single-threaded build of 1M triangles takes a minute or two)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from conftest import random_rays, random_triangles  # noqa: E402
from oracle import oracle  # noqa: E402

KAT = [  # (counter, key, expected output)
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]

SCENE = dict(n=2000, seed=7, max_leaf=8)
RAYS = dict(n=2048, seed=9)


def trace_2k_inputs():
    v, idx = random_triangles(SCENE["n"], seed=SCENE["seed"])
    rays = random_rays(RAYS["n"], seed=RAYS["seed"])
    return v, idx, rays


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def trace_1m_tree():
    import renderbox2_amd as rb
    s = rb.scenes.trace_1m(seed=0)
    v, idx, _, _ = s.geometry()
    b = oracle.Bvh(v, idx)
    nodes, woop, tidx = b.compact2()
    st = b.stats()
    out = {"scene": "scenes.trace_1m(seed=0)", "tris": int(len(idx)), "verts": int(len(v)),
           "sha_verts": sha(v), "sha_idx": sha(idx),
           "sha_nodes": sha(nodes), "sha_woop": sha(woop), "sha_index": sha(tidx),
           "sizes": [int(nodes.size), int(woop.size), int(tidx.size)],
           "stats": {k: st[k] for k in ("n_inner", "n_leaf", "n_tris", "n_child_nodes", "duplicates", "max_depth")},
           "sah": st["sah"]}
    with open(os.path.join(HERE, "trace_1m_tree.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote trace_1m_tree.json:", out["stats"])


def main():
    if sys.argv[1:] == ["1m"]:
        trace_1m_tree()
        return
    with open(os.path.join(HERE, "philox4x32_10_kat.json"), "w") as f:
        json.dump([{"counter": c, "key": k, "out": o} for c, k, o in KAT], f, indent=1)
    v, idx, rays = trace_2k_inputs()
    b = oracle.Bvh(v, idx, max_leaf=SCENE["max_leaf"])
    nodes, woop, tidx = b.compact2()
    ids, t, cnt = oracle.trace_compact2(nodes, woop, tidx, rays, counts=True)
    aid, at = oracle.trace_compact2(nodes, woop, tidx, rays, any_hit=True)
    st = b.stats()
    np.savez_compressed(
        os.path.join(HERE, "trace_2k.npz"),
        sha_nodes=np.frombuffer(hashlib.sha256(np.ascontiguousarray(nodes).tobytes()).digest(), np.uint8),
        sha_woop=np.frombuffer(hashlib.sha256(np.ascontiguousarray(woop).tobytes()).digest(), np.uint8),
        sha_index=np.frombuffer(hashlib.sha256(np.ascontiguousarray(tidx).tobytes()).digest(), np.uint8),
        sizes=np.array([nodes.size, woop.size, tidx.size], np.int64),
        stats=np.array([st[k] for k in ("n_inner", "n_leaf", "n_tris", "n_child_nodes", "duplicates",
                                        "max_depth")], np.int64),
        sah=np.float32(st["sah"]), ids=ids, t=t, counts=cnt, any_ids=aid, any_t=at)
    print("wrote", HERE, "hits", int((ids >= 0).sum()), "of", len(ids))
    trace_1m_tree()


if __name__ == "__main__":
    main()
