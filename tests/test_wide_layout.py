"""The wide (4-child, 64-byte) records the upload builds from the binary
MI355X records (RB_LAYOUT_MI355X_WIDE), checked on the CPU against the
binary records they come from:
* numbering: record 0 is the root and the records follow breadth-first,
  each a binary record at an even depth; a record's children are its
  binary record's children with every inner child replaced by its own two,
  empty leaves left out, in that order; leaf links are the binary ones;
* containment: every decoded plane, the single rounding of p + q * 2^e to
  float32 (what the traversal's fmaf computes), lies on the outer side of
  the exact plane it stands for -- the decoded box holds the binary box,
  which is what makes the wide walk test every leaf the binary walk tests;
* steps fit: the biased exponents are normal (1..254) and every hi step is
  at least its lo step.
The traversal itself is checked against the oracle on the GPU."""
from fractions import Fraction

import numpy as np
import pytest

from conftest import random_triangles

EMPTY = 0x7FFFFFFF


_OVER = Fraction(2) ** 128 - Fraction(2) ** 103  # from here on float32 rounding gives infinity


def _rn32(x):
    """float32 nearest to the rational x (ties to even)."""
    if abs(x) >= _OVER:
        return np.float32(np.inf) if x > 0 else np.float32(-np.inf)
    f = np.float32(float(x))
    best = None
    with np.errstate(over="ignore"):  # the neighbour of +-FLT_MAX is infinite: skipped below
        near = (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf)))
    for c in near:
        if not np.isfinite(c):
            continue
        d = abs(Fraction(float(c)) - x)
        key = (d, int(np.array(c, np.float32).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return best[1]


def _children(bin_nodes, rec):
    """(box, link) of binary record rec's two children: box lo.x hi.x lo.y
    hi.y lo.z hi.z as float32."""
    r = bin_nodes[rec]
    f = r.view(np.float32)
    out = []
    for c in range(2):
        box = [f[4 * c], f[4 * c + 1], f[4 * c + 2], f[4 * c + 3], f[8 + 2 * c], f[9 + 2 * c]]
        out.append((box, int(r[12 + c].astype(np.int32))))
    return out


def _expected(bin_nodes):
    """The wide records' children (box, binary link) in the exporter's order,
    and the breadth-first numbering of the binary records that become wide."""
    order, wid, recs = [0], {0: 0}, []
    h = 0
    while h < len(order):
        kids = []
        for box, link in _children(bin_nodes, order[h]):
            if link >= 0:
                kids += _children(bin_nodes, link)
            else:
                kids.append((box, link))
        kids = [(b, l) for b, l in kids if not (l < 0 and (~l & 31) == 0)]  # empty leaves
        for _, l in kids:
            if l >= 0 and l not in wid:
                wid[l] = len(order)
                order.append(l)
        recs.append(kids)
        h += 1
    return recs, wid


@pytest.mark.parametrize("n,seed,scale", [(3000, 5, 1.0), (2000, 6, 1.0e4), (1500, 7, 1.0e-3), (2000, 13, 1.0e30)])
def test_wide_records_hold_the_binary_boxes(rbm, n, seed, scale):
    from renderbox2_amd import _lib as L
    v, idx = random_triangles(n, seed=seed)
    v = (v * np.float32(scale)).astype(np.float32)
    s = rbm.Scene()
    m = s.add_diffuse_bsdf((0.5, 0.5, 0.5))
    s.add_mesh(v, idx, m)
    b = rbm.BVH(s)
    bn, bt, bx = b.export(L.RB_LAYOUT_MI355X)
    wn, wt, wx = b.export(L.RB_LAYOUT_MI355X_WIDE)
    assert np.array_equal(bt, wt) and np.array_equal(bx, wx)  # same triangles and index
    bin_nodes = bn.reshape(-1, 16)
    wide = wn.reshape(-1, 16)
    recs, wid = _expected(bin_nodes)
    assert len(wide) == len(recs) and len(wide) < 0.6 * len(bin_nodes)
    checked = 0
    for i, kids in enumerate(recs):
        w = wide[i]
        links = w[10:14].astype(np.int32)
        assert len(kids) <= 4
        for c in range(4):
            if c >= len(kids):
                assert links[c] == EMPTY
                continue
            box, l = kids[c]
            assert links[c] == (wid[l] if l >= 0 else l)
        p = w[0:3].view(np.float32)
        eb = [(int(w[3]) >> (8 * a)) & 0xFF for a in range(3)]
        assert all(1 <= e <= 254 for e in eb)
        for c, (box, _) in enumerate(kids):
            # an inverted pair of planes is the box between them (the slab
            # test takes their min and max)
            box = [min(box[k], box[k ^ 1]) if k % 2 == 0 else max(box[k], box[k ^ 1]) for k in range(6)]
            for k in range(6):
                a = k // 2
                q = (int(w[4 + k]) >> (8 * c)) & 0xFF
                dec = _rn32(Fraction(float(p[a])) + q * Fraction(2) ** (eb[a] - 127))
                if k % 2 == 0:
                    assert dec <= box[k], (i, c, k)
                    assert q <= (int(w[5 + k]) >> (8 * c)) & 0xFF
                else:
                    assert dec >= box[k], (i, c, k)
                checked += 1
    assert checked > 500
