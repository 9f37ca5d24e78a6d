"""Leaf occupancy by level of the unrestricted quadtree -- TEST INFRASTRUCTURE ONLY.

Checker of fmm2d_key_leaf_histogram (the parameter proposer's features; no
reference counterpart).  The tree: the root is the bounding square of sources
and targets (common/pts_tree2d.f:494-517), a box holding more than ndiv
SOURCES gets four children (the refinement rule of pts_tree_build,
common/pts_tree2d.f:551-631, idivflag 0) with the reference's centre
arithmetic (child centre = parent centre +- half the child size, point below
/ left when y - cy < 0 / x - cx < 0), and no level restriction.  hist[l] =
sources plus targets in the leaves of level l, deeper levels in the last bin.

Two restatements, pinned against each other (tests/test_oracle.py) on small
inputs: `brute` subdivides box by box; `windows` uses the sorted path keys
(levels 1..20) as the device does.
"""
import numpy as np

DEPTH = 20


def _root(src, trg):
    pts = src if trg is None or trg.shape[1] == 0 else np.concatenate([src, trg], axis=1)
    x0, x1, y0, y1 = pts[0].min(), pts[0].max(), pts[1].min(), pts[1].max()
    bs = max(x1 - x0, y1 - y0)
    return (x0 + x1) / 2, (y0 + y1) / 2, bs


def path_keys(xy, cx0, cy0, bs, depth=DEPTH):
    """40-bit keys: quadrant digit 2*qy + qx per level, level 1 most significant."""
    x, y = xy
    n = x.size
    key = np.zeros(n, np.uint64)
    cx, cy = np.full(n, cx0), np.full(n, cy0)
    bsl = bs
    for _ in range(depth):
        bsl = bsl / 2
        h = 0.5 * bsl
        qx = (x - cx) >= 0
        qy = (y - cy) >= 0
        key = (key << np.uint64(2)) | (qy.astype(np.uint64) << np.uint64(1)) | \
            qx.astype(np.uint64)
        cx = cx + np.where(qx, h, -h)
        cy = cy + np.where(qy, h, -h)
    return key


def _common_levels(a, b, depth=DEPTH):
    x = np.bitwise_xor(a, b)
    out = np.full(x.shape, depth, np.int64)
    nz = x != 0
    bl = np.floor(np.log2(x[nz].astype(np.float64))).astype(np.int64) + 1
    # guard the float rounding of log2 near powers of two
    bl = np.where(np.left_shift(np.uint64(1), (bl - 1).astype(np.uint64)) > x[nz], bl - 1, bl)
    out[nz] = (2 * depth - bl) // 2
    return out


def windows(src, trg, ndiv, nbins=16):
    cx, cy, bs = _root(src, trg)
    ks = np.sort(path_keys(src, cx, cy, bs))
    ns = ks.size
    nt = 0 if trg is None else trg.shape[1]
    lev = np.zeros(ns + nt, np.int64)
    if ns > ndiv:
        w = _common_levels(ks[:ns - ndiv], ks[ndiv:])
        last = ns - ndiv - 1
        for i in range(ns):
            lev[i] = 1 + w[max(0, i - ndiv):min(i, last) + 1].max()
        if nt:
            kt = path_keys(trg, cx, cy, bs)
            pos = np.searchsorted(ks, kt, side="left")
            for t in range(nt):
                js = np.arange(max(0, pos[t] - ndiv - 1), min(pos[t], last) + 1)
                c = np.minimum(w[js], _common_levels(ks[js], np.full(js.size, kt[t], np.uint64)))
                lev[ns + t] = 1 + c.max()
    return np.bincount(np.minimum(lev, nbins - 1), minlength=nbins)[:nbins].astype(np.float64)


def brute(src, trg, ndiv, nbins=16, depth=DEPTH):
    """Box-by-box subdivision (exact definition, small inputs only)."""
    cx, cy, bs = _root(src, trg)
    trg = np.zeros((2, 0)) if trg is None else trg
    hist = np.zeros(nbins)

    def rec(si, ti, cx, cy, bs, l):
        if si.size <= ndiv or l == depth + 1:
            hist[min(l, nbins - 1)] += si.size + ti.size
            return
        h = 0.25 * bs
        sx, sy = src[0, si] - cx >= 0, src[1, si] - cy >= 0
        tx, ty = trg[0, ti] - cx >= 0, trg[1, ti] - cy >= 0
        for qy in (False, True):
            for qx in (False, True):
                rec(si[(sx == qx) & (sy == qy)], ti[(tx == qx) & (ty == qy)],
                    cx + (h if qx else -h), cy + (h if qy else -h), bs / 2, l + 1)

    rec(np.arange(src.shape[1]), np.arange(trg.shape[1]), cx, cy, bs, 0)
    return hist
