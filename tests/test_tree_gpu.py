"""The GPU quadtree against the reference's, array for array.

tests/golden/tree_ref.npz holds the reference's pts_tree_build /
pts_tree_sort / computelists output for three point sets (uniform with
separate targets; Gaussian clusters down to sigma = 3e-4 that force the
level-restriction passes; a perturbed closed curve; exact duplicates that
drive refinement to the 51-level limit) at ndiv 5 and 15.  Every
integer array must match exactly (box numbering, parents, children,
colleagues, point permutation and ranges, lists as sets) and the centres and
box sizes bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("laddr", "ilevel", "iparent", "nchild", "ichild", "ncoll", "coll", "isrc", "isrcse",
        "itarg", "itargse", "centers", "boxsize")


@pytest.mark.parametrize("ndiv", [5, 15])
@pytest.mark.parametrize("name", ["uniform", "cluster", "curve", "deep"])
def test_tree_bit_exact(golden_dir, name, ndiv):
    from fmm2d_amd.tree import build_tree
    g = np.load(f"{golden_dir}/tree_ref.npz")
    src = g[name + "_src"]
    trg = g[name + "_trg"] if name + "_trg" in g.files else None
    t = build_tree(src, trg, ndiv=ndiv, lists=True)
    pre = f"{name}_{ndiv}_"
    assert t["nlevels"] == int(g[pre + "nlevels"])
    assert t["nboxes"] == int(g[pre + "nboxes"])
    for k in KEYS:
        ref = g[pre + k]
        if k in ("itarg", "itargse") and trg is None:
            continue
        assert np.array_equal(t[k], ref), k
    for k in range(1, 5):
        assert np.array_equal(t[f"nlist{k}"], g[pre + f"nlist{k}"]), f"nlist{k}"
        assert np.array_equal(t[f"list{k}"], g[pre + f"list{k}"]), f"list{k}"
