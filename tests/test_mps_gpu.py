"""hfmm2d_mps (helmholtz/hfmm2d_mps.f:19) on the GPU against the reference.

Fixture tests/golden/mps_ref.npz (make_golden.py mps_fixtures): the inputs of
two cases -- the reference's own test (test/helmholtz/test_hfmm2d_mps.f90: 10x10
grid, zk = 2 pi, order 30, eps 0.5e-9; here every source charged) and 300
well-separated clustered centres with nd = 2, zk = 40 + 0.5i, orders 6..20,
expansions in shuffled order with gaps -- the expansions formed by the
reference's h2dformmpc from one point charge each, and the reference's
outputs: hfmm2d_mps and hfmm2d_mps_ndiv with ndiv = n (every pair directly).

Checks, on the local expansions evaluated at each centre's own source (inside
the expansion's radius; this is how the reference's test checks them):
  * equal to the reference's FMM to 1e-11 relative L2;
  * equal to sum_{j != i} c_j H0(k |x_i - x_j|) (scipy, closed form) to eps;
  * the all-pairs mode (opts ndiv = n) equal to the reference's all-pairs
    mode to 1e-12.
The raw coefficients are also compared with the reference's FMM (1e-9: high
orders of a small-scale local expansion are sensitive to rounding but carry
no weight inside its radius).
"""
import numpy as np
import pytest
import scipy.special as sp

import fmm2d_amd
from golden import inputs as I

pytestmark = pytest.mark.gpu


def _case(z, name):
    g = {k: z[f"{name}_{k}"] for k in ("nd", "eps", "zk", "centers", "rscales", "nterms",
                                       "mpole", "impole", "sources", "charges")}
    g["nd"], g["eps"], g["zk"] = int(g["nd"]), float(g["eps"]), complex(g["zk"])
    assert str(z[f"{name}_input_sha256"]) == I.digest(g["centers"], g["rscales"], g["nterms"],
                                                      g["mpole"], g["impole"])
    return g


def _eval_local(loc, g):
    """Local expansions evaluated at the sources: sum_n L_n s^-|n| J_n(k r) e^{i n t}."""
    nd, cen, rs, nt, imp, pts = g["nd"], g["centers"], g["rscales"], g["nterms"], g["impole"], \
        g["sources"]
    out = np.zeros((nd, cen.shape[1]), complex)
    for i in range(cen.shape[1]):
        m = int(nt[i])
        blk = loc[imp[i]:imp[i] + nd * (2 * m + 1)].reshape(2 * m + 1, nd)
        d = pts[:, i] - cen[:, i]
        n = np.arange(-m, m + 1)
        w = sp.jv(n, g["zk"] * np.hypot(*d)) * np.exp(1j * n * np.arctan2(d[1], d[0]))
        out[:, i] = (w / rs[i] ** np.abs(n)) @ blk
    return out


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("name", ["grid", "cluster"])
def test_mps_matches_reference(golden_dir, name):
    z = np.load(f"{golden_dir}/mps_ref.npz")
    g = _case(z, name)
    print(f"hfmm2d_mps case {name}: n={g['centers'].shape[1]} nd={g['nd']}", flush=True)
    loc = fmm2d_amd.hfmm2d_mps(eps=g["eps"], zk=g["zk"], centers=g["centers"],
                               rscales=g["rscales"], nterms=g["nterms"], mpole=g["mpole"],
                               impole=g["impole"], nd=g["nd"])
    ref = z[name + "_local"]
    assert _rel(loc, ref) < 1e-9
    u, uref = _eval_local(loc, g), _eval_local(ref, g)
    assert _rel(u, uref) < 1e-11
    # the potential of all other point charges, from the closed form
    src, ch = g["sources"], g["charges"]
    r = np.hypot(src[0][:, None] - src[0][None, :], src[1][:, None] - src[1][None, :])
    np.fill_diagonal(r, 1.0)
    H = sp.hankel1(0, g["zk"] * r)
    np.fill_diagonal(H, 0.0)
    assert _rel(u, (H @ ch.T).T) < g["eps"]
    # gaps between the expansions' blocks stay zero
    cover = np.zeros(loc.size, bool)
    for i, m in enumerate(g["nterms"]):
        cover[g["impole"][i]:g["impole"][i] + g["nd"] * (2 * int(m) + 1)] = True
    assert np.all(loc[~cover] == 0)


@pytest.mark.parametrize("name", ["grid", "cluster"])
def test_mps_all_pairs_matches_reference(golden_dir, name):
    z = np.load(f"{golden_dir}/mps_ref.npz")
    g = _case(z, name)
    n = g["centers"].shape[1]
    loc = fmm2d_amd.hfmm2d_mps(eps=g["eps"], zk=g["zk"], centers=g["centers"],
                               rscales=g["rscales"], nterms=g["nterms"], mpole=g["mpole"],
                               impole=g["impole"], nd=g["nd"], ndiv=n)
    ref = z[name + "_local_direct"]
    assert _rel(_eval_local(loc, g), _eval_local(ref, g)) < 1e-12
    assert _rel(loc, ref) < 1e-11
