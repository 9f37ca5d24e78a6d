"""fmm2d_l2d_direct_noalloc (no device allocation, static per-device split
array) against fmm2d_l2d_direct on the same inputs, through the C-ABI.

Both increment their outputs; the outputs start from a nonzero pattern so the
increment is checked too.  Cases: two calls in a row, the second smaller (the
array keeps the first call's contents); a call whose split copies do not all
fit the array (fewer splits, so the sums are ordered differently); a call with
one copy larger than the array (no split at all).  Tolerance: 1e-13 relative
L2 per output (FP64; equal split counts give identical sums).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(rng, nd, ns, nt, pg):
    src = rng.random((ns, 2))
    trg = rng.random((nt, 2))
    ch = rng.standard_normal((ns, nd)) + 1j * rng.standard_normal((ns, nd))
    ds = rng.standard_normal((ns, nd)) + 1j * rng.standard_normal((ns, nd))
    dv = rng.standard_normal((ns, 2, nd))
    return src, trg, ch, ds, dv, pg


def _run(L, nat, torch, entry, case, nd):
    src, trg, ch, ds, dv, pg = case
    dev = nat.torch_device()
    ns, nt = src.shape[0], trg.shape[0]
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
         for k, v in dict(src=src, trg=trg, ch=ch, ds=ds, dv=dv).items()}
    base = (np.arange(nd * nt) % 7 + 1.0).astype(np.complex128)
    pot = torch.from_numpy(base.copy()).to(dev)
    grad = torch.from_numpy(np.tile(base, 2)).to(dev)
    hess = torch.from_numpy(np.tile(base, 3)).to(dev)
    ier = getattr(L, entry)(nd, nat.ptr(t["src"]), ns, 1, nat.ptr(t["ch"]), 1, nat.ptr(t["ds"]),
                            nat.ptr(t["dv"]), nat.ptr(t["trg"]), nt, pg, nat.ptr(pot),
                            nat.ptr(grad), nat.ptr(hess), 1e-14, nat.stream_ptr())
    assert ier == 0, (entry, ier)
    torch.cuda.synchronize()
    out = [pot.cpu().numpy()]
    if pg >= 2:
        out.append(grad.cpu().numpy())
    if pg >= 3:
        out.append(hess.cpu().numpy())
    return out


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape", [
    # (nd, ns, nt, pg): split copies fit, then a smaller call
    [(1, 50000, 300, 2), (1, 20000, 100, 2)],
    # 64 wanted splits of 73728 values each do not fit: 28 splits
    [(3, 20000, 4096, 3), (2, 5000, 1000, 1)],
    # one copy (2.4e6 values) exceeds the array: no split
    [(1, 2000, 400000, 3), (1, 3000, 200, 3)],
])
def test_noalloc_matches_allocating_entry(shape):
    import torch
    from fmm2d_amd import _native as nat
    L = nat.lib()
    rng = np.random.default_rng(20261203)
    for nd, ns, nt, pg in shape:
        print(f"noalloc case nd={nd} ns={ns} nt={nt} pg={pg}", flush=True)
        case = _case(rng, nd, ns, nt, pg)
        ref = _run(L, nat, torch, "fmm2d_l2d_direct", case, nd)
        got = _run(L, nat, torch, "fmm2d_l2d_direct_noalloc", case, nd)
        for a, b in zip(got, ref):
            assert np.all(np.isfinite(a))
            assert _rel(a, b) < 1e-13, (nd, ns, nt, pg, _rel(a, b))
