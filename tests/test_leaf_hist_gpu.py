"""fmm2d_key_leaf_histogram (the proposer's features) on the GPU against the
numpy restatement oracle/leaf_hist_ref.windows, which tests/test_oracle.py
pins to box-by-box subdivision.  Integer counts: exact equality."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("via", ["keys", "prepared"])
@pytest.mark.parametrize("case", ["uniform", "cluster", "curve", "targets", "tiny", "win2048",
                                  "win2500"])
def test_key_leaf_histogram_matches_restatement(case, via):
    import torch
    from fmm2d_amd import _native as nat
    from oracle import leaf_hist_ref as H
    rng = np.random.default_rng(7)
    ns, nt, ndiv = 200000, 0, 15
    if case == "uniform":
        src = rng.random((2, ns))
    elif case == "cluster":
        src = 0.3 + 0.002 * rng.standard_normal((2, ns))
        src[:, :20000] = rng.random((2, 20000))
    elif case == "curve":
        t = rng.uniform(0, 2 * np.pi, ns)
        src = np.stack([np.cos(t), np.sin(t)]) + 1e-4 * rng.standard_normal((2, ns))
        ndiv = 37
    elif case == "targets":
        ns, nt = 50000, 30000
        src = 0.5 + 0.05 * rng.standard_normal((2, ns))
    elif case in ("win2048", "win2500"):
        # the largest window held in LDS (ndiv 2048), and one beyond it
        ns, ndiv = 60000, int(case[3:])
        src = 0.5 + 0.01 * rng.standard_normal((2, ns))
    else:
        ns, ndiv = 10, 15
        src = rng.random((2, ns))
    trg = rng.random((2, nt)) * 1.2 - 0.1 if nt else None
    print(f"key leaf histogram case {case}: ns={ns} nt={nt} ndiv={ndiv}", flush=True)
    dev = nat.torch_device()
    s = torch.from_numpy(np.ascontiguousarray(src.T)).to(dev)
    t = torch.from_numpy(np.ascontiguousarray(trg.T)).to(dev) if nt else None
    hist = np.zeros(16)
    bs0 = C.c_double()
    L = nat.lib()
    if via == "keys":
        ier = L.fmm2d_key_leaf_histogram(nat.ptr(s), ns, nat.ptr(t), nt, ndiv, 16,
                                         hist.ctypes.data, C.byref(bs0), nat.stream_ptr())
        assert ier == 0
    else:
        # the keys of fmm2d_prepare (20 levels) give the same counts
        h = C.c_void_p()
        assert L.fmm2d_prepare(nat.ptr(s), ns, nat.ptr(t), nt, C.byref(h), nat.stream_ptr()) == 0
        try:
            ier = L.fmm2d_prepared_leaf_histogram(h, ndiv, 16, hist.ctypes.data, C.byref(bs0),
                                                  nat.stream_ptr())
            assert ier == 0
            h21 = np.zeros(21)
            assert L.fmm2d_prepared_leaf_histogram(h, ndiv, 21, h21.ctypes.data, None,
                                                   nat.stream_ptr()) == 0
            k21 = np.zeros(21)
            assert L.fmm2d_key_leaf_histogram(nat.ptr(s), ns, nat.ptr(t), nt, ndiv, 21,
                                              k21.ctypes.data, None, nat.stream_ptr()) == 0
            assert np.array_equal(h21, k21), (h21, k21)
        finally:
            L.fmm2d_prepare_free(h)
    want = H.windows(src, trg, ndiv)
    assert np.array_equal(hist, want), (hist, want)
    pts = src if trg is None else np.concatenate([src, trg], axis=1)
    assert bs0.value == max(np.ptp(pts[0]), np.ptp(pts[1]))
