"""CPU tests of the host side: the C-ABI library loads and exports every
symbol include/fmm2d_amd.h declares; the Python API mirrors the reference's
argument checks, messages and dispatch quirks without touching the GPU; the
reference FMM fixtures themselves agree with the oracle's direct sum."""
import ctypes
import os
import re

import numpy as np
import pytest

import fmm2d_amd
from fmm2d_amd import _native as nat
from fmm2d_amd import fmm2d as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "fmm2d_amd.h")).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|void)\s+(fmm2d_\w+)\s*\(", txt, flags=re.M)))


def test_header_symbols_exported():
    names = _declared()
    assert len(names) >= 14
    lib = ctypes.CDLL(nat.LIBPATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(nat.SIGNATURES), set(names) ^ set(nat.SIGNATURES)
    assert lib.fmm2d_abi_version() == 1


def test_public_names_match_reference():
    ref = ["hfmm2d", "rfmm2d", "lfmm2d", "cfmm2d", "bhfmm2d", "stfmm2d", "h2ddir", "r2ddir",
           "l2ddir", "c2ddir", "bh2ddir", "st2ddir", "comperr", "Output"]
    for n in ref:
        assert hasattr(fmm2d_amd, n), n
    o = fmm2d_amd.Output()
    for f in ("pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg", "pre", "pretarg"):
        assert getattr(o, f) is None
    assert o.ier == 0


def test_nothing_to_compute(capsys):
    src = np.zeros((2, 4))
    for fn in (lambda: F.lfmm2d(eps=1e-6, sources=src, charges=np.ones(4)),
               lambda: F.cfmm2d(eps=1e-6, sources=src, charges=np.ones(4)),
               lambda: F.hfmm2d(eps=1e-6, zk=1.0, sources=src, charges=np.ones(4)),
               lambda: F.l2ddir(sources=src, targets=src, charges=np.ones(4)),
               lambda: F.h2ddir(zk=1.0, sources=src, targets=src, charges=np.ones(4))):
        out = fn()
        assert out.pot is None and out.pottarg is None
    assert capsys.readouterr().out.count("Nothing to compute") == 5


def test_shape_assertions():
    src = np.zeros((2, 5))
    with pytest.raises(AssertionError, match="first dimension of sources must be 2"):
        F.lfmm2d(eps=1e-6, sources=np.zeros((3, 5)), charges=np.ones(5), pg=1)
    with pytest.raises(AssertionError, match="Charges must be same length"):
        F.lfmm2d(eps=1e-6, sources=src, charges=np.ones(4), pg=1)
    with pytest.raises(AssertionError, match="Charges must be of shape"):
        F.cfmm2d(eps=1e-6, sources=src, charges=np.ones((2, 4)), pg=1, nd=2)
    with pytest.raises(AssertionError, match="dipole vectors must be of shape"):
        F.lfmm2d(eps=1e-6, sources=src, dipstr=np.ones(5), dipvec=np.ones((3, 5)), pg=1)
    with pytest.raises(AssertionError, match="first dimension of targets must be 2"):
        F.l2ddir(sources=src, targets=np.zeros((3, 2)), charges=np.ones(5), pgt=1)


def test_biharmonic_stokes_host_checks(capsys):
    """The reference wrappers' early returns and shape assertions
    (python/fmm2dpy/fmm2d.py:1031-1054, 1172-1228) run before any device work."""
    src = np.zeros((2, 5))
    out = F.bhfmm2d(eps=1e-6, sources=src, charges=np.ones((2, 5), complex))
    assert out.pot is None and "Nothing to compute" in capsys.readouterr().out
    out = F.bh2ddir(sources=src, targets=src, charges=np.ones((2, 5), complex))
    assert out.pottarg is None and "Nothing to compute" in capsys.readouterr().out
    with pytest.raises(AssertionError, match="Charges must be of shape \\[2, number of source\\]"):
        F.bhfmm2d(eps=1e-6, sources=src, charges=np.ones((3, 5), complex), pg=1)
    with pytest.raises(AssertionError, match="Dipole must of shape \\[3, number of sources\\]"):
        F.bhfmm2d(eps=1e-6, sources=src, dipoles=np.ones((2, 5), complex), pg=1)
    with pytest.raises(AssertionError, match="Charges must be of shape \\[nd,2,ns\\]"):
        F.bh2ddir(sources=src, targets=src, charges=np.ones((2, 2, 5), complex), pgt=1, nd=3)
    out = F.stfmm2d(eps=1e-6, sources=src, stoklet=np.ones((2, 5)))
    assert out.pot is None and "set either ifppreg or ifppregtarg" in capsys.readouterr().out
    out = F.stfmm2d(eps=1e-6, sources=src, ifppreg=1)
    assert out.pot is None and "set either stoklet or strslet+strsvec" in capsys.readouterr().out
    out = F.st2ddir(eps=1e-6, sources=src, strslet=np.ones((2, 5)), ifppreg=1)
    assert out.pot is None and "mush be both None" in capsys.readouterr().out
    with pytest.raises(AssertionError, match="stoklet vectors must be of shape"):
        F.stfmm2d(eps=1e-6, sources=src, stoklet=np.ones((3, 5)), ifppreg=1)
    with pytest.raises(AssertionError, match="strsvec vectors must be of shape"):
        F.st2ddir(eps=1e-6, sources=src, strslet=np.ones((2, 5)), strsvec=np.ones((2, 4)),
                  ifppreg=1)


def test_no_cpu_fallback_without_gpu():
    """With no ROCm device the product path fails loudly instead of computing
    on the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(nat.NativeLibraryError):
        F.l2ddir(sources=np.zeros((2, 3)), targets=np.ones((2, 2)), charges=np.ones(3), pgt=1)


def test_comperr_matches_definition():
    rng = np.random.default_rng(0)
    a, b = F.Output(), F.Output()
    b.pot = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    b.grad = rng.standard_normal((2, 30)) + 0j
    a.pot = b.pot * (1 + 1e-6)
    a.grad = b.grad.copy()
    e = F.comperr(ntest=20, out=a, outex=b, pg=2)
    num = np.linalg.norm(b.pot[:20] * 1e-6)
    den = np.sqrt(np.linalg.norm(b.pot[:20]) ** 2 + np.linalg.norm(b.grad[:, :20]) ** 2)
    assert abs(e - num / den) < 1e-12
    # the Cauchy mode compares real parts only
    a.pot = b.pot + 1j
    assert F.comperr(ntest=20, out=a, outex=b, pg=1, cauchy=1) < 1e-15


def test_reference_fixtures_against_oracle_direct_sum(golden_dir):
    """The reference's FMM outputs in fmm_ref.npz are within their eps of the
    oracle's direct sum (the fixtures are physically right, not just stored)."""
    from oracle import oracle as O
    g = np.load(f"{golden_dir}/fmm_ref.npz")
    src, trg = g["src"], g["trg"]
    ch, ds, dv = g["nd1_charge"], g["nd1_dipstr"], g["nd1_dipvec"]
    ex = O.direct("l", 1, src, trg, ch, ds, dv, 3, 1e-16)
    tag = "nd1_l_1e-06_1133_"
    assert np.linalg.norm(g[tag + "gradtarg"] - ex["gradtarg"]) / np.linalg.norm(ex["gradtarg"]) < 1e-6
    exh = O.direct("h", 1, src, trg, ch, None, None, 2, 1e-16, zk=12.0)
    tagh = "nd1_h_1e-10_1022_"
    assert np.linalg.norm(g[tagh + "pottarg"] - exh["pottarg"]) / np.linalg.norm(exh["pottarg"]) < 1e-9


@pytest.mark.parametrize("cname,pytype", [("fmm2d_job", "FmmJob"), ("fmm2d_opts", "FmmOpts")])
def test_struct_layouts(tmp_path, cname, pytype):
    """ctypes FmmJob / FmmOpts have the size and field offsets of struct
    fmm2d_job / fmm2d_opts in the header."""
    cls = getattr(nat, pytype)
    names = [f[0] for f in cls._fields_]
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "fmm2d_amd.h"\nint main(void){\n'
        + f'printf("%zu\\n", sizeof({cname}));\n'
        + "".join(f'printf("%zu\\n", offsetof({cname}, {n}));\n' for n in names)
        + "return 0;}\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                    str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got[0] == ctypes.sizeof(cls)
    assert got[1:] == [getattr(cls, n).offset for n in names]


def test_batch_argument_checks():
    """fmm2d_batch rejects bad arguments before any device work; the Python
    wrapper checks every problem first."""
    L = nat.lib()
    jobs = (nat.FmmJob * 2)()
    assert L.fmm2d_batch(0, None, 4) == 0
    assert L.fmm2d_batch(1, None, 4) == -1
    assert L.fmm2d_batch(-1, jobs, 4) == -1
    assert L.fmm2d_batch(1, jobs, 0) == -1
    assert L.fmm2d_batch(1, jobs, 17) == -1
    src = np.zeros((2, 4))
    ok = dict(kind="l", eps=1e-6, sources=src, charges=np.ones(4), pg=1)
    with pytest.raises(ValueError, match="unknown kind"):
        F.fmm2d_batch([ok, dict(ok, kind="x")])
    with pytest.raises(ValueError, match="pg must be equal to pgt"):
        F.fmm2d_batch([dict(ok, targets=src, pgt=2)])
    with pytest.raises(ValueError, match="needs zk"):
        F.fmm2d_batch([dict(ok, kind="h")])
    with pytest.raises(ValueError, match="nstreams"):
        F.fmm2d_batch([ok], nstreams=0)


def test_library_rules_match_reference():
    """The library's host rules (fmm2d_l2dterms, fmm2d_h2dterms, fmm2d_ndiv) give
    the reference's values, including h2dterms' ier = 13 when no order up to
    50000 meets eps (helmholtz/h2dterms.f:92-93: nterms = 10001)."""
    from oracle import ref_binding as R
    if not R.available():
        pytest.skip("reference library not built (oracle/ref.mk)")
    L = nat.lib()
    v = ctypes.c_int(0)
    for eps in (1e-3, 0.5e-6, 1e-6, 1e-9, 1e-12, 1e-15):
        assert L.fmm2d_l2dterms(eps, ctypes.byref(v)) == 0
        assert v.value == R.l2dterms(eps)[0], eps
        for ns in (100, 10 ** 5, 10 ** 7):
            nd, idf = ctypes.c_int(0), ctypes.c_int(0)
            assert L.fmm2d_ndiv(eps, ns, 0, 0, ctypes.byref(nd), ctypes.byref(idf)) == 0
            assert (nd.value, idf.value) == R.ndiv(eps, ns, 0), (eps, ns)
    for bs, zk, eps in ((1.0, 10.0, 1e-6), (0.25, 100.0 + 2.0j, 1e-9), (2.6, 500.0, 1e-6),
                        (1.0, 6e4, 1e-6), (1.0, 1e5, 1e-6)):
        ier = L.fmm2d_h2dterms(bs, complex(zk).real, complex(zk).imag, eps, ctypes.byref(v))
        assert (v.value, ier) == R.h2dterms(bs, zk, eps), (bs, zk, eps)
    assert R.h2dterms(1.0, 1e5, 1e-6) == (10001, 13)
