"""ctypes binding to the C oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, as the checker.  See oracle.h for
what each function restates and how it is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)
        _lib = C.CDLL(LIBPATH)
        _lib.orc_l2dterms.restype = C.c_int
        _lib.orc_l2dterms.argtypes = [C.c_double]
        _lib.orc_lndiv2d.argtypes = [C.c_double, C.c_int, C.c_int]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def direct(kind, nd, sources, targets, charges=None, dipstr=None, dipvec=None, pgt=1,
           thresh=1e-16, zk=0.0):
    """Same contract as ref_binding.direct: outputs start at zero, Fortran shapes."""
    fam = {"c": 0, "l": 1, "h": 2}[kind]
    src = np.asfortranarray(np.asarray(sources, dtype=np.float64))
    trg = np.asfortranarray(np.asarray(targets, dtype=np.float64))
    ns, nt = src.shape[1], trg.shape[1]
    ch = np.asfortranarray(np.reshape(charges, (nd, ns)).astype(np.complex128)) \
        if charges is not None else np.zeros((nd, 1), np.complex128)
    ds = np.asfortranarray(np.reshape(dipstr, (nd, ns)).astype(np.complex128)) \
        if dipstr is not None else np.zeros((nd, 1), np.complex128)
    dv = np.asfortranarray(np.reshape(dipvec, (nd, 2, ns)).astype(np.float64)) \
        if (dipvec is not None and kind != "c") else np.zeros((nd, 2, 1))
    ng = 1 if kind == "c" else 2
    nh = 1 if kind == "c" else 3
    pot = np.zeros((nd, nt), np.complex128, order="F")
    grad = np.zeros((nd, ng, nt) if ng > 1 else (nd, nt), np.complex128, order="F")
    hess = np.zeros((nd, nh, nt) if nh > 1 else (nd, nt), np.complex128, order="F")
    zk = complex(zk)
    lib().orc_direct(C.c_int(fam), C.c_int(nd), C.c_double(zk.real), C.c_double(zk.imag),
                     _p(src), C.c_int(ns), C.c_int(1 if charges is not None else 0), _p(ch),
                     C.c_int(1 if dipstr is not None else 0), _p(ds), _p(dv), _p(trg),
                     C.c_int(nt), C.c_int(pgt), C.c_double(thresh), _p(pot), _p(grad), _p(hess))
    out = dict(pottarg=pot)
    if pgt >= 2:
        out["gradtarg"] = grad
    if pgt >= 3:
        out["hesstarg"] = hess
    return out


def bh_direct(nd, sources, targets, charges=None, dipoles=None, pgt=1, thresh=1e-16):
    """orc_bh_direct: vel (nd,nt), grad (nd,3,nt) from zero; charges (nd,2,ns),
    dipoles (nd,3,ns) complex."""
    src = np.asfortranarray(np.asarray(sources, dtype=np.float64))
    trg = np.asfortranarray(np.asarray(targets, dtype=np.float64))
    ns, nt = src.shape[1], trg.shape[1]
    ch = np.asfortranarray(np.reshape(charges, (nd, 2, ns)).astype(np.complex128)) \
        if charges is not None else np.zeros((nd, 2, 1), np.complex128)
    dp = np.asfortranarray(np.reshape(dipoles, (nd, 3, ns)).astype(np.complex128)) \
        if dipoles is not None else np.zeros((nd, 3, 1), np.complex128)
    vel = np.zeros((nd, nt), np.complex128, order="F")
    grad = np.zeros((nd, 3, nt), np.complex128, order="F")
    lib().orc_bh_direct(C.c_int(nd), _p(src), C.c_int(ns), C.c_int(charges is not None),
                        _p(ch), C.c_int(dipoles is not None), _p(dp), _p(trg), C.c_int(nt),
                        C.c_int(pgt), C.c_double(thresh), _p(vel), _p(grad))
    out = dict(pottarg=vel)
    if pgt >= 2:
        out["gradtarg"] = grad
    return out


def st_direct(nd, sources, targets, stoklet=None, strslet=None, strsvec=None, thresh=1e-16):
    """orc_st_direct: pot (nd,2,nt), pre (nd,nt), grad (nd,2,2,nt) from zero."""
    src = np.asfortranarray(np.asarray(sources, dtype=np.float64))
    trg = np.asfortranarray(np.asarray(targets, dtype=np.float64))
    ns, nt = src.shape[1], trg.shape[1]

    def arr(a):
        return np.asfortranarray(np.reshape(a, (nd, 2, ns)).astype(np.float64)) \
            if a is not None else np.zeros((nd, 2, 1))
    sk, sl, sv = arr(stoklet), arr(strslet), arr(strsvec)
    pot = np.zeros((nd, 2, nt), order="F")
    pre = np.zeros((nd, nt), order="F")
    grad = np.zeros((nd, 2, 2, nt), order="F")
    lib().orc_st_direct(C.c_int(nd), _p(src), C.c_int(ns), C.c_int(stoklet is not None),
                        _p(sk), C.c_int(strslet is not None), _p(sl), _p(sv), _p(trg),
                        C.c_int(nt), C.c_double(thresh), _p(pot), _p(pre), _p(grad))
    return dict(pottarg=pot, pretarg=pre, gradtarg=grad)


def hank01(z):
    z = np.asarray(z, dtype=np.complex128).ravel()
    h0 = np.zeros_like(z)
    h1 = np.zeros_like(z)
    a = np.zeros(2)
    b = np.zeros(2)
    for k, v in enumerate(z):
        lib().orc_hank01(C.c_double(v.real), C.c_double(v.imag), _p(a), _p(b))
        h0[k] = a[0] + 1j * a[1]
        h1[k] = b[0] + 1j * b[1]
    return h0, h1


def tree(sources, targets, ndiv, idivflag=0, nlmax=51):
    """Tree + lists in the reference's conventions (see ref_binding.tree); list
    rows in the reference's own order."""
    L = lib()
    src = np.ascontiguousarray(np.asarray(sources, dtype=np.float64).T)
    ns = src.shape[0]
    if targets is None:
        trg, nt = np.zeros((1, 2)), 0
    else:
        trg = np.ascontiguousarray(np.asarray(targets, dtype=np.float64).T)
        nt = trg.shape[0]
    nl, nb = C.c_int(0), C.c_int(0)
    h = C.c_void_p()
    L.orc_tree_build(_p(src), C.c_int(ns), _p(trg), C.c_int(nt), C.c_int(idivflag),
                     C.c_int(ndiv), C.c_int(nlmax), C.byref(nl), C.byref(nb), C.byref(h))
    nl, nb = nl.value, nb.value
    caps = (13, 27, 20, 5)
    arr = dict(laddr=np.zeros((nl + 1, 2), np.int32), ilevel=np.zeros(nb, np.int32),
               iparent=np.zeros(nb, np.int32), nchild=np.zeros(nb, np.int32),
               ichild=np.zeros((nb, 4), np.int32), ncoll=np.zeros(nb, np.int32),
               coll=np.zeros((nb, 9), np.int32), centers=np.zeros((nb, 2)),
               boxsize=np.zeros(nl + 1), isrc=np.zeros(max(ns, 1), np.int32),
               isrcse=np.zeros((nb, 2), np.int32), itarg=np.zeros(max(nt, 1), np.int32),
               itargse=np.zeros((nb, 2), np.int32))
    for k in range(4):
        arr[f"nlist{k + 1}"] = np.zeros(nb, np.int32)
        arr[f"list{k + 1}"] = np.zeros((nb, caps[k]), np.int32)
    order = ["laddr", "ilevel", "iparent", "nchild", "ichild", "ncoll", "coll", "centers",
             "boxsize", "isrc", "isrcse", "itarg", "itargse", "nlist1", "list1", "nlist2",
             "list2", "nlist3", "list3", "nlist4", "list4"]
    L.orc_tree_get(h, *[_p(arr[k]) for k in order])
    L.orc_tree_free(h)
    out = dict(nlevels=nl, nboxes=nb)
    for k, v in arr.items():
        out[k] = v.T.copy() if v.ndim == 2 else v
    out["isrc"] = out["isrc"][:ns]
    out["itarg"] = out["itarg"][:nt]
    return out


def l2dterms(eps):
    return lib().orc_l2dterms(eps)


def lndiv2d(eps, ns=0, nt=0):
    return lib().orc_lndiv2d(eps, ns, nt)
