"""ctypes binding to the reference Fortran library built by ``oracle/ref.mk``.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline -- never as the thing measured or shipped.

The library is compiled from the reference's own Fortran sources (flang, see
``oracle/ref.mk``) into ``oracle/_ref/libfmm2d_ref.so``.  Every routine here is
called with Fortran conventions: arguments by reference, column-major arrays,
default 4-byte integers, complex*16 as interleaved (re, im) doubles.

Entry points bound (reference file:line):
  cfmm2d  laplace/cfmm2d.f:23      lfmm2d  laplace/lfmm2d.f:29
  hfmm2d  helmholtz/hfmm2d.f:19    *_ndiv  laplace/lfmm2d_ndiv.f:29,
                                           laplace/cfmm2d_ndiv.f, helmholtz/hfmm2d_ndiv.f:19
  c2d_direct*  laplace/cauchykernels2d.f   l2d_direct* laplace/lapkernels2d.f
  h2d_direct*  helmholtz/helmkernels2d.f   hank103 common/hank103.f:116
  pts_tree_mem/build/sort common/pts_tree2d.f:34,406,1222
  computelists common/tree_routs2d.f:486   l2dterms laplace/l2dterms.f:38
  h2dterms helmholtz/h2dterms.f:35         lndiv2d laplace/lndiv2d.f
  hkrand common/hkrand.f:3
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "_ref", "libfmm2d_ref.so")
_lib = None


def available():
    return os.path.exists(LIBPATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(
                f"{LIBPATH} not built (make -f oracle/ref.mk needs the reference sources)")
        _lib = C.CDLL(LIBPATH)
    return _lib


def _i(v):
    return C.byref(C.c_int(int(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def _z(v):
    a = np.array([complex(v)], dtype=np.complex128)
    return a, a.ctypes.data_as(C.c_void_p)


def _farr(a, dtype):
    """Fortran-ordered contiguous copy; returns (array, pointer)."""
    a = np.asfortranarray(np.asarray(a, dtype=dtype))
    return a, a.ctypes.data_as(C.c_void_p)


def _out(shape, dtype=np.complex128):
    a = np.zeros(shape, dtype=dtype, order="F")
    return a, a.ctypes.data_as(C.c_void_p)


def _src_arrays(nd, ns, charges, dipstr, dipvec, need_dipvec, dtype=np.complex128):
    keep = []
    if charges is None:
        ch, pch = _out((nd, 1), dtype)
        ifc = 0
    else:
        ch, pch = _farr(np.reshape(charges, (nd, ns), order="C"), dtype)
        ifc = 1
    if dipstr is None:
        ds, pds = _out((nd, 1), dtype)
        dv, pdv = _out((nd, 2, 1), np.float64)
        ifd = 0
    else:
        ds, pds = _farr(np.reshape(dipstr, (nd, ns), order="C"), dtype)
        if need_dipvec:
            dv, pdv = _farr(np.reshape(dipvec, (nd, 2, ns), order="C"), np.float64)
        else:
            dv, pdv = _out((nd, 2, 1), np.float64)
        ifd = 1
    keep += [ch, ds, dv]
    return ifc, pch, ifd, pds, pdv, keep


def _fmm(kind, nd, eps, sources, charges=None, dipstr=None, dipvec=None,
         targets=None, ifpgh=0, ifpghtarg=0, zk=None, ndiv=None, idivflag=0,
         ifnear=1):
    """Call cfmm2d/lfmm2d/hfmm2d (or the *_ndiv variants).

    sources (2,ns), targets (2,nt); charges/dipstr (nd,ns) or (ns,);
    dipvec (nd,2,ns) or (2,ns).  Returns dict of outputs with the Fortran
    shapes: pot (nd,n); grad (nd,n) [cauchy] or (nd,2,n); hess (nd,n) or (nd,3,n).
    """
    L = lib()
    sources = np.asarray(sources, dtype=np.float64)
    ns = sources.shape[1]
    src, psrc = _farr(sources, np.float64)
    if targets is None:
        nt = 0
        trg, ptrg = _out((2, 1), np.float64)
    else:
        targets = np.asarray(targets, dtype=np.float64)
        nt = targets.shape[1]
        trg, ptrg = _farr(targets, np.float64)
    need_dv = kind != "c"
    dt = np.float64 if kind == "r" else np.complex128
    ifc, pch, ifd, pds, pdv, keep = _src_arrays(nd, ns, charges, dipstr, dipvec, need_dv, dt)
    gdim = (nd,) if kind == "c" else (nd, 2)
    hdim = (nd,) if kind == "c" else (nd, 3)
    n1 = max(ns, 1)
    t1 = max(nt, 1)
    pot, ppot = _out((nd, n1), dt)
    grad, pgrad = _out(gdim + (n1,), dt)
    hess, phess = _out(hdim + (n1,), dt)
    pt, ppt = _out((nd, t1), dt)
    gt, pgt = _out(gdim + (t1,), dt)
    ht, pht = _out(hdim + (t1,), dt)
    ier = C.c_int(0)
    timeinfo = np.zeros(8)
    args = []
    if kind == "h":
        zka, pzk = _z(zk)
        keep.append(zka)
        head = [_i(nd), _d(eps), pzk, _i(ns), psrc]
    else:
        head = [_i(nd), _d(eps), _i(ns), psrc]
    args = head + [_i(ifc), pch, _i(ifd), pds]
    if kind != "c":
        args.append(pdv)
    args += [_i(0), _i(ifpgh), ppot, pgrad, phess, _i(nt), ptrg, _i(ifpghtarg), ppt, pgt, pht]
    name = {"c": "cfmm2d", "l": "lfmm2d", "h": "hfmm2d", "r": "rfmm2d"}[kind]
    if ndiv is not None:
        args += [_i(ndiv), _i(idivflag), _i(ifnear), timeinfo.ctypes.data_as(C.c_void_p)]
        name += "_ndiv"
    args.append(C.byref(ier))
    getattr(L, name + "_")(*args)
    out = dict(ier=ier.value, timeinfo=timeinfo)
    if ifpgh >= 1:
        out["pot"] = pot
    if ifpgh >= 2:
        out["grad"] = grad
    if ifpgh >= 3:
        out["hess"] = hess
    if ifpghtarg >= 1:
        out["pottarg"] = pt
    if ifpghtarg >= 2:
        out["gradtarg"] = gt
    if ifpghtarg >= 3:
        out["hesstarg"] = ht
    return out


def cfmm2d(**kw):
    return _fmm("c", **kw)


def lfmm2d(**kw):
    return _fmm("l", **kw)


def hfmm2d(**kw):
    return _fmm("h", **kw)


def rfmm2d(**kw):
    return _fmm("r", **kw)


def direct(kind, nd, sources, targets, charges=None, dipstr=None, dipvec=None,
           pgt=1, thresh=1e-16, zk=None):
    """Reference direct kernels c2d_/l2d_/h2d_/r2d_direct{c,d,cd}{p,g,h}; outputs start at 0
    (kind "r": real densities and outputs)."""
    L = lib()
    sources = np.asarray(sources, dtype=np.float64)
    targets = np.asarray(targets, dtype=np.float64)
    ns = sources.shape[1]
    nt = targets.shape[1]
    src, psrc = _farr(sources, np.float64)
    trg, ptrg = _farr(targets, np.float64)
    need_dv = kind != "c"
    dt = np.float64 if kind == "r" else np.complex128
    ifc, pch, ifd, pds, pdv, keep = _src_arrays(nd, ns, charges, dipstr, dipvec, need_dv, dt)
    tag = ("c" if ifc else "") + ("d" if ifd else "") + "pgh"[pgt - 1]
    gdim = (nd,) if kind == "c" else (nd, 2)
    hdim = (nd,) if kind == "c" else (nd, 3)
    pot, ppot = _out((nd, nt), dt)
    grad, pgrad = _out(gdim + (nt,), dt)
    hess, phess = _out(hdim + (nt,), dt)
    args = [_i(nd)]
    if kind == "h":
        zka, pzk = _z(zk)
        keep.append(zka)
        args.append(pzk)
    args += [psrc, _i(ns)]
    if ifc:
        args.append(pch)
    if ifd:
        args.append(pds)
        if need_dv:
            args.append(pdv)
    args += [ptrg, _i(nt), ppot]
    if pgt >= 2:
        args.append(pgrad)
    if pgt >= 3:
        args.append(phess)
    args.append(_d(thresh))
    getattr(L, f"{kind}2d_direct{tag}_")(*args)
    out = dict(pottarg=pot)
    if pgt >= 2:
        out["gradtarg"] = grad
    if pgt >= 3:
        out["hesstarg"] = hess
    return out


def hank103(z, ifexpon=1):
    """H0^(1)(z), H1^(1)(z) via common/hank103.f:116 for an array of complex z."""
    L = lib()
    z = np.asarray(z, dtype=np.complex128).ravel()
    h0 = np.zeros_like(z)
    h1 = np.zeros_like(z)
    zz = np.zeros(1, np.complex128)
    a0 = np.zeros(1, np.complex128)
    a1 = np.zeros(1, np.complex128)
    for k, v in enumerate(z):
        zz[0] = v
        L.hank103_(zz.ctypes.data_as(C.c_void_p), a0.ctypes.data_as(C.c_void_p),
                   a1.ctypes.data_as(C.c_void_p), _i(ifexpon))
        h0[k] = a0[0]
        h1[k] = a1[0]
    return h0, h1


def l2dterms(eps):
    L = lib()
    n = C.c_int(0)
    ier = C.c_int(0)
    L.l2dterms_(_d(eps), C.byref(n), C.byref(ier))
    return n.value, ier.value


def h2dterms(bsize, zk, eps):
    L = lib()
    n = C.c_int(0)
    ier = C.c_int(0)
    zka, pzk = _z(zk)
    L.h2dterms_(_d(bsize), pzk, _d(eps), C.byref(n), C.byref(ier))
    return n.value, ier.value


def ndiv(eps, ns=0, nt=0, helm=False):
    L = lib()
    nd = C.c_int(0)
    idf = C.c_int(0)
    f = L.hndiv2d_ if helm else L.lndiv2d_
    f(_d(eps), _i(ns), _i(nt), _i(1), _i(0), _i(1), _i(0), C.byref(nd), C.byref(idf))
    return nd.value, idf.value


def hkrand_seq(n, seed=None):
    """n successive values of common/hkrand.f:3 (seed None continues the SAVEd state)."""
    L = lib()
    L.hkrand_.restype = C.c_double
    out = np.empty(n)
    if seed is not None:
        out0 = L.hkrand_(_i(seed))
        out[0] = out0
        start = 1
    else:
        start = 0
    for k in range(start, n):
        out[k] = L.hkrand_(_i(0))
    return out


def tree(sources, targets, ndiv, idivflag=0, nlmin=0, nlmax=51, ifunif=0):
    """pts_tree_mem + pts_tree_build + pts_tree_sort (sources and targets).

    Returns dict: nlevels, nboxes, laddr (2,nlevels+1), ilevel, iparent,
    nchild, ichild (4,nboxes), ncoll, coll (9,nboxes), centers (2,nboxes),
    boxsize (nlevels+1,), isrc, isrcse (2,nboxes), itarg, itargse -- all
    integers 1-based as in the reference.
    """
    L = lib()
    sources = np.asarray(sources, dtype=np.float64)
    targets = np.zeros((2, 0)) if targets is None else np.asarray(targets, dtype=np.float64)
    ns = sources.shape[1]
    nt = targets.shape[1]
    src, psrc = _farr(sources, np.float64)
    trg, ptrg = _farr(targets if nt else np.zeros((2, 1)), np.float64)
    nlev = C.c_int(0)
    nbox = C.c_int(0)
    ltree = C.c_int(0)
    L.pts_tree_mem_(psrc, _i(ns), ptrg, _i(nt), _i(idivflag), _i(ndiv), _i(nlmin), _i(nlmax),
                    _i(ifunif), _i(0), C.byref(nlev), C.byref(nbox), C.byref(ltree))
    nl, nb, lt = nlev.value, nbox.value, ltree.value
    itree = np.zeros(lt, dtype=np.int32)
    iptr = np.zeros(8, dtype=np.int32)
    centers = np.zeros((2, nb), order="F")
    boxsize = np.zeros(nl + 1)
    L.pts_tree_build_(psrc, _i(ns), ptrg, _i(nt), _i(idivflag), _i(ndiv), _i(nlmin), _i(nlmax),
                      _i(ifunif), _i(0), _i(nl), _i(nb), _i(lt),
                      itree.ctypes.data_as(C.c_void_p), iptr.ctypes.data_as(C.c_void_p),
                      centers.ctypes.data_as(C.c_void_p), boxsize.ctypes.data_as(C.c_void_p))
    res = dict(nlevels=nl, nboxes=nb, itree=itree, iptr=iptr.copy(), centers=centers,
               boxsize=boxsize)
    p = [int(v) - 1 for v in iptr]
    res["laddr"] = itree[0:2 * (nl + 1)].reshape(nl + 1, 2).T.copy()
    res["ilevel"] = itree[p[1]:p[1] + nb].copy()
    res["iparent"] = itree[p[2]:p[2] + nb].copy()
    res["nchild"] = itree[p[3]:p[3] + nb].copy()
    res["ichild"] = itree[p[4]:p[4] + 4 * nb].reshape(nb, 4).T.copy()
    res["ncoll"] = itree[p[5]:p[5] + nb].copy()
    res["coll"] = itree[p[6]:p[6] + 9 * nb].reshape(nb, 9).T.copy()
    for name, pts, n in (("src", src, ns), ("targ", trg, nt)):
        ix = np.zeros(max(n, 1), dtype=np.int32)
        ixse = np.zeros((2, nb), dtype=np.int32, order="F")
        L.pts_tree_sort_(_i(n), pts.ctypes.data_as(C.c_void_p), itree.ctypes.data_as(C.c_void_p),
                         _i(lt), _i(nb), _i(nl), iptr.ctypes.data_as(C.c_void_p),
                         centers.ctypes.data_as(C.c_void_p), ix.ctypes.data_as(C.c_void_p),
                         ixse.ctypes.data_as(C.c_void_p))
        res["i" + name] = ix[:n]
        res["i" + name + "se"] = ixse
    return res


def lists(tr):
    """computemnlists + computelists on a tree from tree(); 1-based, Fortran layout."""
    L = lib()
    nb = tr["nboxes"]
    nl = tr["nlevels"]
    itree = tr["itree"]
    iptr = tr["iptr"]
    centers = np.asfortranarray(tr["centers"])
    boxsize = tr["boxsize"]
    mn = [13, 27, 20, 5]
    ns = [np.zeros(nb, np.int32) for _ in range(4)]
    ls = [np.zeros((m, nb), np.int32, order="F") for m in mn]
    args = [_i(nl), _i(nb), itree.ctypes.data_as(C.c_void_p), _i(len(itree)),
            iptr.ctypes.data_as(C.c_void_p), centers.ctypes.data_as(C.c_void_p),
            boxsize.ctypes.data_as(C.c_void_p), _i(0)]
    for k in range(4):
        args += [_i(mn[k]), ns[k].ctypes.data_as(C.c_void_p), ls[k].ctypes.data_as(C.c_void_p)]
    L.computelists_(*args)
    return dict(nlist1=ns[0], list1=ls[0], nlist2=ns[1], list2=ls[1],
                nlist3=ns[2], list3=ls[2], nlist4=ns[3], list4=ls[3])


def h2dformmpc(zk, rscale, sources, charges, center, nterms):
    """Multipole expansion of point charges (helmholtz/helmrouts2d.f h2dformmpc),
    nd = 1: coefficients n = -nterms..nterms."""
    L = lib()
    sources = np.asarray(sources, dtype=np.float64)
    ns = sources.shape[1]
    src, psrc = _farr(sources, np.float64)
    ch, pch = _farr(np.asarray(charges, dtype=np.complex128).reshape(1, ns), np.complex128)
    cen, pcen = _farr(np.asarray(center, dtype=np.float64), np.float64)
    zka, pzk = _z(zk)
    mp, pmp = _out((2 * nterms + 1,))
    L.h2dformmpc_(_i(1), pzk, _d(rscale), psrc, _i(ns), pch, pcen, _i(nterms), pmp)
    return mp


def hfmm2d_mps(nd, eps, zk, centers, rscales, nterms, mpole, impole, ndiv=None):
    """helmholtz/hfmm2d_mps.f:19 hfmm2d_mps (or hfmm2d_mps_ndiv.f:70
    hfmm2d_mps_ndiv with the given ndiv; ndiv = nmpole is the reference's
    all-pairs evaluation, hfmm2d_mps_direct).  impole is 1-based (Fortran);
    mpole and the returned local are flat complex arrays of length ntot."""
    L = lib()
    centers = np.asarray(centers, dtype=np.float64)
    n = centers.shape[1]
    cen, pcen = _farr(centers, np.float64)
    rs, prs = _farr(np.asarray(rscales, dtype=np.float64), np.float64)
    mt, pmt = _farr(np.asarray(nterms, dtype=np.int32), np.int32)
    im, pim = _farr(np.asarray(impole, dtype=np.int32), np.int32)
    mp, pmp = _farr(np.asarray(mpole, dtype=np.complex128), np.complex128)
    ntot = mp.shape[0]
    loc, ploc = _out((ntot,))
    zka, pzk = _z(zk)
    ier = C.c_int(0)
    if ndiv is None:
        L.hfmm2d_mps_(_i(nd), _d(eps), pzk, _i(n), pcen, prs, pmt, _i(ntot), pmp, pim, ploc,
                      C.byref(ier))
    else:
        tinfo = np.zeros(8)
        L.hfmm2d_mps_ndiv_(_i(nd), _d(eps), pzk, _i(n), pcen, prs, pmt, _i(ntot), pmp, pim, ploc,
                           _i(ndiv), tinfo.ctypes.data_as(C.c_void_p), C.byref(ier))
    return loc, ier.value


def bh_direct(nd, sources, targets, charges=None, dipoles=None, pgt=1, thresh=1e-16):
    """Reference bh2d_direct{c,d,cd}{p,g} (biharmonic/bhkernels2d.f:26-410): vel (nd,nt),
    grad (nd,3,nt) from zero."""
    L = lib()
    src, psrc = _farr(sources, np.float64)
    trg, ptrg = _farr(targets, np.float64)
    ns, nt = src.shape[1], trg.shape[1]
    keep = []
    args = [_i(nd), psrc, _i(ns)]
    if charges is not None:
        ch, pch = _farr(np.reshape(charges, (nd, 2, ns)), np.complex128)
        keep.append(ch)
        args.append(pch)
    if dipoles is not None:
        dp, pdp = _farr(np.reshape(dipoles, (nd, 3, ns)), np.complex128)
        keep.append(dp)
        args.append(pdp)
    vel, pvel = _out((nd, nt))
    grad, pgrad = _out((nd, 3, nt))
    args += [ptrg, _i(nt), pvel]
    if pgt >= 2:
        args.append(pgrad)
    args.append(_d(thresh))
    tag = ("c" if charges is not None else "") + ("d" if dipoles is not None else "") + \
        "pg"[pgt - 1]
    getattr(L, f"bh2d_direct{tag}_")(*args)
    out = dict(pottarg=vel)
    if pgt >= 2:
        out["gradtarg"] = grad
    return out


def st_direct(nd, sources, targets, stoklet=None, strslet=None, strsvec=None, thresh=1e-16):
    """Reference st2ddirectstokg / st2ddirectstokstrsg (stokes/stokkernels2d.f:24, 169):
    pot (nd,2,nt), pre (nd,nt), grad (nd,2,2,nt) from zero."""
    L = lib()
    src, psrc = _farr(sources, np.float64)
    trg, ptrg = _farr(targets, np.float64)
    ns, nt = src.shape[1], trg.shape[1]

    def arr(a):
        return _farr(np.reshape(a, (nd, 2, ns)) if a is not None else np.zeros((nd, 2, ns)),
                     np.float64)
    sk, psk = arr(stoklet)
    sl, psl = arr(strslet)
    sv, psv = arr(strsvec)
    pot, ppot = _out((nd, 2, nt), np.float64)
    pre, ppre = _out((nd, nt), np.float64)
    grad, pgrad = _out((nd, 2, 2, nt), np.float64)
    if strslet is None:
        L.st2ddirectstokg_(_i(nd), psrc, psk, _i(ns), ptrg, _i(nt), ppot, ppre, pgrad,
                           _d(thresh))
    else:
        L.st2ddirectstokstrsg_(_i(nd), psrc, _i(1 if stoklet is not None else 0), psk, _i(1),
                               psl, psv, _i(ns), ptrg, _i(nt), ppot, ppre, pgrad, _d(thresh))
    return dict(pottarg=pot, pretarg=pre, gradtarg=grad)


def bhfmm2d(nd, eps, sources, charges=None, dipoles=None, targets=None, ifpgh=0,
            ifpghtarg=0):
    """Reference bhfmm2d (biharmonic/bhfmm2d.f:1): pot (nd,ns), grad (nd,3,ns), pottarg,
    gradtarg."""
    L = lib()
    src, psrc = _farr(sources, np.float64)
    ns = src.shape[1]
    if targets is None:
        nt = 0
        trg, ptrg = _out((2, 1), np.float64)
    else:
        trg, ptrg = _farr(targets, np.float64)
        nt = trg.shape[1]
    ifc = 1 if charges is not None else 0
    ifd = 1 if dipoles is not None else 0
    ch, pch = _farr(np.reshape(charges, (nd, 2, ns)) if ifc else np.zeros((nd, 2, ns)),
                    np.complex128)
    dp, pdp = _farr(np.reshape(dipoles, (nd, 3, ns)) if ifd else np.zeros((nd, 3, ns)),
                    np.complex128)
    pot, ppot = _out((nd, ns))
    grad, pgrad = _out((nd, 3, ns))
    hess, phess = _out((nd, 3, ns))
    t1 = max(nt, 1)
    pt, ppt = _out((nd, t1))
    gt, pgt = _out((nd, 3, t1))
    ht, pht = _out((nd, 3, t1))
    ier = C.c_int(0)
    L.bhfmm2d_(_i(nd), _d(eps), _i(ns), psrc, _i(ifc), pch, _i(ifd), pdp, _i(0), _i(ifpgh),
               ppot, pgrad, phess, _i(nt), ptrg, _i(ifpghtarg), ppt, pgt, pht, C.byref(ier))
    return dict(ier=ier.value, pot=pot, grad=grad, pottarg=pt[:, :nt], gradtarg=gt[:, :, :nt])


def stfmm2d(nd, eps, sources, stoklet=None, strslet=None, strsvec=None, targets=None,
            ifppreg=0, ifppregtarg=0):
    """Reference stfmm2d (stokes/stfmm2d.f:27): pot (nd,2,ns), pre (nd,ns),
    grad (nd,2,2,ns) and the target outputs."""
    L = lib()
    src, psrc = _farr(sources, np.float64)
    ns = src.shape[1]
    if targets is None:
        nt = 0
        trg, ptrg = _out((2, 1), np.float64)
    else:
        trg, ptrg = _farr(targets, np.float64)
        nt = trg.shape[1]

    def arr(a):
        return _farr(np.reshape(a, (nd, 2, ns)) if a is not None else np.zeros((nd, 2, ns)),
                     np.float64)
    sk, psk = arr(stoklet)
    sl, psl = arr(strslet)
    sv, psv = arr(strsvec)
    t1 = max(nt, 1)
    pot, ppot = _out((nd, 2, ns), np.float64)
    pre, ppre = _out((nd, ns), np.float64)
    grad, pgrad = _out((nd, 2, 2, ns), np.float64)
    pt, ppt = _out((nd, 2, t1), np.float64)
    prt, pprt = _out((nd, t1), np.float64)
    gt, pgt = _out((nd, 2, 2, t1), np.float64)
    ier = C.c_int(0)
    L.stfmm2d_(_i(nd), _d(eps), _i(ns), psrc, _i(1 if stoklet is not None else 0), psk,
               _i(1 if strslet is not None else 0), psl, psv, _i(ifppreg), ppot, ppre, pgrad,
               _i(nt), ptrg, _i(ifppregtarg), ppt, pprt, pgt, C.byref(ier))
    return dict(ier=ier.value, pot=pot, pre=pre, grad=grad, pottarg=pt[..., :nt],
                pretarg=prt[..., :nt], gradtarg=gt[..., :nt])
