"""Host-side mirror of the reference Python API (python/fmm2dpy/fmm2d.py).

Same public names, keyword-only signatures, assertion messages, dispatch
rules and output shapes as the reference; underneath, every computation is a
call through the C-ABI (include/fmm2d_amd.h) into hand-written gfx950 HIP
kernels.  Inputs may be numpy arrays (copied to the GPU) and outputs are
numpy arrays laid out as f2py returns them (Fortran-ordered).

Reference quirks that are reproduced on purpose (python/fmm2dpy/fmm2d.py):
  * the source-evaluation gate ``iftarg == 0 or pgt != 1 or pgt != 2`` is
    always true (lines 98, 294, 550, 802), so with targets and pg in {1,2,3}
    the sources+targets call decides the outputs and ``pg == pgt`` is asserted;
  * ``hfmm2d`` and ``h2ddir`` have no Hessian (pg/pgt = 3) path;
  * ``l2ddir``/``c2ddir`` with nd == 1 and pgt == 3 reshape ``hesstarg`` but
    leave ``gradtarg`` with its leading density axis (lines 1604-1611, 1717-1725);
  * "Nothing to compute" is printed and an empty Output returned when no
    output is requested.
"""
import ctypes

import numpy as np
import numpy.linalg as la

from . import _native as nat


class Output():
    """Result container (python/fmm2dpy/fmm2d.py:9-18)."""
    pot = None
    grad = None
    hess = None
    pottarg = None
    gradtarg = None
    hesstarg = None
    pre = None
    pretarg = None
    ier = 0


# --------------------------------------------------------------------------
# shared validation (same messages as the reference)
def _nsrc(sources):
    assert sources.shape[0] == 2, "The first dimension of sources must be 2"
    if np.size(np.shape(sources)) == 2:
        return sources.shape[1]
    return 1


def _check_charges(charges, nd, ns):
    if charges is None:
        return 0
    if nd == 1:
        assert charges.shape[0] == ns, "Charges must be same length as second dimension of sources"
    if nd > 1:
        assert charges.shape[0] == nd and charges.shape[1] == ns, \
            "Charges must be of shape [nd,ns] where nd is number of densities, and ns is number of sources"
    return 1


def _check_dipoles(dipstr, dipvec, nd, ns, need_vec=True):
    if need_vec:
        if dipvec is None and dipstr is None:
            return 0
    elif dipstr is None:
        return 0
    if nd == 1:
        assert dipstr.shape[0] == ns, "Dipole strengths must be same length as second dimension of sources"
        if need_vec:
            if ns > 1:
                assert dipvec.shape[0] == 2 and dipvec.shape[1] == ns, \
                    "dipole vectors must be of shape [2,number of sources]"
            else:
                assert dipvec.shape[0] == 2, "dipole vectors must be of shape [2,number of sources]"
    if nd > 1:
        if need_vec:
            assert dipvec.shape[0] == nd and dipvec.shape[1] == 2 and dipvec.shape[2] == ns, \
                "Dipole vectors must be of shape [nd,2,ns] where nd is number of densities, and ns is number of sources"
        assert dipstr.shape[0] == nd and dipstr.shape[1] == ns, \
            "Dipole strengths must be of shape [nd,ns] where nd is number of densities, and ns is number of sources"
    return 1


class _DevInputs:
    """Device copies of the inputs in the C-ABI's (Fortran) layout."""

    def __init__(self, sources, ns, nd, charges, dipstr, dipvec, targets, need_vec, real=False):
        import torch
        self.ns = ns
        self.nd = nd
        src = np.asarray(sources, dtype=np.float64).reshape(2, ns)
        self.src = nat.to_dev_fortran(src, np.float64)
        self.charge = None
        self.dipstr = None
        self.dipvec = None
        dt = np.float64 if real else np.complex128
        if charges is not None:
            self.charge = nat.to_dev_fortran(np.reshape(charges, (nd, ns)), dt)
        if dipstr is not None:
            self.dipstr = nat.to_dev_fortran(np.reshape(dipstr, (nd, ns)), dt)
            if need_vec:
                self.dipvec = nat.to_dev_fortran(np.reshape(dipvec, (nd, 2, ns)), np.float64)
        self.targ = None
        self.nt = 0
        if targets is not None:
            t = np.asarray(targets, dtype=np.float64)
            self.nt = t.shape[1] if t.ndim == 2 else 1
            self.targ = nat.to_dev_fortran(t.reshape(2, self.nt), np.float64)
        self.torch = torch


# --------------------------------------------------------------------------
# FMM drivers
def _fmm(kind, eps, zk, sources, charges, dipstr, dipvec, targets, pg, pgt, nd,
         opts=None):
    """One call of the native driver; returns (dict of f2py-shaped outputs, ier).

    kind: "c" (cfmm2d), "l" (lfmm2d), "h" (hfmm2d) or "r" (rfmm2d, real
    densities and outputs).  pg/pgt are the
    evaluation flags at sources/targets (0 = none); targets may be None.
    """
    ns = _nsrc(sources)
    dev = _DevInputs(sources, ns, nd, charges, dipstr, dipvec,
                     targets if pgt > 0 else None, kind != "c", real=(kind == "r"))
    bufs, ier = _fmm_device(kind, eps, zk, dev, pg, pgt, nd, opts)
    res = {}
    for name, t in bufs.items():
        if t is None:
            continue
        a = nat.fortran_to_numpy(t)
        res[name] = a[0] if nd == 1 else a
    return res, ier


def _fmm_device(kind, eps, zk, dev, pg, pgt, nd, opts=None, bufs=None):
    """The native call of _fmm on inputs already on the device (a _DevInputs);
    returns (dict of device output tensors in the C-ABI layout, ier).  bufs:
    output tensors of an earlier call with the same shapes, reused (the driver
    overwrites them)."""
    ns, nt = dev.ns, dev.nt
    ng = 1 if kind == "c" else 2
    nh = 1 if kind == "c" else 3

    def shapes(n):
        return ((nd, n), (nd, n) if ng == 1 else (nd, ng, n), (nd, n) if nh == 1 else (nd, nh, n))

    if bufs is None:
        import torch
        bufs = {}
        for flag, n, names in ((pg, ns, ("pot", "grad", "hess")),
                               (pgt, nt, ("pottarg", "gradtarg", "hesstarg"))):
            for lvl, (name, shp) in enumerate(zip(names, shapes(n))):
                odt = torch.float64 if kind == "r" else torch.complex128
                bufs[name] = nat.dev_zeros_fortran(shp, odt) if flag > lvl else None
    L = nat.lib()
    ifc = 1 if dev.charge is not None else 0
    ifd = 1 if dev.dipstr is not None else 0
    o = nat.opts_struct(opts)
    head = [nd, float(eps)]
    if kind == "h":
        zk = complex(zk)
        head += [zk.real, zk.imag]
    args = head + [ns, nat.ptr(dev.src), ifc, nat.ptr(dev.charge), ifd, nat.ptr(dev.dipstr)]
    if kind != "c":
        args.append(nat.ptr(dev.dipvec))
    args += [0, pg, nat.ptr(bufs["pot"]), nat.ptr(bufs["grad"]), nat.ptr(bufs["hess"]),
             nt, nat.ptr(dev.targ), pgt, nat.ptr(bufs["pottarg"]), nat.ptr(bufs["gradtarg"]),
             nat.ptr(bufs["hesstarg"]), nat.byref_opts(o), nat.stream_ptr()]
    fn = {"c": L.fmm2d_cfmm2d, "l": L.fmm2d_lfmm2d, "h": L.fmm2d_hfmm2d,
          "r": L.fmm2d_rfmm2d}[kind]
    ier = fn(*args)
    nat.check(ier, kind + "fmm2d")
    if opts is not None:
        # per-call diagnostics written back into the caller's dict
        if "timeinfo" in opts:
            opts["timeinfo"][:] = o.timeinfo_array()
        opts["nlevels"], opts["nboxes"] = int(o.nlevels), int(o.nboxes)
        if o.count_work:
            opts["work"] = np.array(o.work[:], dtype=np.float64)
            opts["near_pairs"], opts["m2l_pairs"] = int(o.near_pairs), int(o.m2l_pairs)
    return bufs, ier


_KIND = {"c": 0, "l": 1, "r": 2, "h": 3}


def fmm2d_batch(problems, nstreams=4):
    """Many independent FMMs in one native call (fmm2d_batch of the C-ABI; the
    reference has no batched driver).

    problems: list of dicts, each with "kind" ("h", "l", "r" or "c") and the
    keyword arguments of hfmm2d / lfmm2d / rfmm2d / cfmm2d (eps, zk, sources,
    charges, dipstr, dipvec, targets, pg, pgt, nd), and optionally "opts" (the
    fmm2d_opts keys of _native.opts_struct: ndiv, precision, nterms, ...; see
    fmm2d_amd.batch_tune for a leaf-capacity tuner that sets them).  When both
    pg and pgt are set they must be equal, as the single-problem calls require.  The problems
    run on nstreams streams at once, largest first; returns one Output per
    problem, in order (ier set on each).
    """
    import torch
    maxpg = {"h": 2, "l": 3, "r": 3, "c": 3}
    for pb in problems:  # argument checks before any device work
        kind = pb.get("kind")
        if kind not in _KIND:
            raise ValueError(f"unknown kind {kind!r}")
        pg, pgt = int(pb.get("pg", 0)), int(pb.get("pgt", 0))
        if not (0 <= pg <= maxpg[kind] and 0 <= pgt <= maxpg[kind]) or (pg == 0 and pgt == 0):
            raise ValueError("pg, pgt: one of them in 1..%d" % maxpg[kind])
        if pg and pgt and pg != pgt:
            raise ValueError("if output is requested at both sources and targets, then pg "
                             "must be equal to pgt")
        if pgt and pb.get("targets") is None:
            raise ValueError("pgt > 0 needs targets")
        if kind == "h" and pb.get("zk") is None:
            raise ValueError("hfmm2d needs zk")
    if not 1 <= int(nstreams) <= 16:
        raise ValueError("nstreams must be in 1..16")
    L = nat.lib()
    jobs, keep = _stage_batch(problems)
    # inputs were staged on torch's stream; the batch runs on streams of its own
    torch.cuda.current_stream().synchronize()
    ier = L.fmm2d_batch(len(problems), jobs, int(nstreams))
    nat.check(ier, "fmm2d_batch")
    outs = []
    for k, (dev, bufs, nd, _) in enumerate(keep):
        out = Output()
        res = {}
        for name, t in bufs.items():
            if t is not None:
                a = nat.fortran_to_numpy(t)
                res[name] = a[0] if nd == 1 else a
        _fill(out, res, int(jobs[k].ier))
        outs.append(out)
    return outs


def _stage_batch(problems):
    """Device inputs and zeroed outputs of checked problems, and their job
    structs: (ctypes array of FmmJob, [(inputs, outputs, nd)])."""
    import torch
    nat.torch_device()
    jobs = (nat.FmmJob * max(len(problems), 1))()
    keep = []
    for k, pb in enumerate(problems):
        kind = pb["kind"]
        nd = int(pb.get("nd", 1))
        pg, pgt = int(pb.get("pg", 0)), int(pb.get("pgt", 0))
        targets = pb.get("targets")
        sources = pb["sources"]
        ns = _nsrc(sources)
        charges, dipstr, dipvec = pb.get("charges"), pb.get("dipstr"), pb.get("dipvec")
        ifc = _check_charges(charges, nd, ns)
        ifd = _check_dipoles(dipstr, dipvec, nd, ns, need_vec=(kind != "c"))
        dev = _DevInputs(sources, ns, nd, charges if ifc else None, dipstr if ifd else None,
                         dipvec, targets if pgt > 0 else None, kind != "c",
                         real=(kind == "r"))
        ng = 1 if kind == "c" else 2
        nh = 1 if kind == "c" else 3
        odt = torch.float64 if kind == "r" else torch.complex128
        bufs = {}
        for flag, n, names in ((pg, ns, ("pot", "grad", "hess")),
                               (pgt, dev.nt, ("pottarg", "gradtarg", "hesstarg"))):
            shp = ((nd, n), (nd, n) if ng == 1 else (nd, ng, n),
                   (nd, n) if nh == 1 else (nd, nh, n))
            for lvl, (name, sh) in enumerate(zip(names, shp)):
                bufs[name] = nat.dev_zeros_fortran(sh, odt) if flag > lvl else None
        j = jobs[k]
        j.kind, j.nd, j.eps = _KIND[kind], nd, float(pb["eps"])
        zk = complex(pb.get("zk", 0.0) or 0.0)
        j.zk_re, j.zk_im = zk.real, zk.imag
        j.ns, j.sources = ns, nat.ptr(dev.src)
        j.ifcharge, j.ifdipole = int(dev.charge is not None), int(dev.dipstr is not None)
        j.charge, j.dipstr = nat.ptr(dev.charge), nat.ptr(dev.dipstr)
        j.dipvec = nat.ptr(dev.dipvec) if kind != "c" else None
        j.ifpgh, j.ifpghtarg = pg, pgt
        j.pot, j.grad, j.hess = (nat.ptr(bufs[n]) for n in ("pot", "grad", "hess"))
        j.nt, j.targ = dev.nt, nat.ptr(dev.targ)
        j.pottarg, j.gradtarg, j.hesstarg = (nat.ptr(bufs[n])
                                             for n in ("pottarg", "gradtarg", "hesstarg"))
        o = None
        if pb.get("opts"):
            o = nat.opts_struct(pb["opts"])
            j.opts = ctypes.pointer(o)
        keep.append((dev, bufs, nd, o))
    return jobs, keep


def _fmm_dispatch(kind, eps, zk, sources, charges, dipstr, dipvec, targets, pg, pgt, nd,
                  maxpg):
    """The reference's branch structure (python/fmm2dpy/fmm2d.py:98-196, 294-454,
    550-710, 802-962): evaluation at sources, at targets, or at both."""
    out = Output()
    ns = _nsrc(sources)
    if pg == 0 and pgt == 0:
        print("Nothing to compute, set either pg or pgt to non-zero")
        return out
    ifcharge = _check_charges(charges, nd, ns)
    ifdipole = _check_dipoles(dipstr, dipvec, nd, ns, need_vec=(kind != "c"))
    if targets is not None:
        assert targets.shape[0] == 2, "The first dimension of targets must be 2"
    if not (ifcharge or ifdipole):
        return out
    ch = charges if ifcharge else None
    ds = dipstr if ifdipole else None
    src_ok = 1 <= pg <= maxpg
    if src_ok and targets is not None:
        msg = ("if both potential or potential at gradient are requested at sources and "
               "targets, then the same pg must be equal to pgt") if kind == "h" else \
              ("if output is requested at both sources and targets, then the same pg must "
               "be equal to pgt")
        if not (1 <= pgt <= maxpg):
            # the reference evaluates at the sources first, then asserts
            res, ier = _fmm(kind, eps, zk, sources, ch, ds, dipvec, None, pg, 0, nd)
            _fill(out, res, ier)
        assert pg == pgt, msg
        res, ier = _fmm(kind, eps, zk, sources, ch, ds, dipvec, targets, pg, pgt, nd)
        _fill(out, res, ier)
    elif src_ok:
        res, ier = _fmm(kind, eps, zk, sources, ch, ds, dipvec, None, pg, 0, nd)
        _fill(out, res, ier)
    elif targets is not None and 1 <= pgt <= maxpg:
        res, ier = _fmm(kind, eps, zk, sources, ch, ds, dipvec, targets, 0, pgt, nd)
        _fill(out, res, ier)
    return out


def _fill(out, res, ier):
    for k, v in res.items():
        setattr(out, k, v)
    out.ier = ier


def hfmm2d(*, eps, zk, sources, charges=None, dipstr=None, dipvec=None,
           targets=None, pg=0, pgt=0, nd=1):
    """Helmholtz FMM (python/fmm2dpy/fmm2d.py:20-198; helmholtz/hfmm2d.f:19).

    u(x) = (i/4) sum_j [c_j H0(k|x-x_j|) + k d_j (v_j.(x-x_j))/|x-x_j| H1(k|x-x_j|)];
    pot and gradient (pg, pgt in {1, 2}) at sources and/or targets.
    """
    return _fmm_dispatch("h", eps, zk, sources, charges, dipstr, dipvec, targets, pg, pgt,
                         nd, maxpg=2)


def lfmm2d(*, eps, sources, charges=None, dipstr=None, dipvec=None,
           targets=None, pg=0, pgt=0, nd=1):
    """Laplace FMM with complex densities (python/fmm2dpy/fmm2d.py:459-712;
    laplace/lfmm2d.f:29).  pg, pgt in {1, 2, 3}."""
    return _fmm_dispatch("l", eps, None, sources, charges, dipstr, dipvec, targets, pg, pgt,
                         nd, maxpg=3)


def rfmm2d(*, eps, sources, charges=None, dipstr=None, dipvec=None,
           targets=None, pg=0, pgt=0, nd=1):
    """Laplace FMM with real densities and real outputs (python/fmm2dpy/fmm2d.py:201-456;
    laplace/rfmm2d.f:27).  pg, pgt in {1, 2, 3}."""
    return _fmm_dispatch("r", eps, None, sources, charges, dipstr, dipvec, targets, pg, pgt,
                         nd, maxpg=3)


def cfmm2d(*, eps, sources, charges=None, dipstr=None,
           targets=None, pg=0, pgt=0, nd=1):
    """Cauchy-kernel FMM (python/fmm2dpy/fmm2d.py:715-964; laplace/cfmm2d.f:23).

    pot = sum c_j log|z-xi_j| + d_j/(z-xi_j) evaluated with analytic (complex
    log) expansions in the far field, exactly as the reference does, so for
    complex charges only Re(pot) of real-charge parts is a log-potential;
    grad = d/dz and hess = d^2/dz^2 are single-valued.
    """
    return _fmm_dispatch("c", eps, None, sources, charges, dipstr, None, targets, pg, pgt,
                         nd, maxpg=3)


# --------------------------------------------------------------------------
# direct evaluators
def _direct(family, zk, sources, targets, charges, dipstr, dipvec, pgt, nd, thresh):
    """Run one direct evaluation; returns (pot, grad, hess) with f2py shapes
    (nd,nt), (nd,nt)|(nd,2,nt), (nd,nt)|(nd,3,nt)."""
    import torch
    ns = _nsrc(sources)
    need_vec = family != "c"
    dev = _DevInputs(sources, ns, nd, charges, dipstr, dipvec, targets, need_vec,
                     real=(family == "r"))
    nt = dev.nt
    ng = 1 if family == "c" else 2
    nh = 1 if family == "c" else 3
    odt = torch.float64 if family == "r" else torch.complex128
    pot = nat.dev_zeros_fortran((nd, nt), odt)
    gshape = (nd, nt) if ng == 1 else (nd, ng, nt)
    hshape = (nd, nt) if nh == 1 else (nd, nh, nt)
    grad = nat.dev_zeros_fortran(gshape, odt) if pgt >= 2 else None
    hess = nat.dev_zeros_fortran(hshape, odt) if pgt >= 3 else None
    L = nat.lib()
    ifc = 1 if dev.charge is not None else 0
    ifd = 1 if dev.dipstr is not None else 0
    common = (nat.ptr(dev.src), ns, ifc, nat.ptr(dev.charge), ifd, nat.ptr(dev.dipstr))
    outs = (nat.ptr(pot), nat.ptr(grad), nat.ptr(hess), float(thresh), nat.stream_ptr())
    if family == "c":
        ier = L.fmm2d_c2d_direct(nd, *common, nat.ptr(dev.targ), nt, pgt, *outs)
    elif family == "l":
        ier = L.fmm2d_l2d_direct(nd, *common, nat.ptr(dev.dipvec), nat.ptr(dev.targ), nt, pgt,
                                 *outs)
    elif family == "r":
        ier = L.fmm2d_r2d_direct(nd, *common, nat.ptr(dev.dipvec), nat.ptr(dev.targ), nt, pgt,
                                 *outs)
    else:
        zk = complex(zk)
        ier = L.fmm2d_h2d_direct(nd, zk.real, zk.imag, *common, nat.ptr(dev.dipvec),
                                 nat.ptr(dev.targ), nt, pgt, *outs)
    nat.check(ier, family + "2d_direct")
    res = [nat.fortran_to_numpy(pot),
           nat.fortran_to_numpy(grad) if grad is not None else None,
           nat.fortran_to_numpy(hess) if hess is not None else None]
    return res, nt


def h2ddir(*, zk, sources, targets, charges=None, dipstr=None, dipvec=None,
           pgt=0, nd=1, thresh=1e-16):
    """Helmholtz direct evaluation (python/fmm2dpy/fmm2d.py:1250-1367).

    u(x) = sum_j c_j (i/4) H0(k|x-x_j|) + dipole terms; pairs with
    |k|*|x-x_j| below thresh are omitted.  pgt in {1, 2}.
    """
    out = Output()
    ns = _nsrc(sources)
    if pgt == 0:
        print("Nothing to compute, set either pg or pgt to non-zero")
        return out
    ifcharge = _check_charges(charges, nd, ns)
    ifdipole = _check_dipoles(dipstr, dipvec, nd, ns)
    assert targets.shape[0] == 2, "The first dimension of targets must be 2"
    if pgt not in (1, 2) or not (ifcharge or ifdipole):
        return _reshape_ddir(out, nd, 0, targets, ndim_g=2)
    (p, g, _), nt = _direct("h", zk, sources, targets, charges if ifcharge else None,
                            dipstr if ifdipole else None, dipvec, pgt, nd, thresh)
    out.pottarg = p
    if pgt == 2:
        out.gradtarg = g
    return _reshape_ddir(out, nd, pgt, targets, ndim_g=2, grad_only_pgt2=True)


def l2ddir(*, sources, targets, charges=None, dipstr=None, dipvec=None,
           pgt=0, nd=1, thresh=1e-16):
    """Laplace direct evaluation (python/fmm2dpy/fmm2d.py:1493-1613).

    u(x) = sum_j c_j log|x-x_j| - d_j v_j.(x-x_j)/|x-x_j|^2 (the reference
    code's normalisation; no 1/(2 pi)).  pgt in {1, 2, 3}.
    """
    out = Output()
    ns = _nsrc(sources)
    if pgt == 0:
        print("Nothing to compute, set either pg or pgt to non-zero")
        return out
    ifcharge = _check_charges(charges, nd, ns)
    ifdipole = _check_dipoles(dipstr, dipvec, nd, ns)
    assert targets.shape[0] == 2, "The first dimension of targets must be 2"
    if pgt not in (1, 2, 3) or not (ifcharge or ifdipole):
        return _reshape_ddir(out, nd, 0, targets, ndim_g=2)
    (p, g, h), nt = _direct("l", None, sources, targets, charges if ifcharge else None,
                            dipstr if ifdipole else None, dipvec, pgt, nd, thresh)
    out.pottarg = p
    if pgt >= 2:
        out.gradtarg = g
    if pgt >= 3:
        out.hesstarg = h
    return _reshape_ddir(out, nd, pgt, targets, ndim_g=2)


def r2ddir(*, sources, targets, charges=None, dipstr=None, dipvec=None,
           pgt=0, nd=1, thresh=1e-16):
    """Real Laplace direct evaluation (python/fmm2dpy/fmm2d.py:1370-1490;
    laplace/rlapkernels2d.f): real densities and outputs, pgt in {1, 2, 3}."""
    out = Output()
    ns = _nsrc(sources)
    if pgt == 0:
        print("Nothing to compute, set either pg or pgt to non-zero")
        return out
    ifcharge = _check_charges(charges, nd, ns)
    ifdipole = _check_dipoles(dipstr, dipvec, nd, ns)
    assert targets.shape[0] == 2, "The first dimension of targets must be 2"
    if pgt not in (1, 2, 3) or not (ifcharge or ifdipole):
        return _reshape_ddir(out, nd, 0, targets, ndim_g=2)
    (p, g, h), nt = _direct("r", None, sources, targets, charges if ifcharge else None,
                            dipstr if ifdipole else None, dipvec, pgt, nd, thresh)
    out.pottarg = p
    if pgt >= 2:
        out.gradtarg = g
    if pgt >= 3:
        out.hesstarg = h
    return _reshape_ddir(out, nd, pgt, targets, ndim_g=2)


def c2ddir(*, sources, targets, charges=None, dipstr=None,
           pgt=0, nd=1, thresh=1e-16):
    """Cauchy direct evaluation (python/fmm2dpy/fmm2d.py:1616-1727).

    u(z) = sum_j c_j log|z-xi_j| + d_j/(z-xi_j); grad = d/dz, hess = d^2/dz^2.
    """
    out = Output()
    ns = _nsrc(sources)
    if pgt == 0:
        print("Nothing to compute, set either pg or pgt to non-zero")
        return out
    ifcharge = _check_charges(charges, nd, ns)
    ifdipole = _check_dipoles(dipstr, None, nd, ns, need_vec=False)
    assert targets.shape[0] == 2, "The first dimension of targets must be 2"
    if pgt not in (1, 2, 3) or not (ifcharge or ifdipole):
        return _reshape_ddir(out, nd, 0, targets, ndim_g=1)
    (p, g, h), nt = _direct("c", None, sources, targets, charges if ifcharge else None,
                            dipstr if ifdipole else None, None, pgt, nd, thresh)
    out.pottarg = p
    if pgt >= 2:
        out.gradtarg = g
    if pgt >= 3:
        out.hesstarg = h
    return _reshape_ddir(out, nd, pgt, targets, ndim_g=1)


def _reshape_ddir(out, nd, pgt, targets, ndim_g, grad_only_pgt2=False):
    """nd == 1 reshapes of the reference (gradtarg only when pgt == 2)."""
    if nd != 1 or pgt == 0:
        return out
    nt = targets.shape[1]
    out.pottarg = out.pottarg.reshape(nt,)
    if pgt == 2:
        out.gradtarg = out.gradtarg.reshape((nt,) if ndim_g == 1 else (2, nt))
    if pgt == 3 and not grad_only_pgt2:
        out.hesstarg = out.hesstarg.reshape((nt,) if ndim_g == 1 else (3, nt))
    return out


# --------------------------------------------------------------------------
def comperr(*, ntest, out, outex, pg=0, pgt=0, nd=1, cauchy=0, bh=0):
    """Relative L2 error over the first ntest points (python/fmm2dpy/fmm2d.py:2017-2139).

    With cauchy=1 the potentials are compared through their real parts only,
    as the reference does.
    """
    r = 0.0
    err = 0.0

    def acc(a, b):
        nonlocal r, err
        r += la.norm(b) ** 2
        err += la.norm(b - a) ** 2

    def sl(x):
        return x[..., 0:ntest]

    comps = 3 if bh == 1 else 2
    for flag, pname, gname, hname in ((pg, "pot", "grad", "hess"),
                                      (pgt, "pottarg", "gradtarg", "hesstarg")):
        if flag <= 0:
            continue
        p, pex = getattr(out, pname), getattr(outex, pname)
        if pname == "pot":
            real = cauchy != 0
        else:
            real = not (cauchy == 0 and bh == 0)
        if nd > 1:
            real = cauchy != 0
        if real:
            acc(sl(p.real).ravel(), sl(pex.real).ravel())
        else:
            acc(sl(p).ravel(), sl(pex).ravel())
        if flag >= 2:
            g, gex = getattr(out, gname), getattr(outex, gname)
            acc(sl(g).ravel(), sl(gex).ravel())
        if flag >= 3:
            h, hex_ = getattr(out, hname), getattr(outex, hname)
            acc(sl(h).ravel(), sl(hex_).ravel())
    return np.sqrt(err / r)


# --------------------------------------------------------------------------
# Helmholtz FMM on multipole sources
def hfmm2d_mps(*, eps, zk, centers, rscales, nterms, mpole, impole=None, nd=1, ndiv=None):
    """Local expansions at every centre due to all other multipole expansions
    (helmholtz/hfmm2d_mps.f:19; with ndiv, helmholtz/hfmm2d_mps_ndiv.f:70 --
    ndiv = number of expansions evaluates every pair directly).

    centers (2, n), rscales (n,), nterms (n,) orders; mpole: flat complex array
    holding expansion i's coefficients (nd, -nterms[i]..nterms[i]) in Fortran
    order (density fastest) from offset impole[i] (0-based; default: the
    expansions one after the other).  Returns local, same layout as mpole
    (entries outside the expansions' blocks are zero).  The reference has no
    Python wrapper for this routine; the arguments mirror its Fortran ones.
    """
    import torch
    cen = np.asarray(centers, dtype=np.float64)
    assert cen.shape[0] == 2, "The first dimension of centers must be 2"
    n = cen.shape[1]
    nt_ = np.asarray(nterms, dtype=np.int32).reshape(n)
    rs = np.asarray(rscales, dtype=np.float64).reshape(n)
    mp = np.asarray(mpole, dtype=np.complex128).ravel()
    if impole is None:
        impole = np.concatenate([[0], np.cumsum(nd * (2 * nt_[:-1].astype(np.int64) + 1))])
    imp = np.asarray(impole, dtype=np.int64).reshape(n)
    assert np.all(imp >= 0) and np.all(imp + nd * (2 * nt_.astype(np.int64) + 1) <= mp.size), \
        "Every expansion must lie inside mpole"
    dev = nat.torch_device()
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
         for k, v in dict(cen=cen.T, rs=rs, nt=nt_, mp=mp, imp=imp + 1).items()}
    loc = torch.zeros_like(d["mp"])
    o = nat.opts_struct({"ndiv": ndiv} if ndiv is not None else None)
    zk = complex(zk)
    L = nat.lib()
    ier = L.fmm2d_hfmm2d_mps(nd, float(eps), zk.real, zk.imag, n, nat.ptr(d["cen"]),
                             nat.ptr(d["rs"]), nat.ptr(d["nt"]), mp.size, nat.ptr(d["mp"]),
                             nat.ptr(d["imp"]), nat.ptr(loc), nat.byref_opts(o),
                             nat.stream_ptr())
    nat.check(ier, "hfmm2d_mps")
    if ier != 0:
        raise RuntimeError(f"fmm2d_amd: hfmm2d_mps returned ier = {ier}")
    return loc.cpu().numpy()


# --------------------------------------------------------------------------
# Biharmonic and Stokes sums (python/fmm2dpy/fmm2d.py:966-1248, 1730-2016)
def _bh_inputs(charges, dipoles, nd, ns):
    """The reference's checks and reshapes (python/fmm2dpy/fmm2d.py:1031-1054)."""
    ifcharge = ifdipole = 0
    if charges is not None:
        if nd == 1 and ns > 1:
            assert charges.shape[0] == 2 and charges.shape[1] == ns, \
                "Charges must be of shape [2, number of source]"
        if nd == 1 and ns == 1:
            assert charges.shape[0] == 2, "Charges must of shape [2, number of sources]"
        if nd > 1:
            assert charges.shape[0] == nd and charges.shape[1] == 2 and charges.shape[2] == ns, \
                "Charges must be of shape [nd,2,ns] where nd is number of densities, and ns is number of sources"
        ifcharge = 1
        charges = np.reshape(charges, [nd, 2, ns])
    else:
        charges = np.zeros([nd, 2, ns], dtype="complex")
    if dipoles is not None:
        if nd == 1 and ns > 1:
            assert dipoles.shape[0] == 3 and dipoles.shape[1] == ns, \
                "Dipole must of shape [3, number of sources]"
        if nd == 1 and ns == 1:
            assert dipoles.shape[0] == 3, "Dipole must of shape [3, number of sources]"
        if nd > 1:
            assert dipoles.shape[0] == nd and dipoles.shape[1] == 3 and dipoles.shape[2] == ns, \
                "Dipole must of shape [nd,2, number of sources]"
        dipoles = np.reshape(dipoles, [nd, 3, ns])
        ifdipole = 1
    else:
        dipoles = np.zeros([nd, 3, ns], dtype="complex")
    return ifcharge, charges, ifdipole, dipoles


def bhfmm2d(*, eps, sources, charges=None, dipoles=None, targets=None, pg=0, pgt=0, nd=1):
    """Biharmonic FMM (python/fmm2dpy/fmm2d.py:966-1085; biharmonic/bhfmm2d.f:1).

    vel(z) = sum_j 2 c_{j,1} log|z-z_j| + c_{j,2} (z-z_j)/conj(z-z_j) + d_{j,1}/(z-z_j)
             + d_{j,3}/conj(z-z_j) + d_{j,2} (z-z_j)/conj(z-z_j)^2;
    grad (3 components): d/dz vel projected on z, on zbar, and d/dzbar vel.
    """
    out = Output()
    ns = _nsrc(sources)
    if pg == 0 and pgt == 0:
        print("Nothing to compute, set either pg or pgt to non-zero")
        return out
    ifcharge, charges, ifdipole, dipoles = _bh_inputs(charges, dipoles, nd, ns)
    if targets is not None:
        assert targets.shape[0] == 2, "The first dimension of targets must be 2"
        nt = np.shape(targets)[1]
    else:
        targets = np.zeros([2, 0], dtype="double")
        nt = 0
    import torch
    src = nat.to_dev_fortran(np.reshape(sources, (2, ns)), np.float64)
    ch = nat.to_dev_fortran(charges, np.complex128)
    dp = nat.to_dev_fortran(dipoles, np.complex128)
    trg = nat.to_dev_fortran(targets, np.float64) if nt > 0 else None
    pot = nat.dev_zeros_fortran((nd, ns), torch.complex128)
    grad = nat.dev_zeros_fortran((nd, 3, ns), torch.complex128)
    pt = nat.dev_zeros_fortran((nd, max(nt, 1)), torch.complex128)
    gt = nat.dev_zeros_fortran((nd, 3, max(nt, 1)), torch.complex128)
    ier = nat.lib().fmm2d_bhfmm2d(nd, float(eps), ns, nat.ptr(src), ifcharge, nat.ptr(ch),
                                  ifdipole, nat.ptr(dp), 0, int(pg), nat.ptr(pot),
                                  nat.ptr(grad), None, nt, nat.ptr(trg), int(pgt) if nt else 0,
                                  nat.ptr(pt), nat.ptr(gt), None, None, nat.stream_ptr())
    nat.check(ier, "bhfmm2d")
    out.pot = nat.fortran_to_numpy(pot)
    out.grad = nat.fortran_to_numpy(grad)
    out.pottarg = nat.fortran_to_numpy(pt)[:, :nt]
    out.gradtarg = nat.fortran_to_numpy(gt)[:, :, :nt]
    out.ier = ier
    out.hess = None
    out.hesstarg = None
    if nd == 1:
        if pgt > 0:
            out.pottarg = out.pottarg.reshape(nt,)
        if pgt == 2:
            out.gradtarg = out.gradtarg.reshape(3, nt)
        if pg > 0:
            out.pot = out.pot.reshape(ns,)
        if pg == 2:
            out.grad = out.grad.reshape(3, ns)
    if pg < 2:
        out.grad = None
    if pgt < 2:
        out.gradtarg = None
    return out


def bh2ddir(*, sources, targets, charges=None, dipoles=None, pgt=0, nd=1, thresh=1e-16):
    """Biharmonic direct evaluation at targets (python/fmm2dpy/fmm2d.py:1730-1843;
    biharmonic/bhkernels2d.f:26-410)."""
    out = Output()
    ns = _nsrc(sources)
    if pgt == 0:
        print("Nothing to compute, set either pg or pgt to non-zero")
        return out
    ifcharge, charges, ifdipole, dipoles = _bh_inputs(charges, dipoles, nd, ns)
    assert targets.shape[0] == 2, "The first dimension of targets must be 2"
    nt = targets.shape[1]
    if not (ifcharge or ifdipole) or pgt not in (1, 2):
        return out
    import torch
    src = nat.to_dev_fortran(np.reshape(sources, (2, ns)), np.float64)
    ch = nat.to_dev_fortran(charges, np.complex128)
    dp = nat.to_dev_fortran(dipoles, np.complex128)
    trg = nat.to_dev_fortran(targets, np.float64)
    vel = nat.dev_zeros_fortran((nd, nt), torch.complex128)
    grad = nat.dev_zeros_fortran((nd, 3, nt), torch.complex128)
    ier = nat.lib().fmm2d_bh2d_direct(nd, nat.ptr(src), ns, ifcharge, nat.ptr(ch), ifdipole,
                                      nat.ptr(dp), nat.ptr(trg), nt, int(pgt), nat.ptr(vel),
                                      nat.ptr(grad), float(thresh), nat.stream_ptr())
    nat.check(ier, "bh2d_direct")
    out.pottarg = nat.fortran_to_numpy(vel)
    if pgt == 2:
        out.gradtarg = nat.fortran_to_numpy(grad)
    if nd == 1:
        out.pottarg = out.pottarg.reshape(nt,)
        if pgt == 2:
            out.gradtarg = out.gradtarg.reshape(3, nt,)
    return out


def _st_inputs(sources, stoklet, strslet, strsvec, targets, nd):
    """The reference's checks and reshapes (python/fmm2dpy/fmm2d.py:1186-1228)."""
    assert sources.shape[0] == 2, "The first dimension of sources must be 2"
    ns = _nsrc(sources)
    if targets is not None:
        assert targets.shape[0] == 2, "The first dimension of targets must be 2"
        nt = targets.shape[1] if np.size(np.shape(targets)) == 2 else 1
    else:
        targets = np.zeros([2, 0], dtype="double")
        nt = 0
    ifstoklet = ifstrslet = 0
    if stoklet is not None:
        if nd == 1 and ns > 1:
            assert stoklet.shape[0] == 2 and stoklet.shape[1] == ns, \
                "stoklet vectors must be of shape [2,number of sources]"
        if nd == 1 and ns == 1:
            assert stoklet.shape[0] == 2, "stoklet vectors must be of shape [2,number of sources]"
        if nd > 1:
            assert stoklet.shape[0] == nd and stoklet.shape[1] == 2 and stoklet.shape[2] == ns, \
                "stoklet vectors must be of shape [nd,2,ns] where nd is number of densities, and ns is number of sources"
        stoklet = np.reshape(stoklet, [nd, 2, ns])
        ifstoklet = 1
    else:
        stoklet = np.zeros([nd, 2, ns], dtype="double")
    if strslet is not None and strsvec is not None:
        if nd == 1 and ns > 1:
            assert strslet.shape[0] == 2 and strslet.shape[1] == ns, \
                "strslet vectors must be of shape [2,number of sources]"
            assert strsvec.shape[0] == 2 and strsvec.shape[1] == ns, \
                "strsvec vectors must be of shape [2,number of sources]"
        if nd == 1 and ns == 1:
            assert strslet.shape[0] == 2, "strslet vectors must be of shape [2,number of sources]"
            assert strsvec.shape[0] == 2, "strsvec vectors must be of shape [2,number of sources]"
        if nd > 1:
            assert strslet.shape[0] == nd and strslet.shape[1] == 2 and strslet.shape[2] == ns, \
                "strslet vectors must be of shape [nd,2,ns] where nd is number of densities, and ns is number of sources"
            assert strsvec.shape[0] == nd and strsvec.shape[1] == 2 and strsvec.shape[2] == ns, \
                "strsvec vectors must be of shape [nd,2,ns] where nd is number of densities, and ns is number of sources"
        strslet = np.reshape(strslet, [nd, 2, ns])
        strsvec = np.reshape(strsvec, [nd, 2, ns])
        ifstrslet = 1
    else:
        strslet = np.zeros([nd, 2, ns], dtype="double")
        strsvec = np.zeros([nd, 2, ns], dtype="double")
    return ns, nt, targets, ifstoklet, stoklet, ifstrslet, strslet, strsvec


def _st_nothing(ifppreg, ifppregtarg, stoklet, strslet, strsvec):
    """The reference's early returns (python/fmm2dpy/fmm2d.py:1172-1184); True when
    there is nothing to compute."""
    if ifppreg == 0 and ifppregtarg == 0:
        print("Nothing to compute, set either ifppreg or ifppregtarg to non-zero")
        return True
    if stoklet is None and strslet is None and strsvec is None:
        print("Nothing to compute, set either stoklet or strslet+strsvec to non-None")
        return True
    if (strslet is not None) != (strsvec is not None):
        print("strslet and strsvec mush be both None or both not None")
        return True
    return False


def _st_trim(out, ifppreg, ifppregtarg):
    """python/fmm2dpy/fmm2d.py:1233-1248 / 2003-2014."""
    if ifppreg < 3:
        out.grad = None
    if ifppregtarg < 3:
        out.gradtarg = None
    if ifppreg < 2:
        out.pre = None
    if ifppregtarg < 2:
        out.pretarg = None
    if ifppreg < 1:
        out.pot = None
    if ifppregtarg < 1:
        out.pottarg = None
    return out


def stfmm2d(*, eps, sources, stoklet=None, strslet=None, strsvec=None, targets=None,
            ifppreg=0, ifppregtarg=0, nd=1):
    """Stokes FMM (python/fmm2dpy/fmm2d.py:1087-1248; stokes/stfmm2d.f:27).

    u_i(x) = sum_m G_ij(x, y_m) sigma_m,j + T_ijk(x, y_m) mu_m,j nu_m,k with
    G_ij = r_i r_j/(2 r^2) - delta_ij log(r)/2, T_ijk = -2 r_i r_j r_k / r^4;
    pressure and velocity gradient on request.  Outputs keep the reference's
    Fortran shapes pot (nd,2,n), pre (nd,n), grad (nd,2,2,n).
    """
    out = Output()
    if _st_nothing(ifppreg, ifppregtarg, stoklet, strslet, strsvec):
        return out
    ns, nt, targets, ifs, stoklet, ift, strslet, strsvec = _st_inputs(
        sources, stoklet, strslet, strsvec, targets, nd)
    import torch
    f64 = torch.float64
    src = nat.to_dev_fortran(np.reshape(sources, (2, ns)), np.float64)
    sk = nat.to_dev_fortran(stoklet, np.float64)
    sl = nat.to_dev_fortran(strslet, np.float64)
    sv = nat.to_dev_fortran(strsvec, np.float64)
    trg = nat.to_dev_fortran(targets, np.float64) if nt > 0 else None
    n1 = max(nt, 1)
    pot, pre = nat.dev_zeros_fortran((nd, 2, ns), f64), nat.dev_zeros_fortran((nd, ns), f64)
    grad = nat.dev_zeros_fortran((nd, 2, 2, ns), f64)
    pt, prt = nat.dev_zeros_fortran((nd, 2, n1), f64), nat.dev_zeros_fortran((nd, n1), f64)
    gt = nat.dev_zeros_fortran((nd, 2, 2, n1), f64)
    ier = nat.lib().fmm2d_stfmm2d(nd, float(eps), ns, nat.ptr(src), ifs, nat.ptr(sk), ift,
                                  nat.ptr(sl), nat.ptr(sv), int(ifppreg), nat.ptr(pot),
                                  nat.ptr(pre), nat.ptr(grad), nt, nat.ptr(trg),
                                  int(ifppregtarg) if nt else 0, nat.ptr(pt), nat.ptr(prt),
                                  nat.ptr(gt), None, nat.stream_ptr())
    nat.check(ier, "stfmm2d")
    out.pot, out.pre, out.grad = (nat.fortran_to_numpy(pot), nat.fortran_to_numpy(pre),
                                  nat.fortran_to_numpy(grad))
    out.pottarg = nat.fortran_to_numpy(pt)[..., :nt]
    out.pretarg = nat.fortran_to_numpy(prt)[..., :nt]
    out.gradtarg = nat.fortran_to_numpy(gt)[..., :nt]
    out.ier = ier
    return _st_trim(out, ifppreg, ifppregtarg)


def st2ddir(*, eps, sources, stoklet=None, strslet=None, strsvec=None, targets=None,
            ifppreg=0, ifppregtarg=0, nd=1, thresh=1e-16):
    """Stokes direct evaluation (python/fmm2dpy/fmm2d.py:1845-2015;
    stokes/stokkernels2d.f:24, 169) at the sources (the self term skipped by
    thresh) and at the targets.

    Two deliberate differences from the reference wrapper: the "nothing to
    compute" test uses ``is None`` (the reference's ``stoklet == None`` raises
    for an array argument), and the evaluation at the sources counts the
    sources (the reference passes the number of targets there, so it returns
    empty source outputs when targets is None).
    """
    out = Output()
    if _st_nothing(ifppreg, ifppregtarg, stoklet, strslet, strsvec):
        return out
    ns, nt, targets, ifs, stoklet, ift, strslet, strsvec = _st_inputs(
        sources, stoklet, strslet, strsvec, targets, nd)
    import torch
    f64 = torch.float64
    L = nat.lib()
    src = nat.to_dev_fortran(np.reshape(sources, (2, ns)), np.float64)
    sk = nat.to_dev_fortran(stoklet, np.float64)
    sl = nat.to_dev_fortran(strslet, np.float64)
    sv = nat.to_dev_fortran(strsvec, np.float64)

    def run(trg_np, n):
        pot, pre = nat.dev_zeros_fortran((nd, 2, n), f64), nat.dev_zeros_fortran((nd, n), f64)
        grad = nat.dev_zeros_fortran((nd, 2, 2, n), f64)
        if n > 0:
            trg = nat.to_dev_fortran(trg_np, np.float64)
            ier = L.fmm2d_st2d_direct(nd, nat.ptr(src), ifs, nat.ptr(sk), 1 if ift else 0,
                                      nat.ptr(sl), nat.ptr(sv), ns, nat.ptr(trg), n,
                                      nat.ptr(pot), nat.ptr(pre), nat.ptr(grad), float(thresh),
                                      nat.stream_ptr())
            nat.check(ier, "st2d_direct")
        return (nat.fortran_to_numpy(pot), nat.fortran_to_numpy(pre),
                nat.fortran_to_numpy(grad))

    out.pot, out.pre, out.grad = run(np.reshape(sources, (2, ns)), ns)
    out.pottarg, out.pretarg, out.gradtarg = run(targets, nt)
    return _st_trim(out, ifppreg, ifppregtarg)
