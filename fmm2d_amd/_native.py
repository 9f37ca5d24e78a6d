"""ctypes binding to libfmm2d_amd.so (the C-ABI declared in include/fmm2d_amd.h).

The library must have been built (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C fmm2d_amd/csrc``).  There is no CPU fallback: every
entry point raises if the library or a ROCm device is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libfmm2d_amd.so")

_lib = None

# (name, restype, argtypes) for every exported symbol of include/fmm2d_amd.h
_vp, _i, _i64, _d = C.c_void_p, C.c_int, C.c_int64, C.c_double
SIGNATURES = {
    "fmm2d_abi_version": (_i, []),
    "fmm2d_c2d_direct": (_i, [_i, _vp, _i64, _i, _vp, _i, _vp, _vp, _i64, _i, _vp, _vp, _vp,
                              _d, _vp]),
    "fmm2d_l2d_direct": (_i, [_i, _vp, _i64, _i, _vp, _i, _vp, _vp, _vp, _i64, _i, _vp, _vp,
                              _vp, _d, _vp]),
    "fmm2d_h2d_direct": (_i, [_i, _d, _d, _vp, _i64, _i, _vp, _i, _vp, _vp, _vp, _i64, _i, _vp,
                              _vp, _vp, _d, _vp]),
    "fmm2d_hank103": (_i, [_vp, _i64, _vp, _vp, _vp]),
}
SIGNATURES["fmm2d_l2d_direct_noalloc"] = SIGNATURES["fmm2d_l2d_direct"]
SIGNATURES["fmm2d_tree_build"] = (_i, [_vp, _i64, _vp, _i64, _i, _i, _i, C.POINTER(_i),
                                         C.POINTER(_i), C.POINTER(_vp), _vp])
SIGNATURES["fmm2d_tree_get"] = (_i, [_vp] * 22)
SIGNATURES["fmm2d_tree_free"] = (None, [_vp])
SIGNATURES["fmm2d_tree_leaf_histogram"] = (_i, [_vp, _i64, _vp, _i64, _i, _i, _i, _vp, _vp,
                                                _vp])
SIGNATURES["fmm2d_key_leaf_histogram"] = (_i, [_vp, _i64, _vp, _i64, _i, _i, _vp, _vp, _vp])
SIGNATURES["fmm2d_prepare"] = (_i, [_vp, _i64, _vp, _i64, C.POINTER(_vp), _vp])
SIGNATURES["fmm2d_prepare_free"] = (None, [_vp])
SIGNATURES["fmm2d_prepared_leaf_histogram"] = (_i, [_vp, _i, _i, _vp, _vp, _vp])
SIGNATURES["fmm2d_l2dterms"] = (_i, [_d, C.POINTER(_i)])
SIGNATURES["fmm2d_h2dterms"] = (_i, [_d, _d, _d, _d, C.POINTER(_i)])
SIGNATURES["fmm2d_ndiv"] = (_i, [_d, _i64, _i64, _i, C.POINTER(_i), C.POINTER(_i)])
SIGNATURES["fmm2d_release_workspace"] = (_i, [])
SIGNATURES["fmm2d_set_workspace_limit"] = (_i, [_i64])
SIGNATURES["fmm2d_set_direct_max_blocks"] = (_i, [_i])
SIGNATURES["fmm2d_rfmm2d"] = (_i, [_i, _d, _i64, _vp, _i, _vp, _i, _vp, _vp, _i, _i, _vp, _vp,
                                   _vp, _i64, _vp, _i, _vp, _vp, _vp, _vp, _vp])
SIGNATURES["fmm2d_r2d_direct"] = (_i, [_i, _vp, _i64, _i, _vp, _i, _vp, _vp, _vp, _i64, _i,
                                       _vp, _vp, _vp, _d, _vp])
_f = C.c_float
SIGNATURES["fmm2d_mlp_nparams"] = (_i64, [_i, _i])
SIGNATURES["fmm2d_mlp_forward"] = (_i, [_i, _i, _vp, _i, _vp, _vp, _vp])
SIGNATURES["fmm2d_mlp_loss_grad"] = (_i, [_i, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp])
SIGNATURES["fmm2d_mlp_adam"] = (_i, [_i64, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _i, _vp])
# FMM drivers: (nd, eps, [zk_re, zk_im,] ns, sources, ifcharge, charge, ifdipole,
# dipstr, [dipvec,] iper, ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg,
# gradtarg, hesstarg, opts, stream)
_FMM_TAIL = [_i64, _vp, _i, _vp, _i, _vp]
_FMM_OUT = [_i, _i, _vp, _vp, _vp, _i64, _vp, _i, _vp, _vp, _vp, _vp, _vp]
SIGNATURES["fmm2d_cfmm2d"] = (_i, [_i, _d] + _FMM_TAIL + _FMM_OUT)
SIGNATURES["fmm2d_lfmm2d"] = (_i, [_i, _d] + _FMM_TAIL + [_vp] + _FMM_OUT)
SIGNATURES["fmm2d_hfmm2d"] = (_i, [_i, _d, _d, _d] + _FMM_TAIL + [_vp] + _FMM_OUT)
SIGNATURES["fmm2d_batch"] = (_i, [_i, _vp, _i])
SIGNATURES["fmm2d_bhfmm2d"] = (_i, [_i, _d] + _FMM_TAIL + _FMM_OUT)
SIGNATURES["fmm2d_bh2d_direct"] = (_i, [_i, _vp, _i64, _i, _vp, _i, _vp, _vp, _i64, _i, _vp,
                                        _vp, _d, _vp])
SIGNATURES["fmm2d_stfmm2d"] = (_i, [_i, _d, _i64, _vp, _i, _vp, _i, _vp, _vp, _i, _vp, _vp,
                                    _vp, _i64, _vp, _i, _vp, _vp, _vp, _vp, _vp])
SIGNATURES["fmm2d_st2d_direct"] = (_i, [_i, _vp, _i, _vp, _i, _vp, _vp, _i64, _vp, _i64, _vp,
                                        _vp, _vp, _d, _vp])
SIGNATURES["fmm2d_hfmm2d_mps"] = (_i, [_i, _d, _d, _d, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                       _vp, _vp])


class FmmOpts(C.Structure):
    """struct fmm2d_opts of include/fmm2d_amd.h."""
    _fields_ = [("ndiv", C.c_int32), ("idivflag", C.c_int32), ("ifnear", C.c_int32),
                ("precision", C.c_int32), ("n_nterms", C.c_int32), ("count_work", C.c_int32),
                ("nterms", C.POINTER(C.c_int32)), ("timeinfo", C.c_double * 16),
                ("nlevels", C.c_int32), ("nboxes", C.c_int32), ("pad1", C.c_int32),
                ("pad2", C.c_int32), ("work", C.c_double * 16), ("near_pairs", C.c_int64),
                ("m2l_pairs", C.c_int64), ("prepared", C.c_void_p)]

    def timeinfo_array(self):
        return np.array(self.timeinfo[:], dtype=np.float64)


class FmmJob(C.Structure):
    """struct fmm2d_job of include/fmm2d_amd.h (one FMM of fmm2d_batch)."""
    _fields_ = [("kind", C.c_int32), ("nd", C.c_int32), ("eps", C.c_double),
                ("zk_re", C.c_double), ("zk_im", C.c_double), ("ns", C.c_int64),
                ("sources", C.c_void_p), ("ifcharge", C.c_int32), ("ifdipole", C.c_int32),
                ("charge", C.c_void_p), ("dipstr", C.c_void_p), ("dipvec", C.c_void_p),
                ("ifpgh", C.c_int32), ("ifpghtarg", C.c_int32), ("pot", C.c_void_p),
                ("grad", C.c_void_p), ("hess", C.c_void_p), ("nt", C.c_int64),
                ("targ", C.c_void_p), ("pottarg", C.c_void_p), ("gradtarg", C.c_void_p),
                ("hesstarg", C.c_void_p), ("opts", C.POINTER(FmmOpts)), ("ier", C.c_int32),
                ("pad", C.c_int32)]


def opts_struct(opts=None):
    """Build an FmmOpts from a dict: ndiv, idivflag, ifnear, precision ("fp64" or
    "mixed"), nterms (per-level list), count_work (0/1), prepared (an
    fmm2d_prepare handle).  Missing keys take the reference defaults."""
    o = FmmOpts()
    opts = opts or {}
    o.ndiv = int(opts.get("ndiv", 0))
    o.idivflag = int(opts.get("idivflag", 0))
    o.ifnear = int(opts.get("ifnear", 1))
    prec = opts.get("precision", "fp64")
    o.precision = {"fp64": 0, "mixed": 1}[prec] if isinstance(prec, str) else int(prec)
    o.count_work = int(opts.get("count_work", 0))
    nt = opts.get("nterms")
    if nt is not None:
        arr = (C.c_int32 * len(nt))(*[int(v) for v in nt])
        o._nterms_keep = arr
        o.nterms = C.cast(arr, C.POINTER(C.c_int32))
        o.n_nterms = len(nt)
    prep = opts.get("prepared")
    if prep is not None:
        o.prepared = getattr(prep, "handle", prep)
    return o


def byref_opts(o):
    return C.byref(o)


class NativeLibraryError(RuntimeError):
    pass


def lib():
    """Load libfmm2d_amd.so once; raise NativeLibraryError if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise NativeLibraryError(
                f"{LIBPATH} is missing: build it with `make -C fmm2d_amd/csrc` "
                "(there is no CPU fallback)")
        # torch bundles its own HIP runtime under the same soname as ROCm's
        # (libamdhip64.so.7); whichever is loaded first serves the whole
        # process.  Importing torch first makes this library bind to torch's
        # runtime, the one that owns the device buffers handed to it -- loaded
        # the other way round, the library's HIP calls find no device.
        import torch  # noqa: F401
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue  # reported by the symbol-export test; calling it raises
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def torch_device():
    """The ROCm device used for all work; raises when no GPU is present."""
    import torch
    if not torch.cuda.is_available():
        raise NativeLibraryError("fmm2d_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def check(ier, what):
    if ier < 0:
        raise RuntimeError(f"fmm2d_amd: {what} failed with code {ier}")
    return ier


# --------------------------------------------------------------- layouts
def to_dev_fortran(a, dtype):
    """Device tensor holding ``a`` in Fortran (column-major) memory order.

    A Fortran array of shape (n1,...,nk) has the memory of a C array of
    shape (nk,...,n1); the returned tensor has that reversed shape.
    """
    import torch
    arr = np.ascontiguousarray(np.transpose(np.asarray(a, dtype=dtype)))
    return torch.from_numpy(arr).to(torch_device())


def dev_zeros_fortran(shape, dtype):
    import torch
    return torch.zeros(tuple(reversed(shape)), dtype=dtype, device=torch_device())


def fortran_to_numpy(t):
    """Inverse of dev_zeros_fortran: a numpy array with the Fortran shape
    (Fortran-ordered strides, as f2py returns)."""
    return np.transpose(t.cpu().numpy())
