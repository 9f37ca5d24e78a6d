"""The device quadtree and interaction lists, exported in the reference's
conventions (common/pts_tree2d.f, common/tree_routs2d.f): 1-based box and
point indices, -1 for "none", Fortran-shaped arrays."""
import ctypes as C

import numpy as np

from . import _native as nat


def build_tree(sources, targets=None, ndiv=15, idivflag=0, nlmax=51, lists=False):
    """pts_tree_mem + pts_tree_build + pts_tree_sort (+ computelists).

    sources (2,ns), targets (2,nt) or None.  Returns a dict with nlevels,
    nboxes, laddr (2,nlevels+1), ilevel, iparent, nchild, ichild (4,nb),
    ncoll, coll (9,nb), centers (2,nb), boxsize, isrc, isrcse (2,nb), itarg,
    itargse and, with lists=True, nlist1..4 / list1..4 (rows sorted).
    """
    import torch
    L = nat.lib()
    src = nat.to_dev_fortran(np.asarray(sources, dtype=np.float64), np.float64)
    ns = src.shape[0]
    if targets is None:
        trg, nt = None, 0
    else:
        trg = nat.to_dev_fortran(np.asarray(targets, dtype=np.float64), np.float64)
        nt = trg.shape[0]
    nl, nb = C.c_int(0), C.c_int(0)
    h = C.c_void_p()
    ier = L.fmm2d_tree_build(nat.ptr(src), ns, nat.ptr(trg), nt, idivflag, ndiv, nlmax,
                             C.byref(nl), C.byref(nb), C.byref(h), nat.stream_ptr())
    nat.check(ier, "fmm2d_tree_build")
    nl, nb = nl.value, nb.value
    try:
        r = dict(nlevels=nl, nboxes=nb)
        shapes = dict(laddr=((nl + 1, 2), np.int32), ilevel=((nb,), np.int32),
                      iparent=((nb,), np.int32), nchild=((nb,), np.int32),
                      ichild=((nb, 4), np.int32), ncoll=((nb,), np.int32),
                      coll=((nb, 9), np.int32), centers=((nb, 2), np.float64),
                      boxsize=((nl + 1,), np.float64), isrc=((max(ns, 1),), np.int32),
                      isrcse=((nb, 2), np.int32), itarg=((max(nt, 1),), np.int32),
                      itargse=((nb, 2), np.int32))
        if lists:
            for k, m in zip(range(1, 5), (13, 27, 20, 5)):
                shapes[f"nlist{k}"] = ((nb,), np.int32)
                shapes[f"list{k}"] = ((nb, m), np.int32)
        bufs = {k: np.zeros(s, dtype=t) for k, (s, t) in shapes.items()}
        order = ["laddr", "ilevel", "iparent", "nchild", "ichild", "ncoll", "coll", "centers",
                 "boxsize", "isrc", "isrcse", "itarg", "itargse", "nlist1", "list1", "nlist2",
                 "list2", "nlist3", "list3", "nlist4", "list4"]
        args = [bufs[k].ctypes.data_as(C.c_void_p) if k in bufs else None for k in order]
        ier = L.fmm2d_tree_get(h, *args)
        nat.check(ier, "fmm2d_tree_get")
    finally:
        L.fmm2d_tree_free(h)
    for k, v in bufs.items():
        r[k] = v.T.copy() if v.ndim == 2 else v
    r["isrc"] = r["isrc"][:ns]
    r["itarg"] = r["itarg"][:nt]
    return r
