"""Write the golden fixtures under tests/golden/ from the reference library.

Runs in the build container only: it needs oracle/_ref/libfmm2d_ref.so,
compiled from the reference Fortran sources by oracle/ref.mk, and mpmath.
Each fixture holds inputs and the reference's outputs; nothing of the
reference's source is stored.

  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import ref_binding as R  # noqa: E402


def direct_fixtures():
    """Direct kernels c2d/l2d/h2d for all 27 variants each, nd in {1, 3}."""
    rng = np.random.default_rng(11)
    out = {}
    ns, nt = 41, 29
    for nd in (1, 3):
        src = rng.random((2, ns))
        trg = rng.random((2, nt))
        trg[:, 0] = src[:, 3]          # coincident pair: excluded by thresh
        trg[:, 1] = src[:, 5] + 1e-17  # within thresh of a source
        ch = rng.standard_normal((nd, ns)) + 1j * rng.standard_normal((nd, ns))
        ds = rng.standard_normal((nd, ns)) + 1j * rng.standard_normal((nd, ns))
        dv = rng.standard_normal((nd, 2, ns))
        pre = f"nd{nd}_"
        out[pre + "src"] = src
        out[pre + "trg"] = trg
        out[pre + "charge"] = ch
        out[pre + "dipstr"] = ds
        out[pre + "dipvec"] = dv
        for kind in "clh":
            zk = 7.5 + 0.25j if kind == "h" else None
            out[pre + "zk"] = np.array(7.5 + 0.25j)
            for ifc, ifd in ((1, 0), (0, 1), (1, 1)):
                for pgt in (1, 2, 3):
                    r = R.direct(kind, nd, src, trg, charges=ch if ifc else None,
                                 dipstr=ds if ifd else None, dipvec=dv if ifd else None,
                                 pgt=pgt, thresh=1e-14, zk=zk)
                    tag = f"{pre}{kind}_{'c' if ifc else ''}{'d' if ifd else ''}_{pgt}_"
                    for k, v in r.items():
                        out[tag + k] = v
    np.savez_compressed(os.path.join(HERE, "direct_ref.npz"), **out)


def hankel_fixture():
    """hank103 and mpmath values of H0, H1 on a complex grid."""
    import mpmath as mp
    mp.mp.dps = 40
    zs = list(np.logspace(-8, 4, 97))
    for r in (0.3, 1.0, 2.49, 2.51, 3.7, 8.0, 16.9, 17.1, 20.0, 55.0):
        for th in np.linspace(0.0, np.pi, 9):
            zs.append(r * np.exp(1j * th))
    for r in (0.5, 2.0, 5.0, 30.0):
        for th in (-0.05, -0.2):
            zs.append(r * np.exp(1j * th))
    z = np.array(zs, dtype=np.complex128)
    h0r, h1r = R.hank103(z)

    def hank(nu, v):
        # mpmath forms H = J + iY, which cancels by ~exp(2 Im z): raise the
        # working precision accordingly
        with mp.workdps(40 + int(abs(v.imag) / 1.1)):
            return complex(mp.hankel1(nu, mp.mpc(v)))

    h0m = np.array([hank(0, v) for v in z])
    h1m = np.array([hank(1, v) for v in z])
    np.savez_compressed(os.path.join(HERE, "hankel.npz"), z=z, h0_ref=h0r, h1_ref=h1r,
                        h0_mp=h0m, h1_mp=h1m)


if __name__ == "__main__":
    direct_fixtures()
    hankel_fixture()
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))
