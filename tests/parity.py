"""Comparison helpers for the parity tests (oracle vs HIP library)."""
import numpy as np

# region selectors in the reference layout (nj+4, ni+4, K): a[j+1, i+1]
def interior(a, ni, nj):
    return a[2:nj + 2, 2:ni + 2]


def iface(a, ni, nj):      # i in 0..ni, j in 1..nj
    return a[2:nj + 2, 1:ni + 2]


def jface(a, ni, nj):      # i in 1..ni, j in 0..nj
    return a[1:nj + 2, 2:ni + 2]


def ghosted(a, ni, nj):
    return a


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    if scale == 0.0:
        scale = 1.0
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


def assert_close(dev, ref, tol, what=""):
    assert dev.shape == ref.shape, (what, dev.shape, ref.shape)
    assert np.all(np.isfinite(dev) == np.isfinite(ref)), f"{what}: finiteness differs"
    m = np.isfinite(ref)
    e = relerr(dev[m], ref[m])
    assert e <= tol, f"{what}: max rel err {e:.3e} > {tol:.1e}"
    return e


def l2rel(dev, ref):
    num = np.sqrt(np.sum((np.asarray(dev) - np.asarray(ref)) ** 2))
    den = np.sqrt(np.sum(np.asarray(ref) ** 2))
    return float(num / den) if den > 0 else float(num)
