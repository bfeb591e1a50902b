"""Shared test configuration.

Markers:
  gpu  — needs an MI355X (runs the HIP library through its C ABI).
Everything unmarked runs on the CPU: the oracle against the reference's own
known-answer drivers and mathematical identities, host-side logic, and the
C-ABI surface of the built library (no device calls).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an AMD Instinct GPU (MI355X)")


U_D = 2.0 ** -53
U_S = 2.0 ** -24


def unit_roundoff(dtype) -> float:
    return U_D if np.dtype(dtype) == np.float64 else U_S


class Mat:
    """A strided matrix view into a 1-D buffer, mirrored on the device.

    Element (i, j) lives at buf[off + i*rs + j*cs] (positive strides). The
    same layout is used for the device copy, so the device pointer of the
    view is base + off*itemsize.
    """

    def __init__(self, m, n, rs=None, cs=None, dtype=np.float64, seed=0, off=0,
                 lo=-1.0, hi=1.0, fill=None):
        if rs is None and cs is None:
            rs, cs = 1, max(m, 1)
        self.m, self.n, self.rs, self.cs, self.off = m, n, rs, cs, off
        self.dtype = np.dtype(dtype)
        size = off + (max(m, 1) - 1) * rs + (max(n, 1) - 1) * cs + 1
        rng = np.random.default_rng(seed)
        self.buf = rng.uniform(lo, hi, size).astype(self.dtype)
        if fill is not None:
            self.buf[:] = fill
        self.dbuf = None

    def view(self, buf=None):
        b = self.buf if buf is None else buf
        es = self.dtype.itemsize
        return np.lib.stride_tricks.as_strided(b[self.off:], shape=(self.m, self.n),
                                               strides=(self.rs * es, self.cs * es))

    def set(self, arr):
        self.view()[...] = arr

    def to_device(self):
        import torch
        self.dbuf = torch.from_numpy(self.buf.copy()).to("cuda")
        return self

    @property
    def ptr(self) -> int:
        return self.dbuf.data_ptr() + self.off * self.dtype.itemsize

    @property
    def hptr(self) -> int:
        return self.buf.ctypes.data + self.off * self.dtype.itemsize

    def device_view(self):
        import torch
        torch.cuda.synchronize()
        return self.view(self.dbuf.cpu().numpy())


def normwise_gemm_err(c_hat, c_ref, a, b, c0, alpha, beta):
    num = np.linalg.norm((c_hat.astype(np.float64) - c_ref.astype(np.float64)).ravel())
    den = abs(alpha) * np.linalg.norm(a.astype(np.float64)) * np.linalg.norm(b.astype(np.float64)) \
        + abs(beta) * np.linalg.norm(c0.astype(np.float64))
    return num / den if den > 0 else num


@pytest.fixture(scope="session")
def blis():
    import blis_amd
    return blis_amd.lib()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


def say(msg: str):
    """Name the case before a GPU step (flushed, so it survives a crash)."""
    print(msg, flush=True)
