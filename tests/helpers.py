"""Shared helpers of the test suite (model builders, fixture loading, error
metrics)."""

import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def golden(kind):
    """All fixtures of one kind ("elastic", "thermal", "prep") as (dims, data)."""
    out = []
    for f in sorted(glob.glob(os.path.join(GOLDEN, f"ref_{kind}_*.npz"))):
        d = dict(np.load(f, allow_pickle=False))
        out.append((tuple(int(v) for v in d["dims"]), d))
    return out


def entry_rel(a, b, floor=0.0):
    """max |a - b| / max(|b|, floor) over all entries."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), floor)
    ok = den > 0
    return float(np.max(np.abs(a - b)[ok] / den[ok])) if ok.any() else float(np.max(np.abs(a - b)))


def rel_norm(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def csr_dense(vals, cols, rng, n):
    A = np.zeros((n, n))
    for i in range(n):
        for s in range(rng[i], rng[i + 1]):
            if cols[s]:
                A[i, cols[s] - 1] += vals[s]
    return A


def csr_sparse(vals, cols, rng, n, m=None):
    import scipy.sparse as sp

    m = n if m is None else m
    rows = np.repeat(np.arange(len(rng) - 1), np.diff(rng))
    msk = np.asarray(cols) > 0
    return sp.csr_matrix((np.asarray(vals)[msk], (rows[msk], np.asarray(cols)[msk] - 1)), shape=(len(rng) - 1, m))


def profile_dense(lt, rng, n):
    A = np.zeros((n, n))
    for i in range(n):
        ln = rng[i + 1] - rng[i]
        A[i, i - ln + 1:i + 1] = lt[rng[i]:rng[i + 1]]
    return A


def refined_solve(A, b, steps=6):
    """Solution of the sparse system A u = b by a sparse LU plus iterative
    refinement whose residuals are formed in extended precision
    (numpy.longdouble, 64-bit mantissa on x86): the result is the FP64 system's
    solution to well below FP64 solver accuracy even where cond(A) eps is
    1e-9."""
    import scipy.sparse.linalg as spl

    Ac = A.tocoo()
    rows, cols, vals = Ac.row, Ac.col, Ac.data.astype(np.longdouble)
    lu = spl.splu(A.tocsc())
    u = lu.solve(np.asarray(b, np.float64))
    bl = np.asarray(b, np.longdouble)
    for _ in range(steps):
        Au = np.zeros(len(b), np.longdouble)
        np.add.at(Au, rows, vals * u.astype(np.longdouble)[cols])
        u = u + lu.solve(np.asarray(bl - Au, np.float64))
    return u


class DeviceBuffer:
    """Device memory for tests of entry points that take device pointers,
    allocated through the HIP runtime the library itself links
    (libamdhip64.so.7), so that one runtime serves the whole process."""

    _hip = None

    @classmethod
    def hip(cls):
        import ctypes as C

        if cls._hip is None:
            from misc_amd import astro

            astro.load_library()  # loads libamdhip64.so.7 as the library's dependency
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        import ctypes as C

        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.nbytes = self.host.nbytes
        p = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(p), max(self.nbytes, 8)) == 0
        self.ptr = p.value
        assert self.hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
