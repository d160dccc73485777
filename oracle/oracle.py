"""ctypes binding of the CPU oracle (oracle/build/libtdbem_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry and bench.py's
cpu_baseline leg as the checker; never by the product library.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libtdbem_oracle.so"

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, _dp, C.c_int, _ip, _ip, C.c_int, _dp, C.c_int, _ip, C.c_double, C.c_double]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_reference_dt.restype = C.c_double
        L.orc_reference_dt.argtypes = [_dp, _ip, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_element_info.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.orc_self_tables.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.orc_dynamat.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_dofmap.argtypes = [C.c_void_p, _lp, _lp, _dp, _ip, _lp, _dp, _ip, _lp, _dp, _lp, _dp, _lp]
        L.orc_kernel.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp]
        L.orc_regular_table.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_needcal.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_quadrature.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, _dp]
        L.orc_assemble.argtypes = [C.c_void_p, C.c_int]
        L.orc_ccsr.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_long),
                               C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_history_rows.argtypes = [C.c_void_p, C.c_int, C.c_long, C.c_long, C.c_int, C.POINTER(C.c_long),
                                       C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_precond_info.argtypes = [C.c_void_p, C.POINTER(C.c_long), C.c_void_p, C.c_void_p]
        L.orc_solve.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_long, C.c_double, C.c_long, _dp, _dp, _ip, _ip]
        L.orc_gmres_ccsr.argtypes = [C.c_long, _ip, _ip, _dp, _dp, C.c_long, C.c_double, _dp, _ip]
        _lib = L
    return _lib


def reference_dt(points, ele, beta, E, v, rho) -> float:
    return lib().orc_reference_dt(np.ascontiguousarray(points, np.float64), np.ascontiguousarray(ele, np.int32),
                                  len(ele), beta, E, v, rho)


class Case:
    """One multi-domain case held by the oracle."""

    def __init__(self, points, ele, bcid, mats, itf, dt, gap):
        self.points = np.ascontiguousarray(points, np.float64)
        self.ele = np.ascontiguousarray(ele, np.int32)
        self.bcid = np.ascontiguousarray(bcid, np.int32)
        self.mats = np.ascontiguousarray(mats, np.float64)
        self.itf = np.ascontiguousarray(np.asarray(itf, np.int32).reshape(-1, 5))
        self.ne = len(self.ele)
        h = lib().orc_create(len(self.points), self.points.ravel(), self.ne, self.ele.ravel(), self.bcid,
                             len(self.mats), self.mats.ravel(), len(self.itf), self.itf.ravel(), dt, gap)
        if not h:
            raise RuntimeError("orc_create failed")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def element_info(self, e):
        c, r, f = np.zeros(3), np.zeros(1), np.zeros(9)
        lib().orc_element_info(self.h, e, c, r, f)
        return c, r[0], f

    def self_tables(self, e):
        u, t1, t2 = np.zeros(9), np.zeros(9), np.zeros(9)
        lib().orc_self_tables(self.h, e, u, t1, t2)
        return u, t1, t2

    def dynamat(self, d):
        out = np.zeros(24)
        lib().orc_dynamat(self.h, d, out)
        return out

    def dofmap(self):
        n = self.ne
        counts = np.zeros(9, np.int64)
        arr = dict(u_block=np.zeros(n, np.int64), u_sign=np.zeros(n), u_known=np.zeros(n, np.int32),
                   t_block=np.zeros(n, np.int64), t_sign=np.zeros(n), t_known=np.zeros(n, np.int32),
                   hu_block=np.zeros(n, np.int64), hu_sign=np.zeros(n), ht_block=np.zeros(n, np.int64),
                   ht_sign=np.zeros(n), preferred=np.zeros(n, np.int64))
        rc = lib().orc_dofmap(self.h, counts, arr["u_block"], arr["u_sign"], arr["u_known"], arr["t_block"],
                              arr["t_sign"], arr["t_known"], arr["hu_block"], arr["hu_sign"], arr["ht_block"],
                              arr["ht_sign"], arr["preferred"])
        arr["counts"] = counts
        arr["ok"] = rc == 0
        return arr

    def kernel(self, d, which, n, ri, nor):
        out = np.zeros(9)
        f = lib().orc_kernel(self.h, d, which, float(n), np.ascontiguousarray(ri, np.float64),
                             np.ascontiguousarray(nor, np.float64), out)
        return f, out

    def regular_table(self):
        pos, ra, nrgv0 = np.zeros(32), np.zeros(16), np.zeros(16)
        lib().orc_regular_table(self.h, pos, ra, nrgv0)
        return pos.reshape(16, 2), ra, nrgv0

    def needcal(self, field, source, typ, n):
        return lib().orc_needcal(self.h, field, source, typ, float(n))

    def quadrature(self, field, source, which, n):
        out = np.zeros(9)
        f = lib().orc_quadrature(self.h, field, source, which, float(n), out)
        return f, out

    def assemble(self, maxn):
        rc = lib().orc_assemble(self.h, maxn)
        if rc < 0:
            raise RuntimeError(f"orc_assemble failed {rc}")

    def ccsr(self, which, step=0):
        nr, nnz = C.c_long(), C.c_long()
        rc = lib().orc_ccsr(self.h, which, step, C.byref(nr), C.byref(nnz), None, None, None)
        if rc < 0:
            raise RuntimeError(f"orc_ccsr failed {rc}")
        rp = np.zeros(nr.value + 1, np.int32)
        col = np.zeros(max(nnz.value, 1), np.int32)
        val = np.zeros(max(nnz.value, 1) * 9)
        lib().orc_ccsr(self.h, which, step, C.byref(nr), C.byref(nnz), rp.ctypes.data, col.ctypes.data,
                       val.ctypes.data)
        return rp, col[:nnz.value], val[:9 * nnz.value].reshape(-1, 3, 3)

    def history_rows(self, which, k, row0, row1):
        """MTS[k] (which 2) or MGS[k] (which 3) restricted to source rows
        row0 .. row1-1 without assembling the rest: (row_ptr, col, blocks)."""
        nnz = C.c_long()
        if lib().orc_history_rows(self.h, k, row0, row1, which, C.byref(nnz), None, None, None) < 0:
            raise RuntimeError("orc_history_rows failed")
        rp = np.zeros(row1 - row0 + 1, np.int32)
        col = np.zeros(max(nnz.value, 1), np.int32)
        val = np.zeros(max(nnz.value, 1) * 9)
        lib().orc_history_rows(self.h, k, row0, row1, which, C.byref(nnz), rp.ctypes.data, col.ctypes.data,
                               val.ctypes.data)
        return rp, col[:nnz.value], val[:9 * nnz.value].reshape(-1, 3, 3)

    def precond_info(self):
        cnt = C.c_long()
        lib().orc_precond_info(self.h, C.byref(cnt), None, None)
        sizes = np.zeros(cnt.value, np.int64)
        rep = np.zeros(self.ne, np.int64)
        lib().orc_precond_info(self.h, C.byref(cnt), sizes.ctypes.data, rep.ctypes.data)
        return sizes, rep

    def solve(self, known_lu, nstep, maxn=None, restart=200, tol=1e-5, maxleaf=10):
        maxn = nstep if maxn is None else maxn
        known = np.ascontiguousarray(known_lu, np.float64).reshape(-1)
        lu = np.zeros((nstep + 1) * self.ne * 3)
        lt = np.zeros_like(lu)
        iters = np.zeros(max(nstep, 1), np.int32)
        flags = np.zeros(max(nstep, 1), np.int32)
        solved = lib().orc_solve(self.h, nstep, maxn, known, restart, tol, maxleaf, lu, lt, iters, flags)
        if solved < 0:
            raise RuntimeError(f"orc_solve failed {solved}")
        return (lu.reshape(nstep + 1, self.ne, 3), lt.reshape(nstep + 1, self.ne, 3), iters[:nstep],
                flags[:nstep], solved)


def gmres_ccsr(row_ptr, col, val, b, restart=200, tol=1e-12):
    x = np.zeros_like(b)
    it = np.zeros(1, np.int32)
    flag = lib().orc_gmres_ccsr(len(row_ptr) - 1, np.ascontiguousarray(row_ptr, np.int32),
                                np.ascontiguousarray(col, np.int32), np.ascontiguousarray(val, np.float64).ravel(),
                                np.ascontiguousarray(b, np.float64), restart, tol, x, it)
    return x, int(it[0]), flag
