"""Python binding of the MI355X multi-domain TDBEM GMRES path.

The compute path is the HIP library ``lib/libtdbem.so`` reached through the
C-ABI in ``include/tdbem_capi.h``; this module only marshals numpy arrays.
There is no CPU fallback: if the library is missing or no GPU is present,
calls raise.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "lib" / "libtdbem.so"

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

EXPORTED = [
    "tdbem_last_error", "tdbem_create", "tdbem_destroy", "tdbem_dofmap_export", "tdbem_assemble_step0",
    "tdbem_assemble_history", "tdbem_ccsr_export", "tdbem_self_tables", "tdbem_solve", "tdbem_solve_batched",
    "tdbem_adjoint", "tdbem_get_materials", "tdbem_set_materials", "tdbem_dyna_gmres_solver",
    "tdbem_timings", "tdbem_timings_n", "tdbem_history_stats", "tdbem_device_synchronize",
    "tdbem_set_option", "tdbem_model_counts", "tdbem_read_model", "tdbem_read_bd", "tdbem_write_model_bin",
    "tdbem_write_bd_bin", "tdbem_write_validation", "tdbem_history_save", "tdbem_history_load",
    "tdbem_run_models", "tdbem_run_models_grouped", "tdbem_release_cached_memory", "tdbem_autotune", "tdbem_dofmap_build",
    "tdbem_builder_create", "tdbem_builder_destroy", "tdbem_builder_add_block", "tdbem_builder_merge",
    "tdbem_builder_info", "tdbem_builder_build", "tdbem_builder_build_sym", "tdbem_ccsr_export_rows", "tdbem_solve_device", "tdbem_solve_batched_device",
    "tdbem_device_alloc", "tdbem_device_free", "tdbem_memcpy", "tdbem_solve_device_noalloc",
]


def build(jobs: int = 8) -> Path:
    """Compile the HIP library for gfx950 in-tree."""
    subprocess.run(["make", "-s", "-j", str(jobs), "-C", str(PKG_DIR)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run build() first")
        L = C.CDLL(str(LIB_PATH))
        L.tdbem_last_error.restype = C.c_char_p
        L.tdbem_create.argtypes = [C.POINTER(C.c_void_p), C.c_long, _dp, C.c_long, _ip, _ip, C.c_int, _dp, C.c_int,
                                   _lp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        L.tdbem_destroy.argtypes = [C.c_void_p]
        L.tdbem_dofmap_export.argtypes = [C.c_void_p, _lp, _lp, _dp, _ip, _lp, _dp, _ip, _lp, _dp, _lp, _dp, _lp]
        L.tdbem_assemble_step0.argtypes = [C.c_void_p, C.c_long]
        L.tdbem_assemble_history.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long)]
        L.tdbem_ccsr_export.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_long),
                                        C.c_void_p, C.c_void_p, C.c_void_p]
        L.tdbem_self_tables.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.tdbem_solve.argtypes = [C.c_void_p, C.c_int, _dp, C.c_long, C.c_double, _dp, _dp, _ip, _ip,
                                  C.POINTER(C.c_int), _dp]
        L.tdbem_solve_batched.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_long, C.c_double, _dp, _dp, _ip,
                                          _ip, _ip, _dp]
        L.tdbem_adjoint.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _ip, _dp, C.c_long, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), _dp, C.c_void_p]
        L.tdbem_get_materials.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), _dp]
        L.tdbem_set_materials.argtypes = [C.c_void_p, _dp]
        L.tdbem_dyna_gmres_solver.argtypes = [C.c_void_p, C.c_long, C.c_double, C.c_long, C.c_int, _dp, _dp, _dp,
                                              C.POINTER(C.c_int)]
        L.tdbem_timings.argtypes = [C.c_void_p, _dp]
        L.tdbem_timings_n.argtypes = [C.c_void_p, C.c_int, _dp]
        L.tdbem_history_stats.argtypes = [C.c_void_p, _lp, C.POINTER(C.c_long)]
        L.tdbem_device_synchronize.argtypes = []
        L.tdbem_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.tdbem_model_counts.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_long),
                                         C.POINTER(C.c_long)]
        L.tdbem_read_model.argtypes = [C.c_char_p, C.c_int, _dp, _ip, C.c_void_p]
        L.tdbem_read_bd.argtypes = [C.c_char_p, C.c_int, C.c_long, C.c_int, C.c_void_p, _ip, _dp]
        L.tdbem_write_model_bin.argtypes = [C.c_char_p, C.c_long, _dp, C.c_long, _ip, C.c_long, C.c_void_p]
        L.tdbem_write_bd_bin.argtypes = [C.c_char_p, C.c_long, C.c_int, _ip, _dp]
        L.tdbem_history_save.argtypes = [C.c_void_p, C.c_char_p]
        L.tdbem_history_load.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int)]
        L.tdbem_write_validation.argtypes = [C.c_void_p, C.c_char_p, C.c_int, _dp, _dp, _ip, _ip, C.c_int, C.c_int]
        _pp = C.POINTER(C.c_void_p)
        L.tdbem_run_models.argtypes = [C.c_int, _pp, C.c_long, _ip, _pp, C.c_long, C.c_double, C.c_double, _pp, _pp,
                                       _pp, _pp, _ip]
        L.tdbem_run_models_grouped.argtypes = [C.c_int, _pp, C.c_long, _ip, _pp, C.c_long, C.c_double, C.c_double,
                                               _pp, _pp, _pp, _pp, _ip, C.POINTER(C.c_int)]
        L.tdbem_autotune.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]
        L.tdbem_solve_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_long, C.c_double, C.c_void_p,
                                         C.c_void_p, _ip, _ip, C.POINTER(C.c_int), _dp]
        L.tdbem_solve_batched_device.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_long, C.c_double,
                                                 C.c_void_p, C.c_void_p, _ip, _ip, _ip, _dp]
        L.tdbem_solve_device_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_long, C.c_double, C.c_void_p,
                                                 C.c_void_p, _ip, _ip, C.POINTER(C.c_int)]
        L.tdbem_device_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
        L.tdbem_device_free.argtypes = [C.c_void_p]
        L.tdbem_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.tdbem_ccsr_export_rows.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_long, C.c_long, C.POINTER(C.c_long),
                                             C.c_void_p, C.c_void_p, C.c_void_p]
        L.tdbem_dofmap_build.argtypes = [C.c_long, _ip, C.c_int, _dp, C.c_int, _lp, _lp, _lp, _dp, _ip, _lp, _dp,
                                         _ip, _lp, _dp, _lp, _dp, _lp]
        L.tdbem_builder_create.argtypes = [C.POINTER(C.c_void_p), C.c_long, C.c_long]
        L.tdbem_builder_destroy.argtypes = [C.c_void_p]
        L.tdbem_builder_add_block.argtypes = [C.c_void_p, C.c_long, C.c_long, C.c_double, C.c_void_p]
        L.tdbem_builder_merge.argtypes = [C.c_void_p, C.c_void_p]
        L.tdbem_builder_info.argtypes = [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.tdbem_builder_build.argtypes = [C.c_void_p, _ip, C.c_void_p, C.c_void_p]
        L.tdbem_builder_build_sym.argtypes = [C.c_void_p, C.c_double, _ip, C.c_void_p, C.c_void_p,
                                              C.POINTER(C.c_int)]
        _lib = L
    return _lib


def dofmap_build(bcid, domain_count=0, mat_rows=None, itf_rows=None):
    """GlobalDofMap::Build on the host (tdbem_dofmap_build): the arrays of
    Case.dofmap() from the element BCIDs and the multi-domain tail, without a
    device.  Raises on the reference's failures (BCIDs other than 123 / 456)."""
    bcid = np.ascontiguousarray(bcid, np.int32)
    n = len(bcid)
    mats = np.ascontiguousarray(mat_rows if mat_rows is not None else np.zeros((1, 4)), np.float64).ravel()
    itf = np.ascontiguousarray(itf_rows if itf_rows is not None else np.zeros((1, 5)), np.int64).ravel()
    ni = -1 if itf_rows is None else len(np.asarray(itf_rows).reshape(-1, 5))
    counts = np.zeros(9, np.int64)
    arr = dict(u_block=np.zeros(n, np.int64), u_sign=np.zeros(n), u_known=np.zeros(n, np.int32),
               t_block=np.zeros(n, np.int64), t_sign=np.zeros(n), t_known=np.zeros(n, np.int32),
               hu_block=np.zeros(n, np.int64), hu_sign=np.zeros(n), ht_block=np.zeros(n, np.int64),
               ht_sign=np.zeros(n), preferred=np.zeros(n, np.int64))
    _check(lib().tdbem_dofmap_build(n, bcid, int(domain_count), mats, ni, itf, counts, arr["u_block"],
                                    arr["u_sign"], arr["u_known"], arr["t_block"], arr["t_sign"], arr["t_known"],
                                    arr["hu_block"], arr["hu_sign"], arr["ht_block"], arr["ht_sign"],
                                    arr["preferred"]), "tdbem_dofmap_build")
    arr["counts"] = counts
    return arr


class DeviceArray:
    """A float64 array in device memory (tdbem_device_alloc), for the
    device-resident solves: upload() from numpy, download() to numpy."""

    def __init__(self, shape):
        self.shape = tuple(int(x) for x in np.atleast_1d(shape))
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        _check(lib().tdbem_device_alloc(max(self.nbytes, 8), C.byref(p)), "tdbem_device_alloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, np.float64)
        d = cls(a.shape)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(lib().tdbem_memcpy(self.ptr, a.ctypes.data, self.nbytes, 0), "tdbem_memcpy")

    def download(self):
        a = np.empty(self.shape)
        _check(lib().tdbem_memcpy(a.ctypes.data, self.ptr, self.nbytes, 1), "tdbem_memcpy")
        return a

    def close(self):
        if getattr(self, "ptr", None):
            lib().tdbem_device_free(self.ptr)
            self.ptr = None

    __del__ = close


class Builder:
    """MultiDomainCCSRBuilder of the library (tdbem_builder_*), host only."""

    def __init__(self, block_rows, block_cols):
        h = C.c_void_p()
        _check(lib().tdbem_builder_create(C.byref(h), block_rows, block_cols), "tdbem_builder_create")
        self.h, self.rows = h, block_rows

    def close(self):
        if getattr(self, "h", None):
            lib().tdbem_builder_destroy(self.h)
            self.h = None

    __del__ = close

    def add_block(self, row, col, sign, block9):
        if block9 is None:
            _check(lib().tdbem_builder_add_block(self.h, row, col, sign, None), "tdbem_builder_add_block")
            return
        b = np.ascontiguousarray(block9, np.float64).ravel()
        _check(lib().tdbem_builder_add_block(self.h, row, col, sign, b.ctypes.data), "tdbem_builder_add_block")

    def merge(self, other):
        _check(lib().tdbem_builder_merge(self.h, other.h), "tdbem_builder_merge")

    def info(self):
        nnz, sk = C.c_long(), C.c_long()
        _check(lib().tdbem_builder_info(self.h, C.byref(nnz), C.byref(sk)), "tdbem_builder_info")
        return nnz.value, sk.value

    def build(self):
        nnz = self.info()[0]
        rp = np.zeros(self.rows + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1) * 9)
        _check(lib().tdbem_builder_build(self.h, rp, col.ctypes.data, val.ctypes.data), "tdbem_builder_build")
        return rp, col[:nnz], val[:9 * nnz].reshape(-1, 3, 3)

    def build_sym(self, tolerance):
        nnz = self.info()[0]
        rp = np.zeros(self.rows + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1) * 6)
        sym = C.c_int()
        _check(lib().tdbem_builder_build_sym(self.h, tolerance, rp, col.ctypes.data, val.ctypes.data,
                                             C.byref(sym)), "tdbem_builder_build_sym")
        return rp, col[:nnz], val[:6 * nnz].reshape(-1, 6), bool(sym.value)


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().tdbem_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def synchronize() -> None:
    _check(lib().tdbem_device_synchronize(), "tdbem_device_synchronize")


def release_cached_memory() -> None:
    """Return the library's cached large device buffers to the device."""
    _check(lib().tdbem_release_cached_memory(), "tdbem_release_cached_memory")


def run_models(cases, knowns, nsteps, maxleaf=10, restart=200, tol=1e-5, budget_bytes=0.0, grouped=False):
    """Whole jobs of independent cases together (tdbem_run_models): step 0,
    history for nsteps[i] steps and the time loop of every case, each on its
    own stream and host thread, with the history bands held at once kept
    within budget_bytes (0: 80 % of the free device memory; the bands are
    released after each case's time loop).  grouped: the cases stacked in
    rounds whose every kernel is one launch for all the round's cases
    (tdbem_run_models_grouped; FP64 bands; run_models.rounds then holds the
    number of rounds).  Returns per case (lu, lt, iters, flags, solved) as
    Case.solve does."""
    n = len(cases)
    if len(knowns) != n or len(nsteps) != n:
        raise ValueError("one load history and one step count per case")
    keep, outs = [], []
    for c, k, ns in zip(cases, knowns, nsteps):
        k = np.ascontiguousarray(k, np.float64).reshape(-1)
        if k.size != (ns + 1) * c.ne * 3:
            raise ValueError("known_lu must hold (nstep+1) x ne x 3 values")
        o = (np.zeros_like(k), np.zeros_like(k), np.zeros(ns, np.int32), np.zeros(ns, np.int32))
        keep.append(k)
        outs.append(o)

    def ptrs(arrs):
        return (C.c_void_p * n)(*[a.ctypes.data for a in arrs])

    ctx = (C.c_void_p * n)(*[c.h.value for c in cases])
    solved = np.zeros(n, np.int32)
    steps = np.asarray(nsteps, np.int32)
    if grouped:
        nr = C.c_int(0)
        _check(lib().tdbem_run_models_grouped(n, ctx, int(maxleaf), steps, ptrs(keep), int(restart), float(tol),
                                              float(budget_bytes), ptrs([o[0] for o in outs]),
                                              ptrs([o[1] for o in outs]), ptrs([o[2] for o in outs]),
                                              ptrs([o[3] for o in outs]), solved, C.byref(nr)),
               "tdbem_run_models_grouped")
        run_models.rounds = nr.value
    else:
        _check(lib().tdbem_run_models(n, ctx, int(maxleaf), steps, ptrs(keep), int(restart), float(tol),
                                      float(budget_bytes), ptrs([o[0] for o in outs]), ptrs([o[1] for o in outs]),
                                      ptrs([o[2] for o in outs]), ptrs([o[3] for o in outs]), solved),
               "tdbem_run_models")
    res = []
    for c, ns, o, sv in zip(cases, nsteps, outs, solved):
        res.append((o[0].reshape(ns + 1, c.ne, 3), o[1].reshape(ns + 1, c.ne, 3), o[2], o[3], int(sv)))
    return res


class Case:
    """One multi-domain case on the GPU (mirrors DynaGMRESSolverMultiDomainCCSR's inputs)."""

    def __init__(self, points, ele, bcid, dt, gap, domain_count=0, mat_rows=None, itf_rows=None,
                 default_mat=(1.0, 0.0, 1.0)):
        self.points = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        self.ele = np.ascontiguousarray(ele, np.int32).reshape(-1, 8)
        self.bcid = np.ascontiguousarray(bcid, np.int32)
        self.ne = len(self.ele)
        mats = np.ascontiguousarray(mat_rows if mat_rows is not None else np.zeros((0, 4)), np.float64).reshape(-1)
        if itf_rows is None:
            nitf, itf = -1, np.zeros(1, np.int64)
        else:
            itf = np.ascontiguousarray(itf_rows, np.int64).reshape(-1)
            nitf = len(itf) // 5
            if itf.size == 0:
                itf = np.zeros(1, np.int64)
        if mats.size == 0:
            mats = np.zeros(1)
        h = C.c_void_p()
        _check(lib().tdbem_create(C.byref(h), len(self.points), self.points.ravel(), self.ne, self.ele.ravel(),
                                  self.bcid, int(domain_count), mats, nitf, itf, float(default_mat[0]),
                                  float(default_mat[1]), float(default_mat[2]), float(dt), float(gap)),
               "tdbem_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().tdbem_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def dofmap(self):
        n = self.ne
        counts = np.zeros(9, np.int64)
        a = dict(u_block=np.zeros(n, np.int64), u_sign=np.zeros(n), u_known=np.zeros(n, np.int32),
                 t_block=np.zeros(n, np.int64), t_sign=np.zeros(n), t_known=np.zeros(n, np.int32),
                 hu_block=np.zeros(n, np.int64), hu_sign=np.zeros(n), ht_block=np.zeros(n, np.int64),
                 ht_sign=np.zeros(n), preferred=np.zeros(n, np.int64))
        _check(lib().tdbem_dofmap_export(self.h, counts, a["u_block"], a["u_sign"], a["u_known"], a["t_block"],
                                         a["t_sign"], a["t_known"], a["hu_block"], a["hu_sign"], a["ht_block"],
                                         a["ht_sign"], a["preferred"]), "tdbem_dofmap_export")
        a["counts"] = counts
        return a

    def assemble_step0(self, maxleaf=10):
        _check(lib().tdbem_assemble_step0(self.h, int(maxleaf)), "tdbem_assemble_step0")

    def assemble_history(self, maxn):
        e = C.c_long()
        _check(lib().tdbem_assemble_history(self.h, int(maxn), C.byref(e)), "tdbem_assemble_history")
        return e.value

    def ccsr(self, which, step=0):
        nr, nnz = C.c_long(), C.c_long()
        _check(lib().tdbem_ccsr_export(self.h, which, step, C.byref(nr), C.byref(nnz), None, None, None),
               "tdbem_ccsr_export")
        rp = np.zeros(nr.value + 1, np.int32)
        col = np.zeros(max(nnz.value, 1), np.int32)
        val = np.zeros(max(nnz.value, 1) * 9)
        _check(lib().tdbem_ccsr_export(self.h, which, step, C.byref(nr), C.byref(nnz), rp.ctypes.data,
                                       col.ctypes.data, val.ctypes.data), "tdbem_ccsr_export")
        return rp, col[:nnz.value], val[:9 * nnz.value].reshape(-1, 3, 3)

    def ccsr_rows(self, which, step, row0, row1):
        """Rows row0 .. row1-1 of MTS[step] (2) / MGS[step] (3) as (row_ptr,
        col, blocks), copying only those rows' bands from the device."""
        nnz = C.c_long()
        _check(lib().tdbem_ccsr_export_rows(self.h, which, step, row0, row1, C.byref(nnz), None, None, None),
               "tdbem_ccsr_export_rows")
        rp = np.zeros(row1 - row0 + 1, np.int32)
        col = np.zeros(max(nnz.value, 1), np.int32)
        val = np.zeros(max(nnz.value, 1) * 9)
        _check(lib().tdbem_ccsr_export_rows(self.h, which, step, row0, row1, C.byref(nnz), rp.ctypes.data,
                                            col.ctypes.data, val.ctypes.data), "tdbem_ccsr_export_rows")
        return rp, col[:nnz.value], val[:9 * nnz.value].reshape(-1, 3, 3)

    def self_tables(self):
        n = 9 * self.ne
        u, t1, t2 = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().tdbem_self_tables(self.h, u, t1, t2), "tdbem_self_tables")
        return u.reshape(-1, 9), t1.reshape(-1, 9), t2.reshape(-1, 9)

    def solve(self, known_lu, nstep, restart=200, tol=1e-5):
        known = np.ascontiguousarray(known_lu, np.float64).reshape(-1)
        if known.size != (nstep + 1) * self.ne * 3:
            raise ValueError("known_lu must hold (nstep+1) x ne x 3 values")
        lu = np.zeros_like(known)
        lt = np.zeros_like(known)
        iters = np.zeros(nstep, np.int32)
        flags = np.zeros(nstep, np.int32)
        solved = C.c_int()
        times = np.zeros(3)
        _check(lib().tdbem_solve(self.h, int(nstep), known, int(restart), float(tol), lu, lt, iters, flags,
                                 C.byref(solved), times), "tdbem_solve")
        return (lu.reshape(nstep + 1, self.ne, 3), lt.reshape(nstep + 1, self.ne, 3), iters, flags, solved.value,
                times)

    def dyna_gmres_solver(self, known_lu, nstep, iterations=200, error=1e-5, maxleaf=10):
        """The whole DynaGMRESSolverMultiDomainCCSR (tdbem_dyna_gmres_solver):
        step 0, history for nstep steps and the time loop.  Returns (result,
        lu, lt) with result 1 when the last solved step's GMRES flag is 0 and
        -1 otherwise."""
        known = np.ascontiguousarray(known_lu, np.float64).reshape(-1)
        if known.size != (nstep + 1) * self.ne * 3:
            raise ValueError("known_lu must hold (nstep+1) x ne x 3 values")
        lu = np.zeros_like(known)
        lt = np.zeros_like(known)
        result = C.c_int()
        _check(lib().tdbem_dyna_gmres_solver(self.h, int(iterations), float(error), int(maxleaf), int(nstep), known,
                                             lu, lt, C.byref(result)), "tdbem_dyna_gmres_solver")
        return result.value, lu.reshape(nstep + 1, self.ne, 3), lt.reshape(nstep + 1, self.ne, 3)

    def solve_device(self, d_known, nstep, d_lu, d_lt, restart=200, tol=1e-5):
        """tdbem_solve_device: DeviceArray inputs / outputs of (nstep+1, ne, 3);
        returns (iters, flags, solved, times)."""
        iters = np.zeros(nstep, np.int32)
        flags = np.zeros(nstep, np.int32)
        solved = C.c_int()
        times = np.zeros(3)
        _check(lib().tdbem_solve_device(self.h, int(nstep), d_known.ptr, int(restart), float(tol), d_lu.ptr, d_lt.ptr,
                                        iters, flags, C.byref(solved), times), "tdbem_solve_device")
        return iters, flags, solved.value, times

    def solve_device_noalloc(self, d_known, nstep, d_lu, d_lt, restart=200, tol=1e-5):
        """tdbem_solve_device_noalloc: as solve_device, with every working
        buffer in the library's fixed scratch array; returns (iters, flags,
        solved)."""
        iters = np.zeros(nstep, np.int32)
        flags = np.zeros(nstep, np.int32)
        solved = C.c_int()
        _check(lib().tdbem_solve_device_noalloc(self.h, int(nstep), d_known.ptr, int(restart), float(tol), d_lu.ptr,
                                                d_lt.ptr, iters, flags, C.byref(solved)), "tdbem_solve_device_noalloc")
        return iters, flags, solved.value

    def solve_batched_device(self, d_known, nrhs, nstep, d_lu, d_lt, restart=200, tol=1e-5):
        """tdbem_solve_batched_device: DeviceArray inputs / outputs of
        (nrhs, nstep+1, ne, 3); returns (iters, flags, solved, times)."""
        iters = np.zeros(nrhs * nstep, np.int32)
        flags = np.zeros(nrhs * nstep, np.int32)
        solved = np.zeros(nrhs, np.int32)
        times = np.zeros(3)
        _check(lib().tdbem_solve_batched_device(self.h, int(nstep), int(nrhs), d_known.ptr, int(restart), float(tol),
                                                d_lu.ptr, d_lt.ptr, iters, flags, solved, times),
               "tdbem_solve_batched_device")
        return iters.reshape(nrhs, nstep), flags.reshape(nrhs, nstep), solved, times

    def solve_batched(self, known_lu, nstep, restart=200, tol=1e-5):
        """known_lu: (nrhs, nstep+1, ne, 3) load histories -> lu, lt of the same
        shape, iters / flags (nrhs, nstep), solved (nrhs,), times."""
        known = np.ascontiguousarray(known_lu, np.float64)
        nrhs = known.shape[0]
        known = known.reshape(-1)
        if known.size != nrhs * (nstep + 1) * self.ne * 3:
            raise ValueError("known_lu must hold nrhs x (nstep+1) x ne x 3 values")
        lu = np.zeros_like(known)
        lt = np.zeros_like(known)
        iters = np.zeros((nrhs, nstep), np.int32)
        flags = np.zeros((nrhs, nstep), np.int32)
        solved = np.zeros(nrhs, np.int32)
        times = np.zeros(3)
        _check(lib().tdbem_solve_batched(self.h, int(nstep), int(nrhs), known, int(restart), float(tol), lu, lt,
                                         iters, flags, solved, times), "tdbem_solve_batched")
        shape = (nrhs, nstep + 1, self.ne, 3)
        return lu.reshape(shape), lt.reshape(shape), iters, flags, solved, times

    def adjoint(self, known_lu, nstep, receivers, obs, restart=200, tol=1e-12, fd_step=1e-6):
        """Misfit J = sum_{K>=1} sum_r |lu_K[r] - obs[K, r]|^2 over the receiver
        elements and its gradient (domains, 3) with respect to (E, v, rho) per
        domain, by the reverse-time adjoint; also returns the forward lu."""
        known = np.ascontiguousarray(known_lu, np.float64).reshape(-1)
        rec = np.ascontiguousarray(receivers, np.int32).reshape(-1)
        ob = np.ascontiguousarray(obs, np.float64).reshape(-1)
        if ob.size != (nstep + 1) * rec.size * 3:
            raise ValueError("obs must hold (nstep+1) x nrec x 3 values")
        nd = len(self.materials())
        J = C.c_double()
        grad = np.zeros(3 * nd)
        lu = np.zeros((nstep + 1) * self.ne * 3)
        _check(lib().tdbem_adjoint(self.h, int(nstep), known, int(rec.size), rec, ob, int(restart), float(tol),
                                   float(fd_step), C.byref(J), grad, lu.ctypes.data_as(C.c_void_p)), "tdbem_adjoint")
        return J.value, grad.reshape(nd, 3), lu.reshape(nstep + 1, self.ne, 3)

    def materials(self):
        """(domains, 3) array of (E, v, rho)."""
        out = np.zeros(3 * 64)
        nd = C.c_int()
        _check(lib().tdbem_get_materials(self.h, 64, C.byref(nd), out), "tdbem_get_materials")
        return out[:3 * nd.value].reshape(nd.value, 3)

    def set_materials(self, evr):
        _check(lib().tdbem_set_materials(self.h, np.ascontiguousarray(evr, np.float64).reshape(-1)),
               "tdbem_set_materials")

    def set_option(self, name, value):
        _check(lib().tdbem_set_option(self.h, name.encode(), float(value)), "tdbem_set_option")

    def history_stats(self, maxn):
        counts = np.zeros(maxn, np.int64)
        npairs = C.c_long()
        _check(lib().tdbem_history_stats(self.h, counts, C.byref(npairs)), "tdbem_history_stats")
        return counts, npairs.value

    def save_history(self, path):
        """Write the assembled history bands to path."""
        _check(lib().tdbem_history_save(self.h, str(path).encode()), "tdbem_history_save")

    def load_history(self, path):
        """Load history bands written for this model; returns the step count."""
        n = C.c_int()
        _check(lib().tdbem_history_load(self.h, str(path).encode(), C.byref(n)), "tdbem_history_load")
        return n.value

    def write_validation(self, directory, U, T, iters, flags, solved, solver=2):
        """validation_state.csv and validation_metrics.txt of a solve (U, T as
        solve() returns them) into directory."""
        nstep = len(iters)
        U = np.ascontiguousarray(U, np.float64).reshape(-1)
        T = np.ascontiguousarray(T, np.float64).reshape(-1)
        if U.size != (nstep + 1) * self.ne * 3 or T.size != U.size:
            raise ValueError("U / T must be (nstep+1, ne, 3)")
        _check(lib().tdbem_write_validation(self.h, str(directory).encode(), nstep, U, T,
                                            np.ascontiguousarray(iters, np.int32),
                                            np.ascontiguousarray(flags, np.int32), int(solved), int(solver)),
               "tdbem_write_validation")

    def autotune(self, cache_path=None):
        """Pick the far-convolution and GMRES-reduction launch shapes by timing
        (tdbem_autotune); returns (far_parts, gmres_blocks, far_ms)."""
        fp, gb, ms = C.c_int(), C.c_int(), C.c_double()
        _check(lib().tdbem_autotune(self.h, cache_path.encode() if cache_path else None, C.byref(fp), C.byref(gb),
                                    C.byref(ms)), "tdbem_autotune")
        return fp.value, gb.value, ms.value

    def timings(self):
        t = np.zeros(45)
        _check(lib().tdbem_timings_n(self.h, 45, t), "tdbem_timings_n")
        out = dict(singular=t[0], cull=t[1], history=t[2], step0=t[3], convolution=t[4], gmres=t[5],
                   conv_far=t[6], conv_far_launches=int(t[7]), quadrature=t[8], adjoint_forward=t[9],
                   adjoint_backward=t[10], adjoint_step0_differences=t[11], adjoint_history_derivatives=t[12],
                   conv_spectral=t[13], conv_spectral_launches=int(t[14]), spectral=bool(t[15]),
                   spectral_setup=t[16], conv_spectral_products=t[17], conv_spectral_flops=t[18],
                   conv_spectral_bytes=t[19], stream_slabs=int(t[20]), stream_far=t[21],
                   quad_evals_t2=int(t[22]), quad_evals_t1u=int(t[23]), quad_evals_regular=int(t[24]),
                   quad_records=int(t[25]), spectral_levels=int(t[42]), quad_pw_entries=int(t[43]),
                   quad_pw_waves=int(t[44]))
        # per level of the spectral part: products seconds, flops, launches, bytes
        out["spectral_level_stats"] = [dict(L=16 << l, products=t[26 + l], flops=t[30 + l], launches=int(t[34 + l]),
                                            bytes=t[38 + l]) for l in range(int(t[42]))]
        return out
