"""Readers for the reference's text inputs: the positional datacard with its
multi-domain tail, the ``.model`` mesh and the ``.bd`` boundary history.

Formats follow DBEM1/DBEM.cpp:637-791 (datacard), DBEM1/DBEM.cpp:849-900
(.model), DBEM1/DBEM.cpp:1010-1085 (.bd) and DBEM1/multidomain.cpp:1305-1362
(multi-domain tail).  Id normalisation (0- or 1-based ids detected by the
presence of a 0) follows DBEM1/multidomain.cpp:500-523,1382-1461.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import List, Optional, Tuple

import numpy as np


@dataclass
class DataCard:
    title: str
    thread_num: int
    flag_dyna: int
    beta: float
    nstep: int
    gap: float
    E: float
    v: float
    rho: float
    amplitude: float
    solver: int
    iterations: int = 200
    error: float = 1e-5
    premaxleaf: int = 10
    input_binary_mode: int = 0
    domain_count: int = 0
    materials: List[Tuple[int, float, float, float]] = field(default_factory=list)
    interfaces: Optional[List[Tuple[int, int, int, int, int]]] = None


def read_datacard(path: Path) -> DataCard:
    tok = Path(path).read_text().split()
    pos = 0

    def nxt():
        nonlocal pos
        if pos >= len(tok):
            return None
        v = tok[pos]
        pos += 1
        return v

    card = DataCard(title=nxt(), thread_num=int(nxt()), flag_dyna=int(nxt()), beta=float(nxt()),
                    nstep=int(nxt()), gap=float(nxt()), E=float(nxt()), v=float(nxt()), rho=float(nxt()),
                    amplitude=float(nxt()), solver=int(nxt()))
    if card.solver != 2:
        raise ValueError("only the GMRES solver (2) datacards are supported")
    card.iterations = int(nxt())
    card.error = float(nxt())
    card.premaxleaf = int(nxt())
    # InputBinaryMode / BatchBoundaryMode (DBEM1/DBEM.cpp:686-700); batch modes are not
    # supported by the multi-domain solver (DBEM1/DBEM.cpp:1227-1231).
    in_mode = int(nxt())
    batch = int(nxt())
    if in_mode not in (0, 1, 2) or batch not in (0,):
        raise ValueError("batch boundary inputs are not supported")
    card.input_binary_mode = in_mode
    dc = nxt()
    if dc is None:
        return card
    card.domain_count = int(dc)
    if card.domain_count <= 0:
        raise ValueError(f"Invalid DomainCount {card.domain_count}")
    for _ in range(card.domain_count):
        row = [nxt() for _ in range(4)]
        if None in row:
            raise ValueError("Invalid material line")
        card.materials.append((int(row[0]), float(row[1]), float(row[2]), float(row[3])))
    ic = nxt()
    if ic is not None:
        n = int(ic)
        if n < 0:
            raise ValueError(f"Invalid InterfaceCount {n}")
        card.interfaces = []
        for _ in range(n):
            row = [nxt() for _ in range(5)]
            if None in row:
                raise ValueError("Invalid interface line")
            card.interfaces.append(tuple(int(x) for x in row))
    return card


def read_model(path: Path) -> Tuple[np.ndarray, np.ndarray]:
    """Return points (np,3) float64 and element point ids (ne,8) 0-based int32."""
    tok = Path(path).read_text().split()
    npnt = int(tok[0])
    pts = np.array([float(x) for x in tok[1:1 + 3 * npnt]], dtype=np.float64).reshape(npnt, 3)
    p = 1 + 3 * npnt
    ne = int(tok[p])
    ele = np.array([int(x) for x in tok[p + 1:p + 1 + 8 * ne]], dtype=np.int64).reshape(ne, 8) - 1
    return pts, ele.astype(np.int32)


def read_bd(path: Path, ne: int, nstep: int) -> Tuple[np.ndarray, np.ndarray]:
    """Return bcid (ne,) int32 and given triplets lu (nstep+1, ne, 3) float64."""
    tok = Path(path).read_text().split()
    bcid = np.array([int(x) for x in tok[:ne]], dtype=np.int32)
    need = (nstep + 1) * ne * 3
    vals = np.array([float(x) for x in tok[ne:ne + need]], dtype=np.float64)
    if vals.size != need:
        raise ValueError("boundary file shorter than NStep+1 steps")
    return bcid, vals.reshape(nstep + 1, ne, 3)


def _norm_domain(i: int, count: int, zero_based: bool) -> int:
    if zero_based and 0 <= i < count:
        return i
    if 1 <= i <= count:
        return i - 1
    return -1


def _norm_element(i: int, ne: int, zero_based: bool) -> int:
    if zero_based:
        return i if 0 <= i < ne else -1
    return i - 1 if 1 <= i <= ne else -1


def normalise_model(card: DataCard, ne: int):
    """Materials (nd,3) ordered by domain and interfaces (ni,5) 0-based, as
    BuildMultiDomainModel produces them (DBEM1/multidomain.cpp:1364-1480)."""
    nd = card.domain_count if card.domain_count > 0 else 1
    mats = np.array([[card.E, card.v, card.rho]] * nd, dtype=np.float64)
    zb = any(m[0] == 0 for m in card.materials)
    for m in card.materials:
        d = _norm_domain(m[0], nd, zb)
        if d < 0:
            raise ValueError(f"Invalid domain material id {m[0]}")
        mats[d] = (m[1], m[2], m[3])
    itf = []
    if card.interfaces is not None:
        zbe = any(t[1] == 0 or t[3] == 0 for t in card.interfaces)
        zbd = any(t[0] == 0 or t[2] == 0 for t in card.interfaces)
        for t in card.interfaces:
            itf.append((_norm_domain(t[0], nd, zbd), _norm_element(t[1], ne, zbe),
                        _norm_domain(t[2], nd, zbd), _norm_element(t[3], ne, zbe), -1 if t[4] == 0 else t[4]))
    elif nd > 1:
        base, extra = divmod(ne, nd)
        begins, b = [], 0
        for d in range(nd):
            begins.append(b)
            b += base + (1 if d < extra else 0)
        for d in range(nd - 1):
            cnt = base + (1 if d < extra else 0)
            itf.append((d, begins[d] + cnt - 1, d + 1, begins[d + 1], -1))
    return mats, np.array(itf, dtype=np.int32).reshape(-1, 5)


def reference_dt(points: np.ndarray, ele: np.ndarray, beta: float, E: float, v: float, rho: float) -> float:
    """dt = Beta * AveEleSize / c1 of the global material line (DBEM1/DBEM.cpp:910-914,2383-2396)."""
    import math
    ave = 0.0
    for e in ele:
        a, b = points[e[0]], points[e[1]]
        ave += math.sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]))
    ave /= len(ele)
    G = E / 2.0 / (1 + v)
    c1 = math.sqrt((2.0 - 2.0 * v) / (1.0 - 2.0 * v) * G / rho)
    return beta * ave / c1


# ---------------------------------------------------------------------------
# Text or DBEMBIN inputs through the C-ABI readers (mode 0 text, 1 binary
# "<path>.bin" required, 2 binary when it exists; DBEM1/DBEM.cpp:828-848,
# 1023-1040) and the DBEMBIN writers (the layout of the reference's
# convert_inputs_to_binary.py).  These touch no device.
# ---------------------------------------------------------------------------
def _lib():
    from . import lib
    return lib()


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {_lib().tdbem_last_error().decode()}")


def load_model(path, mode: int = 0):
    """points (np,3), element ids (ne,8) 0-based int32, infinite-element ids
    (ninf,4) 0-based int64."""
    import ctypes as C
    L = _lib()
    n_p, n_e, n_i = C.c_long(), C.c_long(), C.c_long()
    _ck(L.tdbem_model_counts(str(path).encode(), mode, C.byref(n_p), C.byref(n_e), C.byref(n_i)),
        "tdbem_model_counts")
    pts = np.zeros(3 * n_p.value)
    ele = np.zeros(8 * n_e.value, np.int32)
    inf = np.zeros(max(1, 4 * n_i.value), np.int64)
    _ck(L.tdbem_read_model(str(path).encode(), mode, pts, ele, inf.ctypes.data), "tdbem_read_model")
    return pts.reshape(-1, 3), ele.reshape(-1, 8), inf[:4 * n_i.value].reshape(-1, 4)


def load_bd(path, ne: int, nstep: int, mode: int = 0, expected_bcid=None):
    """bcid (ne,) int32 and given triplets (nstep+1, ne, 3)."""
    L = _lib()
    bcid = np.zeros(ne, np.int32)
    lu = np.zeros((nstep + 1) * ne * 3)
    exp = None if expected_bcid is None else np.ascontiguousarray(expected_bcid, np.int32)
    _ck(L.tdbem_read_bd(str(path).encode(), mode, ne, nstep, None if exp is None else exp.ctypes.data, bcid, lu),
        "tdbem_read_bd")
    return bcid, lu.reshape(nstep + 1, ne, 3)


def write_model_bin(path, points, ele, inf=None):
    L = _lib()
    pts = np.ascontiguousarray(points, np.float64).reshape(-1)
    e = np.ascontiguousarray(ele, np.int32).reshape(-1)
    i = np.zeros(0, np.int64) if inf is None else np.ascontiguousarray(inf, np.int64).reshape(-1)
    _ck(L.tdbem_write_model_bin(str(path).encode(), pts.size // 3, pts, e.size // 8, e, i.size // 4,
                                i.ctypes.data if i.size else None), "tdbem_write_model_bin")


def write_bd_bin(path, bcid, lu):
    L = _lib()
    b = np.ascontiguousarray(bcid, np.int32)
    v = np.ascontiguousarray(lu, np.float64)
    nstep = v.size // (3 * b.size) - 1
    _ck(L.tdbem_write_bd_bin(str(path).encode(), b.size, nstep, b, v.reshape(-1)), "tdbem_write_bd_bin")
