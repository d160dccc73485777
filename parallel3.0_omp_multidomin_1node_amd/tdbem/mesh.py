"""Box meshes of axially stacked sub-domains with constant square elements.

Conventions follow the reference generator (tools/generate_multidomain_rod.py:
75-294,360-575): 8-node quads (corners, then edge midpoints), face order per
domain left, right, y_min, y_max, z_min, z_max with outward windings, point
de-duplication at 12 decimals, interface pairs right face of domain d with
the left face of domain d+1 (normalSign -1), BCID 123 on the left end face of
the first domain and 456 elsewhere.  The stacking axis is x; the loaded
"top" face is the right end face of the last domain, the fixed "bottom" face
the left end face of the first domain.  Loads are given in the element frame
(normal component = 3rd local component), as the reference .bd files hold.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class BoxMesh:
    points: np.ndarray            # (np, 3)
    ele: np.ndarray               # (ne, 8) 0-based point ids
    bcid: np.ndarray              # (ne,)
    domain_count: int
    interfaces: np.ndarray        # (ni, 5) 1-based rows (domA, eleA, domB, eleB, sign)
    top: np.ndarray               # element ids of the loaded end face
    bottom: np.ndarray            # element ids of the fixed end face
    faces: List[Dict[str, List[int]]] = field(default_factory=list)


class _PointTable:
    def __init__(self):
        self.points: List[Tuple[float, float, float]] = []
        self.ids: Dict[Tuple[float, float, float], int] = {}

    def add(self, p):
        key = tuple(round(v, 12) + 0.0 for v in p)
        i = self.ids.get(key)
        if i is None:
            i = len(self.points)
            self.ids[key] = i
            self.points.append(key)
        return i


def _quad(tab, ele, c):
    p0, p1, p2, p3 = c
    mid = lambda a, b: ((a[0] + b[0]) * 0.5, (a[1] + b[1]) * 0.5, (a[2] + b[2]) * 0.5)
    ele.append([tab.add(p0), tab.add(p1), tab.add(p2), tab.add(p3), tab.add(mid(p0, p1)), tab.add(mid(p1, p2)),
                tab.add(mid(p2, p3)), tab.add(mid(p3, p0))])
    return len(ele) - 1


def box_mesh(domain_count: int, length: float, width: float, height: float, nx: int, ny: int, nz: int) -> BoxMesh:
    tab = _PointTable()
    ele: List[List[int]] = []
    faces: List[Dict[str, List[int]]] = []
    dxd = length / domain_count
    y0, y1 = -0.5 * width, 0.5 * width
    z0, z1 = -0.5 * height, 0.5 * height
    for d in range(domain_count):
        x0 = -0.5 * length + d * dxd
        x1 = x0 + dxd
        dx, dy, dz = (x1 - x0) / nx, (y1 - y0) / ny, (z1 - z0) / nz
        f: Dict[str, List[int]] = {k: [] for k in ("left", "right", "y_min", "y_max", "z_min", "z_max")}
        for iy in range(ny):
            for iz in range(nz):
                cy0, cy1, cz0, cz1 = y0 + iy * dy, y0 + (iy + 1) * dy, z0 + iz * dz, z0 + (iz + 1) * dz
                a, b, c, dd = (x0, cy0, cz0), (x0, cy1, cz0), (x0, cy1, cz1), (x0, cy0, cz1)
                f["left"].append(_quad(tab, ele, (a, dd, c, b)))
        for iy in range(ny):
            for iz in range(nz):
                cy0, cy1, cz0, cz1 = y0 + iy * dy, y0 + (iy + 1) * dy, z0 + iz * dz, z0 + (iz + 1) * dz
                e, ff, g, h = (x1, cy0, cz0), (x1, cy1, cz0), (x1, cy1, cz1), (x1, cy0, cz1)
                f["right"].append(_quad(tab, ele, (e, ff, g, h)))
        for ix in range(nx):
            for iz in range(nz):
                cx0, cx1, cz0, cz1 = x0 + ix * dx, x0 + (ix + 1) * dx, z0 + iz * dz, z0 + (iz + 1) * dz
                a, e, h, dd = (cx0, y0, cz0), (cx1, y0, cz0), (cx1, y0, cz1), (cx0, y0, cz1)
                f["y_min"].append(_quad(tab, ele, (a, e, h, dd)))
        for ix in range(nx):
            for iz in range(nz):
                cx0, cx1, cz0, cz1 = x0 + ix * dx, x0 + (ix + 1) * dx, z0 + iz * dz, z0 + (iz + 1) * dz
                b, ff, g, c = (cx0, y1, cz0), (cx1, y1, cz0), (cx1, y1, cz1), (cx0, y1, cz1)
                f["y_max"].append(_quad(tab, ele, (b, c, g, ff)))
        for ix in range(nx):
            for iy in range(ny):
                cx0, cx1, cy0, cy1 = x0 + ix * dx, x0 + (ix + 1) * dx, y0 + iy * dy, y0 + (iy + 1) * dy
                a, e, ff, b = (cx0, cy0, z0), (cx1, cy0, z0), (cx1, cy1, z0), (cx0, cy1, z0)
                f["z_min"].append(_quad(tab, ele, (a, b, ff, e)))
        for ix in range(nx):
            for iy in range(ny):
                cx0, cx1, cy0, cy1 = x0 + ix * dx, x0 + (ix + 1) * dx, y0 + iy * dy, y0 + (iy + 1) * dy
                dd, h, g, c = (cx0, cy0, z1), (cx1, cy0, z1), (cx1, cy1, z1), (cx0, cy1, z1)
                f["z_max"].append(_quad(tab, ele, (dd, h, g, c)))
        faces.append(f)
    itf = []
    for d in range(domain_count - 1):
        for ea, eb in zip(faces[d]["right"], faces[d + 1]["left"]):
            itf.append((d + 1, ea + 1, d + 2, eb + 1, -1))
    ne = len(ele)
    bcid = np.full(ne, 456, np.int32)
    bottom = np.array(faces[0]["left"], np.int64)
    bcid[bottom] = 123
    top = np.array(faces[-1]["right"], np.int64)
    return BoxMesh(points=np.array(tab.points, np.float64), ele=np.array(ele, np.int32), bcid=bcid,
                   domain_count=domain_count, interfaces=np.array(itf, np.int64).reshape(-1, 5), top=top,
                   bottom=bottom, faces=faces)


def elements_per_domain(nx: int, ny: int, nz: int) -> int:
    return 2 * ny * nz + 2 * nx * nz + 2 * nx * ny


def wave_speeds(E: float, v: float, rho: float) -> Tuple[float, float]:
    G = E / 2.0 / (1.0 + v)
    return math.sqrt((2.0 - 2.0 * v) / (1.0 - 2.0 * v) * G / rho), math.sqrt(G / rho)


def ricker(t: np.ndarray, f0: float, t0: Optional[float] = None) -> np.ndarray:
    t0 = 1.0 / f0 if t0 is None else t0
    a = (math.pi * f0 * (t - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def normal_load_history(ne: int, nstep: int, dt: float, elements: Sequence[int], signal: np.ndarray,
                        direction=(0.0, 0.0, 1.0)) -> np.ndarray:
    """(nstep+1, ne, 3) given triplets: signal[k] * direction (element frame) on the listed elements."""
    lu = np.zeros((nstep + 1, ne, 3))
    lu[:, np.asarray(elements, np.int64), :] = signal[:, None, None] * np.asarray(direction, np.float64)[None, None, :]
    return lu


# ---------------------------------------------------------------------------
# Named workloads.  Seeds: numpy default_rng(config index).
# ---------------------------------------------------------------------------
@dataclass
class Workload:
    name: str
    mesh: BoxMesh
    mats: np.ndarray      # (D, 4) rows (id 1-based, E, v, rho)
    dt: float
    nstep: int
    known: np.ndarray     # (nstep+1, ne, 3)
    gap: float = 0.4


def _heaviside(nstep: int, amp: float = 1.0) -> np.ndarray:
    s = np.full(nstep + 1, amp)
    s[0] = 0.0
    return s


def rod_config1(nstep: int = 100) -> Workload:
    """1x1x5 rod of 5 unit-cube domains, h = 0.5 (120 elements), E=20, v=0, rho=20 (c1 = 1),
    fixed end and a unit Heaviside normal traction on the far end, dt = 0.6 h / c1."""
    m = box_mesh(5, 5.0, 1.0, 1.0, 2, 2, 2)
    mats = np.array([[d + 1, 20.0, 0.0, 20.0] for d in range(5)])
    dt = 0.6 * 0.5 / 1.0
    known = normal_load_history(len(m.ele), nstep, dt, m.top, _heaviside(nstep))
    return Workload("rod_1x1x5_h0.5", m, mats, dt, nstep, known)


def reference_rod_case(dt: float, domain_count: int = 5, length: float = 10.0, width: float = 1.0,
                       height: float = 1.0, nx: int = 2, ny: int = 2, nz: int = 2, nstep: int = 100,
                       amplitude: float = 1.0, E: float = 20.0, v: float = 0.0, rho: float = 20.0) -> Workload:
    """The reference generator's multi-domain rod (tools/generate_multidomain_rod.py:148-474, 504-575):
    domain_count stacked prisms of length/domain_count x width x height, faces subdivided nx / ny /
    nz, the left end fixed (BCID 123), a constant normal traction `amplitude` on the right end from
    step 1 (the generator's x-axis load in the element frame).  With the defaults this is the
    reference's rod120_100 case (120 elements, 5 domains, 100 steps); dt is the caller's (the
    reference derives it from the datacard's beta and the mesh)."""
    m = box_mesh(domain_count, length, width, height, nx, ny, nz)
    mats = np.array([[d + 1, E, v, rho] for d in range(domain_count)])
    sig = _heaviside(nstep, amplitude)
    known = normal_load_history(len(m.ele), nstep, dt, m.top, sig)
    return Workload(f"rod{len(m.ele)}_{nstep}", m, mats, dt, nstep, known)


def layered_config2(h: float = 0.1, nstep: int = 400, seed: int = 1) -> Workload:
    """2x2x4 box of 4 stacked 1x2x2 domains, seeded E in [1,10], v in [0.2,0.35], rho in [1,3];
    Ricker normal traction on the top face, bottom fixed, dt = 0.5 h / max c1."""
    n_ax = int(round(1.0 / h))
    n_tr = int(round(2.0 / h))
    m = box_mesh(4, 4.0, 2.0, 2.0, n_ax, n_tr, n_tr)
    rng = np.random.default_rng(seed)
    E = rng.uniform(1.0, 10.0, 4)
    v = rng.uniform(0.2, 0.35, 4)
    rho = rng.uniform(1.0, 3.0, 4)
    c1max = max(wave_speeds(E[d], v[d], rho[d])[0] for d in range(4))
    dt = 0.5 * h / c1max
    f0 = rng.uniform(0.5, 1.5) / (40 * dt)
    sig = ricker(np.arange(nstep + 1) * dt, f0)
    sig[0] = 0.0
    known = normal_load_history(len(m.ele), nstep, dt, m.top, sig)
    mats = np.array([[d + 1, E[d], v[d], rho[d]] for d in range(4)])
    return Workload(f"layered_2x2x4_h{h:g}", m, mats, dt, nstep, known)


def mixed_config3(h: float = 0.125, nstep: int = 600, seed: int = 2) -> Workload:
    """2x2x8 box of 8 stacked 1x2x2 domains (h = 0.125 gives 8192 elements), seeded E in [1,20],
    v in [0.15,0.4], rho in [1,4]; Heaviside normal traction on the top face."""
    n_ax = int(round(1.0 / h))
    n_tr = int(round(2.0 / h))
    m = box_mesh(8, 8.0, 2.0, 2.0, n_ax, n_tr, n_tr)
    rng = np.random.default_rng(seed)
    E = rng.uniform(1.0, 20.0, 8)
    v = rng.uniform(0.15, 0.4, 8)
    rho = rng.uniform(1.0, 4.0, 8)
    c1max = max(wave_speeds(E[d], v[d], rho[d])[0] for d in range(8))
    dt = 0.5 * h / c1max
    known = normal_load_history(len(m.ele), nstep, dt, m.top, _heaviside(nstep))
    mats = np.array([[d + 1, E[d], v[d], rho[d]] for d in range(8)])
    return Workload(f"stacked_2x2x8_h{h:g}", m, mats, dt, nstep, known)


def models_config4ii(n_models: int = 16, nstep: int = 200, seed: int = 4, lo: int = 500,
                     hi: int = 6000) -> List[Workload]:
    """Independent single-domain unit-cube models: element counts drawn uniformly from [lo, hi] and
    rounded to the nearest cube mesh (6 n^2 elements, h = 1/n), seeded E in [1,10], v in [0.2,0.35],
    rho in [1,3], a Heaviside or Ricker normal traction on the top face, bottom fixed,
    dt = 0.5 h / c1 per model."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_models):
        target = int(rng.integers(lo, hi + 1))
        n = max(2, int(round(math.sqrt(target / 6.0))))
        m = box_mesh(1, 1.0, 1.0, 1.0, n, n, n)
        E, v, rho = rng.uniform(1.0, 10.0), rng.uniform(0.2, 0.35), rng.uniform(1.0, 3.0)
        dt = 0.5 / n / wave_speeds(E, v, rho)[0]
        if rng.random() < 0.5:
            sig = _heaviside(nstep, rng.uniform(0.5, 2.0))
        else:
            sig = rng.uniform(0.5, 2.0) * ricker(np.arange(nstep + 1) * dt, rng.uniform(0.5, 1.5) / (40 * dt))
            sig[0] = 0.0
        known = normal_load_history(len(m.ele), nstep, dt, m.top, sig)
        mats = np.array([[1, E, v, rho]])
        out.append(Workload(f"cube_model{i}_n{n}", m, mats, dt, nstep, known))
    return out


def load_cases(w: Workload, nrhs: int = 64, seed: int = 4) -> np.ndarray:
    """(nrhs, nstep+1, ne, 3) seeded load cases on the workload's mesh: each a
    Ricker (centre frequency in [0.5, 1.5] / (40 dt)) or Heaviside pulse on a
    random patch of the top face (the elements within a random radius of a
    random top element), with a random direction (element frame) and an
    amplitude in [0.5, 2]."""
    m = w.mesh
    ne = len(m.ele)
    rng = np.random.default_rng(seed)
    top = np.asarray(m.top, np.int64)
    centre = m.points[m.ele[top, :4]].mean(axis=1)
    span = float(np.ptp(centre, axis=0).max()) or 1.0
    t = np.arange(w.nstep + 1) * w.dt
    out = np.zeros((nrhs, w.nstep + 1, ne, 3))
    for b in range(nrhs):
        c = centre[rng.integers(len(top))]
        radius = rng.uniform(0.15, 0.5) * span
        patch = top[np.linalg.norm(centre - c, axis=1) <= radius]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        amp = rng.uniform(0.5, 2.0)
        if rng.random() < 0.5:
            sig = ricker(t, rng.uniform(0.5, 1.5) / (40 * w.dt))
            sig[0] = 0.0
        else:
            sig = _heaviside(w.nstep)
        out[b] = normal_load_history(ne, w.nstep, w.dt, patch, amp * sig, direction=d)
    return out


@dataclass
class AdjointWorkload:
    workload: Workload
    receivers: np.ndarray   # element ids
    obs_mats: np.ndarray    # materials of the observed run (id, E, v, rho)


def adjoint_config5(h: float = 1.0 / 12.0, nstep: int = 300, seed: int = 5, nrec: int = 32) -> AdjointWorkload:
    """1x1x4 block of 4 unit-cube domains (h = 1/12 gives 3456 elements), seeded E in [1,10], v in
    [0.2,0.35], rho in [1,3], Ricker normal traction on the top face, dt = 0.5 h / max c1; nrec
    seeded receiver elements; observations come from the same run with every material value
    perturbed by a seeded factor in [0.9, 1.1]."""
    n = int(round(1.0 / h))
    m = box_mesh(4, 4.0, 1.0, 1.0, n, n, n)
    rng = np.random.default_rng(seed)
    E = rng.uniform(1.0, 10.0, 4)
    v = rng.uniform(0.2, 0.35, 4)
    rho = rng.uniform(1.0, 3.0, 4)
    c1max = max(wave_speeds(E[d], v[d], rho[d])[0] for d in range(4))
    dt = 0.5 * h / c1max
    f0 = rng.uniform(0.5, 1.5) / (40 * dt)
    sig = ricker(np.arange(nstep + 1) * dt, f0)
    sig[0] = 0.0
    known = normal_load_history(len(m.ele), nstep, dt, m.top, sig)
    mats = np.array([[d + 1, E[d], v[d], rho[d]] for d in range(4)])
    recv = np.sort(rng.choice(len(m.ele), nrec, replace=False))
    obs = mats.copy()
    obs[:, 1:] *= rng.uniform(0.9, 1.1, size=(4, 3))
    return AdjointWorkload(Workload(f"block_1x1x4_h{h:.4g}", m, mats, dt, nstep, known), recv, obs)
