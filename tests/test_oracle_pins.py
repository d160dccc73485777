"""Pins of the CPU oracle against the reference's own data and against
closed forms.  These run without a GPU."""
import math

import numpy as np
import pytest

from conftest import GOLDEN


def _gauss_rows(path, order):
    vals = [float(x.strip().rstrip(",")) for x in path.read_text().split() if x.strip().rstrip(",")]
    return vals[40 * (order - 1): 40 * (order - 1) + order]


def test_gauss_tables_match_reference_files_and_legendre():
    """Orders 3 and 4 of DBEM1/gaussposition / gaussweight (read at offset
    40*(order-1), DBEM1/gaussnumematrix.cpp:17-32) agree with Gauss-Legendre
    to the 15 printed decimals."""
    for order in (3, 4):
        pos = _gauss_rows(GOLDEN / "gaussposition", order)
        w = _gauss_rows(GOLDEN / "gaussweight", order)
        x, wl = np.polynomial.legendre.leggauss(order)
        assert np.max(np.abs(np.array(pos) - x)) < 1e-15
        assert np.max(np.abs(np.array(w) - wl)) < 1e-15
    # the weight string keeps 0.555555555555555, not 5/9 (quirk kept by the oracle)
    assert _gauss_rows(GOLDEN / "gaussweight", 3)[0] == 0.555555555555555


def _small_case(O, E=20.0, v=0.25, rho=2.0, dt=0.05):
    from tdbem import mesh
    m = mesh.box_mesh(2, 2.0, 1.0, 1.0, 2, 2, 2)
    mats = np.array([[E, v, rho], [E, v, rho]])
    itf = m.interfaces.copy()
    itf[:, [0, 1, 2, 3]] -= 1
    return m, O.Case(m.points, m.ele, m.bcid, mats, itf.astype(np.int32), dt, 0.4)


def test_dynamat_closed_forms(oracle_mod):
    m, c = _small_case(oracle_mod, E=20.0, v=0.25, rho=2.0, dt=0.05)
    d = c.dynamat(0)
    G = 20.0 / 2.0 / 1.25
    c1 = math.sqrt((2 - 0.5) / (1 - 0.5) * G / 2.0)
    c2 = math.sqrt(G / 2.0)
    assert d[0] == pytest.approx(G, rel=1e-15)
    assert d[3] == pytest.approx(c1, rel=1e-15)
    assert d[4] == pytest.approx(c2, rel=1e-15)
    assert d[17] == pytest.approx(c1 * 0.05, rel=1e-15)
    assert d[21] == pytest.approx(0.5 * (1 / c2 ** 2 - 1 / c1 ** 2), rel=1e-14)


def test_u_kernel_telescopes_to_kelvin(oracle_mod):
    """Sum over n >= 0 of the time-integrated U(n) equals the static Kelvin
    solution 1/(16 pi (1-v) G r) [(3-4v) delta + r_i r_j / r^2]
    (GetStaticUij, DBEM1/dsquareelement.cpp:577-607): the front windows and
    the between-fronts series sum to one and to (1/c2^2-1/c1^2)/2."""
    m, c = _small_case(oracle_mod, E=20.0, v=0.25, rho=2.0, dt=0.05)
    d = c.dynamat(0)
    G, v = d[0], d[1]
    rng = np.random.default_rng(0)
    for _ in range(5):
        ri = rng.normal(size=3)
        ri *= rng.uniform(0.3, 2.0) / np.linalg.norm(ri)
        r = np.linalg.norm(ri)
        tot = np.zeros(9)
        for n in range(0, 400):
            tot += c.kernel(0, 0, n, ri, np.array([0.0, 0.0, 1.0]))[1]
        kel = (1.0 / (16 * math.pi * (1 - v) * G * r)) * ((3 - 4 * v) * np.eye(3) + np.outer(ri, ri) / r ** 2)
        np.testing.assert_allclose(tot.reshape(3, 3), kel, rtol=1e-11, atol=1e-14)
    # a dropped term would break it: the P-front alone is not the Kelvin solution
    f, u = c.kernel(0, 0, 0, np.array([0.001, 0.0, 0.0]), np.array([0.0, 0.0, 1.0]))
    assert f >= 1


def test_needcal_is_exact_cull(oracle_mod):
    """NeedCal = 0 means no quadrature point of the element sees either
    front or the between-fronts window: brute force over a dense sampling of
    the element (DBEM1/dsquareelement.cpp:1156-1236)."""
    m, c = _small_case(oracle_mod)
    d = c.dynamat(0)
    c1, c2, dt = d[3], d[4], d[5]
    rng = np.random.default_rng(1)
    ne = len(m.ele)
    checked = 0
    for _ in range(60):
        f, s = rng.integers(0, ne // 2, 2)
        centre, radius, _ = c.element_info(int(f))
        src, _, _ = c.element_info(int(s))
        corners = m.points[m.ele[f][:4]]
        for n in range(0, 40):
            if c.needcal(int(f), int(s), 1, n) != 0:
                continue
            checked += 1
            uu = np.linspace(0, 1, 24)
            a, b = np.meshgrid(uu, uu)
            p = (corners[0][None, None] * ((1 - a) * (1 - b))[..., None] + corners[1][None, None] * (a * (1 - b))[..., None]
                 + corners[2][None, None] * (a * b)[..., None] + corners[3][None, None] * ((1 - a) * b)[..., None])
            r = np.linalg.norm(p - src[None, None], axis=-1)
            t0, t1 = n * dt, (n + 1) * dt
            # zero iff the interval [t0, t1) misses [r/c1, r/c2] for every point
            hit = (r / c1 < t1) & (r / c2 > t0)
            assert not hit.any()
    assert checked > 100


def test_dofmap_counts_on_rod_fixture(oracle_mod, rod_fixture):
    """Interfaces allocate (U, T) first, then outer elements; 120 elements,
    16 interface pairs -> 120 unknown / 88 known / 104 history blocks."""
    f = rod_fixture
    O = oracle_mod
    dt = O.reference_dt(f["pts"], f["ele"], f["card"].beta, f["E_global"], f["card"].v, f["rho_global"])
    c = O.Case(f["pts"], f["ele"], f["bcid"], f["mats"], f["itf"], dt, f["card"].gap)
    d = c.dofmap()
    assert d["ok"]
    assert d["counts"].tolist() == [120, 88, 120, 84, 4, 16, 16, 104, 104]
    itf = f["itf"]
    # interface pair i owns unknown blocks 2i (U) and 2i+1 (T); B side traction sign -1
    for i, (da, ea, db, eb, s) in enumerate(itf):
        assert d["u_block"][ea] == d["u_block"][eb] == 2 * i
        assert d["t_block"][ea] == d["t_block"][eb] == 2 * i + 1
        assert d["t_sign"][eb] == s and d["t_sign"][ea] == 1.0
        assert d["preferred"][ea] == 2 * i and d["preferred"][eb] == 2 * i + 1
    assert sorted(d["preferred"].tolist()) == list(range(120))


def test_rod_known_answer(oracle_mod, rod_fixture):
    """DBEM1/analytic_validation.txt: the reference's own run of rod120_100
    reports the loaded-end mean ux at step 100 as 0.75914256602696095."""
    f = rod_fixture
    O = oracle_mod
    card = f["card"]
    dt = O.reference_dt(f["pts"], f["ele"], card.beta, f["E_global"], card.v, f["rho_global"])
    assert dt * 100 == pytest.approx(float(f["expected"]["time"]), rel=1e-13)
    c = O.Case(f["pts"], f["ele"], f["bcid"], f["mats"], f["itf"], dt, card.gap)
    U, T, iters, flags, solved = c.solve(f["lu"], card.nstep, restart=card.iterations, tol=card.error)
    assert solved == card.nstep and not flags.any()
    xs = np.array([c.element_info(e)[0][0] for e in range(len(f["ele"]))])
    sel = [e for e in range(len(f["ele"])) if f["bcid"][e] == 456 and abs(xs[e] - xs.max()) < 1e-8]
    assert sel == [100, 101, 102, 103]
    num = U[100, sel, 0].mean()
    assert abs(num - float(f["expected"]["numeric"])) < 1e-12
    # interface continuity is exact for these axis-aligned frames
    for da, ea, db, eb, s in f["itf"]:
        assert np.array_equal(U[:, ea], U[:, eb])
        assert np.array_equal(T[:, eb], s * T[:, ea])


def test_gmres_matches_dense_solve(oracle_mod):
    rng = np.random.default_rng(3)
    nb = 12
    A = np.eye(3 * nb) * 4 + rng.normal(size=(3 * nb, 3 * nb)) * 0.3
    rp = np.arange(nb + 1, dtype=np.int32) * nb
    col = np.tile(np.arange(nb, dtype=np.int32), nb)
    val = np.zeros((nb * nb, 3, 3))
    for i in range(nb):
        for j in range(nb):
            val[i * nb + j] = A[3 * i:3 * i + 3, 3 * j:3 * j + 3]
    b = rng.normal(size=3 * nb)
    x, it, flag = oracle_mod.gmres_ccsr(rp, col, val, b, restart=60, tol=1e-13)
    assert flag == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_rod_series_config1(oracle_mod):
    """Config-1 rod (1x1x5, h = 0.5, c1 = 1): the loaded-end displacement
    follows the 1-D fixed-free rod series u(L,t) = PL/E - 2P/(EL) sum cos(c k t)/k^2
    (tools/run_rod120_100_analytic_validation.py:165-176) within the
    reference harness tolerance (abs 0.25 or rel 0.3)."""
    from tdbem import io, mesh
    w = mesh.rod_config1(nstep=60)
    m = w.mesh
    card = io.DataCard(title="x", thread_num=1, flag_dyna=2, beta=0.6, nstep=w.nstep, gap=0.4, E=20, v=0, rho=20,
                       amplitude=1, solver=2, domain_count=5,
                       materials=[(int(r[0]), r[1], r[2], r[3]) for r in w.mats.tolist()],
                       interfaces=[tuple(int(x) for x in r) for r in m.interfaces.tolist()])
    mats, itf = io.normalise_model(card, len(m.ele))
    c = oracle_mod.Case(m.points, m.ele, m.bcid, mats, itf, w.dt, 0.4)
    U, T, it, fl, solved = c.solve(w.known, w.nstep, tol=1e-8)
    assert solved == w.nstep
    L, E = 5.0, 20.0
    for k in range(5, w.nstep + 1, 5):
        t = k * w.dt
        s = sum(math.cos(1.0 * ((n + 0.5) * math.pi / L) * t) / ((n + 0.5) * math.pi / L) ** 2 for n in range(4000))
        ana = L / E - 2 / (E * L) * s
        num = U[k, m.top, 0].mean()
        assert abs(num - ana) < 0.25 or abs(num - ana) < 0.3 * abs(ana)


def test_t_kernels_telescope_to_static_traction(oracle_mod):
    """Sum over n of T1(n) + T2(n) equals the Kelvin traction kernel
    T_ij = -1/(8 pi (1-v) r^2) [dr/dn ((1-2v) d_ij + 3 r,i r,j) - (1-2v)(r,i n_j - r,j n_i)]
    with i and j exchanged, the convention the reference notes for Dij..Hij
    (DBEM1/dsquareelement.h:732-893).  A sign slip in any front, window or
    the linear time-shape factor breaks the identity."""
    m, c = _small_case(oracle_mod, E=20.0, v=0.25, rho=2.0, dt=0.05)
    d = c.dynamat(0)
    v = d[1]
    rng = np.random.default_rng(5)
    for _ in range(4):
        ri = rng.normal(size=3)
        ri *= rng.uniform(0.3, 2.0) / np.linalg.norm(ri)
        nor = rng.normal(size=3)
        nor /= np.linalg.norm(nor)
        S = np.zeros(9)
        for n in range(0, 400):
            S += c.kernel(0, 1, n, ri, nor)[1] + c.kernel(0, 2, n, ri, nor)[1]
        r = np.linalg.norm(ri)
        rh = ri / r
        drdn = rh @ nor
        k = -1.0 / (8 * math.pi * (1 - v) * r ** 2)
        Tst = k * (drdn * ((1 - 2 * v) * np.eye(3) + 3 * np.outer(rh, rh))
                   - (1 - 2 * v) * (np.outer(rh, nor) - np.outer(nor, rh)))
        np.testing.assert_allclose(S.reshape(3, 3), Tst.T, rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------------------
# Regular-path quadrature (quirks Q1 and Q2).  The regular 4x4 rule
# (IntDynaUij / IntDynaTij, DBEM1/dsquareelement.cpp:1955-2009,2233-2290) is
# taken when a whole element stays between the P and S fronts for a whole step
# (NeedCal = 2).  It weights every point by the 8-node serendipity shape
# function N0(s1, s2; gap) of corner node 0 (m_NRGV[0], DBEM1/quadinfo.h:
# 149-172, DBEM1/quadinfo.cpp:92-97) and sums T1(n) + T2(n) whatever typeT is.
# The pins below restate that integral independently of the oracle: the
# textbook serendipity function, the element map from the mesh corners, and
# Gauss-Legendre rules from numpy, with only the point kernels (themselves
# pinned by the telescoping identities above) taken from the oracle.
# ---------------------------------------------------------------------------
_GAP = 0.4


def _serendipity(s1, s2, a):
    """The eight textbook 8-node serendipity shape functions on [-a, a]^2,
    nodes ordered corners (-,-), (+,-), (+,+), (-,+), then the edge midpoints
    (0,-), (+,0), (0,+), (-,0) (DBEM1/quadinfo.h:80-107)."""
    x, y = s1 / a, s2 / a
    out = []
    for xi, yi in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
        out.append(0.25 * (1 + x * xi) * (1 + y * yi) * (x * xi + y * yi - 1))
    out.append(0.5 * (1 - x * x) * (1 - y))
    out.append(0.5 * (1 + x) * (1 - y * y))
    out.append(0.5 * (1 - x * x) * (1 + y))
    out.append(0.5 * (1 - x) * (1 - y * y))
    return np.array(out)


def _regular_probe_case(O):
    """Two 2x1x1 domains with v = 0.45 (c1/c2 = 3.3) and dt = 0.02: many
    element / source / step triples lie wholly between the fronts."""
    from tdbem import mesh
    m = mesh.box_mesh(2, 4.0, 1.0, 1.0, 8, 2, 2)
    mats = np.array([[10.0, 0.45, 1.0], [10.0, 0.45, 1.0]])
    itf = m.interfaces.copy()
    itf[:, [0, 1, 2, 3]] -= 1
    return m, O.Case(m.points, m.ele, m.bcid, mats, itf.astype(np.int32), 0.02, _GAP)


def _regular_triples(c, ne, nstep=30):
    half = ne // 2
    return [(f, s, k) for f in range(half) for s in range(0, half, 3) for k in range(1, nstep)
            if c.needcal(f, s, 1, k) == 2]


def _element_rule(m, c, f, s, k, which, x, w, weight):
    """sum over the tensor rule (x, w) of weight(s1, s2) * J * K(n = k) on the
    element's bilinear map from its mesh corners (flat square elements, so the
    8-node geometry reduces to it); which 'U' -> U, 'T' -> T1 + T2, 'T1', 'T2'."""
    cor = m.points[m.ele[f][:4]]
    src = c.element_info(s)[0]
    tot = np.zeros(9)
    for i, s2 in enumerate(x):
        for j, s1 in enumerate(x):
            a, b = 0.5 * (1 + s1), 0.5 * (1 + s2)
            p = cor[0] * (1 - a) * (1 - b) + cor[1] * a * (1 - b) + cor[2] * a * b + cor[3] * (1 - a) * b
            d1 = 0.5 * (-(1 - b) * cor[0] + (1 - b) * cor[1] + b * cor[2] - b * cor[3])
            d2 = 0.5 * (-(1 - a) * cor[0] - a * cor[1] + a * cor[2] + (1 - a) * cor[3])
            nn = np.cross(d1, d2)
            J = np.linalg.norm(nn)
            nn = nn / J
            K = np.zeros(9)
            if which == "U":
                K = c.kernel(0, 0, k, p - src, nn)[1]
            if which in ("T", "T1"):
                K = K + c.kernel(0, 1, k, p - src, nn)[1]
            if which in ("T", "T2"):
                K = K + c.kernel(0, 2, k, p - src, nn)[1]
            tot += K * J * w[i] * w[j] * weight(s1, s2)
    return tot


def test_regular_weights_are_serendipity_n0(oracle_mod):
    """Q1: the oracle's regular weights are N0(s1, s2; gap) times the 4x4
    Gauss weight at the 16 Gauss points, with N0 the textbook serendipity
    corner function (nodal at the eight nodes at +-gap, partition of unity).
    Weight 1 (the patch-wise convention) differs by O(1)."""
    m, c = _regular_probe_case(oracle_mod)
    nodes = np.array([(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0)], float) * _GAP
    for j, (a, b) in enumerate(nodes):
        np.testing.assert_allclose(_serendipity(a, b, _GAP), np.eye(8)[j], atol=1e-15)
    rng = np.random.default_rng(7)
    for a, b in rng.uniform(-1, 1, size=(20, 2)):
        assert abs(_serendipity(a, b, _GAP).sum() - 1.0) < 1e-14
    pos, ra, nrgv0 = c.regular_table()
    x, w = np.polynomial.legendre.leggauss(4)
    np.testing.assert_allclose(pos[:, 0], np.tile(x, 4), atol=1e-15)
    np.testing.assert_allclose(pos[:, 1], np.repeat(x, 4), atol=1e-15)
    np.testing.assert_allclose(ra, np.outer(w, w).ravel(), atol=1e-15)
    n0 = np.array([_serendipity(a, b, _GAP)[0] for a, b in pos])
    np.testing.assert_allclose(nrgv0, n0 * ra, rtol=1e-14, atol=1e-16)
    assert _serendipity(0.0, 0.0, _GAP)[0] == -0.25
    assert np.abs(nrgv0 - ra).max() > 0.1


def test_regular_quadrature_matches_independent_rules(oracle_mod):
    """On triples with NeedCal = 2 the oracle's judged U block equals the 4x4
    rule of N0 * w * J * U(n) and its judged T block (typeT 1 and 2 alike)
    the 4x4 rule of N0 * w * J * (T1(n) + T2(n)), restated here from the mesh
    corners (<= 1e-13 of the block); both are within 3e-4 of a 24x24 rule of
    the same integrand (the 4x4 rule's own error), and a weight of 1, T1 alone
    or T2 alone would miss by more than 5 %."""
    m, c = _regular_probe_case(oracle_mod)
    trip = _regular_triples(c, len(m.ele))
    assert len(trip) > 1000
    rng = np.random.default_rng(11)
    x4, w4 = np.polynomial.legendre.leggauss(4)
    x24, w24 = np.polynomial.legendre.leggauss(24)
    n0 = lambda a, b: _serendipity(a, b, _GAP)[0]
    one = lambda a, b: 1.0
    for t in rng.choice(len(trip), 5, replace=False):
        f, s, k = trip[t]
        flag, qu = c.quadrature(f, s, 0, k)
        assert flag == 2
        flag, qt = c.quadrature(f, s, 1, k)
        assert flag == 2
        su, st = np.abs(qu).max(), np.abs(qt).max()
        assert np.abs(qu - _element_rule(m, c, f, s, k, "U", x4, w4, n0)).max() <= 1e-13 * su
        assert np.abs(qt - _element_rule(m, c, f, s, k, "T", x4, w4, n0)).max() <= 1e-13 * st
        if c.needcal(f, s, 2, k) == 2:
            assert np.array_equal(c.quadrature(f, s, 2, k)[1], qt)
        assert np.abs(qu - _element_rule(m, c, f, s, k, "U", x24, w24, n0)).max() <= 3e-4 * su
        assert np.abs(qt - _element_rule(m, c, f, s, k, "T", x24, w24, n0)).max() <= 3e-4 * st
        assert np.abs(qu - _element_rule(m, c, f, s, k, "U", x4, w4, one)).max() > 0.05 * su
        assert np.abs(qt - _element_rule(m, c, f, s, k, "T", x4, w4, one)).max() > 0.05 * st
        assert np.abs(qt - _element_rule(m, c, f, s, k, "T1", x4, w4, n0)).max() > 0.05 * st
        assert np.abs(qt - _element_rule(m, c, f, s, k, "T2", x4, w4, n0)).max() > 0.05 * st


def test_history_blocks_double_the_regular_t(oracle_mod):
    """Q2 at the assembly level: where both T judges of a history entry take the
    regular path, MTS[k] receives the typeT = 2 call and then the typeT = 1 call
    (DBEM1/multidomain.cpp:2053-2064), each T1(k) + T2(k), so the stored block is
    twice the regular T block (times the history DOF sign); MGS[k] holds the
    regular U block (DBEM1/multidomain.cpp:1137-1141 transposes; both blocks are
    symmetric here)."""
    m, c = _regular_probe_case(oracle_mod)
    trip = _regular_triples(c, len(m.ele))
    both = [t for t in trip if c.needcal(t[0], t[1], 2, t[2]) == 2]
    assert len(both) > 100
    c.assemble(29)
    d = c.dofmap()
    rng = np.random.default_rng(13)
    for t in rng.choice(len(both), 6, replace=False):
        f, s, k = both[t]
        for which_m, blk, sgn, which_q, factor in ((2, "hu_block", "hu_sign", 1, 2.0),
                                                   (3, "ht_block", "ht_sign", 0, 1.0)):
            rp, col, val = c.ccsr(which_m, k)
            hit = [i for i in range(rp[s], rp[s + 1]) if col[i] == d[blk][f]]
            assert len(hit) == 1
            q = c.quadrature(f, s, which_q, k)[1].reshape(3, 3)
            assert np.array_equal(val[hit[0]], factor * d[sgn][f] * q.T), (which_m, f, s, k)
