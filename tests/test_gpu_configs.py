"""The benchmark workloads at their stated sizes on the GPU, each checked
against the oracle where the oracle can follow (per-row history blocks,
small-case gradients) and otherwise through size-independent properties:

  1  rod 1x1x5, h = 0.5, 120 elements, 100 steps: every step against the oracle
  2  layered 2x2x4, h = 0.1, 6400 elements, 400 steps
  3  stacked 2x2x8, h = 0.125, 8192 elements, 600 steps (FP64 and mixed)
  4i 64 load cases on the config-2 mesh, one batch
  4ii 16 independent cube models of 500-6000 elements, 200 steps, together
  5  adjoint on the 1x1x4 block, h = 1/12, 3456 elements, 300 steps

The history blocks of sampled source rows are compared with the oracle's
assembly of the same rows (orc_history_rows): the pattern bit for bit, every
3x3 block to 1e-10 of its largest entry (the default quadrature sums its
points in another order than the reference), or of the row's largest entry
for blocks that are rounding noise (the self block of step 1 holds the
singular T2 table, which vanishes for a flat element up to ~1e-14 whose
digits depend on the summation order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tdbem():
    import tdbem
    if not tdbem.LIB_PATH.exists():
        pytest.fail("HIP library not built")
    return tdbem


def _gpu_case(w, **opts):
    tdbem = _tdbem()
    m = w.mesh
    c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=m.domain_count, mat_rows=w.mats,
                   itf_rows=m.interfaces)
    for k, v in opts.items():
        c.set_option(k, v)
    return c


def _oracle_case(O, w):
    m = w.mesh
    itf = m.interfaces.copy()
    itf[:, [0, 1, 2, 3]] -= 1
    return O.Case(m.points, m.ele, m.bcid, w.mats[:, 1:].copy(), itf.astype(np.int32), w.dt, w.gap)


def _blocks_close(a, b, rtol=1e-10):
    """Same pattern; each block within rtol of max(its largest entry, the
    largest entry of its row)."""
    assert np.array_equal(a[0], b[0]), "row_ptr"
    assert np.array_equal(a[1], b[1]), "columns"
    if len(b[2]) == 0:
        return 0.0
    vb = b[2].reshape(len(b[2]), -1)
    own = np.abs(vb).max(1)
    rowmax = np.zeros(len(b[0]) - 1)
    for r in range(len(b[0]) - 1):
        if b[0][r + 1] > b[0][r]:
            rowmax[r] = own[b[0][r]:b[0][r + 1]].max()
    rowof = np.repeat(np.arange(len(b[0]) - 1), np.diff(b[0]))
    rel = np.abs(a[2].reshape(len(b[2]), -1) - vb).max(1) / np.maximum(np.maximum(own, rowmax[rowof]), 1e-300)
    assert rel.max() <= rtol, rel.max()
    return float(rel.max())


def _sampled_rows_vs_oracle(O, w, gc, steps, rows_per_domain=2, seed=0):
    """History rows of sampled sources at sampled steps, GPU vs oracle;
    returns (blocks compared, regular-path entries among them)."""
    oc = _oracle_case(O, w)
    m = w.mesh
    nd = m.domain_count
    n = len(m.ele) // nd
    rng = np.random.default_rng(seed)
    rows = []
    for d in range(nd):
        rows.append(d * n)
        rows.extend(int(x) for x in d * n + rng.integers(1, n, rows_per_domain - 1))
    blocks = regular = 0
    worst = 0.0
    for s in rows:
        for k in steps:
            for which in (2, 3):
                a, b = gc.ccsr_rows(which, k, s, s + 1), oc.history_rows(which, k, s, s + 1)
                worst = max(worst, _blocks_close(a, b))
                blocks += len(b[1])
        # how many of this row's entries take the regular 4x4 path (NeedCal 2)
        d = s // n
        for f in range(d * n, d * n + n, 7):
            for k in steps:
                regular += oc.needcal(f, s, 1, k) == 2
    print(f"{w.name}: {len(rows)} rows x {len(steps)} steps, {blocks} blocks, worst {worst:.2e}, "
          f"regular-path triples sampled {regular}", flush=True)
    return blocks, regular


def _properties(w, gc, U, T):
    """Zero load gives zero, the response is linear in the load, interface
    displacements agree across the two sides."""
    U2, T2, *_ = gc.solve(2.0 * w.known, w.nstep, tol=1e-12)
    assert np.abs(U2 - 2 * U).max() <= 1e-9 * np.abs(U).max()
    assert np.abs(T2 - 2 * T).max() <= 1e-9 * np.abs(T).max()
    Z, ZT, *_ = gc.solve(np.zeros_like(w.known), w.nstep, tol=1e-12)
    assert not Z.any() and not ZT.any()
    for da, ea, db, eb, s in w.mesh.interfaces:
        assert np.abs(U[:, ea - 1] - U[:, eb - 1]).max() <= 1e-12 * np.abs(U).max()


def test_config1_rod_every_step_against_the_oracle(oracle_mod):
    """Config 1 (1x1x5 rod, h = 0.5, 120 elements, 100 steps): MTS / MGS of
    every step and the displacements and tractions of every step against the
    oracle."""
    from tdbem import mesh
    w = mesh.rod_config1(nstep=100)
    print("case:", w.name, "all 100 steps", flush=True)
    gc = _gpu_case(w)
    oc = _oracle_case(oracle_mod, w)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    oc.assemble(w.nstep)
    for k in range(1, w.nstep + 1):
        for which in (2, 3):
            _blocks_close(gc.ccsr(which, k), oc.ccsr(which, k))
    U, T, it, fl, solved, _ = gc.solve(w.known, w.nstep, tol=1e-12)
    OU, OT, oit, ofl, osolved = oc.solve(w.known, w.nstep, tol=1e-12)
    gc.close()
    assert solved == osolved == w.nstep and not fl.any()
    assert np.abs(U - OU).max() <= 1e-9 * np.abs(OU).max()
    assert np.abs(T - OT).max() <= 1e-9 * np.abs(OT).max()


def test_config2_full_size(oracle_mod):
    """Config 2 (layered 2x2x4, h = 0.1, 6400 elements, 400 steps): sampled
    history rows against the oracle (regular-path entries among them), the
    solution's properties, MGS symmetry, and the time-blocked convolution
    against the per-step sum."""
    from tdbem import mesh
    w = mesh.layered_config2(h=0.1, nstep=400)
    print("case:", w.name, len(w.mesh.ele), "elements", flush=True)
    gc = _gpu_case(w)
    gc.assemble_step0(10)
    entries = gc.assemble_history(w.nstep)
    assert entries > 4e8
    blocks, regular = _sampled_rows_vs_oracle(oracle_mod, w, gc, steps=(1, 2, 3, 17, 64, 150, 300, 400))
    assert blocks > 5000 and regular > 0
    for k in (1, 40, 200):
        rp, col, val = gc.ccsr_rows(3, k, 0, 64)
        assert np.array_equal(val, np.transpose(val, (0, 2, 1)))
    U, T, it, fl, solved, _ = gc.solve(w.known, w.nstep, tol=1e-12)
    assert solved == w.nstep and not fl.any()
    _properties(w, gc, U, T)
    # the per-step sum reassociates the tiled one; each step's GMRES stops at
    # ||r|| <= 1e-12 ||b|| on an iterate that rounding may change, and the
    # difference carries through 400 steps of history (observed 1.6e-9)
    gc.set_option("conv_tiled", 0)
    U0, T0, *_ = gc.solve(w.known, w.nstep, tol=1e-12)
    gc.close()
    eu, et = np.abs(U - U0).max() / np.abs(U0).max(), np.abs(T - T0).max() / np.abs(T0).max()
    print("tiled vs per-step convolution: u", eu, "t", et, flush=True)
    assert eu <= 1e-8 and et <= 1e-8


def test_config3_full_size_fp64_and_mixed(oracle_mod):
    """Config 3 (stacked 2x2x8, h = 0.125, 8192 elements, 600 steps): sampled
    history rows against the oracle; the mixed modes against the FP64 run
    (all at GMRES tol 1e-12): FP32 + bf16 bands (mode 2) within 1e-6 relative
    displacement and traction error, FP32 bands (mode 1) within 5e-6 (it
    stays above 1e-6 on this case: the FP32 rounding of the blocks is
    amplified about 40x over 600 steps)."""
    from tdbem import mesh
    w = mesh.mixed_config3(h=0.125, nstep=600)
    print("case:", w.name, len(w.mesh.ele), "elements", flush=True)
    gc = _gpu_case(w)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    blocks, regular = _sampled_rows_vs_oracle(oracle_mod, w, gc, steps=(1, 2, 5, 33, 120, 300, 600), rows_per_domain=1)
    assert blocks > 2000
    U, T, it, fl, solved, _ = gc.solve(w.known, w.nstep, tol=1e-12)
    gc.close()
    assert solved == w.nstep and not fl.any()
    for da, ea, db, eb, s in w.mesh.interfaces:
        assert np.abs(U[:, ea - 1] - U[:, eb - 1]).max() <= 1e-12 * np.abs(U).max()
    err = {}
    for mixed in (1, 2):
        c = _gpu_case(w, mixed=mixed)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        Um, Tm, itm, flm, sm, _ = c.solve(w.known, w.nstep, tol=1e-12)
        c.close()
        assert sm == w.nstep
        err[mixed] = (np.abs(Um - U).max() / np.abs(U).max(), np.abs(Tm - T).max() / np.abs(T).max())
    print("mixed-precision error vs FP64 (u, t):", err, flush=True)
    assert max(err[2]) <= 1e-6
    assert max(err[1]) <= 5e-6


def test_config4i_64_load_cases_on_the_config2_mesh():
    """Config 4(i): 64 seeded load cases on the config-2 mesh solved as one
    batch; each sampled case equals its own single solve to rounding, as
    amplified by that case's own conditioning: both solves stop at
    ||r|| <= 1e-12 ||b|| but may stop on different iterations, the batch
    reassociates sums (its far part in the frequency domain), and the
    difference carries through 400 steps of history.  The bound is 200 x
    the change a 1e-15 relative perturbation of the case's load causes in
    its single solve (the batch differs like a 5e-14 perturbation: 25-60 x
    observed; case 63 responds 130 x more than cases 0 and 31), and at most
    1e-6."""
    from tdbem import mesh
    w = mesh.layered_config2(h=0.1, nstep=400)
    cases = mesh.load_cases(w, 64)
    print("case:", w.name, "64 load cases", flush=True)
    gc = _gpu_case(w)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    BU, BT, bit, bfl, bsolved, _ = gc.solve_batched(cases, w.nstep, tol=1e-12)
    assert (bsolved == w.nstep).all() and not bfl.any()
    for b in (0, 31, 63):
        U, T, it, fl, solved, _ = gc.solve(cases[b], w.nstep, tol=1e-12)
        assert solved == w.nstep
        Up, Tp, *_ = gc.solve(cases[b] * (1 + 1e-15), w.nstep, tol=1e-12)
        pu = np.abs(Up / (1 + 1e-15) - U).max() / np.abs(U).max()
        pt = np.abs(Tp / (1 + 1e-15) - T).max() / np.abs(T).max()
        eu, et = np.abs(BU[b] - U).max() / np.abs(U).max(), np.abs(BT[b] - T).max() / np.abs(T).max()
        print("case", b, "batch vs single: u", eu, "t", et, "response to a 1e-15 load change: u", pu, "t", pt,
              flush=True)
        assert eu <= min(1e-6, 200 * max(pu, 1e-14)) and et <= min(1e-6, 200 * max(pt, 1e-14)), b
    gc.close()


def test_config4ii_16_models_together():
    """Config 4(ii): 16 cube models of 600-5766 elements, 200 steps each,
    through tdbem_run_models (their bands, 1.3 TB in all, held within the
    device budget) and stacked in rounds (tdbem_run_models_grouped, one launch
    per kernel for a round's models): every model's grouped results equal the
    concurrent ones bit for bit, and the smallest, a middle and the largest
    model equal their own single jobs bit for bit."""
    tdbem = _tdbem()
    from tdbem import mesh
    ws = mesh.models_config4ii()
    sizes = [len(w.mesh.ele) for w in ws]
    print("case: 16 models", sizes, flush=True)
    cs = [_gpu_case(w) for w in ws]
    res = tdbem.run_models(cs, [w.known for w in ws], [w.nstep for w in ws])
    for c in cs:
        c.close()
    assert all(r[4] == w.nstep for r, w in zip(res, ws))
    print("case: 16 models grouped", flush=True)
    cs = [_gpu_case(w) for w in ws]
    resg = tdbem.run_models(cs, [w.known for w in ws], [w.nstep for w in ws], grouped=True)
    for c in cs:
        c.close()
    print("rounds:", tdbem.run_models.rounds, flush=True)
    for i, (a, b) in enumerate(zip(res, resg)):
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x), np.asarray(y)), i
    order = np.argsort(sizes)
    for i in (order[0], order[len(order) // 2], order[-1]):
        w = ws[i]
        c = _gpu_case(w)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        U, T, it, fl, solved, _ = c.solve(w.known, w.nstep)
        c.close()
        assert np.array_equal(U, res[i][0]) and np.array_equal(T, res[i][1]), i
        assert np.array_equal(it, res[i][2]) and solved == res[i][4], i
    tdbem.release_cached_memory()


def _misfit(U, recv, obs):
    return float(((U[1:, recv, :] - obs[1:]) ** 2).sum())


def test_config5_adjoint_gradient():
    """Config 5 (1x1x4 block, h = 1/12, 3456 elements, 300 steps, 32
    receivers, observations from materials perturbed by +-10 %): the adjoint
    gradient against central differences of the misfit (relative step 1e-5)
    along each domain's (E, rho) -> s (E, rho) direction, which keeps every
    wave front in place, so the misfit is smooth along it: within 1e-5 of the
    largest derivative.  Single components change the wave speeds; at this
    size some of the ~1e10 (quadrature point, lag) pairs cross a front window
    within any finite step, so the misfit's differences along them carry the
    kernel's jumps (the reference's piecewise kernels) and are not compared
    here; they are compared on small cases with the band structure frozen."""
    from tdbem import mesh
    a = mesh.adjoint_config5()
    w = a.workload
    print("case:", w.name, len(w.mesh.ele), "elements, adjoint", flush=True)

    def solved(mats, tol=1e-13):
        ww = mesh.Workload(w.name, w.mesh, mats, w.dt, w.nstep, w.known)
        c = _gpu_case(ww)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        U, *_ = c.solve(w.known, w.nstep, tol=tol)
        c.close()
        return U

    obs = solved(a.obs_mats)[:, a.receivers, :]
    c = _gpu_case(w)
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    J, g, lu = c.adjoint(w.known, w.nstep, a.receivers, obs, tol=1e-13, fd_step=1e-6)
    assert abs(J - _misfit(lu, a.receivers, obs)) <= 1e-12 * J
    evr = c.materials()
    ad, fd = np.zeros(4), np.zeros(4)
    for d in range(4):
        h = 1e-5
        Js = []
        for sgn in (1, -1):
            e = evr.copy()
            e[d, [0, 2]] *= 1 + sgn * h
            Js.append(_misfit(solved(np.column_stack([w.mats[:, 0], e])), a.receivers, obs))
        fd[d] = (Js[0] - Js[1]) / (2 * h)
        ad[d] = evr[d, 0] * g[d, 0] + evr[d, 2] * g[d, 2]
    print("dJ/ds adjoint", ad, "differences", fd, flush=True)
    assert np.abs(ad - fd).max() <= 1e-5 * np.abs(fd).max()
    c.close()
    assert np.isfinite(g).all()


def test_adjoint_against_an_fp64_reference_gradient(oracle_mod):
    """An FP64 reference gradient computed without the library: central
    differences (relative step 1e-5) of the oracle's misfit along each
    domain's (E, rho) -> s (E, rho) direction on the config-5 family at
    h = 1/4 (384 elements, 40 steps).  The GPU adjoint's directional
    derivatives agree within 1e-5 of the largest."""
    from tdbem import mesh
    a = mesh.adjoint_config5(h=0.25, nstep=40, nrec=16)
    w = a.workload
    print("case:", w.name, len(w.mesh.ele), "elements, FP64 reference gradient", flush=True)

    def oracle_u(mats):
        ww = mesh.Workload(w.name, w.mesh, mats, w.dt, w.nstep, w.known)
        U, *_ = _oracle_case(oracle_mod, ww).solve(w.known, w.nstep, tol=1e-13)
        return U

    obs = oracle_u(a.obs_mats)[:, a.receivers, :]
    ref = np.zeros(4)
    for d in range(4):
        Js = []
        for sgn in (1, -1):
            mats = w.mats.copy()
            mats[d, [1, 3]] *= 1 + sgn * 1e-5
            Js.append(_misfit(oracle_u(mats), a.receivers, obs))
        ref[d] = (Js[0] - Js[1]) / 2e-5
    c = _gpu_case(w)
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    J, g, _ = c.adjoint(w.known, w.nstep, a.receivers, obs, tol=1e-13, fd_step=1e-6)
    c.close()
    J0 = _misfit(oracle_u(w.mats), a.receivers, obs)
    assert abs(J - J0) <= 1e-9 * J0
    ad = w.mats[:, 1] * g[:, 0] + w.mats[:, 3] * g[:, 2]
    print("dJ/ds adjoint", ad, "oracle differences", ref, flush=True)
    assert np.abs(ad - ref).max() <= 1e-5 * np.abs(ref).max()
