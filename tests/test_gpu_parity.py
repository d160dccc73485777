"""GPU parity of the HIP path (through the C-ABI) against the CPU oracle.

Integer / pattern data must be bit-exact; FP64 blocks are compared bitwise
where the arithmetic order is the reference's (the lane-per-entry history
quadrature, DOF map) and to 1e-10 relative per 3x3 block otherwise (the
default wave-per-pair quadrature); solutions to the stated tolerances."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tdbem():
    import tdbem
    if not tdbem.LIB_PATH.exists():
        pytest.fail("HIP library not built")
    return tdbem


def _rod_cases(oracle_mod, f):
    tdbem = _tdbem()
    card = f["card"]
    dt = oracle_mod.reference_dt(f["pts"], f["ele"], card.beta, f["E_global"], card.v, f["rho_global"])
    oc = oracle_mod.Case(f["pts"], f["ele"], f["bcid"], f["mats"], f["itf"], dt, card.gap)
    rows = np.array([[m[0], m[1], m[2], m[3]] for m in card.materials], float)
    gc = tdbem.Case(f["pts"], f["ele"], f["bcid"], dt, card.gap, domain_count=card.domain_count, mat_rows=rows,
                    itf_rows=np.array(card.interfaces), default_mat=(f["E_global"], card.v, f["rho_global"]))
    return oc, gc, dt


@pytest.fixture(scope="module")
def rod_pair(oracle_mod, rod_fixture):
    print("case: rod120_100 fixture", flush=True)
    oc, gc, dt = _rod_cases(oracle_mod, rod_fixture)
    nstep = rod_fixture["card"].nstep
    gc.assemble_step0(10)
    gc.assemble_history(nstep)
    oc.assemble(nstep)
    return oc, gc, rod_fixture


def test_dofmap_bitexact(rod_pair):
    oc, gc, f = rod_pair
    od, gd = oc.dofmap(), gc.dofmap()
    for k in gd:
        assert np.array_equal(od[k], gd[k]), k


def test_self_tables(oracle_mod, rod_pair, rod_fixture):
    """Singular self tables: the default wave-per-element kernel (tree-ordered
    sums) within 1e-13 of the table's largest entry; the lane-per-element
    kernel (quad_staged 0, reference summation order) bit for bit.  The T2
    table of a flat element is zero up to rounding (entries of 1e-16 whose
    digits depend on the summation order), so it is compared against the T1
    table's scale."""
    oc, gc, f = rod_pair
    ref = np.array([oc.self_tables(e) for e in range(len(f["ele"]))])
    scale = {"U": np.abs(ref[:, 0]).max(), "T1": np.abs(ref[:, 1]).max(), "T2": np.abs(ref[:, 1]).max()}
    for name, got, want in zip(("U", "T1", "T2"), gc.self_tables(), (ref[:, 0], ref[:, 1], ref[:, 2])):
        err = np.abs(got - want).max()
        print(name, "largest difference", err, "scale", scale[name], flush=True)
        assert err <= 1e-13 * scale[name]
    _, gx, _ = _rod_cases(oracle_mod, rod_fixture)
    gx.set_option("quad_staged", 0)
    u, t1, t2 = gx.self_tables()
    gx.close()
    assert np.array_equal(u, ref[:, 0]) and np.array_equal(t1, ref[:, 1]) and np.array_equal(t2, ref[:, 2])


def _bands_close(x, y, rtol=1e-10):
    """Same CCSR pattern bit for bit; every 3x3 block within rtol of its own
    largest entry."""
    assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]), "pattern"
    if len(y[2]) == 0:
        return 0.0
    vy = y[2].reshape(len(y[2]), -1)
    err = np.abs(x[2].reshape(len(y[2]), -1) - vy).max(1)
    rel = err / np.maximum(np.abs(vy).max(1), 1e-300)
    assert rel.max() <= rtol, rel.max()
    return float(rel.max())


def _same_ccsr(a, b, rtol):
    assert np.array_equal(a[0], b[0]), "row_ptr"
    assert np.array_equal(a[1], b[1]), "col"
    scale = max(np.abs(b[2]).max(), 1e-300)
    assert np.abs(a[2] - b[2]).max() <= rtol * scale


def test_step0_matrices(rod_pair):
    oc, gc, f = rod_pair
    for which in (0, 1):
        _same_ccsr(gc.ccsr(which), oc.ccsr(which), 1e-12)


def test_history_blocks_all_steps(rod_pair):
    oc, gc, f = rod_pair
    for k in range(1, f["card"].nstep + 1):
        for which in (2, 3):
            _bands_close(gc.ccsr(which, k), oc.ccsr(which, k))


def test_history_exact_path_bitexact(oracle_mod, rod_fixture):
    """The lane-per-entry quadrature (quad_staged 0) keeps the reference's
    arithmetic order: MTS / MGS bands equal the oracle's bit for bit."""
    oc, gc, dt = _rod_cases(oracle_mod, rod_fixture)
    nstep = rod_fixture["card"].nstep
    gc.assemble_step0(10)
    gc.set_option("quad_staged", 0)
    gc.assemble_history(nstep)
    oc.assemble(nstep)
    for k in range(1, nstep + 1):
        for which in (2, 3):
            a, b = gc.ccsr(which, k), oc.ccsr(which, k)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (which, k)
            assert np.array_equal(a[2], b[2]), (which, k)
    gc.close()


def test_rod_known_answer_on_gpu(rod_pair):
    """Reference-held known answer (DBEM1/analytic_validation.txt) on the GPU path
    with the reference's GMRES settings (restart 200, tol 1e-5)."""
    oc, gc, f = rod_pair
    card = f["card"]
    U, T, it, fl, solved, times = gc.solve(f["lu"], card.nstep, restart=card.iterations, tol=card.error)
    assert solved == card.nstep and not fl.any()
    num = U[100, [100, 101, 102, 103], 0].mean()
    assert abs(num - float(f["expected"]["numeric"])) < 1e-12
    OU, OT, oit, ofl, osolved = oc.solve(f["lu"], card.nstep, restart=card.iterations, tol=card.error)
    assert np.array_equal(it, oit)
    assert np.abs(U - OU).max() <= 1e-10 * np.abs(OU).max()
    assert np.abs(T - OT).max() <= 1e-10 * np.abs(OT).max()
    for da, ea, db, eb, s in f["itf"]:
        assert np.array_equal(U[:, ea], U[:, eb])


def _workload_cases(oracle_mod, w):
    from tdbem import io
    tdbem = _tdbem()
    m = w.mesh
    card = io.DataCard(title="x", thread_num=1, flag_dyna=2, beta=0.5, nstep=w.nstep, gap=w.gap, E=1.0, v=0.0,
                       rho=1.0, amplitude=1, solver=2, domain_count=m.domain_count,
                       materials=[(int(r[0]), r[1], r[2], r[3]) for r in w.mats.tolist()],
                       interfaces=[tuple(int(x) for x in r) for r in m.interfaces.tolist()])
    mats, itf = io.normalise_model(card, len(m.ele))
    oc = oracle_mod.Case(m.points, m.ele, m.bcid, mats, itf, w.dt, w.gap)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=m.domain_count, mat_rows=w.mats,
                    itf_rows=m.interfaces)
    return oc, gc


@pytest.mark.parametrize("which", ["rod_config1", "layered_small"])
def test_generated_cases_match_oracle(oracle_mod, which):
    from tdbem import mesh
    if which == "rod_config1":
        w = mesh.rod_config1(nstep=40)
    else:
        w = mesh.layered_config2(h=0.5, nstep=40)
    print("case:", w.name, len(w.mesh.ele), "elements", flush=True)
    oc, gc = _workload_cases(oracle_mod, w)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    oc.assemble(w.nstep)
    for which_m in (0, 1):
        _same_ccsr(gc.ccsr(which_m), oc.ccsr(which_m), 1e-12)
    for k in range(1, w.nstep + 1):
        for which_m in (2, 3):
            _bands_close(gc.ccsr(which_m, k), oc.ccsr(which_m, k))
    U, T, it, fl, solved, _ = gc.solve(w.known, w.nstep, tol=1e-12)
    OU, OT, oit, ofl, osolved = oc.solve(w.known, w.nstep, tol=1e-12)
    assert solved == osolved == w.nstep
    assert np.abs(U - OU).max() <= 1e-9 * np.abs(OU).max()
    assert np.abs(T - OT).max() <= 1e-9 * np.abs(OT).max()


def test_regular_path_exercised_and_matched(oracle_mod):
    """A case where whole elements sit between the P and S fronts for a whole
    step (NeedCal = 2, regular path with the N0(gap) weight and T1+T2 sum):
    history blocks still bit-exact."""
    from tdbem import mesh
    m = mesh.box_mesh(2, 4.0, 1.0, 1.0, 8, 2, 2)
    mats = np.array([[1, 10.0, 0.45, 1.0], [2, 10.0, 0.45, 1.0]])
    known = mesh.normal_load_history(len(m.ele), 30, 0.02, m.top, np.r_[0.0, np.ones(30)])
    w = mesh.Workload("regular_path_probe", m, mats, 0.02, 30, known)
    oc, gc = _workload_cases(oracle_mod, w)
    regular = 0
    ne = len(m.ele)
    half = ne // 2
    for f in range(0, half, 3):
        for s in range(0, half, 5):
            for k in range(1, 30):
                regular += oc.needcal(f, s, 1, k) == 2
    assert regular > 0
    gc.assemble_step0(10)
    oc.assemble(w.nstep)
    for staged in (0, 1):
        gc.set_option("quad_staged", staged)
        gc.assemble_history(w.nstep)
        for k in range(1, w.nstep + 1):
            for which_m in (2, 3):
                a, b = gc.ccsr(which_m, k), oc.ccsr(which_m, k)
                if staged:
                    _bands_close(a, b)
                else:
                    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (which_m, k)
                    assert np.array_equal(a[2], b[2]), (which_m, k)


def test_full_size_properties():
    """Size-independent properties on a larger layered box (no oracle run):
    zero load gives zero response, the response is linear in the load, MGS
    blocks are symmetric and interface displacements agree across sides."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.25, nstep=60)
    m = w.mesh
    print("case:", w.name, len(m.ele), "elements", flush=True)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    U, T, it, fl, solved, _ = gc.solve(w.known, w.nstep, tol=1e-12)
    assert solved == w.nstep and not fl.any()
    U2, T2, *_ = gc.solve(2.0 * w.known, w.nstep, tol=1e-12)
    assert np.abs(U2 - 2 * U).max() <= 1e-9 * np.abs(U).max()
    Z, ZT, *_ = gc.solve(np.zeros_like(w.known), w.nstep, tol=1e-12)
    assert not Z.any() and not ZT.any()
    for k in (1, 10, 30):
        rp, col, val = gc.ccsr(3, k)
        assert np.array_equal(val, np.transpose(val, (0, 2, 1)))
    for da, ea, db, eb, s in m.interfaces:
        assert np.abs(U[:, ea - 1] - U[:, eb - 1]).max() <= 1e-12 * np.abs(U).max()


def test_tiled_convolution_matches_per_step_sum(oracle_mod):
    """The time-blocked convolution (far part per 8- or 16-step tile + near
    part per step) reassociates the reference's per-step history sum; both give
    the same solution to rounding.  The two one-pair-per-item 16-step far
    kernels agree bit for bit; the packed one (default) only reorders their
    final sums."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=45)
    m = w.mesh
    print("case:", w.name, "tiled vs per-step", flush=True)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    tiled = {}
    for tile in (16, 8):
        gc.set_option("conv_tile", tile)
        tiled[tile] = gc.solve(w.known, w.nstep, tol=1e-13)[:2]
    # 16-step tiles: the pre-filtered, prefetching far kernel (default) and the
    # one that stages and filters raw states add the same terms in the same order
    gc.set_option("conv_tile", 16)
    gc.set_option("conv_packed", 0)
    Up, Tp = gc.solve(w.known, w.nstep, tol=1e-13)[:2]
    gc.set_option("conv_filtered", 0)
    Ur, Tr = gc.solve(w.known, w.nstep, tol=1e-13)[:2]
    assert np.array_equal(Ur, Up) and np.array_equal(Tr, Tp)
    assert np.abs(tiled[16][0] - Up).max() <= 1e-13 * np.abs(Up).max()
    assert np.abs(tiled[16][1] - Tp).max() <= 1e-13 * np.abs(Tp).max()
    gc.set_option("conv_filtered", 1)
    gc.set_option("conv_packed", 1)
    gc.set_option("conv_tiled", 0)
    U0, T0, *_ = gc.solve(w.known, w.nstep, tol=1e-13)
    for tile, (U1, T1) in tiled.items():
        assert np.abs(U1 - U0).max() <= 1e-11 * np.abs(U0).max(), tile
        assert np.abs(T1 - T0).max() <= 1e-11 * np.abs(T0).max(), tile


def test_device_driven_gmres_matches_host_driven():
    """The device-driven GMRES (state, convergence test and stopping flags in
    device memory; iterations replayed from hipGraphs or launched directly)
    gives the host-driven solve's iterates bit for bit: same solution, same
    iteration counts and flags, also when a step stops at the restart length
    (flag 1) and the solve ends there."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=30)
    m = w.mesh
    print("case:", w.name, "device vs host GMRES", flush=True)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    for restart, tol in ((200, 1e-13), (200, 1e-5), (2, 1e-13)):
        out = {}
        for mode in (0, 1, 2):
            gc.set_option("gmres_device", mode)
            out[mode] = gc.solve(w.known, w.nstep, restart=restart, tol=tol)
        for mode in (1, 2):
            for a, b in zip(out[0][:5], out[mode][:5]):
                assert np.array_equal(np.asarray(a), np.asarray(b)), (restart, tol, mode)
        if restart == 2:
            assert out[0][4] < w.nstep and out[0][3][out[0][4] - 1] == 1
        else:
            assert out[0][4] == w.nstep and not out[0][3].any()
    gc.set_option("gmres_device", 1)
    gc.close()


def test_models_run_together_match_single_jobs():
    """Independent models with their own sizes, materials and time steps run
    together (tdbem_run_models: one host thread and stream per model, all
    contexts alive at once) give each model's own job bit for bit; so does a
    run whose memory budget admits one model's bands at a time."""
    tdbem = _tdbem()
    from tdbem import mesh
    ws = []
    for i, (n, E, v, rho) in enumerate(((4, 3.0, 0.25, 2.0), (6, 8.0, 0.3, 1.2), (5, 1.5, 0.22, 2.7))):
        m = mesh.box_mesh(1, 1.0, 1.0, 1.0, n, n, n)
        dt = 0.5 / n / mesh.wave_speeds(E, v, rho)[0]
        nstep = 20 + 5 * i
        sig = mesh.ricker(np.arange(nstep + 1) * dt, 0.8 / (20 * dt))
        sig[0] = 0.0
        known = mesh.normal_load_history(len(m.ele), nstep, dt, m.top, sig)
        ws.append(mesh.Workload(f"cube{n}", m, np.array([[1, E, v, rho]]), dt, nstep, known))
    print("case: three cube models run together", flush=True)

    def case(w):
        m = w.mesh
        return tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=1, mat_rows=w.mats,
                          itf_rows=m.interfaces)

    alone = []
    for w in ws:
        c = case(w)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        alone.append(c.solve(w.known, w.nstep, tol=1e-10)[:5])
        c.close()
    for budget in (0.0, 1.0):
        cs = [case(w) for w in ws]
        got = tdbem.run_models(cs, [w.known for w in ws], [w.nstep for w in ws], tol=1e-10, budget_bytes=budget)
        for c in cs:
            c.close()
        for w, a, g in zip(ws, alone, got):
            assert g[4] == w.nstep
            for x, y in zip(a, g):
                assert np.array_equal(np.asarray(x), np.asarray(y)), (w.name, budget)


def test_autotuner_choices_and_cache(tmp_path):
    """The autotuner picks one of the candidate launch shapes, the solve after
    tuning agrees with the default solve to rounding, every far-part shape
    and reduction partition does too, and a second context of the same shape
    takes the stored choice from the cache without timing."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.25, nstep=60)
    m = w.mesh
    print("case:", w.name, "autotuner", flush=True)

    def case():
        c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats,
                       itf_rows=m.interfaces)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        return c

    c = case()
    U0, T0, *_ = c.solve(w.known, w.nstep, tol=1e-13)
    cache = str(tmp_path / "tune.txt")
    fp, gb, ms = c.autotune(cache)
    assert fp in (0, 2, 4, 8) and gb in (8, 16, 32, 64, 128) and ms > 0.0
    U1, T1, *_ = c.solve(w.known, w.nstep, tol=1e-13)
    assert np.abs(U1 - U0).max() <= 1e-12 * np.abs(U0).max()
    for parts, blocks in ((2, 8), (8, 128), (1, 16)):
        c.set_option("conv_packed", 1)
        c.set_option("far_parts", parts)
        c.set_option("gmres_blocks", blocks)
        U2, T2, *_ = c.solve(w.known, w.nstep, tol=1e-13)
        assert np.abs(U2 - U0).max() <= 1e-12 * np.abs(U0).max(), (parts, blocks)
        assert np.abs(T2 - T0).max() <= 1e-12 * np.abs(T0).max(), (parts, blocks)
    c.close()
    c2 = case()
    assert c2.autotune(cache) == (fp, gb, 0.0)
    c2.close()
    assert len(open(cache).read().strip().splitlines()) == 1


def test_packed_far_convolution_on_long_rows():
    """Rows of 256 pairs (eight groups of 32 per row, two per wavefront) and
    bands up to the full history: the packed far kernel and the one-pair-per-
    item kernel give the same solution to rounding, FP64 and FP32 + bf16."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.25, nstep=70)
    m = w.mesh
    print("case:", w.name, "packed vs one pair per item", flush=True)
    for mixed in (0, 2):
        gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats,
                        itf_rows=m.interfaces)
        gc.set_option("mixed", mixed)
        gc.assemble_step0(10)
        gc.assemble_history(w.nstep)
        U1, T1, it1, *_ = gc.solve(w.known, w.nstep, tol=1e-13)
        gc.set_option("conv_packed", 0)
        U0, T0, it0, *_ = gc.solve(w.known, w.nstep, tol=1e-13)
        gc.close()
        assert np.abs(U1 - U0).max() <= 1e-12 * np.abs(U0).max(), mixed
        assert np.abs(T1 - T0).max() <= 1e-12 * np.abs(T0).max(), mixed


@pytest.mark.parametrize("which", ["regular_path_probe", "layered_small"])
def test_staged_quadrature_matches_per_entry(oracle_mod, which):
    """The lane-per-entry quadrature equals the oracle bit for bit; the default
    wave-per-pair quadrature (basis-tensor form, closed-form window integrals,
    fused multiply-adds) has the same band pattern and agrees to 1e-10 per
    block, and both give the same solution to 1e-12."""
    from tdbem import mesh
    if which == "regular_path_probe":
        m = mesh.box_mesh(2, 4.0, 1.0, 1.0, 8, 2, 2)
        mats = np.array([[1, 10.0, 0.45, 1.0], [2, 10.0, 0.45, 1.0]])
        known = mesh.normal_load_history(len(m.ele), 30, 0.02, m.top, np.r_[0.0, np.ones(30)])
        w = mesh.Workload("regular_path_probe", m, mats, 0.02, 30, known)
    else:
        w = mesh.layered_config2(h=0.5, nstep=40)
    print("case:", w.name, "staged vs per-entry quadrature", flush=True)
    oc, gc = _workload_cases(oracle_mod, w)
    gc.assemble_step0(10)
    gc.set_option("quad_staged", 0)
    gc.assemble_history(w.nstep)
    ref = [(gc.ccsr(2, k), gc.ccsr(3, k)) for k in range(1, w.nstep + 1)]
    U0, T0, *_ = gc.solve(w.known, w.nstep, tol=1e-12)
    gc.set_option("quad_staged", 1)
    gc.assemble_history(w.nstep)
    U1, T1, *_ = gc.solve(w.known, w.nstep, tol=1e-12)
    oc.assemble(w.nstep)
    worst = 0.0
    for k in range(1, w.nstep + 1):
        for i, which_m in enumerate((2, 3)):
            a, b, o = gc.ccsr(which_m, k), ref[k - 1][i], oc.ccsr(which_m, k)
            assert np.array_equal(b[0], o[0]) and np.array_equal(b[1], o[1]), (which_m, k)
            assert np.array_equal(b[2], o[2]), (which_m, k)
            worst = max(worst, _bands_close(a, b))
    print("worst block-relative difference", worst, flush=True)
    assert np.abs(U1 - U0).max() <= 1e-12 * np.abs(U0).max()
    assert np.abs(T1 - T0).max() <= 1e-12 * np.abs(T0).max()


def test_mixed_precision_error_against_fp64():
    """Mixed precision (blocks computed in FP64 and stored in FP32, mode 1, or
    as FP32 + bf16 residual, mode 2; FP32-stored inner GMRES under an FP64
    refinement residual): stored blocks are the FP64 blocks rounded to FP32
    (mode 1) or within 2^-30 of them (mode 2), and displacements and tractions
    agree with the FP64 run to <= 1e-6 relative (mode 1) and <= 1e-8 (mode 2),
    all solved to tol 1e-12."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.25, nstep=60)
    m = w.mesh
    print("case:", w.name, "mixed vs fp64", flush=True)
    out = {}
    for mixed in (0, 1, 2):
        gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats,
                        itf_rows=m.interfaces)
        gc.set_option("mixed", mixed)
        gc.assemble_step0(10)
        gc.assemble_history(w.nstep)
        U, T, it, fl, solved, _ = gc.solve(w.known, w.nstep, tol=1e-12)
        assert solved == w.nstep and not fl.any()
        out[mixed] = (U, T, [gc.ccsr(which, k) for which in (2, 3) for k in (1, 7, 30)])
        gc.close()
    U0, T0, B0 = out[0]
    for mixed, bound in ((1, 1e-6), (2, 1e-8)):
        U1, T1, B1 = out[mixed]
        for a, b in zip(B0, B1):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            if mixed == 1:
                assert np.array_equal(a[2].astype(np.float32).astype(np.float64), b[2])
            else:
                assert np.all(np.abs(a[2] - b[2]) <= 2.0 ** -30 * np.abs(a[2]))
        eu = np.abs(U1 - U0).max() / np.abs(U0).max()
        et = np.abs(T1 - T0).max() / np.abs(T0).max()
        print("mixed", mixed, "vs fp64: max rel err U", eu, "T", et, flush=True)
        assert eu <= bound and et <= bound, mixed


@pytest.mark.parametrize("nrhs,mixed", [(3, 0), (67, 0), (5, 2)])
def test_batched_load_cases_match_single_solves(nrhs, mixed):
    """Load cases solved as one batch (shared bands, lane-per-case convolution,
    lockstep GMRES) give each case's single-solve result: the same GMRES
    iteration counts and displacements / tractions to 1e-9 relative
    (tol 1e-12; the batch sums its dot products and the far convolution in
    another order).  67 cases exercise a second, partial group of lanes."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=30)
    m = w.mesh
    cases = mesh.load_cases(w, nrhs)
    print("case:", w.name, "batched", nrhs, "mixed", mixed, flush=True)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    gc.set_option("mixed", mixed)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    BU, BT, bit, bfl, bsolved, _ = gc.solve_batched(cases, w.nstep, tol=1e-12)
    assert (bsolved == w.nstep).all() and not bfl.any()
    for b in sorted({0, nrhs // 2, nrhs - 1}):
        U, T, it, fl, solved, _ = gc.solve(cases[b], w.nstep, tol=1e-12)
        assert solved == w.nstep
        assert np.abs(BU[b] - U).max() <= 1e-9 * np.abs(U).max(), b
        assert np.abs(BT[b] - T).max() <= 1e-9 * np.abs(T).max(), b
        assert np.abs(bit[b] - it).max() <= 1, b


def test_batched_cases_stop_where_their_single_solves_stop():
    """A restart length at which some load cases' GMRES stops with flag 1 at
    some step while the others finish: in the batch each stopped case keeps
    the flags, iteration counts and solved step of its single solve, with
    zero outputs after that step (the reference's loop ends there,
    DBEM1/multidomain.cpp:3884-3885), and the others go on to their single
    solves' results (1e-9 relative at tol 1e-12).  Six seeded cases, case 0
    again with its load delayed by ten steps (zero right-hand sides, solved
    in 0 iterations, until then) and a case with no load at all.  The batch
    sums its dot products in another order, so its iteration counts may
    differ by one: the restart is chosen from the single solves' counts at
    restart 200 such that every step a case runs needs at most restart - 1
    iterations, and the step where it stops at least restart + 2."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=20)
    m = w.mesh
    seeded = mesh.load_cases(w, 6)
    delayed = np.zeros_like(seeded[0])
    delayed[10:] = seeded[0][:-10]
    cases = np.concatenate([seeded, delayed[None], np.zeros_like(delayed)[None]])
    nrhs = len(cases)
    print("case:", w.name, "batched", nrhs, "with stopping GMRES", flush=True)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    probe = [gc.solve(cases[b], w.nstep, restart=200, tol=1e-12) for b in range(nrhs)]
    assert all(p[4] == w.nstep and not p[3].any() for p in probe)
    counts = np.array([p[2] for p in probe])
    chosen = None
    for r in range(2, 40):
        stops, ok = [], True
        for n in counts:
            over = np.nonzero(n > r)[0]
            k = int(over[0]) if len(over) else len(n)
            ok = ok and bool((n[:k] <= r - 1).all()) and (k == len(n) or n[k] >= r + 2)
            stops.append(k < len(n))
        if ok and any(stops) and not all(stops):
            chosen = r
            break
    assert chosen is not None, counts
    print("restart", chosen, "stopping cases", [b for b in range(nrhs) if stops[b]], flush=True)
    BU, BT, bit, bfl, bsolved, _ = gc.solve_batched(cases, w.nstep, restart=chosen, tol=1e-12)
    for b in range(nrhs):
        U, T, it, fl, solved, _ = gc.solve(cases[b], w.nstep, restart=chosen, tol=1e-12)
        assert bsolved[b] == solved and bool(fl.any()) == stops[b], b
        assert np.array_equal(bfl[b], fl), b
        assert np.abs(bit[b] - it).max() <= 1, b
        if stops[b]:
            assert bit[b][solved - 1] == it[solved - 1], b
        assert np.abs(BU[b] - U).max() <= 1e-9 * np.abs(U).max(), b
        assert np.abs(BT[b] - T).max() <= 1e-9 * np.abs(T).max(), b
        assert not BU[b][solved + 1:].any() and not BT[b][solved + 1:].any(), b
    # every output entry is written: the same calls over arrays filled with junk
    known = np.ascontiguousarray(cases, np.float64).reshape(-1)
    lu, lt = np.full_like(known, np.nan), np.full_like(known, np.nan)
    its, fls = np.full((nrhs, w.nstep), -7, np.int32), np.full((nrhs, w.nstep), -7, np.int32)
    sol = np.full(nrhs, -7, np.int32)
    tdbem._check(tdbem.lib().tdbem_solve_batched(gc.h, w.nstep, nrhs, known, chosen, 1e-12, lu, lt, its, fls, sol,
                                                 np.zeros(3)), "tdbem_solve_batched")
    assert np.array_equal(lu.reshape(BU.shape), BU) and np.array_equal(lt.reshape(BT.shape), BT)
    assert np.array_equal(its, bit) and np.array_equal(fls, bfl) and np.array_equal(sol, bsolved)
    U, T, it, fl, solved, _ = gc.solve(cases[0], w.nstep, restart=chosen, tol=1e-12)
    k0 = np.ascontiguousarray(cases[0]).reshape(-1)
    lu, lt = np.full_like(k0, np.nan), np.full_like(k0, np.nan)
    its, fls, sv = np.full(w.nstep, -7, np.int32), np.full(w.nstep, -7, np.int32), tdbem.C.c_int(-7)
    tdbem._check(tdbem.lib().tdbem_solve(gc.h, w.nstep, k0, chosen, 1e-12, lu, lt, its, fls, tdbem.C.byref(sv),
                                         np.zeros(3)), "tdbem_solve")
    assert np.array_equal(lu.reshape(U.shape), U) and np.array_equal(lt.reshape(T.shape), T)
    assert np.array_equal(its, it) and np.array_equal(fls, fl) and sv.value == solved
    gc.close()


@pytest.mark.parametrize("levels,mixed,maxn,mfma,nstep", [(1, 0, 60, 1, 60), (1, 2, 37, 0, 60), (2, 0, 37, 1, 60),
                                                          (3, 2, 50, 0, 60), (3, 0, 140, 1, 140), (3, 2, 140, 1, 140),
                                                          (3, 0, 140, 0, 140), (3, 2, 97, 1, 140)])
def test_batched_spectral_far_part_matches_lag_by_lag(levels, mixed, maxn, mfma, nstep):
    """The frequency-domain far part against the lag-by-lag far part (8-step
    tiles) on 20 load cases (a partial 64-case tile).  Spectral levels of
    sub-filter length 16, 32, ... (lags 17-32 at 16, 33-64 at 32, ..., the
    last level the rest up to maxn; 2L-point spectra of the sub-filters and
    of the 1-2-1 combined state windows, L-step tiles), the products on the
    FP64 matrix cores or with vector FMAs: the same solution to 1e-9 relative
    at GMRES tol 1e-12 (the two sum in another order).  With the history
    truncated at maxn < nstep the blocks past maxn stay out of both; maxn 37,
    50 and 137 end inside a level, and levels that would start past maxn are
    left out."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.25, nstep=nstep)
    m = w.mesh
    cases = mesh.load_cases(w, 20)
    print("case:", w.name, "spectral far part, levels", levels, "mixed", mixed, "maxn", maxn, "mfma", mfma,
          "nstep", nstep, flush=True)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    gc.set_option("batched_fft_levels", levels)
    gc.set_option("mixed", mixed)
    gc.set_option("batched_fft_mfma", mfma)
    gc.assemble_step0(10)
    gc.assemble_history(maxn)
    res = {}
    for fft in (1, 0):
        gc.set_option("batched_fft", fft)
        BU, BT, bit, bfl, bsolved, _ = gc.solve_batched(cases, w.nstep, tol=1e-12)
        assert (bsolved == w.nstep).all() and not bfl.any()
        res[fft] = (BU, BT)
        if fft:
            tm = gc.timings()
            expect = max(l for l in range(1, levels + 1) if (16 << (l - 1)) + 1 <= maxn)
            assert tm["spectral"] and tm["spectral_levels"] == expect, tm["spectral_levels"]
            for st in tm["spectral_level_stats"]:
                assert st["launches"] == (w.nstep - 2) // st["L"] + 1, st
    gc.close()
    eu = np.abs(res[1][0] - res[0][0]).max() / np.abs(res[0][0]).max()
    et = np.abs(res[1][1] - res[0][1]).max() / np.abs(res[0][1]).max()
    print("spectral vs lag-by-lag: u", eu, "t", et, flush=True)
    assert eu <= 1e-9 and et <= 1e-9


@pytest.mark.gpu
def test_batched_spectral_products_launch_shapes_agree():
    """The matrix-core spectral products as one-wavefront workgroups (default:
    the eight wavefronts of a row group and slot dispatched together on one
    XCD, the row groups in descending order of work) and as eight-wavefront
    workgroups that meet at a barrier every few items (option
    batched_fft_pace): each wavefront sums the same items in the same order,
    so the two give the same solution bit for bit, over two spectral levels
    and 20 load cases."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.25, nstep=80)
    m = w.mesh
    cases = mesh.load_cases(w, 20)
    print("case:", w.name, "spectral products launch shapes", flush=True)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    gc.set_option("batched_fft_levels", 2)
    gc.set_option("mixed", 2)
    gc.set_option("batched_fft", 1)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    res = []
    for pace in (0, 3):
        gc.set_option("batched_fft_pace", pace)
        out = gc.solve_batched(cases, w.nstep, tol=1e-12)
        assert (out[4] == w.nstep).all() and not out[3].any()
        assert gc.timings()["spectral"]
        res.append(out[:5])
    gc.close()
    for x, y in zip(*res):
        assert np.array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [0, 2])
def test_batched_far_matrix_cores_match_vector_fmas(mixed):
    """The batched far part on the FP64 matrix cores (row groups of five
    sources, the state fixed per k-step; columns split into 4, 1 or 3 partial
    sums) and with vector FMAs (one source row per workgroup, a register
    window of states) give the same batch to rounding: 1e-10 relative on
    every case's displacements and tractions, on a mesh whose domains end
    with partial row groups and 70 load cases (a partial second group of
    64)."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.25, nstep=40)
    m = w.mesh
    cases = mesh.load_cases(w, 70)
    print("case:", w.name, "elements", len(m.ele), "mixed", mixed, flush=True)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    gc.set_option("mixed", mixed)
    gc.set_option("batched_fft", 0)  # the lag-by-lag kernels
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    out = []
    for mf, parts in ((0, 1), (1, 4), (1, 1), (1, 3)):
        gc.set_option("batched_mfma", mf)
        gc.set_option("batched_far_parts", parts)
        BU, BT, bit, bfl, bsolved, _ = gc.solve_batched(cases, w.nstep, tol=1e-12)
        assert (bsolved == w.nstep).all() and not bfl.any()
        out.append((BU, BT))
    gc.close()
    U0, T0 = out[0]
    for U1, T1 in out[1:]:
        assert np.abs(U1 - U0).max() <= 1e-10 * np.abs(U0).max()
        assert np.abs(T1 - T0).max() <= 1e-10 * np.abs(T0).max()


def _adjoint_setup():
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=25)
    m = w.mesh
    rng = np.random.default_rng(5)
    recv = np.array(sorted(set(list(m.top) + list(rng.choice(len(m.ele), 8, replace=False)))), np.int32)

    def case(mats):
        c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=mats, itf_rows=m.interfaces)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        return c

    pert = w.mats.copy()
    pert[:, 1:] *= 1 + 0.1 * rng.uniform(-1, 1, size=(4, 3))
    c = case(pert)
    U, *_ = c.solve(w.known, w.nstep, tol=1e-12)
    c.close()
    return w, recv, U[:, recv, :], case


def test_adjoint_gradient_along_wave_speed_preserving_directions():
    """The reverse-time adjoint against central differences of the misfit
    along (E, rho) -> s (E, rho) per domain: this direction keeps every wave
    front in place, so the misfit is smooth along it and the two agree to
    1e-6 of the largest directional derivative.  The adjoint leaves the case
    as it was (a later solve is bitwise unchanged)."""
    w, recv, obs, case = _adjoint_setup()
    print("case:", w.name, "adjoint", flush=True)
    c = case(w.mats)
    U0, *_ = c.solve(w.known, w.nstep, tol=1e-12)
    J, g, lu = c.adjoint(w.known, w.nstep, recv, obs, tol=1e-12, fd_step=1e-6)
    U1, *_ = c.solve(w.known, w.nstep, tol=1e-12)
    assert np.array_equal(U0, U1)
    assert abs(J - float(((lu[1:, recv, :] - obs[1:]) ** 2).sum())) <= 1e-12 * J
    evr = c.materials()
    c.close()
    fd = np.zeros(4)
    ad = np.zeros(4)
    for d in range(4):
        h = 1e-6
        ep, em = evr.copy(), evr.copy()
        ep[d, [0, 2]] *= 1 + h
        em[d, [0, 2]] *= 1 - h
        Js = []
        for e in (ep, em):
            rows = np.column_stack([w.mats[:, 0], e])
            cc = case(rows)
            U, *_ = cc.solve(w.known, w.nstep, tol=1e-12)
            cc.close()
            Js.append(float(((U[1:, recv, :] - obs[1:]) ** 2).sum()))
        fd[d] = (Js[0] - Js[1]) / (2 * h)
        ad[d] = evr[d, 0] * g[d, 0] + evr[d, 2] * g[d, 2]
    print("dJ/ds adjoint", ad, "fd", fd, flush=True)
    assert np.abs(ad - fd).max() <= 1e-6 * np.abs(fd).max()


def test_adjoint_time_blocked_transposed_term():
    """The backward sweep's time-blocked transposed history term (lags > 8
    once per block of 8 steps, the shorter lags per step) reassociates the
    per-step sum over every lag: same misfit and gradient to rounding."""
    w, recv, obs, case = _adjoint_setup()
    print("case:", w.name, "adjoint tiled vs per-step", flush=True)
    c = case(w.mats)
    J1, g1, _ = c.adjoint(w.known, w.nstep, recv, obs, tol=1e-12, fd_step=1e-4)
    c.set_option("conv_tiled", 0)
    J0, g0, _ = c.adjoint(w.known, w.nstep, recv, obs, tol=1e-12, fd_step=1e-4)
    c.close()
    print("J", J1, J0, "gradient difference", np.abs(g1 - g0).max() / np.abs(g0).max(), flush=True)
    assert abs(J1 - J0) <= 1e-12 * J0
    assert np.abs(g1 - g0).max() <= 1e-9 * np.abs(g0).max()


def test_adjoint_device_driven_gmres_is_bit_identical():
    """The forward and backward GMRES driven from the device (hipGraph
    replay, default) give the host-driven adjoint's misfit and gradient bit
    for bit."""
    w, recv, obs, case = _adjoint_setup()
    print("case:", w.name, "adjoint device vs host GMRES", flush=True)
    c = case(w.mats)
    J1, g1, u1 = c.adjoint(w.known, w.nstep, recv, obs, tol=1e-12, fd_step=1e-6)
    c.set_option("gmres_device", 0)
    J0, g0, u0 = c.adjoint(w.known, w.nstep, recv, obs, tol=1e-12, fd_step=1e-6)
    c.close()
    assert J1 == J0
    assert np.array_equal(g1, g0) and np.array_equal(u1, u0)


def test_adjoint_analytic_gradient_matches_small_step_differences():
    """Analytic history-block derivatives (default): every (E, v, rho)
    component against central differences of the misfit at a relative step of
    1e-7 on the frozen structure (no quadrature point changes its front or
    window case within such a step): agreement to 1e-6 of the largest
    component (observed ~1e-8); and against the operator-difference mode at
    1e-6 to 1e-5 (observed ~3e-7)."""
    w, recv, obs, case = _adjoint_setup()
    print("case:", w.name, "analytic adjoint vs small-step differences", flush=True)
    c = case(w.mats)
    J, ga, _ = c.adjoint(w.known, w.nstep, recv, obs, tol=1e-12, fd_step=1e-6)
    c.set_option("adjoint_analytic", 0)
    J2, gf, _ = c.adjoint(w.known, w.nstep, recv, obs, tol=1e-12, fd_step=1e-6)
    c.set_option("adjoint_analytic", 1)
    assert J == J2
    evr = c.materials()
    c.set_option("freeze_structure", 1)
    gfd = np.zeros((4, 3))
    for d in range(4):
        for p in range(3):
            dl = 1e-7 * max(abs(evr[d, p]), 0.1)
            Js = []
            for sgn in (1, -1):
                e = evr.copy()
                e[d, p] += sgn * dl
                c.set_materials(e)
                c.assemble_step0(10)
                c.assemble_history(w.nstep)
                U, *_ = c.solve(w.known, w.nstep, tol=1e-13)
                Js.append(float(((U[1:, recv, :] - obs[1:]) ** 2).sum()))
            gfd[d, p] = (Js[0] - Js[1]) / (2 * dl)
    c.close()
    print("analytic\n", ga, "\noperator differences\n", gf, "\nmisfit differences\n", gfd, flush=True)
    assert np.abs(ga - gfd).max() <= 1e-6 * np.abs(gfd).max()
    assert np.abs(ga - gf).max() <= 1e-5 * np.abs(gf).max()


def test_adjoint_gradient_matches_same_step_differences():
    """Operator-difference mode (adjoint_analytic 0), all (E, v, rho)
    components: the reference's time-integrated kernels jump
    where a quadrature point crosses a front window, so the misfit is only
    piecewise smooth in the wave speeds and its derivative depends on the
    difference step.  With the operators differenced at the same relative
    step (1e-3) as the misfit, the adjoint gradient agrees with the misfit's
    central differences to second order (1e-2 of the largest component)."""
    w, recv, obs, case = _adjoint_setup()
    c = case(w.mats)
    c.set_option("adjoint_analytic", 0)
    J, g, lu = c.adjoint(w.known, w.nstep, recv, obs, tol=1e-12, fd_step=1e-3)
    evr = c.materials()
    # the misfit's differences on the same frozen structure (bands and judges
    # of the unperturbed materials) the adjoint differences its operators on
    c.set_option("freeze_structure", 1)
    gfd = np.zeros((4, 3))
    for d in range(4):
        for p in range(3):
            dl = 1e-3 * max(abs(evr[d, p]), 0.1)
            Js = []
            for sgn in (1, -1):
                e = evr.copy()
                e[d, p] += sgn * dl
                c.set_materials(e)
                c.assemble_step0(10)
                c.assemble_history(w.nstep)
                U, *_ = c.solve(w.known, w.nstep, tol=1e-12)
                Js.append(float(((U[1:, recv, :] - obs[1:]) ** 2).sum()))
            gfd[d, p] = (Js[0] - Js[1]) / (2 * dl)
    c.close()
    print("grad adjoint\n", g, "\ngrad fd\n", gfd, flush=True)
    assert np.abs(g - gfd).max() <= 1e-2 * np.abs(gfd).max()


_AUTOGRAD_SCRIPT = r"""
import sys, json
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
import tdbem
from tdbem import mesh
from tdbem.autograd import MisfitProblem, misfit
w = mesh.layered_config2(h=0.5, nstep=12)
m = w.mesh
recv = np.array(m.top, np.int32)
mats = w.mats.copy()
c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=mats * [1, 1.05, 1, 0.97],
               itf_rows=m.interfaces)
c.assemble_step0(10); c.assemble_history(w.nstep)
U, *_ = c.solve(w.known, w.nstep, tol=1e-12); c.close()
prob = MisfitProblem(m.points, m.ele, m.bcid, w.dt, w.gap, mats[:, 0], m.interfaces, w.known, w.nstep, recv,
                     U[:, recv, :])
theta = torch.tensor(mats[:, 1:], dtype=torch.float64, requires_grad=True)
J = misfit(theta, prob)
(2.0 * J).backward()
J2, g2 = prob.evaluate(mats[:, 1:])
print(json.dumps({"J": float(J), "J2": J2, "grad": theta.grad.numpy().tolist(), "g2": g2.tolist()}))
"""


def test_autograd_function_returns_the_adjoint_gradient(tmp_path):
    """torch.autograd.Function over the C-ABI adjoint (run in its own process so
    torch's HIP runtime and the library's are not mixed with the other tests):
    backward of 2 J gives twice the adjoint gradient."""
    import json
    import subprocess
    import sys
    script = tmp_path / "autograd_check.py"
    script.write_text(_AUTOGRAD_SCRIPT)
    pkg = str(Path(__file__).resolve().parents[1] / "parallel3.0_omp_multidomin_1node_amd")
    print("case: autograd subprocess", flush=True)
    out = subprocess.run([sys.executable, str(script), pkg], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert r["J"] == r["J2"]
    assert np.allclose(np.array(r["grad"]), 2.0 * np.array(r["g2"]), rtol=1e-12, atol=0.0)


def test_validation_outputs_of_a_solve(tmp_path):
    """validation_state.csv / validation_metrics.txt in the reference's layout
    (DBEM1/DBEM.cpp:434-561, DBEM1/multidomain.cpp:388-416): the CSV holds the
    solved global-frame displacements and tractions exactly (%.17g), and the
    metrics agree with what this test recomputes from the solution, the
    exported matrices and the materials."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=20)
    m = w.mesh
    print("case:", w.name, "validation outputs", flush=True)
    gc = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    gc.assemble_step0(10)
    gc.assemble_history(w.nstep)
    U, T, it, fl, solved, _ = gc.solve(w.known, w.nstep, tol=1e-10)
    gc.write_validation(tmp_path, U, T, it, fl, solved)
    A = gc.ccsr(0)
    mts = [len(gc.ccsr(2, k)[1]) for k in range(1, w.nstep + 1)]
    mgs = [len(gc.ccsr(3, k)[1]) for k in range(1, w.nstep + 1)]
    gc.close()
    ne = len(m.ele)
    lines = (tmp_path / "validation_state.csv").read_text().splitlines()
    assert lines[0] == "step,element,localNode,nodeId,domain,surfaceType,bcid,x,y,z,ux,uy,uz,tx,ty,tz"
    rows = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]])
    assert rows.shape == ((w.nstep + 1) * ne, 16)
    assert np.array_equal(rows[:, 0], np.repeat(np.arange(w.nstep + 1), ne))
    assert np.array_equal(rows[:, 1], np.tile(np.arange(ne), w.nstep + 1)) and np.array_equal(rows[:, 1], rows[:, 3])
    assert np.array_equal(rows[:, 10:13], U.reshape(-1, 3)) and np.array_equal(rows[:, 13:16], T.reshape(-1, 3))
    assert np.array_equal(rows[:ne, 6], m.bcid)
    itf = np.asarray(m.interfaces)
    surf = np.zeros(ne)
    surf[itf[:, 1] - 1] = 1
    surf[itf[:, 3] - 1] = 1
    assert np.array_equal(rows[:ne, 5], surf)
    assert np.array_equal(rows[:ne, 4], np.repeat(np.arange(4), ne // 4))
    kv = dict(ln.split("=", 1) for ln in (tmp_path / "validation_metrics.txt").read_text().splitlines())
    assert kv["NStep"] == str(w.nstep) and kv["ElementCount"] == str(ne) and kv["DomainCount"] == "4"
    assert float(kv["Dt"]) == w.dt and kv["InterfaceCount"] == str(len(itf))
    for d in range(4):
        E, v, rho = w.mats[d, 1:]
        G = E / 2.0 / (1.0 + v)
        c1 = np.sqrt((2.0 - 2.0 * v) / (1.0 - 2.0 * v) * G / rho)
        vals = kv["DomainMaterial[%d]" % d].split()
        assert abs(float(vals[1]) - c1) <= 1e-15 * c1 and abs(float(vals[5]) - c1 * w.dt) <= 1e-15 * c1 * w.dt
    ea, eb, sgn = itf[:, 1] - 1, itf[:, 3] - 1, itf[:, 4]
    for k in (0, 7, w.nstep):
        du = np.abs(U[k, ea] - U[k, eb]).max()
        tb = np.abs(T[k, eb] - sgn[:, None] * T[k, ea]).max()
        assert float(kv["InterfaceStep[%d].MaxAbsU" % k]) == du
        assert float(kv["InterfaceStep[%d].MaxAbsTBalance" % k]) == tb
    assert kv["MultiDomainMetricsValid"] == "1" and kv["MultiDomainFinalFlag"] == "0"
    assert kv["MultiDomainSolvedSteps"] == str(w.nstep)
    assert kv["MultiDomainTotalInnerIterations"] == str(int(np.sum(it)))
    assert kv["MultiDomainTotalOuterIterations"] == str(int(np.sum(np.asarray(it) > 0)))
    assert kv["MultiDomainMTS0NonZeroBlocks"] == str(len(A[1]))
    assert abs(float(kv["MultiDomainMTSAverageBlocks"]) - np.mean(mts)) <= 1e-6
    assert abs(float(kv["MultiDomainMGSAverageBlocks"]) - np.mean(mgs)) <= 1e-6
    rp, col = A[0], A[1]
    diag = sum(int(np.sum(col[rp[r]:rp[r + 1]] == r)) for r in range(len(rp) - 1))
    assert kv["MultiDomainMatrixStructureValid"] == "1" and kv["MultiDomainMatrixZeroRows"] == "0"
    assert kv["MultiDomainMatrixDiagonalBlocks"] == str(diag)
    assert kv["MultiDomainInterfaceDofInvariantValid"] == "1"


def test_history_persistence(tmp_path):
    """Saved history bands loaded into a fresh case of the same model give the
    same MTS / MGS matrices and the same solution bit for bit (FP64 and the
    FP32 + bf16 mode); a case with other materials or another mesh refuses the
    file."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=30)
    m = w.mesh
    print("case:", w.name, "history save / load", flush=True)

    def case(mats=w.mats, mm=m, mixed=0):
        c = tdbem.Case(mm.points, mm.ele, mm.bcid, w.dt, w.gap, domain_count=4, mat_rows=mats,
                       itf_rows=mm.interfaces)
        c.set_option("mixed", mixed)
        c.assemble_step0(10)
        return c

    for mixed in (0, 2):
        path = tmp_path / ("hist%d.bin" % mixed)
        a = case(mixed=mixed)
        a.assemble_history(w.nstep)
        U0, T0, it0, *_ = a.solve(w.known, w.nstep, tol=1e-12)
        a.save_history(path)
        ref = [(a.ccsr(2, k), a.ccsr(3, k)) for k in (1, 9, w.nstep)]
        a.close()
        b = case()
        assert b.load_history(path) == w.nstep
        for (x2, x3), k in zip(ref, (1, 9, w.nstep)):
            for x, y in ((x2, b.ccsr(2, k)), (x3, b.ccsr(3, k))):
                assert all(np.array_equal(u, v) for u, v in zip(x, y)), (mixed, k)
        U1, T1, it1, *_ = b.solve(w.known, w.nstep, tol=1e-12)
        b.close()
        assert np.array_equal(U0, U1) and np.array_equal(T0, T1) and np.array_equal(it0, it1), mixed
    other = w.mats.copy()
    other[0, 1] *= 1.01
    c = case(mats=other)
    with pytest.raises(RuntimeError, match="other materials"):
        c.load_history(tmp_path / "hist0.bin")
    c.close()
    m2 = mesh.layered_config2(h=0.25, nstep=30).mesh
    c = case(mm=m2)
    with pytest.raises(RuntimeError, match="another model"):
        c.load_history(tmp_path / "hist0.bin")
    c.close()


def test_single_domain_gates():
    """The reference's D = 1 gates (tools/run_multidomain_validation.py:
    996-1030): (i) a single-domain model read without the multi-domain
    datacard tail and the same model with an explicit one-domain tail give the
    same solution bit for bit; (ii) a homogeneous bar solved as one domain and
    as two domains with an interface agree on the shared outer elements
    (reference tolerances 1e-7 absolute / 1e-5 relative; here 1e-9 of the
    largest displacement) over steps whose time windows end well before the
    load's P wave can reach the interface."""
    tdbem = _tdbem()
    from tdbem import mesh
    E, v, rho = 10.0, 0.25, 1.0
    c1, _ = mesh.wave_speeds(E, v, rho)
    h = 0.25
    dt = 0.5 * h / c1
    nstep = 6
    # the P front of the loaded face (x = 2) stays far from the interface at
    # x = 0 over every step's time window, numerical precursors included
    assert c1 * dt * (nstep + 1) <= 1.0
    one = mesh.box_mesh(1, 4.0, 1.0, 1.0, 16, 4, 4)
    two = mesh.box_mesh(2, 4.0, 1.0, 1.0, 8, 4, 4)
    print("case: single-domain gates", len(one.ele), len(two.ele), "elements", flush=True)

    def run(mm, domain_count, mats, itf):
        known = mesh.normal_load_history(len(mm.ele), nstep, dt, mm.top, np.r_[0.0, np.ones(nstep)],
                                         direction=(0.0, 0.0, 1.0))
        c = tdbem.Case(mm.points, mm.ele, mm.bcid, dt, 0.4, domain_count=domain_count, mat_rows=mats,
                       itf_rows=itf, default_mat=(E, v, rho))
        c.assemble_step0(10)
        c.assemble_history(nstep)
        U, T, it, fl, solved, _ = c.solve(known, nstep, tol=1e-12)
        c.close()
        assert solved == nstep and not fl.any()
        return U, T

    U0, T0 = run(one, 0, None, None)
    U1, T1 = run(one, 1, np.array([[1, E, v, rho]]), np.zeros((0, 5), np.int64))
    assert np.array_equal(U0, U1) and np.array_equal(T0, T1)
    U2, T2 = run(two, 2, np.array([[1, E, v, rho], [2, E, v, rho]]), two.interfaces)

    def centres(mm):
        return mm.points[mm.ele[:, :4]].mean(axis=1)

    c_one, c_two = centres(one), centres(two)
    itf = set((two.interfaces[:, 1] - 1).tolist()) | set((two.interfaces[:, 3] - 1).tolist())
    idx = {tuple(np.round(c, 9)): i for i, c in enumerate(c_one)}
    pairs = [(idx[tuple(np.round(c, 9))], j) for j, c in enumerate(c_two) if j not in itf]
    assert len(pairs) == len(one.ele)
    a = np.array([p[0] for p in pairs])
    b = np.array([p[1] for p in pairs])
    scale = np.abs(U0).max()
    assert scale > 0
    diff = np.abs(U0[:, a] - U2[:, b]).max()
    print("outer displacement difference", diff, "of", scale, flush=True)
    assert diff <= 1e-9 * scale
    assert np.abs(T0[:, a] - T2[:, b]).max() <= 1e-9 * np.abs(T0).max()
