"""The drop-in boundary on the GPU: the whole-solver entry point
tdbem_dyna_gmres_solver and the C++ drop-in header's
DynaGMRESSolverMultiDomainCCSR against the reference's known answer and the
oracle, the reference's GMRES-flag stop, and a history file refused after
the loading case's cull."""
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG

pytestmark = pytest.mark.gpu

KNOWN_UX = "0.75914256602696095"


def _tdbem():
    import tdbem
    if not tdbem.LIB_PATH.exists():
        pytest.fail("HIP library not built")
    return tdbem


def _rod(oracle_mod, f):
    tdbem = _tdbem()
    card = f["card"]
    dt = oracle_mod.reference_dt(f["pts"], f["ele"], card.beta, f["E_global"], card.v, f["rho_global"])
    rows = np.array([[m[0], m[1], m[2], m[3]] for m in card.materials], float)
    gc = tdbem.Case(f["pts"], f["ele"], f["bcid"], dt, card.gap, domain_count=card.domain_count, mat_rows=rows,
                    itf_rows=np.array(card.interfaces), default_mat=(f["E_global"], card.v, f["rho_global"]))
    oc = oracle_mod.Case(f["pts"], f["ele"], f["bcid"], f["mats"], f["itf"], dt, card.gap)
    return gc, oc


def test_dyna_gmres_solver_entry(oracle_mod, rod_fixture):
    """tdbem_dyna_gmres_solver on the reference's rod120_100 inputs: result 1
    and the reference's loaded-end mean ux at step 100
    (DBEM1/analytic_validation.txt); with restart 1 and an unreachable
    tolerance the first step's GMRES stops with flag 1 and the result is -1
    (DBEM1/multidomain.cpp:3884-3885,3925), as the oracle's loop stops."""
    print("case: rod120_100 through tdbem_dyna_gmres_solver", flush=True)
    f = rod_fixture
    card = f["card"]
    gc, oc = _rod(oracle_mod, f)
    res, U, T = gc.dyna_gmres_solver(f["lu"], card.nstep, card.iterations, card.error, card.premaxleaf)
    assert res == 1
    sel = [100, 101, 102, 103]
    assert abs(U[100, sel, 0].mean() - float(KNOWN_UX)) < 1e-12
    gc.close()
    gc, _ = _rod(oracle_mod, f)
    res, U, T = gc.dyna_gmres_solver(f["lu"], card.nstep, 1, 1e-300, card.premaxleaf)
    assert res == -1
    gc.close()
    gc, _ = _rod(oracle_mod, f)
    gc.assemble_step0(card.premaxleaf)
    gc.assemble_history(card.nstep)
    GU, GT, git, gfl, gsolved, _ = gc.solve(f["lu"], card.nstep, restart=1, tol=1e-300)
    OU, OT, oit, ofl, osolved = oc.solve(f["lu"], card.nstep, restart=1, tol=1e-300)
    assert gsolved == osolved == 1
    assert gfl[0] == ofl[0] != 0 and git[0] == oit[0]
    assert not GU[1:].any() and not OU[1:].any()
    gc.close()
    with pytest.raises(RuntimeError, match="invalid arguments"):
        _rod(oracle_mod, f)[0].dyna_gmres_solver(f["lu"], card.nstep, 0, card.error, card.premaxleaf)


def test_dropin_header_solver_known_answer():
    """The compiled C++ caller of include/tdbem_multidomain.hpp reads the
    reference's rod120_100 inputs and gets result 1 and the reference's
    loaded-end mean ux at step 100 through DynaGMRESSolverMultiDomainCCSR."""
    exe = PKG / "lib" / "rod_dropin"
    print("case: rod120_100 through the C++ drop-in header", flush=True)
    r = subprocess.run([str(exe), str(GOLDEN / "rod120_100" / "BEM_DATACARD.DAT"),
                        str(GOLDEN / "rod120_100" / "rod120_100.model"), str(GOLDEN / "rod120_100" / "rod120_100.bd"),
                        "--global-E", "200"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [x for x in r.stdout.splitlines() if x.startswith("result=")][-1]
    kv = dict(p.split("=") for p in line.split())
    assert kv["result"] == "1" and kv["elements"] == "4"
    assert abs(float(kv["mean_ux_last"]) - float(KNOWN_UX)) < 1e-12


def test_history_file_refused_after_the_cull_drops_the_bands(tmp_path):
    """A history file whose header matches (element and pair counts, dt, gap,
    materials) but whose bands differ from the loading case's cull (perturbed
    geometry) is refused, and the case is left with no history: a solve then
    raises instead of reading bands of another layout."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=20)
    m = w.mesh
    print("case:", w.name, "history refusal after the cull", flush=True)

    def case(points):
        c = tdbem.Case(points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
        c.assemble_step0(10)
        return c

    a = case(m.points)
    a.assemble_history(w.nstep)
    a.save_history(tmp_path / "h.bin")
    a.close()
    pts = m.points.copy()
    pts[:, 0] *= 1.5                      # the same counts, longer distances: other bands
    b = case(pts)
    b.assemble_history(w.nstep)           # a history of its own, resident before the load
    b.solve(w.known, w.nstep, tol=1e-10)
    with pytest.raises(RuntimeError, match="bands differ"):
        b.load_history(tmp_path / "h.bin")
    with pytest.raises(RuntimeError, match="history not assembled"):
        b.solve(w.known, w.nstep, tol=1e-10)
    b.assemble_history(w.nstep)           # and it assembles again normally
    U, *_ = b.solve(w.known, w.nstep, tol=1e-10)
    assert np.isfinite(U).all()
    b.close()


def test_device_resident_solves_equal_the_host_buffer_solves():
    """tdbem_solve_device / tdbem_solve_batched_device (inputs already in
    device memory, outputs left there) give the host-buffer calls' results
    bit for bit, step 0 and a GMRES-stopped case included."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=20)
    m = w.mesh
    print("case:", w.name, "device-resident solves", flush=True)
    c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    U, T, it, fl, solved, _ = c.solve(w.known, w.nstep, tol=1e-10)
    dk = tdbem.DeviceArray.from_host(w.known)
    du, dt_ = tdbem.DeviceArray(w.known.shape), tdbem.DeviceArray(w.known.shape)
    it2, fl2, solved2, _ = c.solve_device(dk, w.nstep, du, dt_, tol=1e-10)
    assert solved2 == solved and np.array_equal(it2, it) and np.array_equal(fl2, fl)
    assert np.array_equal(du.download(), U) and np.array_equal(dt_.download(), T)
    cases = mesh.load_cases(w, 5)
    dk = tdbem.DeviceArray.from_host(cases)
    du, dt_ = tdbem.DeviceArray(cases.shape), tdbem.DeviceArray(cases.shape)
    for restart, tol in ((200, 1e-10), (1, 1e-12)):   # the second stops every case at step 1
        BU, BT, bit, bfl, bsolved, _ = c.solve_batched(cases, w.nstep, restart=restart, tol=tol)
        bit2, bfl2, bsolved2, _ = c.solve_batched_device(dk, 5, w.nstep, du, dt_, restart=restart, tol=tol)
        assert np.array_equal(bsolved2, bsolved) and np.array_equal(bit2, bit) and np.array_equal(bfl2, bfl)
        assert np.array_equal(du.download(), BU) and np.array_equal(dt_.download(), BT)
    assert (bsolved == 1).all()
    c.close()
