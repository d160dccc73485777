"""Grouped models (tdbem_run_models_grouped): independent models of uneven
sizes stacked so that every kernel of the assembly and of the time loop is one
launch for all of them.  Each model keeps its own materials, time step, DOF
numbering, preconditioner leaves and GMRES, so its results must equal its own
job bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tdbem():
    import tdbem

    tdbem.lib()
    return tdbem


def _models(nstep_of=lambda i: 24):
    from tdbem import mesh

    ws = []
    for i, (n, E, v, rho) in enumerate(((4, 3.0, 0.25, 2.0), (6, 8.0, 0.3, 1.2), (3, 1.5, 0.22, 2.7))):
        m = mesh.box_mesh(1, 1.0, 1.0, 1.0, n, n, n)
        dt = 0.5 / n / mesh.wave_speeds(E, v, rho)[0]
        nstep = nstep_of(i)
        sig = mesh.ricker(np.arange(nstep + 1) * dt, 0.8 / (20 * dt))
        sig[0] = 0.0
        known = mesh.normal_load_history(len(m.ele), nstep, dt, m.top, sig)
        ws.append(mesh.Workload(f"cube{n}", m, np.array([[1, E, v, rho]]), dt, nstep, known))
    # a four-domain layered model with interfaces and its own seeded materials
    ws.insert(1, mesh.layered_config2(h=0.5, nstep=nstep_of(3)))
    return ws


def _case(tdbem, w):
    m = w.mesh
    return tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=len(w.mats), mat_rows=w.mats,
                      itf_rows=m.interfaces)


def _alone(tdbem, ws, tol, restart=200):
    out = []
    for w in ws:
        c = _case(tdbem, w)
        c.assemble_step0(10)
        c.assemble_history(w.nstep)
        out.append(c.solve(w.known, w.nstep, restart=restart, tol=tol)[:5])
        c.close()
    return out


def test_grouped_models_match_their_own_jobs():
    """Four models (cubes of 96, 216 and 54 elements with their own materials
    and time steps, and a four-domain layered box with interfaces) in one
    round, then with a budget that admits one model per round: every model's
    displacements, tractions, GMRES iteration counts, flags and solved steps
    equal its own job's bit for bit."""
    tdbem = _tdbem()
    ws = _models()
    print("case: grouped models", [len(w.mesh.ele) for w in ws], flush=True)
    alone = _alone(tdbem, ws, 1e-10)
    for budget, rounds in ((0.0, 1), (1.0, len(ws))):
        cs = [_case(tdbem, w) for w in ws]
        got = tdbem.run_models(cs, [w.known for w in ws], [w.nstep for w in ws], tol=1e-10, budget_bytes=budget,
                               grouped=True)
        for c in cs:
            c.close()
        assert tdbem.run_models.rounds == rounds
        for w, a, g in zip(ws, alone, got):
            assert g[4] == w.nstep
            for x, y in zip(a, g):
                assert np.array_equal(np.asarray(x), np.asarray(y)), (w.name, budget)


def test_grouped_models_split_rounds_by_step_count_and_refuse_mixed_bands():
    """Models with different step counts go to different rounds (same results
    as their own jobs); a context asking for mixed-precision bands is refused
    with an error that says so."""
    tdbem = _tdbem()
    ws = _models(lambda i: (18, 22, 18, 18)[i])  # in model order: 18, 18 (layered), 22, 18
    print("case: grouped models, step counts", [w.nstep for w in ws], flush=True)
    alone = _alone(tdbem, ws, 1e-10)
    cs = [_case(tdbem, w) for w in ws]
    got = tdbem.run_models(cs, [w.known for w in ws], [w.nstep for w in ws], tol=1e-10, grouped=True)
    assert tdbem.run_models.rounds == 3  # (0, 1), (2), (3): rounds keep the given order
    for w, a, g in zip(ws, alone, got):
        for x, y in zip(a, g):
            assert np.array_equal(np.asarray(x), np.asarray(y)), w.name
    cs[2].set_option("mixed", 1)
    with pytest.raises(RuntimeError, match="mixed"):
        tdbem.run_models(cs, [w.known for w in ws], [w.nstep for w in ws], tol=1e-10, grouped=True)
    for c in cs:
        c.close()


def test_grouped_models_stop_where_their_own_jobs_stop():
    """A short restart length at which some models' GMRES stops with flag 1
    (restart reached) at some step while the others finish: in the grouped
    run each stopped model keeps its flags, iteration counts, solved steps
    and zero outputs from its stopping step on, as its own job leaves them,
    and the others go on to equal their own jobs bit for bit."""
    tdbem = _tdbem()
    ws = _models(lambda i: 16)
    print("case: grouped models with stopping GMRES", [len(w.mesh.ele) for w in ws], flush=True)
    for restart in range(2, 12):
        alone = _alone(tdbem, ws, 1e-12, restart)
        stopped = [a[4] < w.nstep or bool(np.asarray(a[3]).any()) for a, w in zip(alone, ws)]
        if any(stopped) and not all(stopped):
            break
    else:
        pytest.fail("no restart length mixes stopped and finished models")
    print("restart", restart, "stopped", stopped, flush=True)
    cs = [_case(tdbem, w) for w in ws]
    got = tdbem.run_models(cs, [w.known for w in ws], [w.nstep for w in ws], restart=restart, tol=1e-12,
                           grouped=True)
    for c in cs:
        c.close()
    for w, a, g in zip(ws, alone, got):
        for x, y in zip(a, g):
            assert np.array_equal(np.asarray(x), np.asarray(y)), (w.name, restart)
