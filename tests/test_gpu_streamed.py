"""Streamed history bands (matrix-free regeneration when the bands exceed the
device budget): the blocks of lags <= 16 stay resident, the far part of every
16-step tile regenerates the rest one slab of source rows at a time.  The
reference holds every MTS[k] / MGS[k] instead (DBEM1/multidomain.cpp:3718-3790)
and reads them every step (3840-3893), so a streamed solve must give the
resident solve's answer."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tdbem():
    import tdbem
    if not tdbem.LIB_PATH.exists():
        pytest.fail("HIP library not built")
    return tdbem


def _case(tdbem, w, **opts):
    m = w.mesh
    c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=m.domain_count, mat_rows=w.mats,
                   itf_rows=m.interfaces)
    for k, v in opts.items():
        c.set_option(k, v)
    return c


@pytest.mark.parametrize("staged", [1, 0])
def test_streamed_history_matches_resident_bands(staged):
    """On the layered 4-domain box (h = 0.25, 1024 elements, 60 steps) with a
    budget that holds the near blocks and a fraction of the rest: at least two
    slabs; the solution bit for bit that of resident bands from the same
    quadrature kernels (staged, or lane per entry with quad_staged 0: the same
    blocks, convolved in the same order), the same iteration counts; within
    1e-10 relative of the other quadrature's resident solve; the consumers
    that need resident bands refuse."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.25, nstep=60)
    print("case:", w.name, "streamed history, quad_staged", staged, flush=True)
    ref = _case(tdbem, w, quad_staged=staged)
    ref.assemble_step0(10)
    entries = ref.assemble_history(w.nstep)
    U0, T0, it0, fl0, solved0, _ = ref.solve(w.known, w.nstep, tol=1e-12)
    ref.close()
    other = _case(tdbem, w, quad_staged=1 - staged)
    other.assemble_step0(10)
    other.assemble_history(w.nstep)
    U1, T1, it1, *_ = other.solve(w.known, w.nstep, tol=1e-12)
    other.close()
    budget_gb = 0.6 * entries * 121 / 1e9
    # quad_staged also selects the singular self-table kernel, as for `ref`
    st = _case(tdbem, w, quad_staged=staged, history_stream=1, history_budget_gb=budget_gb)
    st.assemble_step0(10)
    st.assemble_history(w.nstep)
    slabs = st.timings()["stream_slabs"]
    print("entries", entries, "budget GB", budget_gb, "slabs", slabs, flush=True)
    assert slabs >= 2
    U, T, it, fl, solved, _ = st.solve(w.known, w.nstep, tol=1e-12)
    assert solved == w.nstep and not np.any(fl)
    print("streamed vs resident:", np.abs(U - U0).max(), np.abs(T - T0).max(), flush=True)
    assert np.array_equal(U, U0) and np.array_equal(T, T0)
    assert np.array_equal(it, it0)
    assert np.abs(U - U1).max() <= 1e-10 * np.abs(U1).max()
    assert np.abs(T - T1).max() <= 1e-10 * np.abs(T1).max()
    with pytest.raises(RuntimeError, match="resident history bands"):
        st.solve_batched(np.stack([w.known, w.known]), w.nstep, tol=1e-12)
    with pytest.raises(RuntimeError, match="resident history bands"):
        st.ccsr_rows(2, 1, 0, 3)
    st.close()


def test_streamed_history_refused_or_chosen_by_budget():
    """history_stream 0 with a budget below the bands refuses the assembly;
    -1 (the default) streams when the bands exceed the budget and keeps them
    resident when they fit."""
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=30)
    c = _case(tdbem, w)
    c.assemble_step0(10)
    entries = c.assemble_history(w.nstep)
    assert c.timings()["stream_slabs"] == 0
    small = 0.7 * entries * 121 / 1e9
    c.set_option("history_stream", 0)
    c.set_option("history_budget_gb", small)
    with pytest.raises(RuntimeError, match="history_stream"):
        c.assemble_history(w.nstep)
    c.set_option("history_stream", -1)
    c.assemble_history(w.nstep)
    assert c.timings()["stream_slabs"] >= 1
    U, T, it, fl, solved, _ = c.solve(w.known, w.nstep, tol=1e-12)
    assert solved == w.nstep
    c.close()
