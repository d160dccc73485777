"""The solve without device allocation (tdbem_solve_device_noalloc) against
the device-resident solve it is a variant of (tdbem_solve_device), through the
C-ABI: bit-identical displacements, tractions, iteration counts and flags on
two calls in a row, the second with a smaller input, and the context's own
solves unchanged afterwards; shapes beyond the scratch array are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tdbem():
    import tdbem
    if not tdbem.LIB_PATH.exists():
        pytest.fail("HIP library not built")
    return tdbem


def test_solve_without_allocation_matches_the_device_resident_solve():
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=40)
    m = w.mesh
    print("case:", w.name, len(m.ele), "elements, solve without allocation", flush=True)
    c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    cases = mesh.load_cases(w, 2)
    for nstep, known in ((40, cases[0]), (25, cases[1][:26])):    # the second call smaller
        dk = tdbem.DeviceArray.from_host(known)
        ru, rt = tdbem.DeviceArray(known.shape), tdbem.DeviceArray(known.shape)
        na_u, na_t = tdbem.DeviceArray(known.shape), tdbem.DeviceArray(known.shape)
        it0, fl0, s0, _ = c.solve_device(dk, nstep, ru, rt, tol=1e-10)
        it1, fl1, s1 = c.solve_device_noalloc(dk, nstep, na_u, na_t, tol=1e-10)
        assert s1 == s0 == nstep
        assert np.array_equal(it1, it0) and np.array_equal(fl1, fl0)
        assert np.array_equal(na_u.download(), ru.download()), nstep
        assert np.array_equal(na_t.download(), rt.download()), nstep
    # the context's own workspace is back: a host-buffer solve still gives the same result
    U, T, it, fl, solved, _ = c.solve(cases[1][:26], 25, tol=1e-10)
    assert np.array_equal(U, na_u.download()) and np.array_equal(it, it1)
    c.close()


def test_solve_without_allocation_refuses_what_does_not_fit():
    tdbem = _tdbem()
    from tdbem import mesh
    w = mesh.layered_config2(h=0.5, nstep=20)
    m = w.mesh
    print("case:", w.name, len(m.ele), "elements, oversize request", flush=True)
    c = tdbem.Case(m.points, m.ele, m.bcid, w.dt, w.gap, domain_count=4, mat_rows=w.mats, itf_rows=m.interfaces)
    c.assemble_step0(10)
    c.assemble_history(w.nstep)
    dk = tdbem.DeviceArray.from_host(w.known)
    du, dt_ = tdbem.DeviceArray(w.known.shape), tdbem.DeviceArray(w.known.shape)
    # a restart of 3000 needs a 3001 x 3001 Givens array alone: 72 MB
    with pytest.raises(RuntimeError, match="bytes of scratch"):
        c.solve_device_noalloc(dk, w.nstep, du, dt_, restart=3000, tol=1e-10)
    it, fl, solved = c.solve_device_noalloc(dk, w.nstep, du, dt_, restart=200, tol=1e-10)
    assert solved == w.nstep
    c.set_option("mixed", 2)
    c.assemble_history(20)
    with pytest.raises(RuntimeError, match="FP64 history bands"):
        c.solve_device_noalloc(dk, 20, du, dt_, tol=1e-10)
    c.close()
