"""Host-side boundary checks that need no GPU: the library's
MultiDomainCCSRBuilder and GlobalDofMap reached through the C-ABI
(tdbem_builder_*, tdbem_dofmap_build), the reference's failure paths for
them and for the datacard tail, and the C++ drop-in header
(include/tdbem_multidomain.hpp) in a compiled caller."""
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG

DROPIN = PKG / "lib" / "rod_dropin"


def _tdbem():
    import tdbem
    if not tdbem.LIB_PATH.exists():
        tdbem.build()
    return tdbem


def test_builder_transposes_signs_and_drops_like_the_reference():
    """AddBlock (DBEM1/multidomain.cpp:1120-1143): the column-major input is
    stored transposed times the sign; null, zero-sign, all-zero and
    max|.| <= 1e-14 blocks are skipped, non-finite ones skipped and counted;
    an entry that cancels to <= 1e-14 is erased and a later add starts from
    zero; Build lists rows and columns ascending (1182-1214)."""
    t = _tdbem()
    rng = np.random.default_rng(0)
    b = t.Builder(4, 5)
    B = rng.normal(size=9)
    b.add_block(2, 3, -2.0, B)
    rp, col, val = b.build()
    assert rp.tolist() == [0, 0, 0, 1, 1] and col.tolist() == [3]
    np.testing.assert_array_equal(val[0], -2.0 * B.reshape(3, 3).T)
    for blk, sign in ((None, 1.0), (B, 0.0), (np.zeros(9), 1.0), (np.full(9, 1e-14), 1.0)):
        b.add_block(0, 0, sign, blk)
    b.add_block(4, 0, 1.0, B)          # out of range rows / columns are ignored
    b.add_block(0, 5, 1.0, B)
    assert b.info() == (1, 0)
    b.add_block(0, 0, 1.0, np.full(9, 2e-14))
    assert b.info() == (2, 0)
    bad = B.copy()
    bad[4] = np.nan
    b.add_block(1, 1, 1.0, bad)
    bad[4] = np.inf
    b.add_block(1, 2, 1.0, bad)
    assert b.info() == (2, 2)
    b.add_block(2, 3, 2.0, B)          # cancels: erased
    assert b.info()[0] == 1
    b.add_block(2, 3, 1.0, B)          # restarts from zero
    rp, col, val = b.build()
    assert rp.tolist() == [0, 1, 1, 2, 2] and col.tolist() == [0, 3]
    np.testing.assert_array_equal(val[1], B.reshape(3, 3).T)
    # MergeFrom adds in (row, col) order
    c = t.Builder(4, 5)
    c.add_block(3, 4, 1.0, B)
    c.add_block(0, 0, -1.0, np.full(9, 2e-14))  # cancels the entry of b after the merge
    b.merge(c)
    rp, col, val = b.build()
    assert rp.tolist() == [0, 0, 0, 1, 2] and col.tolist() == [3, 4]


def test_builder_symmetric_form():
    """BuildSym (DBEM1/multidomain.cpp:1216-1269): (b00 b01 b02 b11 b12 b22)
    per block; symmetric unless an off-diagonal pair differs by more than
    max(tolerance, 1e-14 + 1e-10 * max|block|)."""
    t = _tdbem()
    S = np.array([[4.0, 1.0, 2.0], [1.0, 5.0, 3.0], [2.0, 3.0, 6.0]])
    b = t.Builder(2, 2)
    b.add_block(0, 1, 1.0, S.T.ravel())
    rp, col, val, sym = b.build_sym(1e-10)
    assert sym and val[0].tolist() == [4.0, 1.0, 2.0, 5.0, 3.0, 6.0]
    A = S.copy()
    A[0, 1] += 5e-10 * 6.0            # inside 1e-10 * 6 only for the looser tolerance
    b2 = t.Builder(2, 2)
    b2.add_block(1, 1, 1.0, A.T.ravel())
    assert not b2.build_sym(1e-10)[3]
    assert b2.build_sym(1e-8)[3]


def test_dofmap_build_matches_oracle_and_fails_on_mixed_bcid(oracle_mod, rod_fixture):
    """GlobalDofMap on the host equals the oracle's map of the reference's
    rod fixture bit for bit; a BCID other than 123 / 456 makes both refuse,
    as DBEM1/multidomain.cpp:850-855 does."""
    t = _tdbem()
    f = rod_fixture
    card = f["card"]
    rows = np.array([[m[0], m[1], m[2], m[3]] for m in card.materials], float)
    d = t.dofmap_build(f["bcid"], card.domain_count, rows, np.array(card.interfaces))
    oc = oracle_mod.Case(f["pts"], f["ele"], f["bcid"], f["mats"], f["itf"], 0.1, card.gap)
    od = oc.dofmap()
    assert od["ok"]
    for k in d:
        assert np.array_equal(d[k], od[k]), k
    bad = f["bcid"].copy()
    bad[7] = 126
    with pytest.raises(RuntimeError, match="GlobalDofMap::Build failed"):
        t.dofmap_build(bad, card.domain_count, rows, np.array(card.interfaces))
    assert not oracle_mod.Case(f["pts"], f["ele"], bad, f["mats"], f["itf"], 0.1, card.gap).dofmap()["ok"]


def _run_dropin(*args):
    assert DROPIN.exists(), "build() makes lib/rod_dropin"
    return subprocess.run([str(DROPIN), *map(str, args)], capture_output=True, text=True, timeout=120)


def test_dropin_header_host_path():
    """A C++ caller of the drop-in header reads the reference's rod120_100
    inputs, builds GlobalDofMap (120 / 88 / 104 blocks) and round-trips a
    MultiDomainCCSRBuilder block, with no device."""
    r = _run_dropin(GOLDEN / "rod120_100" / "BEM_DATACARD.DAT", GOLDEN / "rod120_100" / "rod120_100.model",
                    GOLDEN / "rod120_100" / "rod120_100.bd", "--global-E", 200, "--host-only")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dofmap unknown=120 known=88 equation=120 historyU=104 historyT=104" in r.stdout
    assert "builder nnz=1 first=-1,-4,-7" in r.stdout


@pytest.mark.parametrize("tail,message", [
    ("0\n", "Invalid DomainCount 0"),
    ("2\n1 20 0 20\n2 20 0\n", "Invalid material line"),
    ("1\n1 20 0 20\n-1\n", "Invalid InterfaceCount -1"),
    ("2\n1 20 0 20\n2 20 0 20\n1\n1 5 2\n", "Invalid interface line"),
])
def test_malformed_datacard_tail_is_refused(tmp_path, tail, message):
    """ReadOptionalMultiDomainInput returns -1 on a malformed tail
    (DBEM1/multidomain.cpp:1315-1359): the drop-in's reader and the Python
    reader both refuse, naming the same line."""
    from tdbem import io
    head = (GOLDEN / "rod120_100" / "BEM_DATACARD.DAT").read_text().split()[:16]
    card = tmp_path / "BEM_DATACARD.DAT"
    card.write_text("\n".join(head) + "\n" + tail)
    with pytest.raises(ValueError, match=message.split(" line")[0]):
        io.read_datacard(card)
    r = _run_dropin(card, GOLDEN / "rod120_100" / "rod120_100.model", GOLDEN / "rod120_100" / "rod120_100.bd",
                    "--host-only")
    assert r.returncode == 1 and message in r.stdout, r.stdout
