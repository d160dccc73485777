"""Host-only checks: the C-ABI library loads and exports every entry point
declared in include/tdbem_capi.h; input readers and the box generator."""
import ctypes
import re

import numpy as np
import pytest

from conftest import GOLDEN, PKG, REPO


def test_capi_exports_every_declared_symbol():
    import tdbem
    if not tdbem.LIB_PATH.exists():
        tdbem.build()
    header = (REPO / "include" / "tdbem_capi.h").read_text()
    declared = sorted(set(re.findall(r"\b(tdbem_[a-z0-9_]+)\s*\(", header)))
    assert declared, "no declarations found"
    lib = ctypes.CDLL(str(tdbem.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(tdbem.EXPORTED) == declared


def test_datacard_and_model_readers():
    from tdbem import io
    card = io.read_datacard(GOLDEN / "rod120_100" / "BEM_DATACARD.DAT")
    assert (card.nstep, card.beta, card.gap, card.iterations, card.error, card.premaxleaf) == (100, 0.8, 0.4, 200, 1e-5, 10)
    assert card.domain_count == 5 and len(card.interfaces) == 16
    pts, ele = io.read_model(GOLDEN / "rod120_100" / "rod120_100.model")
    assert pts.shape == (286, 3) and ele.shape == (120, 8) and ele.min() == 0
    bcid, lu = io.read_bd(GOLDEN / "rod120_100" / "rod120_100.bd", 120, card.nstep)
    assert (bcid == 123).sum() == 4 and lu.shape == (101, 120, 3)
    assert np.all(lu[0] == 0) and np.all(lu[1, 100:104, 2] == 1.0)
    mats, itf = io.normalise_model(card, 120)
    assert itf[0].tolist() == [0, 4, 1, 24, -1]


def test_box_generator_counts():
    from tdbem import mesh
    for (D, L, W, H, nx, ny, nz) in [(5, 5.0, 1.0, 1.0, 2, 2, 2), (4, 4.0, 2.0, 2.0, 2, 4, 4), (8, 8.0, 2.0, 2.0, 1, 2, 2)]:
        m = mesh.box_mesh(D, L, W, H, nx, ny, nz)
        assert len(m.ele) == D * mesh.elements_per_domain(nx, ny, nz)
        assert len(m.interfaces) == (D - 1) * ny * nz
        assert (m.bcid == 123).sum() == ny * nz
    # configuration sizes quoted for the workloads
    assert 4 * mesh.elements_per_domain(10, 20, 20) == 6400
    assert 8 * mesh.elements_per_domain(8, 16, 16) == 8192
