import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
PKG = REPO / "parallel3.0_omp_multidomin_1node_amd"
GOLDEN = Path(__file__).resolve().parent / "golden"
for p in (str(REPO), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    if not O.LIB_PATH.exists():
        O.build()
    return O


@pytest.fixture(scope="session")
def rod_fixture():
    """The reference's shipped rod120_100 case (DBEM1_val/rod120_100_analytic/...)
    with the polluted global material (E=200, rho=20) of its analytic-validation run."""
    from tdbem import io
    card = io.read_datacard(GOLDEN / "rod120_100" / "BEM_DATACARD.DAT")
    pts, ele = io.read_model(GOLDEN / "rod120_100" / "rod120_100.model")
    bcid, lu = io.read_bd(GOLDEN / "rod120_100" / "rod120_100.bd", len(ele), card.nstep)
    mats, itf = io.normalise_model(card, len(ele))
    expected = {}
    for line in (GOLDEN / "rod120_100" / "analytic_validation.txt").read_text().splitlines():
        k, v = line.split("=", 1)
        expected[k] = v
    return dict(card=card, pts=pts, ele=ele, bcid=bcid, lu=lu, mats=mats, itf=itf, expected=expected,
                E_global=200.0, rho_global=20.0)
