"""Reference input / output formats through the C-ABI (host-only entry
points, no device needed): DBEMBIN v1 binary model and boundary files and the
text files they mirror, with the reference's binary-mode choice.

The expected bytes are built here with ``struct`` from the format the
reference reads (DBEM1/DBEM.cpp:89-330) and its converter writes
(convert_inputs_to_binary.py): an 8-byte magic "DBEMBIN\\0", uint32 version 1,
kind (1 model, 2 boundary), endian mark 0x01020304 and a reserved 0, then
little-endian int64 counts with float64 coordinates and int64 0-based node
ids (model), or int64 EleNum and NStep, int32 BCIDs and float64 triplets
(boundary)."""
import struct

import numpy as np
import pytest


def _lib_or_skip():
    import tdbem
    if not tdbem.LIB_PATH.exists():
        pytest.skip("HIP library not built")
    return tdbem


def _expected_model_bytes(pts, ele, inf):
    b = struct.pack("<8sIIII", b"DBEMBIN\0", 1, 1, 0x01020304, 0)
    b += struct.pack("<q", len(pts)) + struct.pack("<%dd" % pts.size, *pts.reshape(-1))
    b += struct.pack("<q", len(ele)) + struct.pack("<%dq" % ele.size, *ele.reshape(-1).tolist())
    b += struct.pack("<q", len(inf)) + (struct.pack("<%dq" % inf.size, *inf.reshape(-1).tolist()) if len(inf) else b"")
    return b


def _expected_bd_bytes(bcid, lu):
    ne, nstep = len(bcid), lu.shape[0] - 1
    b = struct.pack("<8sIIII", b"DBEMBIN\0", 1, 2, 0x01020304, 0)
    b += struct.pack("<qq", ne, nstep) + struct.pack("<%di" % ne, *bcid.tolist())
    return b + struct.pack("<%dd" % lu.size, *lu.reshape(-1))


def test_rod_fixture_binary_round_trip(tmp_path, rod_fixture):
    """The reference's shipped rod120_100 inputs: text read through the C++
    reader equals the Python reader; the DBEMBIN files written have exactly the
    reference layout; reading them back (mode 1, and mode 2 with the .bin
    present) gives the same arrays."""
    _lib_or_skip()
    from tdbem import io
    from conftest import GOLDEN
    model = GOLDEN / "rod120_100" / "rod120_100.model"
    bd = GOLDEN / "rod120_100" / "rod120_100.bd"
    f = rod_fixture
    ne, nstep = len(f["ele"]), f["card"].nstep
    pts, ele, inf = io.load_model(model, 0)
    assert np.array_equal(pts, f["pts"]) and np.array_equal(ele, f["ele"]) and inf.shape == (0, 4)
    bcid, lu = io.load_bd(bd, ne, nstep, 0)
    assert np.array_equal(bcid, f["bcid"]) and np.array_equal(lu, f["lu"])
    tm = tmp_path / "rod.model"
    tb = tmp_path / "rod.bd"
    io.write_model_bin(str(tm) + ".bin", pts, ele)
    io.write_bd_bin(str(tb) + ".bin", bcid, lu)
    assert (tmp_path / "rod.model.bin").read_bytes() == _expected_model_bytes(pts, ele.astype(np.int64), inf)
    assert (tmp_path / "rod.bd.bin").read_bytes() == _expected_bd_bytes(bcid, lu)
    for mode in (1, 2):
        p2, e2, _ = io.load_model(tm, mode)
        b2, l2 = io.load_bd(tb, ne, nstep, mode, expected_bcid=bcid)
        assert np.array_equal(p2, pts) and np.array_equal(e2, ele)
        assert np.array_equal(b2, bcid) and np.array_equal(l2, lu)


def test_binary_mode_choice_and_errors(tmp_path):
    """Mode 2 falls back to the text file when no .bin exists, mode 1 requires
    it; a boundary file for another size, a changed BCID or a bad header is
    refused with a message."""
    tdbem = _lib_or_skip()
    from tdbem import io
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1, 1, size=(10, 3))
    ele = rng.integers(0, 10, size=(2, 8)).astype(np.int32)
    text = tmp_path / "m.model"
    lines = [str(len(pts))] + [" ".join("%.17g" % v for v in p) for p in pts] + [str(len(ele))]
    lines += [" ".join(str(int(i) + 1) for i in e) for e in ele]
    text.write_text("\n".join(lines) + "\n")
    p0, e0, _ = io.load_model(text, 2)          # no .bin: text
    assert np.array_equal(p0, pts) and np.array_equal(e0, ele)
    with pytest.raises(RuntimeError, match="missing required binary"):
        io.load_model(text, 1)
    # a .bin with different contents wins in mode 2
    io.write_model_bin(str(text) + ".bin", pts * 2.0, ele)
    p2, _, _ = io.load_model(text, 2)
    assert np.array_equal(p2, pts * 2.0)
    p0, _, _ = io.load_model(text, 0)
    assert np.array_equal(p0, pts)
    bcid = np.array([123, 456], np.int32)
    lu = rng.normal(size=(4, 2, 3))
    bd = tmp_path / "b.bd"
    io.write_bd_bin(str(bd) + ".bin", bcid, lu)
    with pytest.raises(RuntimeError, match="size mismatch"):
        io.load_bd(bd, 2, 5, 1)
    with pytest.raises(RuntimeError, match="boundary type changed"):
        io.load_bd(bd, 2, 3, 1, expected_bcid=np.array([123, 123], np.int32))
    raw = bytearray((tmp_path / "b.bd.bin").read_bytes())
    raw[12] = 1                                  # kind 1 (model) instead of 2
    (tmp_path / "b.bd.bin").write_bytes(bytes(raw))
    with pytest.raises(RuntimeError, match="header mismatch"):
        io.load_bd(bd, 2, 3, 1)
    assert tdbem.lib().tdbem_last_error()


def test_generator_reproduces_the_reference_rod_case():
    """tdbem.mesh's box generator against the reference generator's own output, the shipped
    rod120_100 case (tools/generate_multidomain_rod.py:148-575 wrote it): the point table, the
    8-node elements, the BCIDs, the interface rows of the datacard and the boundary history are
    identical, bit for bit."""
    from conftest import GOLDEN
    from tdbem import io, mesh
    G = GOLDEN / "rod120_100"
    card = io.read_datacard(G / "BEM_DATACARD.DAT")
    pts, ele = io.read_model(G / "rod120_100.model")
    bcid, lu = io.read_bd(G / "rod120_100.bd", len(ele), card.nstep)
    w = mesh.reference_rod_case(dt=1.0, nstep=card.nstep, amplitude=card.amplitude)
    m = w.mesh
    assert m.points.shape == pts.shape and np.array_equal(m.points, pts)
    assert np.array_equal(m.ele, ele)
    assert np.array_equal(m.bcid, bcid)
    assert np.array_equal(np.asarray(m.interfaces), np.asarray(card.interfaces))
    assert np.array_equal(w.known, lu)
    assert card.domain_count == m.domain_count

