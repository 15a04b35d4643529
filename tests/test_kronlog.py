"""The KronLogger dump format reader/writer (dmrgpp_amd.kronlog)."""
import os

import numpy as np
import pytest

from dmrgpp_amd import kronlog, problems
from oracle import oracle
from tests import fixtures


def _same(p, q):
    assert list(p.L) == list(q.L) and list(p.R) == list(q.R) and p.nC == q.nC
    assert set(p.blocks) == set(q.blocks)
    for k, (a, b) in p.blocks.items():
        np.testing.assert_array_equal(a, q.blocks[k][0])
        np.testing.assert_array_equal(b, q.blocks[k][1])


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_write_read_round_trip_is_exact(tmp_path, dtype):
    prob = problems.random_problem([3, 1, 4], [2, 5, 1], nC=3, density=0.7, seed=4, dtype=dtype)
    y = np.random.default_rng(1).uniform(-1, 1, prob.n).astype(dtype)
    path = tmp_path / "kron.txt"
    kronlog.write(str(path), prob, y)
    y2, entries = kronlog.read(str(path))
    np.testing.assert_array_equal(y2, y)
    _same(prob, kronlog.to_problem(entries))


DUMP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kron_data31_0_ref.txt")


def test_reads_the_reference_dump_bit_exact():
    # The reference's own dump TestSuite/inputs/kron_data31_0_ref.txt (KronLogger.hh:133-174,
    # written by the input31 run: Heisenberg L=8), kept as a data fixture, read by kronlog.read:
    # every value equals the independent conversion kron_data31.json (tests/golden/make_kron31.py)
    # bit for bit, including the dense / sparse (MatrixMarket) format of every block.
    d = fixtures.kron31_entries()
    y, entries = kronlog.read(DUMP)
    assert y.size == 70 and len(entries) == len(d["entries"])
    np.testing.assert_array_equal(y, np.array(d["y"]))
    for e, want in zip(entries, d["entries"]):
        for k in ("nc", "np", "out", "in", "ic", "applyAdjoint", "performTranspose", "offsetX",
                  "offsetY"):
            assert e[k] == want[k], k
        for tag in ("A", "B"):
            blk = want[tag]
            M = np.array(blk["data"]).reshape((blk["rows"], blk["cols"]), order="F")
            assert e[f"{tag}_format"] == blk["format"]
            np.testing.assert_array_equal(e[tag], M)
    formats = {e["A_format"] for e in entries} | {e["B_format"] for e in entries}
    assert formats == {"dense", "sparse"}
    # The replayed problem is the fixture's full problem and the ED-pinned L=8 sector.
    prob = kronlog.to_problem(entries, name="kron31")
    full, _, _, _ = fixtures.kron31_problem()
    _same(full, prob)
    off = prob.offsets
    for e in entries:
        assert off[e["out"]] == e["offsetX"] and off[e["in"]] == e["offsetY"]
    H = oracle.kron_form(prob)
    assert abs(np.linalg.eigvalsh(0.5 * (H + H.T))[0] - (-3.374932598687897)) < 1e-9


@pytest.mark.gpu
def test_gpu_replays_the_reference_dump():
    # The dump read by kronlog.read, through the HIP matvec (FP64) against the oracle and the
    # explicit Kronecker form, and its ground state by the device Lanczos against ED.
    import dmrgpp_amd as dm

    y, entries = kronlog.read(DUMP)
    prob = kronlog.to_problem(entries, name="kron31")
    want = oracle.plugin_apply(prob, y)
    np.testing.assert_allclose(want, oracle.kron_form(prob) @ y, atol=1e-13)
    op = dm.BatchedGemm(prob, "fp64")
    dx, dy = dm.DeviceArray.from_host(y), dm.DeviceArray(prob.n)
    op.apply_Htarget(dx, dy)
    op.sync()
    got = dy.download()
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    init = dm.DeviceArray.from_host(np.random.default_rng(1234).uniform(-1, 1, prob.n))
    dz = dm.DeviceArray(prob.n)
    r = op.lanczos_ground(init, dz)
    assert abs(r["energy"] - (-3.374932598687897)) < 1e-10 * 3.4
    op.close()


def test_reader_rejects_adjoint_and_malformed(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("# nc, total, x\n1 1 0 0 0 1 0 0 0\n## A: dense\n1 1\n2.0\n## B: dense\n1 1\n3.0\n")
    _, entries = kronlog.read(str(p))
    with pytest.raises(ValueError):
        kronlog.to_problem(entries)
    p.write_text("# nc, total, x\n1 1 0 0 0 0 0 0 0\n## A: dense\n1 2\n2.0\n## B: dense\n1 1\n3.0\n")
    with pytest.raises(ValueError):
        kronlog.read(str(p))
