"""The KronLogger dump format reader/writer (dmrgpp_amd.kronlog)."""
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


def _emit_reference_layout(path, d):
    # The kron_data31 entries in the dump's own layout: dense rows, or MatrixMarket coordinate
    # for the blocks the reference stored sparse, lower-part entries transposed.
    with open(path, "w") as fh:
        fh.write("## KronLogger: Hello from ctor\nVector\n%d\n" % len(d["y"]))
        for v in d["y"]:
            fh.write(repr(v) + "\n")
        for e in d["entries"]:
            fh.write("\n# nc, total, outPatch, inPatch, ic, applyAdjoint, performTranspose, "
                     "offsetX, offsetY\n")
            fh.write(" ".join(str(e[k]) for k in ("nc", "np", "out", "in", "ic", "applyAdjoint",
                                                   "performTranspose", "offsetX", "offsetY"))
                     + "\n")
            for tag in ("A", "B"):
                blk = e[tag]
                M = np.array(blk["data"]).reshape((blk["rows"], blk["cols"]), order="F")
                fh.write(f"## {tag}: {blk['format']}\n")
                if blk["format"] == "dense":
                    fh.write(f"{M.shape[0]} {M.shape[1]}\n")
                    for r in range(M.shape[0]):
                        fh.write(" ".join(repr(float(v)) for v in M[r]) + "\n")
                else:
                    nz = np.argwhere(M != 0)
                    fh.write("%%MatrixMarket matrix coordinate real general\n")
                    fh.write(f"{M.shape[0]} {M.shape[1]} {len(nz)}\n")
                    for r, c in nz:
                        fh.write(f"{r + 1} {c + 1} {float(M[r, c])!r}\n")


def test_reads_the_kron31_dump_layout(tmp_path):
    d = fixtures.kron31_entries()
    path = tmp_path / "kron_data31.txt"
    _emit_reference_layout(path, d)
    y, entries = kronlog.read(str(path))
    assert len(entries) == len(d["entries"]) and y.size == 70
    prob = kronlog.to_problem(entries, name="kron31")
    full, _, y_ref, _ = fixtures.kron31_problem()
    _same(full, prob)
    np.testing.assert_array_equal(y, y_ref)
    # The replayed problem is the ED-pinned Heisenberg L=8 sector.
    H = oracle.kron_form(prob)
    assert abs(np.linalg.eigvalsh(0.5 * (H + H.T))[0] - (-3.374932598687897)) < 1e-9


def test_reader_rejects_adjoint_and_malformed(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("# nc, total, x\n1 1 0 0 0 1 0 0 0\n## A: dense\n1 1\n2.0\n## B: dense\n1 1\n3.0\n")
    _, entries = kronlog.read(str(p))
    with pytest.raises(ValueError):
        kronlog.to_problem(entries)
    p.write_text("# nc, total, x\n1 1 0 0 0 0 0 0 0\n## A: dense\n1 2\n2.0\n## B: dense\n1 1\n3.0\n")
    with pytest.raises(ValueError):
        kronlog.read(str(p))
