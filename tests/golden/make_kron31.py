"""Converts the reference's dumped Kron problem TestSuite/inputs/kron_data31_0_ref.txt (written by
KronLogger, dmrg/Engine/MatrixVectorKron/KronLogger.hh:112-174) into kron_data31.json.

The dump holds the vector y of one H*y product (70 values) and, for every (outPatch, inPatch, ic)
visited by KronConnections::doTask, the A and B blocks used (dense, or MatrixMarket coordinate
for sparse blocks). Entries with performTranspose=1 log the stored lower-part blocks
xc(inPatch, outPatch), yc(inPatch, outPatch). Blocks are written column major.

Run once in a tree that has the reference mounted:
    python tests/golden/make_kron31.py /root/reference/TestSuite/inputs/kron_data31_0_ref.txt
"""
import json
import os
import sys


def parse(path):
    lines = [ln.rstrip("\n") for ln in open(path)]
    i = 0
    y = None
    entries = []

    def read_block(i):
        head = lines[i]
        assert head.startswith("## A:") or head.startswith("## B:"), head
        kind = head.split(":")[1].strip()
        i += 1
        if kind == "dense":
            rows, cols = (int(t) for t in lines[i].split())
            i += 1
            mat = [[0.0] * cols for _ in range(rows)]
            for r in range(rows):
                vals = [float(t) for t in lines[i].split()]
                assert len(vals) == cols
                mat[r] = vals
                i += 1
        else:
            assert lines[i].startswith("%%MatrixMarket"), lines[i]
            i += 1
            rows, cols, nnz = (int(t) for t in lines[i].split())
            i += 1
            mat = [[0.0] * cols for _ in range(rows)]
            for _ in range(nnz):
                r, c, v = lines[i].split()
                mat[int(r) - 1][int(c) - 1] = float(v)
                i += 1
        data = [mat[r][c] for c in range(cols) for r in range(rows)]
        return i, {"rows": rows, "cols": cols, "format": kind, "data": data}

    while i < len(lines):
        ln = lines[i]
        if ln == "Vector":
            n = int(lines[i + 1])
            y = [float(lines[i + 2 + k]) for k in range(n)]
            i += 2 + n
            continue
        if ln.startswith("# nc, total"):
            f = [int(t) for t in lines[i + 1].split()]
            i += 2
            i, a = read_block(i)
            i, b = read_block(i)
            entries.append({"nc": f[0], "np": f[1], "out": f[2], "in": f[3], "ic": f[4],
                            "applyAdjoint": f[5], "performTranspose": f[6],
                            "offsetX": f[7], "offsetY": f[8], "A": a, "B": b})
            continue
        i += 1
    return y, entries


def main():
    src = sys.argv[1]
    y, entries = parse(src)
    out = {"source": "TestSuite/inputs/kron_data31_0_ref.txt", "y": y, "entries": entries}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kron_data31.json")
    with open(dst, "w") as fh:
        json.dump(out, fh)
    print(f"wrote {dst}: n={len(y)} entries={len(entries)}")


if __name__ == "__main__":
    main()
