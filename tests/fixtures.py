"""Loaders that turn the committed golden fixtures into KronProblems."""
import json
import os

import numpy as np

from dmrgpp_amd.problem import KronProblem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def batchedgemm_fixture():
    """test_BatchedGemm.cpp:181-232: 1 patch, 1 operator, A 2x2, B 3x3."""
    d = json.load(open(os.path.join(GOLDEN, "batchedgemm_fixture.json")))
    A = np.array(d["A"]).reshape((2, 2), order="F")
    B = np.array(d["B"]).reshape((3, 3), order="F")
    prob = KronProblem(L=[2], R=[3], nC=1, blocks={(0, 0, 0): (A, B)}, name="test_BatchedGemm")
    return prob, np.array(d["vin"]), np.array(d["expected"]), d["tolerance"]


def kron_connections_fixture():
    """test_KronConnections.cpp:132-178: two 1x1 complex patches, one connection."""
    c = complex
    blocks = {
        (0, 0, 0): (np.array([[c(1, 1)]]), np.array([[c(2, -1)]])),
        (0, 1, 0): (np.array([[c(2, 1)]]), np.array([[c(1, 2)]])),
        (1, 0, 0): (np.array([[c(1, -2)]]), np.array([[c(3, 1)]])),
        (1, 1, 0): (np.array([[c(-1, 1)]]), np.array([[c(2, 2)]])),
    }
    prob = KronProblem(L=[1, 1], R=[1, 1], nC=1, blocks=blocks, dtype=np.complex128,
                       name="test_KronConnections")
    return prob


def kron31_entries():
    return json.load(open(os.path.join(GOLDEN, "kron_data31.json")))


def _mat(blk):
    return np.array(blk["data"]).reshape((blk["rows"], blk["cols"]), order="F")


def kron31_problem():
    """The Kron problem dumped by KronLogger for input31 (Heisenberg L=8).

    Returns (full_problem, lower_problem, y): the full problem has every (I,J) block (upper
    blocks are the transposes of the logged lower ones, as KronConnections' performTranspose
    applies them); the lower problem holds only the stored (out >= in) blocks, for the
    useLowerPart restatement.
    """
    d = kron31_entries()
    ents = d["entries"]
    npch = ents[0]["np"]
    nC = ents[0]["nc"]
    L = [0] * npch
    R = [0] * npch
    full, lower = {}, {}
    for e in ents:
        A, B = _mat(e["A"]), _mat(e["B"])
        o, i, k = e["out"], e["in"], e["ic"]
        if e["performTranspose"]:
            full[(o, i, k)] = (A.T.copy(), B.T.copy())
            lower[(i, o, k)] = (A, B)
        else:
            full[(o, i, k)] = (A, B)
            lower[(o, i, k)] = (A, B)
            L[o], R[o] = A.shape[0], B.shape[0]
            L[i], R[i] = A.shape[1], B.shape[1]
    y = np.array(d["y"])
    return (KronProblem(L=L, R=R, nC=nC, blocks=full, name="kron_data31"),
            KronProblem(L=L, R=R, nC=nC, blocks=lower, name="kron_data31-lower"), y, ents)
