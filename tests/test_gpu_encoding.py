"""Problem encoding on the device: a bag of str problems goes to the GPU as
its joined text (latin-1 bytes or UTF-32 code points) and a symbol table;
a kernel turns it into codes.  Results must equal the host-encoded path
(sequences given as lists of symbols) bit for bit, and a symbol outside the
alphabet must raise the reference encoder's error for the first offending
problem (pyalign/problems/factory.py:103-153).
"""

import zlib

import numpy as np
import pytest

import oracle as O
import pyalign_amd as pa
from pyalign_amd import solve as solve_mod
from tests.test_gpu_matrices import assert_matrices
from tests.test_gpu_parity import dense, rounded

pytestmark = pytest.mark.gpu

ALPHABETS = {
    "byte": "ACGT",                              # ByteAlphabetEncoder (latin-1 str alphabet)
    "dict20": list("ACDEFGHIKLMNPQRSTVWY"),      # AlphabetEncoder, latin-1 text
    "greek": list("αβγδεζηθ"),                   # AlphabetEncoder, 4-byte code points
}


def factory(name, rng, direction="maximize"):
    alphabet = ALPHABETS[name]
    K = len(alphabet)
    t = rng.integers(-3, 5, size=(K, K)).astype(np.float32) / 2
    t = np.triu(t) + np.triu(t, 1).T
    lookup = {(x, y): t[i, j] for i, x in enumerate(alphabet) for j, y in enumerate(alphabet)}
    return pa.alphabetic(alphabet, lambda x, y: lookup[(x, y)], direction=direction)


def bag(rng, name, n, lo, hi):
    alphabet = list(ALPHABETS[name])
    return [("".join(rng.choice(alphabet, int(rng.integers(lo, hi + 1)))),
             "".join(rng.choice(alphabet, int(rng.integers(lo, hi + 1))))) for _ in range(n)]


def no_host_encoding(monkeypatch):
    def refuse(*a, **k):
        raise AssertionError("the bag was encoded on the host")
    monkeypatch.setattr(solve_mod, "_encode_bag", refuse)


@pytest.mark.parametrize("name", list(ALPHABETS))
@pytest.mark.parametrize("codomain", [pa.Score, pa.Solution])
def test_indexed_text_equals_host_encoding(name, codomain, monkeypatch):
    rng = np.random.default_rng(zlib.crc32(repr((name, str(codomain), "enc")).encode()))
    pf = factory(name, rng)
    seqs = bag(rng, name, 300, 0, 90)
    solver = pa.LocalSolver(gap_cost=pa.AffineGapCost(3, 1), codomain=codomain)
    host = solver.solve([pf.new_problem(list(a), list(b)) for a, b in seqs])
    with monkeypatch.context() as m:
        no_host_encoding(m)
        dev = solver.solve([pf.new_problem(a, b) for a, b in seqs])
    for (a, b), h, d in zip(seqs, host, dev):
        if codomain is pa.Score:
            assert np.float32(h).tobytes() == np.float32(d).tobytes(), (a, b)
        else:
            assert np.float32(h.score).tobytes() == np.float32(d.score).tobytes(), (a, b)
            assert np.array_equal(np.asarray(h.path), np.asarray(d.path)), (a, b)
    # and every tenth problem against the oracle
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    goal = O.GOAL_SCORE if codomain is pa.Score else O.GOAL_ONE
    for k in range(0, len(seqs), 10):
        a, b = seqs[k]
        if not a or not b:  # (empty problems: tests/test_gpu_parity.py)
            continue
        ref = O.OracleResult(O.make_config(locality=O.LOCAL, gap=("affine", 3, 1), goal=goal,
                                           lim_s=lims[0], lim_t=lims[1]), dense(pf.new_problem(list(a), list(b))))
        got = dev[k] if codomain is pa.Score else dev[k].score
        assert np.float32(got).tobytes() == ref.score.tobytes(), (k, a, b)
        if codomain is pa.Solution:
            assert np.array_equal(np.asarray(dev[k].path, dtype=np.int32), ref.path()), (k, a, b)


def test_indexed_text_against_oracle_and_matrices(monkeypatch):
    rng = np.random.default_rng(41)
    pf = factory("greek", rng)
    seqs = bag(rng, "greek", 40, 1, 30)
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    probs = [pf.new_problem(a, b) for a, b in seqs]
    with monkeypatch.context() as m:
        no_host_encoding(m)
        sols = pa.GlobalSolver(gap_cost=pa.AffineGapCost(2, 0.5), codomain=pa.Solution).solve(probs)
    for p, s in zip(probs, sols):
        ref = O.OracleResult(O.make_config(locality=O.GLOBAL, gap=("affine", 2, 0.5), goal=O.GOAL_ONE,
                                           lim_s=lims[0], lim_t=lims[1]), dense(p))
        assert np.float32(s.score).tobytes() == ref.score.tobytes()
        assert np.array_equal(np.asarray(s.path, dtype=np.int32), ref.path())
    # Solution matrices re-solve one problem from its slice of the raw text
    p, s = probs[3], sols[3]
    ref = O.OracleResult(O.make_config(locality=O.GLOBAL, gap=("affine", 2, 0.5), goal=O.GOAL_ONE,
                                       lim_s=rounded(len(p.s)), lim_t=rounded(len(p.t))), dense(p))
    one = pa.GlobalSolver(gap_cost=pa.AffineGapCost(2, 0.5), codomain=pa.Solution).solve(p)
    assert_matrices(one, ref, (p.s, p.t))


@pytest.mark.parametrize("text", ["latin", "wide"])
def test_binary_text_equals_host_encoding(text, monkeypatch):
    rng = np.random.default_rng(43 if text == "latin" else 44)
    alphabet = list("ACGT") if text == "latin" else list("ACGT") + ["λ", "Ж"]
    seqs = [("".join(rng.choice(alphabet, int(rng.integers(0, 70)))),
             "".join(rng.choice(alphabet, int(rng.integers(0, 70))))) for _ in range(200)]
    pf = pa.binary(eq=2, ne=-1)
    solver = pa.SemiglobalSolver(gap_cost=pa.LinearGapCost(1), codomain=pa.Solution)
    host = solver.solve([pf.new_problem(np.frombuffer(a.encode("utf-32-le"), np.uint32),
                                        np.frombuffer(b.encode("utf-32-le"), np.uint32)) for a, b in seqs])
    with monkeypatch.context() as m:
        no_host_encoding(m)
        dev = solver.solve([pf.new_problem(a, b) for a, b in seqs])
    for (a, b), h, d in zip(seqs, host, dev):
        assert np.float32(h.score).tobytes() == np.float32(d.score).tobytes(), (a, b)
        assert np.array_equal(np.asarray(h.path), np.asarray(d.path)), (a, b)
    # and every tenth problem against the oracle
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    for k in range(0, len(seqs), 10):
        a, b = seqs[k]
        if not a or not b:
            continue
        S = np.where(np.array([ord(x) for x in a])[:, None] == np.array([ord(y) for y in b])[None, :], 2, -1)
        ref = O.OracleResult(O.make_config(locality=O.SEMIGLOBAL, gap=("linear", 1), goal=O.GOAL_ONE,
                                           lim_s=lims[0], lim_t=lims[1]), S.reshape(len(a), len(b)).astype(np.float32))
        assert np.float32(dev[k].score).tobytes() == ref.score.tobytes(), (k, a, b)
        assert np.array_equal(np.asarray(dev[k].path, dtype=np.int32), ref.path()), (k, a, b)


def test_symbol_outside_alphabet_errors():
    rng = np.random.default_rng(45)
    solver = pa.GlobalSolver(gap_cost=pa.LinearGapCost(1), codomain=pa.Score)
    # byte encoder: the reference's message (factory.py:121-122)
    pf = factory("byte", rng)
    with pytest.raises(ValueError, match="^item not in alphabet$"):
        solver.solve([pf.new_problem("ACGT", "ACGT"), pf.new_problem("ACXT", "AC")])
    # dict encoder: the first offending problem's symbol, s before t
    pf = factory("dict20", rng)
    probs = [pf.new_problem("ACD", "EFG"), pf.new_problem("ACD", "EFBG"), pf.new_problem("AZC", "EF")]
    with pytest.raises(ValueError, match="^'B' is not in alphabet$"):
        solver.solve(probs)
    probs = [pf.new_problem("ACD", "EFG"), pf.new_problem("AOD", "EFBG")]
    with pytest.raises(ValueError, match="^'O' is not in alphabet$"):
        solver.solve(probs)
    # the solver stays usable after the error
    assert solver.solve([pf.new_problem("ACD", "ACD")])[0] == pytest.approx(
        solver.solve([pf.new_problem(list("ACD"), list("ACD"))])[0])


def test_ragged_codes_beyond_the_table_raise():
    """Host uint8 codes of the ragged indexed form are checked on the device
    (an identity code table) before any DP work: a code >= K raises the
    reference's error for its side; valid codes solve as before."""
    from pyalign_amd import algorithm
    rng = np.random.default_rng(46)
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Score), locality=algorithm.Locality.LOCAL, gap_cost=pa.AffineGapCost(3, 1)))
    solver = algorithm.m16.create_solver(64, 64, opts)
    ls = np.array([10, 20, 5], np.int32)
    lt = np.array([7, 30, 9], np.int32)
    cs = rng.integers(0, 4, int(ls.sum())).astype(np.uint8)
    ct = rng.integers(0, 4, int(lt.sum())).astype(np.uint8)
    os_ = np.array([0, 10, 30], np.int64)
    ot = np.array([0, 7, 37], np.int64)
    table = np.eye(4, dtype=np.float32) * 3 - 1
    ok = solver.solve_ragged("score", "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=os_, off_t=ot, table=table)
    for k in range(3):
        S = table[cs[os_[k]:os_[k] + ls[k]]][:, ct[ot[k]:ot[k] + lt[k]]]
        ref = O.solve(S, locality=O.LOCAL, gap=("affine", 3, 1), lim_s=64, lim_t=64)
        assert np.float32(ok[k]).tobytes() == ref.score.tobytes()
    bad = cs.copy()
    bad[12] = 4
    with pytest.raises(ValueError, match="out of bounds index in a"):
        solver.solve_ragged("score", "indexed", ls, lt, codes_s=bad, codes_t=ct, off_s=os_, off_t=ot, table=table)
    bad = ct.copy()
    bad[40] = 200
    with pytest.raises(ValueError, match="out of bounds index in b"):
        solver.solve_ragged("score", "indexed", ls, lt, codes_s=cs, codes_t=bad, off_s=os_, off_t=ot, table=table)
    again = solver.solve_ragged("score", "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=os_, off_t=ot, table=table)
    assert np.array_equal(np.asarray(again, np.float32).view(np.uint32), np.asarray(ok, np.float32).view(np.uint32))
