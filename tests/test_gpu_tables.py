"""Indexed-form tables on the GPU: rectangular K0 x K1 tables in bag solves
(s codes index rows, t codes columns, as indexed_matrix_form checks them,
pyalign/algorithm/factory.h:602-614) and alphabets above 128 symbols (the
reference's AlphabetEncoder takes any size, pyalign/problems/factory.py:133-153).
Every score, path and all-optimal enumeration is compared with the oracle.
"""

import typing

import numpy as np
import pytest
import torch

import oracle as O
import pyalign_amd as pa
from pyalign_amd import _lib, algorithm
from pyalign_amd.problems.instance import IndexedMatrixProblem

pytestmark = pytest.mark.gpu

SOLVERS = {"local": (pa.LocalSolver, O.LOCAL), "global": (pa.GlobalSolver, O.GLOBAL),
           "semiglobal": (pa.SemiglobalSolver, O.SEMIGLOBAL)}


def rounded(x):
    return pa.solve.SizeRounder()(x)


class RectProblem(IndexedMatrixProblem):
    """An indexed problem over a K0 x K1 table: s symbols and t symbols are
    drawn from different alphabets."""

    def __init__(self, table, a, b):
        super().__init__((len(a), len(b)), a, b, direction="maximize")
        self._table = table

    def similarity_lookup_table(self):
        return self._table

    def build_index_sequences(self, a, b):
        a[:] = self.s
        b[:] = self.t


def _oracle(S, loc, gap, goal, lims, direction=O.MAXIMIZE):
    return O.OracleResult(O.make_config(locality=loc, direction=direction, gap=gap, goal=goal,
                                        lim_s=lims[0], lim_t=lims[1]), S)


def _all(ref):
    out = []
    while True:
        x = ref.all_next()
        if x is None:
            return out
        out.append(x)


def _check_bag(problems, mats, loc, gap, gap_obj, lims, direction=O.MAXIMIZE):
    cls, oloc = SOLVERS[loc]
    scores = cls(gap_cost=gap_obj, codomain=pa.Score).solve(problems)
    sols = cls(gap_cost=gap_obj, codomain=pa.Solution).solve(problems)
    alls = cls(gap_cost=gap_obj, codomain=typing.List[pa.Alignment]).solve(problems)
    for k, S in enumerate(mats):
        ref = _oracle(S, oloc, gap, O.GOAL_SCORE, lims, direction)
        assert np.float32(scores[k]).tobytes() == ref.score.tobytes(), k
        ref = _oracle(S, oloc, gap, O.GOAL_ONE, lims, direction)
        if ref.has_seed:
            assert np.float32(sols[k].score).tobytes() == ref.score.tobytes(), k
            assert np.array_equal(np.asarray(sols[k].path, dtype=np.int32), ref.path()), k
        expect = _all(_oracle(S, oloc, gap, O.GOAL_ALL, lims, direction))[:100]
        assert len(alls[k]) == len(expect), k
        for g, (score, _, edges) in zip(alls[k], expect):
            assert np.float32(g.score).tobytes() == np.float32(score).tobytes()
            assert sorted(g.edges.tolist()) == sorted(edges.tolist())


@pytest.mark.parametrize("shape", [(4, 6), (6, 4)])
@pytest.mark.parametrize("loc", list(SOLVERS))
def test_rectangular_table_bag_matches_oracle(shape, loc):
    rng = np.random.default_rng(shape[0] * 10 + len(loc))
    K0, K1 = shape
    table = rng.integers(-3, 5, size=shape).astype(np.float32)
    problems, mats = [], []
    for _ in range(24):
        a = rng.integers(0, K0, int(rng.integers(0, 30))).astype(np.uint32)
        b = rng.integers(0, K1, int(rng.integers(0, 30))).astype(np.uint32)
        problems.append(RectProblem(table, a, b))
        mats.append(table[np.ix_(a, b)].astype(np.float32))
    lims = (rounded(max(len(p.s) for p in problems)), rounded(max(len(p.t) for p in problems)))
    gap = ("affine", 3, 1)
    _check_bag(problems, mats, loc, gap, pa.AffineGapCost(3, 1), lims)
    # a problem alone scores as it does in the bag
    one = SOLVERS[loc][0](gap_cost=pa.AffineGapCost(3, 1), codomain=pa.Score)
    bag = one.solve(problems)
    for k in (0, 5, 11):
        assert np.float32(one.solve(problems[k])).tobytes() == np.float32(bag[k]).tobytes()


def test_rectangular_table_codes_checked_per_side():
    rng = np.random.default_rng(3)
    table = rng.integers(-3, 5, size=(4, 6)).astype(np.float32)
    good_t = np.array([5, 4, 0], np.uint32)  # t codes may reach K1 - 1 = 5
    problems = [RectProblem(table, np.array([3, 1], np.uint32), good_t) for _ in range(3)]
    pa.GlobalSolver(gap_cost=pa.LinearGapCost(1), codomain=pa.Score).solve(problems)
    bad_s = [RectProblem(table, np.array([4], np.uint32), good_t) for _ in range(3)]  # s code 4 >= K0
    with pytest.raises(ValueError, match="out of bounds index in a"):
        pa.GlobalSolver(gap_cost=pa.LinearGapCost(1), codomain=pa.Score).solve(bad_s)
    table2 = table.T.copy()  # 6 x 4: t codes below 4
    bad_t = [RectProblem(table2, np.array([5], np.uint32), np.array([4], np.uint32)) for _ in range(3)]
    with pytest.raises(ValueError, match="out of bounds index in b"):
        pa.GlobalSolver(gap_cost=pa.LinearGapCost(1), codomain=pa.Score).solve(bad_t)


def test_rectangular_table_ragged_entry_checks_on_device():
    """Host uint8 codes of the ragged entry point are checked against each
    side's table size by the device encoder."""
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Score), locality=algorithm.Locality.GLOBAL, gap_cost=pa.LinearGapCost(1)))
    solver = algorithm.m16.create_solver(8, 8, opts)
    table = np.arange(12, dtype=np.float32).reshape(3, 4)
    ls = np.array([2, 2], np.int32)
    lt = np.array([2, 2], np.int32)
    kw = dict(off_s=np.array([0, 2]), off_t=np.array([0, 2]), table=table)
    got = solver.solve_ragged("score", "indexed", ls, lt, codes_s=np.array([0, 2, 1, 1], np.uint8),
                              codes_t=np.array([3, 0, 2, 3], np.uint8), **kw)
    for k, (a, b) in enumerate((([0, 2], [3, 0]), ([1, 1], [2, 3]))):
        ref = _oracle(table[np.ix_(a, b)], O.GLOBAL, ("linear", 1), O.GOAL_SCORE, (8, 8))
        assert np.float32(got[k]).tobytes() == ref.score.tobytes()
    with pytest.raises(ValueError, match="out of bounds index in a"):
        solver.solve_ragged("score", "indexed", ls, lt, codes_s=np.array([0, 3, 1, 1], np.uint8),
                            codes_t=np.array([3, 0, 2, 3], np.uint8), **kw)
    with pytest.raises(ValueError, match="out of bounds index in b"):
        solver.solve_ragged("score", "indexed", ls, lt, codes_s=np.array([0, 2, 1, 1], np.uint8),
                            codes_t=np.array([3, 0, 2, 4], np.uint8), **kw)


def _word_factory(rng, n_words):
    words = [f"w{k:03d}" for k in range(n_words)]
    tab = rng.integers(-4, 4, size=(n_words, n_words))
    tab = (np.triu(tab) + np.triu(tab, 1).T).astype(np.float32)
    np.fill_diagonal(tab, rng.integers(4, 12, size=n_words).astype(np.float32))
    index = {w: k for k, w in enumerate(words)}
    pf = pa.alphabetic(words, lambda x, y: tab[index[x], index[y]])
    return words, tab, index, pf


@pytest.mark.parametrize("loc", list(SOLVERS))
@pytest.mark.parametrize("gap", [("linear", 2), ("affine", 6, 1)], ids=str)
def test_word_alphabet_300_symbols(loc, gap):
    """A 300-word alphabet: the table is read from global memory with uint16
    codes (it does not fit in LDS)."""
    rng = np.random.default_rng(300 + len(loc) + len(gap))
    words, tab, index, pf = _word_factory(rng, 300)
    seqs = [(list(rng.choice(words, int(rng.integers(0, 40)))), list(rng.choice(words, int(rng.integers(0, 40)))))
            for _ in range(30)]
    problems = [pf.new_problem(a, b) for a, b in seqs]
    mats = [tab[np.ix_([index[x] for x in a], [index[y] for y in b])].astype(np.float32) for a, b in seqs]
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    gap_obj = pa.LinearGapCost(gap[1]) if gap[0] == "linear" else pa.AffineGapCost(gap[1], gap[2])
    _check_bag(problems, mats, loc, gap, gap_obj, lims)


def test_word_alphabet_long_and_matrices():
    """Long problems over a 300-word alphabet (several strips, row buffers)
    and their Solution matrices."""
    rng = np.random.default_rng(77)
    words, tab, index, pf = _word_factory(rng, 300)
    seqs = [(list(rng.choice(words, int(rng.integers(300, 700)))), list(rng.choice(words, int(rng.integers(300, 700)))))
            for _ in range(4)]
    problems = [pf.new_problem(a, b) for a, b in seqs]
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    sols = pa.LocalSolver(gap_cost=pa.AffineGapCost(6, 1), codomain=pa.Solution).solve(problems)
    for (a, b), s in zip(seqs, sols):
        S = tab[np.ix_([index[x] for x in a], [index[y] for y in b])].astype(np.float32)
        ref = _oracle(S, O.LOCAL, ("affine", 6, 1), O.GOAL_ONE, lims)
        assert np.float32(s.score).tobytes() == ref.score.tobytes()
        assert np.array_equal(np.asarray(s.path, dtype=np.int32), ref.path())
    small = [pf.new_problem(list(rng.choice(words, 12)), list(rng.choice(words, 9))) for _ in range(2)]
    for p in small:
        s = pa.GlobalSolver(gap_cost=pa.AffineGapCost(6, 1), codomain=pa.Solution).solve(p)
        S = p.matrix.astype(np.float32)
        ref = _oracle(S, O.GLOBAL, ("affine", 6, 1), O.GOAL_ONE, (16, 16))
        assert np.array_equal(s.values, ref.values())


def test_single_character_alphabet_of_200_symbols_from_text():
    """200 one-character symbols: the joined text is encoded on the device
    into uint16 codes (table above 128 symbols)."""
    rng = np.random.default_rng(200)
    alphabet = [chr(0x100 + k) for k in range(200)]
    tab = rng.integers(-4, 6, size=(200, 200)).astype(np.float32)
    index = {c: k for k, c in enumerate(alphabet)}
    pf = pa.alphabetic(alphabet, lambda x, y: tab[index[x], index[y]])
    seqs = [("".join(rng.choice(alphabet, int(rng.integers(1, 60)))), "".join(rng.choice(alphabet, int(rng.integers(1, 60)))))
            for _ in range(40)]
    problems = [pf.new_problem(a, b) for a, b in seqs]
    lims = (rounded(max(len(a) for a, _ in seqs)), rounded(max(len(b) for _, b in seqs)))
    got = pa.SemiglobalSolver(gap_cost=pa.AffineGapCost(5, 1), codomain=pa.Solution).solve(problems)
    for (a, b), s in zip(seqs, got):
        S = tab[np.ix_([index[x] for x in a], [index[y] for y in b])].astype(np.float32)
        ref = _oracle(S, O.SEMIGLOBAL, ("affine", 5, 1), O.GOAL_ONE, lims)
        assert np.float32(s.score).tobytes() == ref.score.tobytes()
        assert np.array_equal(np.asarray(s.path, dtype=np.int32), ref.path())
    with pytest.raises(ValueError, match="not in alphabet"):
        pa.SemiglobalSolver(gap_cost=pa.AffineGapCost(5, 1), codomain=pa.Score).solve(
            problems[:3] + [pf.new_problem("A" + alphabet[0], alphabet[1])])


def test_wide_table_reference_shaped_call():
    """solve_indexed_for_* (the reference's per-batch entry) with K = 300."""
    rng = np.random.default_rng(5)
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Solution), locality=algorithm.Locality.LOCAL,
        gap_cost=pa.AffineGapCost(4, 1)))
    solver = algorithm.m16.create_solver(64, 64, opts)
    B = solver.batch_size
    sim = rng.integers(-3, 6, size=(300, 260)).astype(np.float32)
    a = rng.integers(0, 300, size=(B, 40)).astype(np.uint32)
    b = rng.integers(0, 260, size=(B, 50)).astype(np.uint32)
    length = np.stack([rng.integers(0, 41, B), rng.integers(0, 51, B)]).astype(np.uint16)
    res = solver.solve_indexed_for_solution(a, b, sim, length)
    for k in range(0, B, 3):
        S = sim[np.ix_(a[k, :length[0, k]], b[k, :length[1, k]])]
        ref = _oracle(S, O.LOCAL, ("affine", 4, 1), O.GOAL_ONE, (64, 64))
        assert np.float32(res[k].score).tobytes() == ref.score.tobytes()
        if ref.has_seed:
            assert np.array_equal(np.asarray(res[k].path, dtype=np.int32), ref.path())
    with pytest.raises(ValueError, match="out of bounds index in b"):
        solver.solve_indexed_for_score(a, b + 260, sim, length)


def test_wide_table_device_resident_uint16_codes():
    """Codes already in HBM as uint16 (code_bytes 2), scores to a device
    buffer."""
    rng = np.random.default_rng(6)
    K = 1000
    tab = rng.integers(-4, 6, size=(K, K)).astype(np.float32)
    n = 200
    ls = rng.integers(1, 120, n).astype(np.int32)
    lt = rng.integers(1, 120, n).astype(np.int32)
    cs = rng.integers(0, K, int(ls.sum())).astype(np.int16)  # < 32768: same bits as uint16
    ct = rng.integers(0, K, int(lt.sum())).astype(np.int16)
    off_s = np.concatenate([[0], np.cumsum(ls)[:-1]]).astype(np.int64)
    off_t = np.concatenate([[0], np.cumsum(lt)[:-1]]).astype(np.int64)
    dcs = torch.from_numpy(cs).cuda()
    dct = torch.from_numpy(ct).cuda()
    scores = torch.empty(n, dtype=torch.float32, device="cuda")
    opts = algorithm.m16.create_options(dict(
        solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE, return_dup=False,
        codomain=pa.solve.Codomain(pa.Score), locality=algorithm.Locality.LOCAL, gap_cost=pa.AffineGapCost(7, 1)))
    solver = algorithm.m16.create_solver(128, 128, opts)
    rg, keep = algorithm.make_ragged("indexed", ls, lt, codes_s=dcs.data_ptr(), codes_t=dct.data_ptr(),
                                     off_s=off_s, off_t=off_t, table=tab, on_device=True, code_bytes=2)
    torch.cuda.synchronize()
    _lib.check(_lib.lib().pa_solve_ragged_scores_device(solver._h, rg, _lib.PA_SCHED_AUTO, scores.data_ptr()))
    _lib.check(_lib.lib().pa_solver_sync(solver._h))
    got = scores.cpu().numpy()
    for k in range(0, n, 5):
        S = tab[np.ix_(cs[off_s[k]:off_s[k] + ls[k]].astype(np.int64), ct[off_t[k]:off_t[k] + lt[k]].astype(np.int64))]
        ref = _oracle(S, O.LOCAL, ("affine", 7, 1), O.GOAL_SCORE, (128, 128))
        assert got[k].tobytes() == ref.score.tobytes(), k
    # uint8 device codes cannot address a table above 128 symbols
    rg8, keep8 = algorithm.make_ragged("indexed", ls, lt, codes_s=dcs.data_ptr(), codes_t=dct.data_ptr(),
                                       off_s=off_s, off_t=off_t, table=tab, on_device=True, code_bytes=1)
    with pytest.raises(ValueError, match="must be uint16"):
        _lib.check(_lib.lib().pa_solve_ragged_scores_device(solver._h, rg8, _lib.PA_SCHED_AUTO, scores.data_ptr()))
    del keep, keep8
