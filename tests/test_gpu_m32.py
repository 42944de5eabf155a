"""The m32 index path on the GPU: a side of 16,000 or more makes the solver
cache pick Options<float, int32_t> (pyalign/solve.py:184-187), whose index
arrays are int32 with INT32_MIN as the no-traceback sentinel
(pyalign/algorithm/solver.h:182-185).  Compared with the oracle at 32-bit
indices: score bytes, path, alignment, Solution values and traceback matrix.
"""

import numpy as np
import pytest

import oracle as O
import pyalign_amd as pa

pytestmark = pytest.mark.gpu

INT32_MIN = np.iinfo(np.int32).min


def _pair(rng, ls, lt):
    letters = np.array(list("ACGT"))
    return "".join(rng.choice(letters, ls)), "".join(rng.choice(letters, lt))


@pytest.mark.parametrize("loc", ["global", "semiglobal"])
def test_m32_long_side_score_and_solution(loc):
    rng = np.random.default_rng(16500 + len(loc))
    a, b = _pair(rng, 16500, 300)
    pf = pa.binary(eq=1, ne=-1)
    problem = pf.new_problem(a, b)
    cls = pa.GlobalSolver if loc == "global" else pa.SemiglobalSolver
    oloc = O.GLOBAL if loc == "global" else O.SEMIGLOBAL
    # extend 0.5 is dyadic: the global border vector w(k) at the non-power-of-two
    # limit 16,500 is exact under any linspace evaluation order
    gap = pa.AffineGapCost(2.0, 0.5)
    lims = (16500, 512)
    S = np.where(np.frombuffer(a.encode(), np.uint8)[:, None] == np.frombuffer(b.encode(), np.uint8)[None, :],
                 1.0, -1.0).astype(np.float32)

    score = cls(gap_cost=gap, codomain=pa.Score).solve(problem)
    ref = O.OracleResult(O.make_config(locality=oloc, gap=("affine", 2.0, 0.5), goal=O.GOAL_SCORE,
                                       index_bits=32, lim_s=lims[0], lim_t=lims[1]), S)
    assert np.float32(score).tobytes() == ref.score.tobytes()

    sol = cls(gap_cost=gap, codomain=pa.Solution).solve(problem)
    ref = O.OracleResult(O.make_config(locality=oloc, gap=("affine", 2.0, 0.5), goal=O.GOAL_ONE,
                                       index_bits=32, lim_s=lims[0], lim_t=lims[1]), S)
    assert np.float32(sol.score).tobytes() == ref.score.tobytes()
    path = np.asarray(sol.path)
    assert path.dtype == np.int32
    assert np.array_equal(path, ref.path())
    al = sol.alignment
    assert al.s_to_t.dtype == np.int32 and al.t_to_s.dtype == np.int32
    assert sorted(al.edges.tolist()) == sorted(ref.edges().tolist())

    values = sol.values
    assert values.shape == (3, 16501, 301) and values.dtype == np.float32
    want = ref.values()
    assert np.array_equal(values.view(np.uint32), want.view(np.uint32))
    tb = sol.traceback("matrix")
    assert tb.dtype == np.int32 and tb.shape == (3, 16501, 301, 2)
    assert (tb == INT32_MIN).any()  # the 32-bit sentinel, where a cell has no predecessor
    assert np.array_equal(tb, ref.traceback_matrix())
