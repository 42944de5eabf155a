"""The plain global alignment workload at full size: 4,096 pairs of 100 x 100
uniform ACGT strings, Equality(eq=1, ne=-1), LinearGapCost(0.2), codomain
Solution (the reference's CPU-runnable configuration).  Every score must
equal the oracle's bit for bit and every path must be the oracle's
tie-broken path, with the general (dense) and the alphabetic forms."""

import numpy as np
import pytest

import oracle as O
import pyalign_amd as pa

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["general", "alphabetic", "binary"])
def test_global_solution_batch_full_size(form):
    import bench
    pairs = bench.c1_inputs()
    if form == "general":
        pf = pa.general(pa.Equality(eq=1, ne=-1), direction="maximize")
    elif form == "alphabetic":
        pf = pa.alphabetic("ACGT", pa.Equality(eq=1, ne=-1), direction="maximize")
    else:
        pf = pa.binary(eq=1, ne=-1)
    probs = [pf.new_problem(a, b) for a, b in pairs]
    sols = pa.GlobalSolver(gap_cost=pa.LinearGapCost(0.2), codomain=pa.Solution).solve(probs)
    assert len(sols) == 4096
    cfg = O.make_config(locality=O.GLOBAL, gap=("linear", 0.2), goal=O.GOAL_ONE, lim_s=128, lim_t=128)
    for (a, b), sol in zip(pairs, sols):
        S = np.where(np.frombuffer(a.encode(), np.uint8)[:, None] == np.frombuffer(b.encode(), np.uint8)[None, :],
                     1.0, -1.0).astype(np.float32)
        ref = O.OracleResult(cfg, S)
        assert np.float32(sol.score).tobytes() == ref.score.tobytes(), (a, b)
        assert np.array_equal(np.asarray(sol.path, dtype=np.int32), ref.path()), (a, b)
