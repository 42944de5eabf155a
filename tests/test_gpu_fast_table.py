"""The score-only local maximising kernel with its LDS table of row-symbol
pairs (two rows' similarities per lookup) against the same kernel with the
single-symbol table (PYALIGN_AMD_NO_PAIR_TABLE): scores bit for bit, and both
against the oracle on a sample, for a bag of uneven lengths like the 65,536-
pair workload's (affine 10 / 1, 20-letter integer table) and for linear gaps."""

import numpy as np
import pytest

import oracle as O
import pyalign_amd as pa
from tests.test_gpu_band import dense_of, lengths, make_solver, ragged_codes, sym_table

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gap", [("affine", 10, 1), ("linear", 2)], ids=str)
def test_pair_table_equals_single_symbol_table(gap, monkeypatch):
    rng = np.random.default_rng(91 if gap[0] == "affine" else 92)
    ls, lt = lengths(rng, 3000, 20, 400, 20, 400)
    cs, ct, os_, ot = ragged_codes(rng, ls, lt, 20)
    table = sym_table(rng, 20, -4, 12)
    out = []
    for pair in (True, False):
        with monkeypatch.context() as m:
            if not pair:
                m.setenv("PYALIGN_AMD_NO_PAIR_TABLE", "1")
            solver = make_solver("local", gap, "maximize", pa.Score, 512, 512)
        out.append(np.asarray(solver.solve_ragged("score", "indexed", ls, lt, codes_s=cs, codes_t=ct, off_s=os_,
                                                  off_t=ot, table=table), dtype=np.float32))
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))
    for k in list(range(0, len(ls), 97)):
        S = dense_of("indexed", k, cs, ct, os_, ot, ls, lt, table)
        ref = O.OracleResult(O.make_config(locality=O.LOCAL, gap=gap, goal=O.GOAL_SCORE, lim_s=512, lim_t=512), S)
        assert out[0][k].tobytes() == ref.score.tobytes(), k
