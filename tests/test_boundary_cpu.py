"""CPU tests of the C ABI boundary and the host-side API (no GPU needed)."""

import ctypes
import os
import re
import typing

import numpy as np
import pytest

from pyalign_amd import _lib, algorithm
import pyalign_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "pyalign_hip.h")).read()
    return sorted(set(re.findall(r"\b(pa_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, name


def test_no_device_here_fails_loudly():
    if _lib.lib().pa_device_count() > 0:
        pytest.skip("a GPU is visible")
    solver = pa.LocalSolver(gap_cost=pa.LinearGapCost(1), codomain=pa.Score)
    pf = pa.binary(eq=1, ne=-1)
    with pytest.raises(RuntimeError, match="no HIP device"):
        solver.solve(pf.new_problem("ACGT", "AGT"))


def _opts(**kw):
    base = dict(solver=algorithm.Type.ALIGNMENT, batch=True, direction=algorithm.Direction.MAXIMIZE,
                return_dup=False, codomain=pa.solve.Codomain(pa.Score), locality=algorithm.Locality.LOCAL,
                gap_cost=pa.LinearGapCost(1))
    base.update(kw)
    return base


def test_general_gap_costs_accepted_mixed_kinds_rejected():
    o = algorithm.m16.create_options(_opts(gap_cost=pa.LogarithmicGapCost(3, 1)))
    assert o.gap_kind == "general"
    algorithm.m16.create_options(_opts(gap_cost={"s": pa.ExponentialGapCost(2, 1),
                                                 "t": pa.LogarithmicGapCost(3, 1)}))
    # the reference reads an absent general cost here (factory.h:1126-1133)
    with pytest.raises(ValueError, match="different kinds"):
        algorithm.m16.create_options(_opts(gap_cost={"s": pa.LinearGapCost(1), "t": pa.AffineGapCost(3, 1)}))
    with pytest.raises(ValueError, match="different kinds"):
        algorithm.m16.create_options(_opts(gap_cost={"s": pa.LogarithmicGapCost(3, 1), "t": None}))


def test_general_gap_cost_vector_length_checked():
    """check_gap_tensor_shape (solver.h:3047-3054): costs(n) must return n entries."""
    class Short(pa.gaps.GapCost):
        def costs(self, n):
            return np.zeros(n - 1, dtype=np.float32)

    o = algorithm.m16.create_options(_opts(gap_cost=Short()))
    with pytest.raises(ValueError, match="expected gap cost tensor length of 9, got 8"):
        algorithm.m16.create_solver(8, 8, o)


def test_asymmetric_gap_dict_accepted():
    algorithm.m16.create_options(_opts(gap_cost={"s": pa.AffineGapCost(3, 1), "t": pa.AffineGapCost(5, 2)}))
    algorithm.m16.create_options(_opts(gap_cost={"s": pa.LinearGapCost(1), "t": pa.LinearGapCost(2)}))
    algorithm.m16.create_options(_opts(gap_cost=None))
    algorithm.m16.create_options(_opts(gap_cost=pa.ConstantGapCost(0)))


def test_implementation_length_limit():
    o = algorithm.m16.create_options(_opts())
    # int16 index: maximum supported length is 32767 >> 1 (solver.h:496-503)
    algorithm.m16.create_solver(16383, 10, o)
    with pytest.raises(RuntimeError, match="exceeds maximum supported sequence length in this implementation 16383"):
        algorithm.m16.create_solver(16384, 10, o)
    o32 = algorithm.m32.create_options(_opts())
    algorithm.m32.create_solver(20000, 10, o32)


def test_batch_sizes():
    o = algorithm.m16.create_options(_opts())
    assert algorithm.m16.create_solver(8, 8, o).batch_size == 64
    o1 = algorithm.m16.create_options(_opts(batch=False))
    assert algorithm.m16.create_solver(8, 8, o1).batch_size == 1


def test_enums_match_reference_values():
    assert algorithm.Direction.__members__["MAXIMIZE"] == 1
    assert algorithm.Direction.__members__["MINIMIZE"] == 0
    assert [m.name for m in algorithm.Locality] == ["LOCAL", "GLOBAL", "SEMIGLOBAL"]
    assert [m.name for m in algorithm.Detail] == ["SCORE", "ALIGNMENT", "SOLUTION"]
    assert [m.name for m in algorithm.Count] == ["ONE", "ALL"]
    assert [m.name for m in algorithm.Type] == ["ALIGNMENT", "DTW"]


def test_codomain_keys_and_errors():
    assert pa.solve.Codomain(pa.Score).key == (algorithm.Detail.SCORE, algorithm.Count.ONE, None, "optimal")
    assert pa.solve.Codomain(typing.List[pa.Alignment]).key == (
        algorithm.Detail.ALIGNMENT, algorithm.Count.ALL, "list", "optimal")
    with pytest.raises(ValueError):
        pa.solve.Codomain(int)


def test_size_rounder_and_cache_limits():
    r = pa.solve.SizeRounder()
    assert [r(x) for x in (0, 1, 3, 100, 1000, 1023, 1024, 1500)] == [1, 1, 4, 128, 1024, 1024, 1024, 1500]
    cache = pa.solve.SolverCache(_opts())
    cache.ensure(100, 40)
    assert (cache._max_lim_s, cache._max_lim_t) == (128, 64)
    cache.ensure(20, 3000)
    assert (cache._max_lim_s, cache._max_lim_t) == (128, 3000)  # limits never shrink


def test_morton_batches_cover_every_problem_once():
    rng = np.random.default_rng(1)
    pf = pa.binary()
    probs = [pf.new_problem("A" * int(a), "C" * int(b)) for a, b in rng.integers(0, 40, size=(200, 2))]
    bag = pa.ProblemBag(probs)
    seen = []
    for batch in bag.batches(16):
        assert len(batch.problems) <= 16
        for p in batch.problems:
            assert p.shape[0] <= batch.shape[0] and p.shape[1] <= batch.shape[1]
        seen += list(batch.indices)
    assert sorted(seen) == list(range(200))
    assert pa.problems.instance.interleave2(3, 5) == 0b100111


def test_encoders():
    enc = pa.ByteAlphabetEncoder("ACGT", "latin1")
    codes = enc.encode("GATTACA")
    assert "".join(enc.alphabet[c] for c in codes) == "GATTACA"
    with pytest.raises(ValueError, match="item not in alphabet"):
        enc.encode("GATXACA")
    flat, off, lens = enc.encode_many(["GA", "", "TTC"])
    assert lens.tolist() == [2, 0, 3] and off.tolist() == [0, 2, 2]
    e2 = pa.AlphabetEncoder(["x1", "y2", "z3"])
    with pytest.raises(ValueError, match="is not in alphabet"):
        e2.encode(["x1", "q"])
    ie = pa.IntegerEncoder()
    assert ie.encode("AB").tolist() == [65, 66]
    with pytest.raises(ValueError):
        ie.encode(np.zeros(3, dtype=np.float32))


def test_bag_checks():
    a = pa.binary(direction="maximize").new_problem("AC", "A")
    b = pa.binary(direction="minimize").new_problem("AC", "A")
    with pytest.raises(ValueError, match="same direction"):
        pa.ProblemBag([a, b])


def test_gap_cost_tables():
    assert pa.AffineGapCost(10, 1).to_special_case() == {"affine": (9, 1)}
    assert pa.LinearGapCost(0.2).to_special_case() == {"linear": 0.2}
    assert pa.ConstantGapCost(0).to_special_case() == {"linear": 0}
    assert pa.ConstantGapCost(2).to_special_case() == {}
    np.testing.assert_allclose(pa.AffineGapCost(3, 1).costs(4), [0, 3, 4, 5])
    with pytest.raises(RuntimeError):
        pa.AffineGapCost(1, 2)


def test_char_lut_matches_encoders():
    """The device encoder's symbol tables give the host encoders' codes."""
    from pyalign_amd.problems.factory import AlphabetEncoder, ByteAlphabetEncoder
    enc = ByteAlphabetEncoder("ACGT", "latin1")
    lut = enc.char_lut(1)
    for x in "ACGT":
        assert lut[ord(x)] == enc.encode(x)[0]
    assert lut[ord("X")] == 255 and enc.char_lut(4) is None
    enc = AlphabetEncoder(list("ACDEFGHIKLMNPQRSTVWY"))
    for width in (1, 4):
        lut = enc.char_lut(width)
        for x in "ACDEFGHIKLMNPQRSTVWY":
            assert lut[ord(x)] == enc.encode([x])[0]
        assert lut[ord("B")] == 255
    enc = AlphabetEncoder(list("αβγ"))
    lut = enc.char_lut(4)
    assert [lut[ord(x)] for x in "αβγ"] == enc.encode(list("αβγ"))
    assert AlphabetEncoder(["ab", "c"]).char_lut(4) is None  # multi-character symbols stay on the host


def test_text_payload_layout():
    from pyalign_amd import solve as S
    pf = pa.alphabetic(list("ACDEFGHIKLMNPQRSTVWY"), lambda x, y: float(x == y))
    probs = [pf.new_problem("ACD", "EF"), pf.new_problem("", "GHIK"), pf.new_problem("W", "")]
    ls = np.array([3, 0, 1], np.int32)
    lt = np.array([2, 4, 0], np.int32)
    payload, text = S._text_payload(probs, ls, lt, True)
    assert payload["raw_width"] == 1
    assert bytes(payload["codes_s"]) == b"ACDW" and bytes(payload["codes_t"]) == b"EFGHIK"
    assert payload["off_s"].tolist() == [0, 3, 3] and payload["off_t"].tolist() == [0, 2, 6]
    probs = [pf.new_problem(list("ACD"), "EF")]
    assert S._text_payload(probs, ls[:1], lt[:1], True) == (None, None)


def test_collector_pause_is_shared_by_concurrent_builds():
    """The cyclic collector stays paused until the last concurrent result
    build ends, and one that was off before stays off."""
    import gc
    import threading
    from pyalign_amd import solve as S

    assert gc.isenabled()
    inside, release = threading.Event(), threading.Event()

    def other():
        with S._collector_paused():
            inside.set()
            release.wait(5)

    th = threading.Thread(target=other)
    th.start()
    assert inside.wait(5)
    with S._collector_paused():
        assert not gc.isenabled()
    assert not gc.isenabled()  # the other thread's build is still running
    release.set()
    th.join(5)
    assert gc.isenabled()
    gc.disable()
    try:
        with S._collector_paused():
            assert not gc.isenabled()
        assert not gc.isenabled()
    finally:
        gc.enable()
