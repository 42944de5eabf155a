"""Solvers and codomains (the pyalign.solve surface, pyalign/solve.py:1-621).

Same classes, arguments, results and errors as the reference.  One change
of strategy, not of results: a bag of several problems is solved with one
native call over the whole bag (``Solver.solve_ragged``), where the
reference loops over SIMD-sized batches in Python (solve.py:554-560).
Each problem's cells depend only on its own region (solver.h:2593-2705),
so the grouping cannot change any result; the solver limits, which do
affect global affine borders, are chosen exactly as the reference's
SolverCache does.
"""

import contextlib
import itertools
import time
import typing
from collections.abc import Sequence
from functools import cached_property

import numpy as np
from operator import attrgetter, is_, methodcaller

from . import algorithm
from .gaps import GapCost, ConstantGapCost
from .problems import ProblemBag, Form
from .problems.instance import Problem
from .problems.factory import AlphabetEncoder, AlphabetProblem, SeriesProblem


class Solution:
    """The optimal solution of one problem: score, alignment, path."""

    def __init__(self, problem, solver, solution):
        self._problem = problem
        self._solver = solver
        self._solution = solution
        self._traceback = {}

    @property
    def problem(self):
        return self._problem

    @property
    def score(self):
        return self._solution.score

    @cached_property
    def alignment(self):
        if self._solution.alignment is None:
            return None
        return Alignment(self._problem, self._solver, self._solution.alignment)

    @property
    def shape(self):
        return self._solution.values.shape

    @cached_property
    def values(self):
        return self._solution.values

    def traceback(self, form="matrix"):
        if form not in self._traceback:
            if form == "matrix":
                self._traceback[form] = self._solution.traceback_as_matrix
            elif form == "edges":
                self._traceback[form] = self._solution.traceback_as_edges
            else:
                return ValueError(form)
        return self._traceback[form]

    @cached_property
    def path(self):
        return self._solution.path

    @property
    def complexity(self):
        return self._solution.algorithm


class Alignment:
    """A specific alignment that is part of a solution."""

    def __init__(self, problem, solver, alignment):
        self._problem = problem
        self._solver = solver
        self._alignment = alignment

    @property
    def problem(self):
        return self._problem

    @property
    def solver(self):
        return self._solver

    @property
    def score(self):
        return self._alignment.score

    @property
    def s_to_t(self):
        return self._alignment.s_to_t

    @property
    def t_to_s(self):
        return self._alignment.t_to_s

    @cached_property
    def edges(self):
        s_to_t = self.s_to_t
        i = np.nonzero(s_to_t >= 0)[0]
        return np.column_stack([i, s_to_t[i]])


class Score:
    pass


def next_power_of_2(x):
    return 1 if x == 0 else 2 ** (int(x) - 1).bit_length()


class SizeRounder:
    """Limits below `limit` round up to a power of two; larger stay exact."""

    def __init__(self, limit=1024):
        self._limit = limit

    def __call__(self, x):
        return x if x >= self._limit else next_power_of_2(x)


class SolverCache:
    """Native solvers keyed by (direction, batch); limits only ever grow, and
    growing them drops every cached solver (solve.py:156-196)."""

    def __init__(self, options):
        self._options = options
        self._max_lim_s = 0
        self._max_lim_t = 0
        self._solvers = {}
        self._rounder = SizeRounder()

    def ensure(self, len_s, len_t):
        lim_s = max(self._max_lim_s, self._rounder(len_s))
        lim_t = max(self._max_lim_t, self._rounder(len_t))
        if lim_s > self._max_lim_s or lim_t > self._max_lim_t:
            self._max_lim_s = lim_s
            self._max_lim_t = lim_t
            self._solvers = {}

    def get(self, len_s, len_t, direction, batch):
        self.ensure(len_s, len_t)
        key = (direction, batch)
        solver = self._solvers.get(key)
        if solver is None:
            options = self._options.copy()
            options["direction"] = algorithm.Direction.__members__[direction.upper()]
            options["batch"] = batch
            options["return_dup"] = options.get("return_dup", False)
            bits = algorithm.m16 if (self._max_lim_s < 16000 and self._max_lim_t < 16000) else algorithm.m32
            solver = bits.create_solver(self._max_lim_s, self._max_lim_t, bits.create_options(options))
            self._solvers[key] = solver
        return solver


class Codomain:
    _1 = {Score, Alignment, Solution}
    _n = dict(
        [(typing.Iterator[x], (x, "all", "iterator")) for x in [Score, Alignment, Solution]] +
        [(typing.List[x], (x, "all", "list")) for x in [Score, Alignment, Solution]])

    def __init__(self, type):
        self._type = type
        if type in Codomain._1:
            self._base_type = type
            self._count = "one"
            container_type = None
        else:
            entry = Codomain._n.get(type)
            if entry is None:
                raise ValueError(f"illegal codomain type '{type}'")
            self._base_type, self._count, container_type = entry
        self._optimal = True
        self._key = (self.detail, self.count, container_type) + (("optimal",) if self._optimal else ())

    @property
    def type(self):
        return self._type

    def __str__(self):
        return str(self.type)

    @property
    def detail(self):
        return algorithm.Detail.__members__[self._base_type.__name__.upper()]

    @property
    def count(self):
        return algorithm.Count.__members__[self._count.upper()]

    @property
    def key(self):
        return self._key


class NoTimings:
    @contextlib.contextmanager
    def measure(self, name):
        yield


class Timings:
    """Accumulates wall time of the "prepare" and "solve" phases."""

    def __init__(self, solver):
        self._solver = solver
        self._timings = {}

    def __enter__(self):
        self._solver._timings = self
        return self

    def __exit__(self, type, value, traceback):
        self._solver._timings = NoTimings()

    @contextlib.contextmanager
    def measure(self, name):
        t0 = time.perf_counter_ns()
        yield
        self._timings[name] = self._timings.get(name, 0) + (time.perf_counter_ns() - t0)

    def get(self):
        return self._timings

    def _ipython_display_(self):
        for k, t in self._timings.items():
            print(f"{k}: {t / 1000:.1f} µs")


def chunks(items, n):
    for i in range(0, len(items), n):
        yield items[i:i + n]


class Iterator:
    def __init__(self, problem, solver, iterator):
        self._problem = problem
        self._solver = solver
        self._iterator = iterator

    def is_exhausted(self):
        return self._iterator.next() is None

    def __iter__(self):
        while True:
            x = self._iterator.next()
            if x is None:
                break
            yield self._element_class(self._problem, self._solver, x)


class AlignmentIterator(Iterator):
    _element_class = Alignment


class SolutionIterator(Iterator):
    _element_class = Solution


max_results = 100


def make_list_factory(iterator):
    def to_list(*args, **kwargs):
        it = iterator(*args, **kwargs)
        r = list(itertools.islice(it, max_results))
        if not it.is_exhausted():
            raise RuntimeError(f"more than {max_results} results, please use an iterator")
        return r

    return to_list


def solver_variants(prefix):
    data = {
        (algorithm.Detail.SCORE, algorithm.Count.ONE, None, "optimal"): ("for_score", lambda _1, _2, x: x),
        (algorithm.Detail.ALIGNMENT, algorithm.Count.ONE, None, "optimal"): ("for_alignment", Alignment),
        (algorithm.Detail.ALIGNMENT, algorithm.Count.ALL, "iterator", "optimal"): (
            "for_alignment_iterator", AlignmentIterator),
        (algorithm.Detail.ALIGNMENT, algorithm.Count.ALL, "list", "optimal"): (
            "for_alignment_iterator", make_list_factory(AlignmentIterator)),
        (algorithm.Detail.SOLUTION, algorithm.Count.ONE, None, "optimal"): ("for_solution", Solution),
        (algorithm.Detail.SOLUTION, algorithm.Count.ALL, "iterator", "optimal"): (
            "for_solution_iterator", SolutionIterator),
        (algorithm.Detail.SOLUTION, algorithm.Count.ALL, "list", "optimal"): (
            "for_solution_iterator", make_list_factory(SolutionIterator)),
    }
    return {k: (f"{prefix}_{v1}", v2) for k, (v1, v2) in data.items()}


def _variant(forms_cls, codomain):
    variant = forms_cls._solvers.get(codomain.key)
    if variant is None:
        raise ValueError(f"codomain {codomain} is currently not supported")
    return variant


class MatrixForm:
    """Dense per-lane similarity matrices (solve.py:342-379)."""
    _solvers = solver_variants("solve")

    def __init__(self, solver, codomain, bag):
        self._solver = solver
        batch_size = solver.batch_size
        max_shape = bag.max_shape
        self._matrix = np.zeros((max_shape[0], max_shape[1], batch_size), dtype=bag.dtype)
        self._len = np.zeros((2, batch_size), dtype=np.uint16)
        name, self._construct = _variant(MatrixForm, codomain)
        self._solve = getattr(solver, name)

    def prepare(self, batch):
        shape = batch.shape
        matrix = self._matrix[:shape[0], :shape[1], :]
        for k, p in enumerate(batch.problems):
            p.build_matrix(matrix[:, :, k])
            self._len[0, k], self._len[1, k] = p.shape

    def solve(self, batch):
        shape = batch.shape
        r = self._solve(self._matrix[:shape[0], :shape[1], :], self._len)
        return [self._construct(p, self._solver, x) for p, x in zip(batch.problems, r)]


class AbstractIndexedMatrixForm:
    def prepare(self, batch):
        shape = batch.shape
        for k, p in enumerate(batch.problems):
            p.build_index_sequences(self._a[k, :shape[0]], self._b[k, :shape[1]])
            self._len[0, k], self._len[1, k] = p.shape


class IndexedMatrixForm(AbstractIndexedMatrixForm):
    """Index sequences into a shared lookup table (solve.py:399-432)."""
    _solvers = solver_variants("solve_indexed")

    def __init__(self, solver, codomain, bag):
        self._solver = solver
        batch_size = solver.batch_size
        max_shape = bag.max_shape
        self._a = np.zeros((batch_size, max_shape[0]), dtype=np.uint32)
        self._b = np.zeros((batch_size, max_shape[1]), dtype=np.uint32)
        self._len = np.zeros((2, batch_size), dtype=np.uint16)
        self._sim = _shared_table(bag)
        name, self._construct = _variant(IndexedMatrixForm, codomain)
        self._solve = getattr(solver, name)

    def solve(self, batch):
        shape = batch.shape
        r = self._solve(self._a[:, :shape[0]], self._b[:, :shape[1]], self._sim, self._len)
        return [self._construct(p, self._solver, x) for p, x in zip(batch.problems, r)]


class BinaryMatrixForm(AbstractIndexedMatrixForm):
    """Raw symbol values compared for equality (solve.py:435-469)."""
    _solvers = solver_variants("solve_binary")

    def __init__(self, solver, codomain, bag):
        self._solver = solver
        batch_size = solver.batch_size
        max_shape = bag.max_shape
        self._a = np.zeros((batch_size, max_shape[0]), dtype=np.uint32)
        self._b = np.zeros((batch_size, max_shape[1]), dtype=np.uint32)
        self._len = np.zeros((2, batch_size), dtype=np.uint16)
        self._sim = _shared_eq_ne(bag)
        name, self._construct = _variant(BinaryMatrixForm, codomain)
        self._solve = getattr(solver, name)

    def solve(self, batch):
        shape = batch.shape
        r = self._solve(self._a[:, :shape[0]], self._b[:, :shape[1]], *self._sim, self._len)
        return [self._construct(p, self._solver, x) for p, x in zip(batch.problems, r)]


def _all_are(values, obj):
    """Whether every value is obj (identity; the loop runs in C)."""
    return all(map(is_, values, itertools.repeat(obj)))


def _all_have(problems, name, obj, classes=()):
    """Whether every problem's attribute name is obj; a missing attribute is
    False.  classes: the problems' classes; where all of them inherit the
    alphabet problems' encoder property, its stored value is read directly."""
    if name == "encoder" and classes and all(getattr(c, "encoder", None) is AlphabetProblem.encoder
                                             for c in classes):
        name = "_encoder"
    try:
        return _all_are(map(attrgetter(name), problems), obj)
    except AttributeError:
        return False


def _shared_table(bag):
    sim = bag.problems[0].similarity_lookup_table()
    # (alphabet problems return their stored table: read it without a call each)
    if all(getattr(c, "similarity_lookup_table", None) is AlphabetProblem.similarity_lookup_table
           for c in bag.classes):
        same = _all_are(map(attrgetter("_matrix"), bag.problems), sim)
    else:
        same = _all_are(map(methodcaller("similarity_lookup_table"), bag.problems), sim)
    if not same:
        raise ValueError("similarity table must be identical for all problems in a batch")
    return sim


def _shared_eq_ne(bag):
    sim = bag.problems[0].binary_similarity_values()
    if not _all_are(map(methodcaller("binary_similarity_values"), bag.problems), sim):
        raise ValueError("similarity must be identical for all problems in a batch")
    return sim


def _offsets(lens):
    off = np.zeros(len(lens), dtype=np.int64)
    if len(lens) > 1:
        np.cumsum(lens[:-1], out=off[1:])
    return off


def _encode_bag(problems, which):
    """Flat codes, offsets and lengths of every problem's s (which=0) or t."""
    enc = getattr(problems[0], "encoder", None)
    if enc is not None and hasattr(enc, "encode_many") and _all_have(problems, "encoder", enc):
        return enc.encode_many(list(map(attrgetter("s" if which == 0 else "t"), problems)))
    lens = np.array([p.shape[which] for p in problems], dtype=np.int64)
    off = _offsets(lens)
    flat = np.zeros(int(lens.sum()), dtype=np.uint32)
    for p, o, n in zip(problems, off, lens):
        a = np.zeros(p.shape[0], dtype=np.uint32)
        b = np.zeros(p.shape[1], dtype=np.uint32)
        p.build_index_sequences(a, b)
        flat[o:o + n] = a if which == 0 else b
    return flat, off, lens


# Python's cyclic collector paused while result objects are built (shared
# with the library layer, which builds the core results)
_collector_paused = algorithm.collector_paused


def _joined_text(problems, classes=()):
    """The bag's s and t texts joined (latin-1 bytes when every symbol fits,
    else UTF-32-LE code points) for encoding on the device, or None when a
    sequence is not a str.  classes: the problems' classes; where none of
    them overrides the s and t properties the stored sequences are read
    directly (no property call per problem)."""
    plain = bool(classes) and all(getattr(c, "s", None) is Problem.s and getattr(c, "t", None) is Problem.t
                                  for c in classes)
    get_s, get_t = (attrgetter("_s"), attrgetter("_t")) if plain else (attrgetter("s"), attrgetter("t"))
    try:  # join accepts only str items: lists or arrays of symbols go to the host encoders
        js = "".join(map(get_s, problems))
        jt = "".join(map(get_t, problems))
    except TypeError:
        return None
    try:
        return js, jt, js.encode("latin-1"), jt.encode("latin-1"), 1
    except UnicodeEncodeError:
        return js, jt, js.encode("utf-32-le"), jt.encode("utf-32-le"), 4


def _text_payload(problems, ls, lt, indexed, classes=()):
    """Raw-symbol payload of a bag of str problems (the device encodes it):
    indexed problems need their shared encoder's symbol table (char_lut)."""
    lut = None
    enc = getattr(problems[0], "encoder", None)
    if indexed:
        if not hasattr(enc, "char_lut") or not _all_have(problems, "encoder", enc, classes):
            return None, None
    text = _joined_text(problems, classes)
    if text is None:
        return None, None
    js, jt, rs, rt, width = text
    if indexed:
        lut = enc.char_lut(width)
        if lut is None:
            return None, None
    payload = dict(codes_s=np.frombuffer(rs, dtype=np.uint8), codes_t=np.frombuffer(rt, dtype=np.uint8),
                   off_s=_offsets(ls.astype(np.int64)), off_t=_offsets(lt.astype(np.int64)), raw_width=width)
    if lut is not None:
        payload["lut"] = lut
    return payload, (js, jt, enc)


def _bad_symbol_error(solver, text, off_s, off_t):
    """The encoder's error for the first problem (s before t) holding a
    symbol outside the alphabet (problems/factory.py:103-153)."""
    js, jt, enc = text
    ps, pt = solver.last_bad_symbol()
    cands = []
    if ps >= 0:
        cands.append((int(np.searchsorted(off_s, ps, side="right")) - 1, 0, js[ps]))
    if pt >= 0:
        cands.append((int(np.searchsorted(off_t, pt, side="right")) - 1, 1, jt[pt]))
    if not cands or not isinstance(enc, AlphabetEncoder):
        return ValueError("item not in alphabet")
    return ValueError(f"'{min(cands)[2]}' is not in alphabet")


_KIND = {"for_score": "score", "for_alignment": "alignment", "for_alignment_iterator": "alignment_iterator",
         "for_solution": "solution", "for_solution_iterator": "solution_iterator"}


class Solver:
    """A solver that obtains solutions to alignment problems."""

    _matrix_c = {
        Form.MATRIX_FORM: MatrixForm,
        Form.INDEXED_MATRIX_FORM: IndexedMatrixForm,
        Form.BINARY_MATRIX_FORM: BinaryMatrixForm,
    }

    def __init__(self, gap_cost: GapCost = None, codomain=Solution, **kwargs):
        self._codomain = Codomain(Solution if codomain is None else codomain)
        if gap_cost is None:
            gap_cost = ConstantGapCost(0)
        self._options = dict(gap_cost=gap_cost, codomain=self._codomain, **kwargs)
        self._cache = SolverCache(self._options)
        self._timings = NoTimings()
        max_len_s = self._options.get("max_len_s")
        max_len_t = self._options.get("max_len_t")
        if max_len_s and max_len_t:
            self._cache.ensure(max_len_s, max_len_t)

    @property
    def gap_cost(self):
        return self._options["gap_cost"]

    def to_codomain(self, codomain, return_dup=False):
        kwargs = self._options.copy()
        kwargs["codomain"] = codomain
        kwargs["return_dup"] = return_dup
        return Solver(**kwargs)

    @property
    def codomain(self):
        return self._codomain.type

    @property
    def batch_size(self):
        return self._cache.get(1, 1, direction="maximize", batch=True).batch_size

    def timings(self):
        return Timings(self)

    def _solve_bag(self, bag):
        max_shape = bag.max_shape
        solver = self._cache.get(max_shape[0], max_shape[1], direction=bag.direction, batch=len(bag) > 1)
        matrix_c = Solver._matrix_c.get(bag.form)
        if matrix_c is None:
            raise ValueError(bag.form)
        if len(bag) > 1:
            return self._solve_bag_at_once(solver, matrix_c, bag)
        form_solver = matrix_c(solver, self._codomain, bag)
        result = [None] * len(bag.problems)
        for batch in bag.batches(solver.batch_size):
            with self._timings.measure("prepare"):
                form_solver.prepare(batch)
            with self._timings.measure("solve"):
                for i, r in zip(batch.indices, form_solver.solve(batch)):
                    result[i] = r
        return result

    def _solve_bag_at_once(self, solver, matrix_c, bag):
        name, construct = _variant(matrix_c, self._codomain)
        kind = _KIND[name[name.index("for_"):]]
        problems = bag.problems
        with self._timings.measure("prepare"):
            ls = bag.shapes[:, 0].astype(np.int32)
            lt = bag.shapes[:, 1].astype(np.int32)
            text = None
            if bag.form in (Form.INDEXED_MATRIX_FORM, Form.BINARY_MATRIX_FORM) and problems:
                payload, text = _text_payload(problems, ls, lt, bag.form == Form.INDEXED_MATRIX_FORM,
                                              bag.classes)
            if text is not None:
                # str sequences: the device encodes the joined text
                if bag.form == Form.INDEXED_MATRIX_FORM:
                    payload["table"] = np.asarray(_shared_table(bag), dtype=np.float32)
                    form = "indexed"
                else:
                    payload["eq"], payload["ne"] = _shared_eq_ne(bag)
                    form = "binary"
            elif bag.form == Form.INDEXED_MATRIX_FORM:
                table = _shared_table(bag)
                cs, os_, _ = _encode_bag(problems, 0)
                ct, ot, _ = _encode_bag(problems, 1)
                if cs.size and cs.max() >= table.shape[0]:
                    raise ValueError("out of bounds index in a")
                if ct.size and ct.max() >= table.shape[1]:
                    raise ValueError("out of bounds index in b")
                code_t = np.uint16 if max(table.shape) > 128 else np.uint8
                payload = dict(codes_s=cs.astype(code_t), codes_t=ct.astype(code_t), off_s=os_, off_t=ot,
                               table=np.array(table, dtype=np.float32))  # (a copy: the caller's table may change)
                form = "indexed"
            elif bag.form == Form.BINARY_MATRIX_FORM:
                eq, ne = _shared_eq_ne(bag)
                cs, os_, _ = _encode_bag(problems, 0)
                ct, ot, _ = _encode_bag(problems, 1)
                payload = dict(codes_s=cs, codes_t=ct, off_s=os_, off_t=ot, eq=eq, ne=ne)
                form = "binary"
            elif all(isinstance(p, SeriesProblem) for p in problems) and 1 <= problems[0].dim <= 4 and \
                    all(p.dim == problems[0].dim for p in problems):
                xs = np.concatenate([p.s for p in problems]) if problems else np.zeros((0, 1), np.float32)
                ys = np.concatenate([p.t for p in problems]) if problems else np.zeros((0, 1), np.float32)
                payload = dict(codes_s=xs.ravel(), codes_t=ys.ravel(), off_s=_offsets(ls.astype(np.int64)),
                               off_t=_offsets(lt.astype(np.int64)), dim=problems[0].dim)
                form = "series"
            else:
                sizes = ls.astype(np.int64) * lt.astype(np.int64)
                dense_off = _offsets(sizes)
                dense = np.zeros(int(sizes.sum()), dtype=bag.dtype)
                for p, o, a, b in zip(problems, dense_off, ls, lt):
                    p.build_matrix(dense[o:o + a * b].reshape(a, b))
                payload = dict(dense=dense, dense_off=dense_off)
                form = "dense"
        with self._timings.measure("solve"):
            try:
                # every payload array above was built for this call
                r = solver.solve_ragged(kind, form, ls, lt, payload_owned=True, **payload)
            except ValueError as e:
                if text is None or "not in alphabet" not in str(e):
                    raise
                raise _bad_symbol_error(solver, text, payload["off_s"], payload["off_t"]) from None
            if kind == "score":  # the construct is the identity
                return list(r)
            with _collector_paused():
                return [construct(p, solver, x) for p, x in zip(problems, r)]

    def _solve_problem(self, problem):
        return self._solve_bag(ProblemBag([problem]))[0]

    def solve(self, x):
        if isinstance(x, Sequence):
            return self._solve_bag(ProblemBag(list(x)))
        return self._solve_problem(x)


class LocalSolver(Solver):
    """Optimal local alignments (Smith-Waterman for linear gap costs)."""

    def __init__(self, gap_cost: GapCost = None, **kwargs):
        super().__init__(solver=algorithm.Type.ALIGNMENT, locality=algorithm.Locality.LOCAL,
                         gap_cost=gap_cost, **kwargs)


class GlobalSolver(Solver):
    """Optimal global alignments (Needleman-Wunsch for linear gap costs)."""

    def __init__(self, gap_cost: GapCost = None, **kwargs):
        super().__init__(solver=algorithm.Type.ALIGNMENT, locality=algorithm.Locality.GLOBAL,
                         gap_cost=gap_cost, **kwargs)


class SemiglobalSolver(Solver):
    """Optimal semiglobal alignments."""

    def __init__(self, gap_cost: GapCost = None, **kwargs):
        super().__init__(solver=algorithm.Type.ALIGNMENT, locality=algorithm.Locality.SEMIGLOBAL,
                         gap_cost=gap_cost, **kwargs)


class ElasticSolver(Solver):
    """Dynamic time warping (DTW) for elastic alignment problems."""

    def __init__(self, **kwargs):
        super().__init__(solver=algorithm.Type.DTW, **kwargs)
