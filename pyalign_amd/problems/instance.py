"""Problem forms and bags (pyalign/problems/instance.py:1-275)."""

import enum
from itertools import chain
from operator import attrgetter

import numpy as np


class Form(enum.Enum):
    MATRIX_FORM = 0
    INDEXED_MATRIX_FORM = 1
    BINARY_MATRIX_FORM = 2


def interleave2(x, y):
    """Morton code of (x, y): bits of x in the even positions, of y in the odd."""
    code = 0
    for b in range(32):
        code |= ((int(x) >> b) & 1) << (2 * b)
        code |= ((int(y) >> b) & 1) << (2 * b + 1)
    return code


class Problem:
    """An alignment problem between sequences s and t of shape (|s|, |t|)."""

    def __init__(self, shape, s=None, t=None, direction="maximize", dtype=np.float32):
        self._shape = tuple(shape)
        self._direction = direction
        self._dtype = dtype
        if s is not None and shape[0] != len(s):
            raise ValueError(f"sequence s [{len(s)}] does not match matrix shape {shape}")
        if t is not None and shape[1] != len(t):
            raise ValueError(f"sequence t [{len(t)}] does not match matrix shape {shape}")
        self._s = s
        self._t = t

    @property
    def shape(self):
        return self._shape

    @property
    def direction(self):
        return self._direction

    @property
    def dtype(self):
        return self._dtype

    @property
    def s(self):
        return self._s

    @property
    def t(self):
        return self._t

    def build_matrix(self, out):
        raise NotImplementedError()

    @property
    def matrix(self):
        m = np.empty(self.shape, dtype=self._dtype)
        self.build_matrix(m)
        return m

    @property
    def form(self):
        raise NotImplementedError()

    def __str__(self):
        return str(self.matrix)


class MatrixProblem(Problem):
    """A problem posed as a dense matrix M[i, j] of affinities/distances."""

    @property
    def form(self):
        return Form.MATRIX_FORM


class IndexedMatrixProblem(Problem):
    """A problem posed as a lookup table M and index vectors A, B."""

    def build_matrix(self, out):
        sim = self.similarity_lookup_table()
        a = np.empty((self.shape[0],), dtype=np.uint32)
        b = np.empty((self.shape[1],), dtype=np.uint32)
        self.build_index_sequences(a, b)
        out[:, :] = sim[np.ix_(a, b)]

    def similarity_lookup_table(self):
        raise NotImplementedError()

    def build_index_sequences(self, a, b):
        raise NotImplementedError()

    @property
    def form(self):
        return Form.INDEXED_MATRIX_FORM


class BinaryMatrixProblem(Problem):
    def binary_similarity_values(self):
        raise NotImplementedError()

    @property
    def form(self):
        return Form.BINARY_MATRIX_FORM


class ProblemBatch:
    def __init__(self, problems, indices, master_shape):
        shapes = np.array([p.shape for p in problems])
        self._max_shape = (np.max(shapes[:, 0]), np.max(shapes[:, 1]))
        self._shape = master_shape
        self._problems = problems
        self._indices = indices

    @property
    def shape(self):
        return self._shape

    @property
    def problems(self):
        return self._problems

    @property
    def indices(self):
        return self._indices


class ProblemBag:
    """Problems solved together: one direction and dtype, the lowest form."""

    def __init__(self, problems):
        self._problems = problems
        classes = set(map(type, problems))
        self._classes = classes
        # where no class overrides the properties, the stored values are read
        # directly (no property call per problem)
        plain = all(getattr(c, name, None) is getattr(Problem, name) for c in classes
                    for name in ("direction", "dtype", "shape"))
        get = (lambda name: attrgetter("_" + name)) if plain else attrgetter
        # list.count compares in C (identity first): every item equal to the first
        self._direction = problems[0].direction
        if list(map(get("direction"), problems)).count(self._direction) != len(problems):
            raise ValueError("problems in a bag need to have same direction")
        self._dtype = problems[0].dtype
        if list(map(get("dtype"), problems)).count(self._dtype) != len(problems):
            raise ValueError("problems in a bag need to have same dtype")
        # a problem's form is fixed by its class: one lookup per class
        reps = [problems[0]] if len(classes) == 1 else [next(p for p in problems if type(p) is c) for c in classes]
        self._form = Form(min(p.form.value for p in reps))
        shapes = np.fromiter(chain.from_iterable(map(get("shape"), problems)), dtype=np.int64,
                             count=2 * len(problems)).reshape(len(problems), 2)
        self._shapes = shapes
        self._max_shape = (np.max(shapes[:, 0]), np.max(shapes[:, 1]))

    def __len__(self):
        return len(self._problems)

    @property
    def problems(self):
        return self._problems

    @property
    def max_shape(self):
        return self._max_shape

    @property
    def shapes(self):
        """int64 [n, 2]: every problem's (|s|, |t|)."""
        return self._shapes

    @property
    def direction(self):
        return self._direction

    @property
    def dtype(self):
        return self._dtype

    @property
    def form(self):
        return self._form

    @property
    def classes(self):
        """The set of the problems' classes."""
        return self._classes

    def batches(self, batch_size):
        """Greedy batches: the largest remaining Morton code is the master
        shape, filled with problems that fit inside it (instance.py:230-275)."""
        problems = self._problems
        if batch_size == 1:
            for i, p in enumerate(problems):
                yield ProblemBatch([p], [i], p.shape)
            return
        code = [interleave2(*p.shape) for p in problems]
        order = sorted(range(len(problems)), key=lambda i: code[i])
        used = [False] * len(problems)
        while True:
            while order and used[order[-1]]:
                order.pop()
            if not order:
                break
            m = order.pop()
            master = problems[m].shape
            members = [m]
            for i in reversed(order):
                if used[i]:
                    continue
                s = problems[i].shape
                if s[0] <= master[0] and s[1] <= master[1]:
                    members.append(i)
                    used[i] = True
                    if len(members) == batch_size:
                        break
            yield ProblemBatch([problems[i] for i in members], members, master)
