"""Problem factories and symbol encoders (pyalign/problems/factory.py:1-257).

Besides the per-problem encoders of the reference, every encoder offers
``encode_many(seqs)`` that encodes a whole bag into one flat code array
with offsets, which feeds the solver's one-call-per-bag path.
"""

import codecs
from typing import Sequence, Union

import numpy as np

from .function import Function
from .instance import BinaryMatrixProblem, IndexedMatrixProblem, MatrixProblem


class Problem(MatrixProblem):
    """A problem induced by a binary function f(x, y) (general form)."""

    def __init__(self, f, s, t, **kwargs):
        super().__init__((len(s), len(t)), s, t, **kwargs)
        self._f = f

    def build_matrix(self, out):
        f = self._f
        for i, x in enumerate(self.s):
            for j, y in enumerate(self.t):
                out[i, j] = f(x, y)


class ProblemFactory:
    """Problems built from a function w(x, y) of affinity (or distance)."""

    def __init__(self, w: callable, direction="maximize", dtype=np.float32):
        if w is None and direction == "minimize":
            from scipy.spatial.distance import euclidean
            w = euclidean
        self._w = w
        self._direction = direction
        self._dtype = dtype

    def new_problem(self, s, t):
        return Problem(self._w, s, t, direction=self._direction, dtype=self._dtype)


general = ProblemFactory


class AlphabetProblem(IndexedMatrixProblem):
    def __init__(self, matrix, encoder, *args, **kwargs):
        self._matrix = matrix
        self._encoder = encoder
        super().__init__(*args, **kwargs)

    def similarity_lookup_table(self):
        return self._matrix

    def build_index_sequences(self, a, b):
        self._encoder.encode(self._s, out=a)
        self._encoder.encode(self._t, out=b)

    @property
    def encoder(self):
        return self._encoder


def _try_encoding(s, encoding):
    if not isinstance(s, str):
        return False
    try:
        s.encode(encoding)
    except UnicodeEncodeError:
        return False
    return True


def _unique_in_order(alphabet):
    seen = {}
    for x in alphabet:
        seen.setdefault(x, len(seen))
    return tuple(seen)


class ByteAlphabetEncoder:
    """Codes for a small single-byte alphabet via a 256-entry lookup.

    The symbol order is an internal detail (results depend only on the
    table entry for each symbol pair).
    """

    def __init__(self, alphabet, encoding):
        self._encoding = encoding
        self._encoder = codecs.getencoder(encoding)
        self._alphabet = _unique_in_order(alphabet)
        assert len(self._alphabet) < 0x10 - 1
        index = np.full(256, 0xFF, dtype=np.uint8)
        for i, k in enumerate(self._alphabet):
            index[k.encode(encoding)[0]] = i
        self._index = index

    def encode(self, s, out=None):
        if not isinstance(s, str):
            s = "".join(s)
        r = self._index[np.frombuffer(self._encoder(s)[0], dtype=np.uint8)]
        if r.size and np.max(r) == 0xFF:
            raise ValueError("item not in alphabet")
        if out is None:
            return r
        out[:len(r)] = r

    def char_lut(self, width):
        """Code of every 1-byte symbol (255: not in the alphabet), for
        encoding joined latin-1 text on the device; None for 4-byte text,
        which this encoder cannot encode."""
        return self._index if width == 1 and self._encoding == "latin1" else None

    def encode_many(self, seqs):
        seqs = [s if isinstance(s, str) else "".join(s) for s in seqs]
        lens = np.fromiter((len(s) for s in seqs), dtype=np.int64, count=len(seqs))
        r = self._index[np.frombuffer(self._encoder("".join(seqs))[0], dtype=np.uint8)]
        if r.size and np.max(r) == 0xFF:
            raise ValueError("item not in alphabet")
        off = np.zeros(len(seqs), dtype=np.int64)
        if len(seqs) > 1:
            np.cumsum(lens[:-1], out=off[1:])
        return r, off, lens

    @property
    def alphabet(self):
        return self._alphabet


class AlphabetEncoder:
    """Codes for an arbitrary alphabet of hashable symbols."""

    def __init__(self, alphabet):
        self._alphabet = _unique_in_order(alphabet)
        self._ids = {k: i for i, k in enumerate(self._alphabet)}

    def encode(self, s, out=None):
        ids = self._ids
        try:
            if out is None:
                return [ids[x] for x in s]
            if out.shape[0] != len(s) or len(out.shape) != 1:
                raise ValueError(f"expected shape ({len(s)},), got {out.shape}")
            for i, x in enumerate(s):
                out[i] = ids[x]
        except KeyError as e:
            raise ValueError(f"'{e.args[0]}' is not in alphabet")

    def char_lut(self, width):
        """Code of every symbol value of `width`-byte text (latin-1 bytes or
        code points; 255: not in the alphabet) when every alphabet symbol
        is one character; else None."""
        if not all(isinstance(x, str) and len(x) == 1 for x in self._alphabet) or len(self._alphabet) > 255:
            return None
        points = [ord(x) for x in self._alphabet]
        if width == 1 and max(points) > 0xFF:
            points = [p for p in points if p <= 0xFF]
        size = 256 if width == 1 else max(points) + 1
        if size > 1 << 21:
            return None
        lut = np.full(size, 0xFF, dtype=np.uint8)
        for x in self._alphabet:
            if ord(x) < size:
                lut[ord(x)] = self._ids[x]
        return lut

    def encode_many(self, seqs):
        ids = self._ids
        lens = np.fromiter((len(s) for s in seqs), dtype=np.int64, count=len(seqs))
        try:
            flat = np.fromiter((ids[x] for s in seqs for x in s), dtype=np.int64, count=int(lens.sum()))
        except KeyError as e:
            raise ValueError(f"'{e.args[0]}' is not in alphabet")
        off = np.zeros(len(seqs), dtype=np.int64)
        if len(seqs) > 1:
            np.cumsum(lens[:-1], out=off[1:])
        return flat, off, lens

    @property
    def alphabet(self):
        return self._alphabet


class AlphabetProblemFactory:
    """Problems over a small fixed alphabet: a symbol-pair lookup table plus
    index sequences (the indexed form)."""

    def __init__(self, alphabet: Union[str, Sequence], w: callable, direction="maximize", dtype=np.float32):
        if _try_encoding(alphabet, "latin1"):
            self._encoder = ByteAlphabetEncoder(alphabet, "latin1")
        else:
            self._encoder = AlphabetEncoder(alphabet)
        symbols = self._encoder.alphabet
        n = len(symbols)
        self._matrix = np.zeros((n, n), dtype=np.float32)
        if isinstance(w, Function):
            w.build_matrix(self._encoder, self._matrix)
        else:
            for i, x in enumerate(symbols):
                for j, y in enumerate(symbols):
                    self._matrix[i, j] = w(x, y)
        self._direction = direction
        self._dtype = dtype

    def new_problem(self, s, t):
        return AlphabetProblem(self._matrix, self._encoder, (len(s), len(t)), s, t,
                               direction=self._direction, dtype=self._dtype)


alphabetic = AlphabetProblemFactory


class IntegerEncoder:
    """Raw symbol values: code points of a str, or an integer array."""

    _supported_dtype = {np.uint8, np.int8, np.uint16, np.int16, np.uint32, np.int32}

    def __init__(self):
        pass

    @staticmethod
    def _values(s):
        if isinstance(s, str):
            return np.frombuffer(s.encode("utf-32le"), dtype=np.uint32)
        if isinstance(s, np.ndarray):
            if s.dtype.type not in IntegerEncoder._supported_dtype:
                raise ValueError(f"array dtype {s.dtype} is not supported")
            return s
        raise ValueError(f"expected str or numpy.array, got {type(s)}")

    def encode(self, s, out=None):
        arr = self._values(s)
        if out is None:
            return arr
        if out.shape != arr.shape:
            raise ValueError(f"expected shape {arr.shape}, got {out.shape}")
        out[:] = arr

    def encode_many(self, seqs):
        if all(isinstance(s, str) for s in seqs):
            lens = np.fromiter((len(s) for s in seqs), dtype=np.int64, count=len(seqs))
            flat = np.frombuffer("".join(seqs).encode("utf-32le"), dtype=np.uint32)
        else:
            parts = [self._values(s) for s in seqs]
            lens = np.array([len(x) for x in parts], dtype=np.int64)
            flat = np.concatenate(parts).astype(np.uint32) if parts else np.zeros(0, np.uint32)
        off = np.zeros(len(seqs), dtype=np.int64)
        if len(seqs) > 1:
            np.cumsum(lens[:-1], out=off[1:])
        return flat, off, lens


class BinaryProblem(BinaryMatrixProblem):
    def __init__(self, sim, encoder, *args, **kwargs):
        self._sim = sim
        self._encoder = encoder
        super().__init__(*args, **kwargs)

    def binary_similarity_values(self):
        return self._sim  # eq, ne

    def build_index_sequences(self, a, b):
        self._encoder.encode(self._s, out=a)
        self._encoder.encode(self._t, out=b)

    @property
    def encoder(self):
        return self._encoder


class BinaryProblemFactory:
    """Problems scored eq for equal symbols and ne otherwise (binary form)."""

    def __init__(self, eq=1, ne=-1, direction="maximize", dtype=np.float32):
        self._encoder = IntegerEncoder()
        self._sim = (eq, ne)
        self._direction = direction
        self._dtype = dtype

    def new_problem(self, s, t):
        return BinaryProblem(self._sim, self._encoder, (len(s), len(t)), s, t,
                             direction=self._direction, dtype=self._dtype)


binary = BinaryProblemFactory


class SeriesProblem(MatrixProblem):
    """Elastic alignment of two vector series x[n, d], y[m, d] under the
    squared Euclidean cost sum_k (x[i,k] - y[j,k])^2, accumulated over k in
    order in float32 (no fused multiply-add)."""

    def __init__(self, x, y, **kwargs):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        if x.ndim == 1:
            x = x[:, None]
        if y.ndim == 1:
            y = y[:, None]
        if x.shape[1] != y.shape[1]:
            raise ValueError(f"series dimensions differ: {x.shape[1]} vs {y.shape[1]}")
        super().__init__((x.shape[0], y.shape[0]), x, y, **kwargs)

    @property
    def dim(self):
        return self._s.shape[1]

    def build_matrix(self, out):
        x, y = self._s, self._t
        c = None
        for k in range(x.shape[1]):
            d = x[:, k][:, None] - y[:, k][None, :]
            c = d * d if c is None else c + d * d
        out[:, :] = c


class SeriesProblemFactory:
    """Problems over vector series with squared Euclidean costs (for DTW,
    ``ElasticSolver``).  The cost is evaluated on the GPU from the series;
    no dense matrix is built for bags."""

    def __init__(self, direction="minimize", dtype=np.float32):
        self._direction = direction
        self._dtype = dtype

    def new_problem(self, x, y):
        return SeriesProblem(x, y, direction=self._direction, dtype=self._dtype)


series = SeriesProblemFactory
