"""Problems and problem factories (the pyalign.problems surface)."""

from .instance import (Form, Problem as _BaseProblem, MatrixProblem, IndexedMatrixProblem,
                       BinaryMatrixProblem, ProblemBatch, ProblemBag)
from .function import Function, Coalesced, Dict, Equality
from .factory import (Problem, ProblemFactory, general, AlphabetProblem, ByteAlphabetEncoder,
                      AlphabetEncoder, AlphabetProblemFactory, alphabetic, IntegerEncoder,
                      BinaryProblem, BinaryProblemFactory, binary, SeriesProblem, SeriesProblemFactory,
                      series)

__all__ = [
    "Form", "MatrixProblem", "IndexedMatrixProblem", "BinaryMatrixProblem", "ProblemBatch", "ProblemBag",
    "Function", "Coalesced", "Dict", "Equality", "Problem", "ProblemFactory", "general", "AlphabetProblem",
    "ByteAlphabetEncoder", "AlphabetEncoder", "AlphabetProblemFactory", "alphabetic", "IntegerEncoder",
    "BinaryProblem", "BinaryProblemFactory", "binary", "SeriesProblem", "SeriesProblemFactory", "series",
]
