"""Unpickling restricted to plain data and NumPy arrays.

`Grid.Save` and the reference's molecule cache write pickles (reference
`qmtools/qmtools.py:290-314`, `scripts/train-singleshot.py:45-71`).  Reading them
with `pickle.load` would run whatever callable a file names; this reader
resolves only the globals NumPy needs to rebuild arrays, dtypes and scalars
and refuses every other one with `pickle.UnpicklingError`.
"""
from __future__ import annotations

import io
import pickle

_NUMPY_GLOBALS = {
    ("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"), ("numpy._core.multiarray", "scalar"),
    ("numpy.core.numeric", "_frombuffer"), ("numpy._core.numeric", "_frombuffer"),
    ("numpy", "ndarray"), ("numpy", "dtype"),
}


class ArrayOnlyUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _NUMPY_GLOBALS:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(f"refused pickle global {module}.{name}: only NumPy arrays are accepted")


def load(path):
    """The object pickled in `path`, built from plain data and NumPy arrays only."""
    with open(path, "rb") as fh:
        return ArrayOnlyUnpickler(io.BytesIO(fh.read())).load()
