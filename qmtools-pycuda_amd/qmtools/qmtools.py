"""Drop-in module `qmtools.qmtools` (reference `qmtools/qmtools.py`)."""
from .basis import MAXAOC, MAXAOS, BasisSet  # noqa: F401
from .grid import Grid, make_float3  # noqa: F401
from .molecule import ANG2BOR, Molecule  # noqa: F401

MOLFAT = 2.0
