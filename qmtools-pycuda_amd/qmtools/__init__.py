"""qmtools: MI355X-native grid hot path (density, Hartree, V_ne) behind the
reference's `qmtools.qmtools` entry points."""
from .qmtools import *  # noqa: F401,F403
