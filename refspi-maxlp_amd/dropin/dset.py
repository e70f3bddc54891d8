"""Drop-in module ``dset`` of the reference (same import name): re-exports
``refspi_amd.dset.Dataset`` so ``from dset import Dataset`` resolves to the MI355X path."""
import os as _os
import sys as _sys

_sys.path.insert(0, _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))

from refspi_amd.dset import *  # noqa: F401,F403,E402
from refspi_amd.dset import Dataset  # noqa: E402,F401
