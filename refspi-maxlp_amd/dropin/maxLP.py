"""Drop-in module ``maxLP`` of the reference (same import name): re-exports
``refspi_amd.maxlp.MaxLPhaser`` so ``from maxLP import MaxLPhaser`` resolves to the MI355X path."""
import os as _os
import sys as _sys

_sys.path.insert(0, _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))

from refspi_amd.maxlp import *  # noqa: F401,F403,E402
from refspi_amd.maxlp import MaxLPhaser  # noqa: E402,F401
