"""Drop-in module ``calc_frc`` of the reference (same import name): re-exports
``refspi_amd.frc.FRC`` so ``from calc_frc import FRC`` resolves to this package."""
import os as _os
import sys as _sys

_sys.path.insert(0, _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))

from refspi_amd.frc import FRC  # noqa: E402,F401
