#!/usr/bin/env python
"""Drop-in ``recon.py`` of the reference: ``python recon.py NUM_ITER -c config.ini``."""
import os as _os
import sys as _sys

_sys.path.insert(0, _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))

from refspi_amd.recon import main  # noqa: E402

if __name__ == '__main__':
    main()
