"""CPU oracle for the FWI3D wave-propagation path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package.  The product (fwi3d_amd) never does.
"""
from .oracle import *  # noqa: F401,F403
