"""fwi3d_amd -- MI355X-native wave-propagation path of FWI3D.

Native parts (built by ``python -m fwi3d_amd.build``):
  * libfwi3d_hip.so   hand-written gfx950 HIP kernels behind the C ABI of
                      include/fwi3d.h;
  * libfwi3d_host.so  the C++ Modeling / Migration / Inversion classes and
                      the modeling_main / migration_main / inversion_main
                      drivers (bin/), reading the reference's parameters.txt
                      and file formats.

``fwi3d_amd.lib`` is a thin ctypes binding of the C ABI (used by tests and
bench.py); ``fwi3d_amd.configs`` writes the synthetic workloads.
"""
from . import lib  # noqa: F401
from .lib import Context, FwiError  # noqa: F401
