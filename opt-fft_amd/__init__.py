"""MI355X-native batched small-volume 3D FFT (gfx950 HIP kernels behind a C ABI).

The package directory name is not a Python identifier; load it by path with
``load_package()`` from ``__graft_entry__`` or ``tests/conftest.py``.
"""
from . import hip  # noqa: F401
from . import fft3d  # noqa: F401
from . import serving  # noqa: F401
