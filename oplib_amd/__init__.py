"""oplib_amd -- MI355X-native operator-method pricing (OPLib's hot path).

The product is two shared libraries built from ``oplib_amd/csrc``:

* ``libophip.so``  -- the C-ABI drop-in for OPLib's ``opcuda`` (HIP, gfx950);
* ``libopmodel.so`` -- the host model layer (C++), calling only that C-ABI.

``oplib_amd.ophip`` and ``oplib_amd.opmodel`` are thin ctypes bindings used
by tests and the benchmark.  Loading fails loudly when a library is missing:
there is no CPU fallback.
"""
from .ophip import lib as ophip_lib  # noqa: F401
