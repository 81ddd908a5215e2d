"""argweaver_amd -- ARGweaver's threading HMM on AMD MI355X (gfx950).

The native library ``libargweaver_amd.so`` (built in this directory by
``make -C argweaver_amd``) exports the reference's ``arghmm_*`` C ABI for
the threading path; :mod:`argweaver_amd.argweaverc` binds it through ctypes
the way the reference's ``argweaver/argweaverc.py`` binds ``libargweaver.so``.

There is no CPU fallback: every call that runs the HMM executes HIP
kernels, and loading fails loudly if the library has not been built.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libargweaver_amd.so")

_lib = None


def load_library():
    """Load libargweaver_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libargweaver_amd.so is missing; build it with "
                "`make -C %s` (hipcc, gfx950)" % _HERE)
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    return _lib


def device_available():
    """True when a HIP device is visible to the library."""
    return bool(load_library().awamd_device_available())


from . import argweaverc  # noqa: E402,F401
