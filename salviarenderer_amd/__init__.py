"""MI355X-native implementation of the Salvia rasterization pipeline.

`hip_lib()` loads the product library (libsrx_hip.so, HIP kernels for gfx950) and raises if it
is not built; there is no CPU fallback in the product path.
"""
import os

from . import api

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(PKG_DIR, "lib", "libsrx_hip.so")


def hip_lib() -> "api.Lib":
    if not os.path.exists(HIP_LIB_PATH):
        raise RuntimeError(f"HIP backend not built: {HIP_LIB_PATH} is missing (run __graft_entry__.build())")
    lib = api.Lib(HIP_LIB_PATH)
    if lib.backend != "hip":
        raise RuntimeError(f"{HIP_LIB_PATH} is not the HIP backend")
    return lib
