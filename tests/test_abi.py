"""The C-ABI libraries load and export every entry point include/srx.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "srx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\*\s]+?)\b(srx_\w+)\s*\(", text, re.M)))


def test_header_declares_the_renderer_surface():
    names = declared_functions()
    for n in ("srx_create_renderer", "srx_draw", "srx_draw_index", "srx_clear_color", "srx_clear_depth_stencil",
              "srx_set_render_targets", "srx_set_viewport", "srx_create_sampler", "srx_surface_resolve", "srx_flush"):
        assert n in names
    assert len(names) >= 50


@pytest.mark.parametrize("which", ["hip", "cpu"])
def test_library_exports_every_symbol(which, cpu_lib):
    if which == "hip":
        from salviarenderer_amd import HIP_LIB_PATH as path
        if not os.path.exists(path):
            pytest.fail("libsrx_hip.so is not built")
    else:
        path = cpu_lib.path
    dll = ctypes.CDLL(path, mode=os.RTLD_LOCAL)
    missing = [n for n in declared_functions() if not hasattr(dll, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from salviarenderer_amd import api

    assert set(declared_functions()) == set(api.PROTOTYPES)
