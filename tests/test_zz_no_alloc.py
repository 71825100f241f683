"""srx_draw_noalloc (every scratch buffer in one fixed per-GPU region, no allocation) through the
C ABI: byte for byte what srx_draw produces, on two calls in a row, the second smaller; a draw
that does not fit is refused without touching the targets."""
import numpy as np
import pytest

from salviarenderer_amd import api as A
from salviarenderer_amd import scenes as S

pytestmark = pytest.mark.gpu


def frame(lib, noalloc, **kw):
    r = lib.renderer()
    f = S.setup_plain(r, **kw)
    f.draw = (lambda: r.draw_noalloc(0, f.n_tris)) if noalloc else f.draw
    f.render()
    r.flush()
    out = {"color": f.color.read().copy(), "ds": f.ds.read().copy()}
    if f.resolved is not None:
        out["resolved"] = f.resolved.read().copy()
    return r, f, out


def same(a, b):
    for k in b:
        d = int((a[k] != b[k]).sum())
        assert d == 0, f"{k} differs in {d} bytes"


def test_noalloc_equals_draw_twice_second_smaller(hip_lib):
    big = dict(width=512, height=512, n_tris=2000)
    small = dict(width=256, height=192, n_tris=600, samples=4, seed=7)
    want_big = frame(hip_lib, False, **big)[2]
    want_small = frame(hip_lib, False, **small)[2]
    # one renderer per scene, calls back to back on the same GPU region
    got_big = frame(hip_lib, True, **big)[2]
    got_small = frame(hip_lib, True, **small)[2]
    same(got_big, want_big)
    same(got_small, want_small)
    # and on one renderer: the region keeps the last call's contents, the next call must not care
    r = hip_lib.renderer()
    f1 = S.setup_plain(r, **big)
    r.clear_color(f1.color, (0.0, 0.0, 0.0, 0.0))
    r.clear_depth_stencil(f1.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
    r.draw_noalloc(0, f1.n_tris)
    r.flush()
    same({"color": f1.color.read(), "ds": f1.ds.read()}, want_big)


def test_noalloc_refuses_what_does_not_fit(hip_lib):
    r = hip_lib.renderer()
    f = S.setup_plain(r, width=1024, height=1024, n_tris=400_000)
    r.clear_color(f.color, (0.25, 0.5, 0.75, 1.0))
    r.clear_depth_stencil(f.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
    r.flush()
    before = (f.color.read().copy(), f.ds.read().copy())
    with pytest.raises(A.SrxError) as e:
        r.draw_noalloc(0, f.n_tris)
    assert e.value.code == A.OUT_OF_MEMORY
    r.flush()
    assert np.array_equal(f.color.read(), before[0]) and np.array_equal(f.ds.read(), before[1])
    # the renderer keeps working, with and without the region
    small = S.setup_plain(r, width=128, height=128, n_tris=64)
    small.draw = lambda: r.draw_noalloc(0, small.n_tris)
    small.render()
    r.flush()
    got = small.color.read().copy()
    small.draw = lambda: r.draw(0, small.n_tris)
    small.render()
    r.flush()
    assert np.array_equal(got, small.color.read())
