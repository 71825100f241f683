"""HIP pipeline against the CPU oracle on the same seeded scenes, through the C ABI.

Coverage, depth, stencil and per-pixel primitive order must be bit-exact; UNORM8 colours may
differ by at most 1 LSB (they are compared exactly as well, and the count of differing bytes is
reported).
"""
import numpy as np
import pytest

from salviarenderer_amd import api as A
from salviarenderer_amd import scenes as S

pytestmark = pytest.mark.gpu


def render(lib, setup, **kw):
    """Renders the scene twice: without statistics queries (opaque early-Z draws then take the
    order-independent visibility kernel) and with them (in-order kernel, exact ps_invocations)."""
    r = lib.renderer()
    f = setup(r, **kw)
    f.render()
    r.flush()
    fast = {"color": f.color.read().copy(), "ds": f.ds.read().copy()}
    if f.resolved is not None:
        fast["resolved"] = f.resolved.read().copy()
    q = r.create_query(A.QUERY_PIPELINE_STATISTICS)
    qi = r.create_query(A.QUERY_INTERNAL_STATISTICS)
    r.begin(q)
    r.begin(qi)
    f.render()
    r.end(q)
    r.end(qi)
    r.flush()
    st = r.get_data(q)[1]
    sti = r.get_data(qi)[1]
    out = {"color": f.color.read().copy(), "ds": f.ds.read().copy(),
           "ps_invocations": st.ps_invocations, "cprimitives": st.cprimitives,
           "backend_input_pixels": sti.backend_input_pixels}
    if f.resolved is not None:
        out["resolved"] = f.resolved.read().copy()
    out["fast"] = fast
    return out


def compare(h, c, color_exact=True):
    _compare(h, c, color_exact)
    # the statistics-free path must give the same images
    f = dict(h["fast"])
    for k in ("ps_invocations", "cprimitives", "backend_input_pixels"):
        f[k] = c[k]
    _compare(f, c, color_exact)


def _compare(h, c, color_exact=True):
    print(f"\n  coverage-bearing depth/stencil bytes differing: {int((h['ds'] != c['ds']).sum())}", flush=True)
    assert np.array_equal(h["ds"], c["ds"]), "depth/stencil differs"
    diff = np.abs(h["color"].astype(np.int32) - c["color"].astype(np.int32))
    print(f"  colour bytes differing: {int((diff > 0).sum())} of {diff.size}, max {int(diff.max())}", flush=True)
    assert diff.max() <= 1
    if color_exact:
        assert np.array_equal(h["color"], c["color"])
    if "resolved" in c:
        rd = np.abs(h["resolved"].astype(np.int32) - c["resolved"].astype(np.int32))
        assert rd.max() <= 1
    for k in ("ps_invocations", "cprimitives", "backend_input_pixels"):
        assert h[k] == c[k], (k, h[k], c[k])


def test_plain_small_config(hip_lib, cpu_lib):
    """512x512, 2,000 Gouraud triangles, 1 sample, depth LESS."""
    kw = dict(width=512, height=512, n_tris=2000)
    compare(render(hip_lib, S.setup_plain, **kw), render(cpu_lib, S.setup_plain, **kw))


def test_plain_msaa4_cull_back(hip_lib, cpu_lib):
    kw = dict(width=256, height=192, n_tris=600, samples=4, cull=A.CULL_BACK, seed=7)
    compare(render(hip_lib, S.setup_plain, **kw), render(cpu_lib, S.setup_plain, **kw))


def test_plain_msaa2(hip_lib, cpu_lib):
    kw = dict(width=200, height=130, n_tris=300, samples=2, seed=11)
    compare(render(hip_lib, S.setup_plain, **kw), render(cpu_lib, S.setup_plain, **kw))


@pytest.mark.parametrize("camera", sorted(S.MESH_CAMERAS))
def test_textured_mesh_small(hip_lib, cpu_lib, camera):
    kw = dict(width=320, height=176, samples=4, grid=48, tex_size=256, camera=camera)
    compare(render(hip_lib, S.setup_mesh, **kw), render(cpu_lib, S.setup_mesh, **kw), color_exact=False)


@pytest.mark.parametrize("camera", sorted(S.MESH_CAMERAS))
def test_textured_mesh_full_config(hip_lib, cpu_lib, camera):
    """1920x1080 4xMSAA, 522,242 triangles, 2048^2 texture, 16x anisotropic."""
    compare(render(hip_lib, S.setup_mesh, camera=camera), render(cpu_lib, S.setup_mesh, camera=camera),
            color_exact=False)


def render_multipass(lib, stats, **kw):
    r = lib.renderer()
    f = S.MultipassFrame(r, **kw)
    if stats:
        q = r.create_query(A.QUERY_PIPELINE_STATISTICS)
        qi = r.create_query(A.QUERY_INTERNAL_STATISTICS)
        r.begin(q)
        r.begin(qi)
    f.render()
    out = {}
    if stats:
        r.end(q)
        r.end(qi)
        st, sti = r.get_data(q)[1], r.get_data(qi)[1]
        out.update(ps_invocations=st.ps_invocations, cprimitives=st.cprimitives,
                   backend_input_pixels=sti.backend_input_pixels)
    r.flush()
    out.update(color=f.color.read().copy(), ds=f.ds.read().copy(), shadow=f.shadow.read().copy(),
               resolved=f.resolved.read().copy())
    return out


def assert_multipass_not_vacuous(out):
    """Every pass of the frame must leave a mark: blended samples (alpha < 255: only the
    translucent pass writes such alpha over the opaque clear), stencil 1 under the mirror, and a
    shadow map that is not all cleared."""
    alpha = out["color"].reshape(-1, 4)[:, 3]
    frac = float((alpha < 255).mean())
    print(f"\n  blended colour samples: {frac:.3f}", flush=True)
    assert frac > 0.15, frac
    stencil = np.ascontiguousarray(out["ds"]).view(np.uint32).reshape(-1, 2)[:, 1]
    assert 0.02 < float((stencil == 1).mean()) < 0.5
    shadow_depth = np.ascontiguousarray(out["shadow"]).view(np.float32).reshape(-1, 2)[:, 0]
    assert float((shadow_depth < 1.0).mean()) > 0.3


@pytest.mark.parametrize("kw", [dict(width=320, height=180, samples=4, shadow_size=256, n_tris=3000),
                                dict(width=640, height=360, samples=2, shadow_size=512, n_tris=20000),
                                dict()],
                         ids=["small", "medium", "full_config"])
def test_multipass_shadow_blend_stencil_mirror(hip_lib, cpu_lib, kw):
    """Shadow map (depth only, split tiles) -> PCF-shaded opaque pass -> translucent blend pass ->
    stencil mirror mark -> mirrored redraw, bit-exact on both GPU paths."""
    want = render_multipass(cpu_lib, True, **kw)
    assert_multipass_not_vacuous(want)
    for stats in (False, True):
        got = render_multipass(hip_lib, stats, **kw)
        for k in ("shadow", "ds", "color", "resolved"):
            d = int((got[k] != want[k]).sum())
            assert d == 0, f"stats={stats}: {k} differs in {d} bytes"
        if stats:
            for k in ("ps_invocations", "cprimitives", "backend_input_pixels"):
                assert got[k] == want[k], (k, got[k], want[k])
