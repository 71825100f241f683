"""The benchmark scenes on the oracle: each configuration draws what it claims to draw."""
import numpy as np

from salviarenderer_amd import scenes as S
from test_gpu_parity import assert_multipass_not_vacuous


def test_multipass_every_pass_leaves_a_mark(cpu_lib):
    """The translucent triangles are drawn (blended samples exist), the mirror marks stencil and
    the shadow map holds depth: a regression guard for passes that silently draw nothing."""
    r = cpu_lib.renderer()
    f = S.MultipassFrame(r, width=320, height=180, samples=4, shadow_size=256, n_tris=3000)
    f.render()
    r.flush()
    assert_multipass_not_vacuous({"color": f.color.read(), "ds": f.ds.read(), "shadow": f.shadow.read()})


def test_multipass_translucent_binding_selects_the_translucent_triangles(cpu_lib):
    """A non-indexed draw ignores its start (index_fetcher.cpp:58-62,103-105): the translucent
    pass must select its triangles by the vertex-buffer offset.  Drawing only that pass over an
    opaque black clear leaves alpha < 1 exactly where translucent triangles land."""
    from salviarenderer_amd import api as A

    r = cpu_lib.renderer()
    f = S.MultipassFrame(r, width=160, height=90, samples=1, shadow_size=64, n_tris=400)
    r.set_render_targets([f.color], f.ds)
    r.set_viewport(0, 0, f.width, f.height, 0.0, 1.0)
    r.clear_color(f.color, (0.0, 0.0, 0.0, 1.0))
    r.clear_depth_stencil(f.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
    f._bind_geometry(f.vb, f.I, f.L, first_tri=f.n_opaque)
    r.set_pixel_shader(A.PS_GOURAUD)
    r.set_blend_shader(A.BLEND_ON)
    r.set_depth_stencil_state(f.dss_translucent, 0)
    r.draw(0, f.n_trans)
    r.flush()
    alpha = f.color.read().reshape(-1, 4)[:, 3]
    touched = alpha < 255
    assert touched.any()
    # blending src alpha a over dst alpha 1 gives 1 + (a - 1) * a < 1 for a in [0.2, 0.8]
    assert alpha[touched].max() < 255 and alpha[touched].min() >= int(255 * (1 + (0.2 - 1) * 0.8)) - 1
    assert np.count_nonzero(touched) > 0.1 * alpha.size
