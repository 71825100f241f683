"""The reference's sample shaders (ColorizedTriangle, ComplexMesh, StandardShadowMap's ESM,
StencilMirror's cup, TextureAndBlending's box) on the HIP library against the oracle.

These shaders call sqrt, pow, exp and log: the oracle evaluates them with the host's libm, the
kernels with the device math library, so colours may differ by the last bit before conversion.
The bar: depth/stencil bit-exact; UNORM8 colour within 1 LSB; FP32 colour within 1e-5 relative.
Each comparison prints how many values differ and by how much."""
import numpy as np
import pytest

from salviarenderer_amd import api as A
from salviarenderer_amd import scenes as S

FP32_REL = 1e-5


def render(lib, name, stats=False, **kw):
    r = lib.renderer()
    f = S.SampleShaderFrame(r, name, **kw)
    q = None
    if stats:
        q = r.create_query(A.QUERY_PIPELINE_STATISTICS)
        r.begin(q)
    f.render()
    out = {}
    if q is not None:
        r.end(q)
        out["ps_invocations"] = r.get_data(q)[1].ps_invocations
    r.flush()
    out["color"] = f.color.read().copy()
    out["ds"] = f.ds.read().copy()
    if f.resolved is not None:
        out["resolved"] = f.resolved.read().copy()
    if name == "ssm":
        out["shadow"] = f.shadow.read().copy()
    return out


def test_sample_scenes_draw_something_on_the_oracle(cpu_lib):
    for name in S.SAMPLE_SHADERS:
        out = render(cpu_lib, name)
        c = out["color"].reshape(-1, 4)[:, :3].astype(np.float64)
        covered = (out["ds"].view(np.float32).reshape(-1, 2)[:, 0] < 1.0).mean()
        assert covered > 0.5, name
        assert c.std(0).min() > 10, (name, c.std(0))
    ssm = render(cpu_lib, "ssm")
    assert 0.05 < (ssm["shadow"].view(np.float32).reshape(-1, 2)[:, 0] < 1.0).mean() < 0.95


def compare_unorm8(name, got, want):
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    print(f"\n  {name}: UNORM8 bytes differing {int((d > 0).sum())} of {d.size}, max {int(d.max())}", flush=True)
    assert d.max() <= 1, name


def compare_fp32(name, got, want):
    g = got.view(np.float32).astype(np.float64)
    w = want.view(np.float32).astype(np.float64)
    rel = np.abs(g - w) / np.maximum(np.abs(w), 1e-30)
    rel[(g == w)] = 0.0
    print(f"\n  {name}: FP32 values differing {int((g != w).sum())} of {g.size}, max relative {rel.max():.2e}",
          flush=True)
    assert rel.max() <= FP32_REL, name


@pytest.mark.gpu
@pytest.mark.parametrize("name", S.SAMPLE_SHADERS)
@pytest.mark.parametrize("samples", [1, 4])
def test_sample_shader_unorm8(hip_lib, cpu_lib, name, samples):
    want = render(cpu_lib, name, stats=True, samples=samples)
    for stats in (False, True):
        got = render(hip_lib, name, stats=stats, samples=samples)
        assert np.array_equal(got["ds"], want["ds"]), (name, "depth/stencil")
        if name == "ssm":
            assert np.array_equal(got["shadow"], want["shadow"])
        compare_unorm8(name, got["color"], want["color"])
        if samples > 1:
            compare_unorm8(name + " resolved", got["resolved"], want["resolved"])
        if stats:
            assert got["ps_invocations"] == want["ps_invocations"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", S.SAMPLE_SHADERS)
def test_sample_shader_fp32(hip_lib, cpu_lib, name):
    want = render(cpu_lib, name, fmt=A.PF_RGBA32F)
    got = render(hip_lib, name, fmt=A.PF_RGBA32F)
    assert np.array_equal(got["ds"], want["ds"]), name
    compare_fp32(name, got["color"], want["color"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", S.SAMPLE_SHADERS)
def test_sample_shader_full_screen_1080p(hip_lib, cpu_lib, name):
    """The sample shaders at 1920x1080 over a 256x256-vertex grid (130k triangles)."""
    kw = dict(width=1920, height=1080, grid=256, tex_size=512)
    want = render(cpu_lib, name, **kw)
    got = render(hip_lib, name, **kw)
    assert np.array_equal(got["ds"], want["ds"]), name
    compare_unorm8(name, got["color"], want["color"])
