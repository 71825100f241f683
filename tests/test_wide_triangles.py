"""Long thin triangles over large targets, on the HIP library against the oracle.

Slivers that cross the whole target reach far more 64x64 tiles than they have records; some are
cut by the near or far plane.  Their records are counted and emitted by whole waves (more than 32
64x64 tiles in the bounding box, or a clipped triangle), which these cases exercise on the
order-independent path (opaque early-Z draws), the in-order path (blending, statistics queries)
and with 4 samples.  Depth/stencil must be bit-exact, colours within 1 LSB (and are reported).
"""
import numpy as np
import pytest

from salviarenderer_amd import api as A

from features import Draw

W, H = 1536, 1024


def slivers(n, seed, *, alpha=None):
    """n thin triangles: two vertices up to 4 clip units apart along a random direction, the third
    beside their midpoint; a third of them have vertex depths in [-0.5, 1.5] (near/far clipping)."""
    rng = np.random.default_rng(seed)
    p0 = rng.uniform(-1.6, 1.6, (n, 2))
    ang = rng.uniform(0, 2 * np.pi, n)
    d = np.stack([np.cos(ang), np.sin(ang)], 1) * rng.uniform(0.5, 4.0, (n, 1))
    p1 = p0 + d
    nrm = np.stack([-d[:, 1], d[:, 0]], 1) / np.linalg.norm(d, axis=1, keepdims=True)
    p2 = 0.5 * (p0 + p1) + nrm * np.exp(rng.uniform(np.log(0.002), np.log(0.05), (n, 1)))
    xy = np.stack([p0, p1, p2], 1)
    clipped = rng.uniform(0, 1, n) < 1 / 3
    z = np.where(clipped[:, None], rng.uniform(-0.5, 1.5, (n, 3)), rng.uniform(0.1, 0.9, (n, 3)))
    pos = np.concatenate([xy, z[..., None], np.ones((n, 3, 1))], -1).reshape(-1, 4)
    col = rng.uniform(0, 1, (n * 3, 4))
    if alpha is not None:
        col[:, 3] = np.repeat(rng.uniform(alpha[0], alpha[1], n), 3)
    return pos.astype(np.float32), col.astype(np.float32)


def render(lib, draws, samples, stats):
    r = lib.renderer()
    ct = r.create_tex2d(W, H, samples, A.PF_RGBA8)
    dt = r.create_tex2d(W, H, samples, A.PF_RG32F)
    color, ds = ct.surface(0), dt.surface(0)
    r.set_render_targets([color], ds)
    r.set_viewport(0, 0, W, H, 0.0, 1.0)
    r.clear_color(color, (0.05, 0.1, 0.15, 0.2))
    r.clear_depth_stencil(ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
    if stats:
        q = r.create_query(A.QUERY_PIPELINE_STATISTICS)
        r.begin(q)
    keep, tex_cache = [], {}
    for d in draws:
        keep += d.issue(r, tex_cache)
    out = {}
    if stats:
        r.end(q)
        st = r.get_data(q)[1]
        out.update(cprimitives=st.cprimitives, ps_invocations=st.ps_invocations)
    r.flush()
    out["color"] = color.read().copy()
    out["ds"] = ds.read().copy()
    return out


CASES = {
    "opaque": lambda: [Draw(*slivers(400, 901))],
    "blend_over_opaque": lambda: [Draw(*slivers(300, 902)),
                                  Draw(*slivers(300, 903, alpha=(0.2, 0.8)), depth_write=False, blend=A.BLEND_ON)],
}


@pytest.mark.gpu
@pytest.mark.parametrize("samples", [1, 4])
@pytest.mark.parametrize("stats", [False, True], ids=["fast", "stats"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_slivers_match_oracle(hip_lib, cpu_lib, name, samples, stats):
    h = render(hip_lib, CASES[name](), samples, stats)
    c = render(cpu_lib, CASES[name](), samples, stats)
    covered = float((c["ds"].view(np.float32).reshape(-1, 2)[:, 0] < 1.0).mean())
    print(f"\n  {name} {samples}x: covered {covered:.3f}, depth/stencil bytes differing "
          f"{int((h['ds'] != c['ds']).sum())}", flush=True)
    assert np.array_equal(h["ds"], c["ds"])
    diff = np.abs(h["color"].astype(np.int32) - c["color"].astype(np.int32))
    print(f"  colour bytes differing: {int((diff > 0).sum())} of {diff.size}, max {int(diff.max())}", flush=True)
    assert diff.max() <= 1
    if stats:
        assert h["cprimitives"] == c["cprimitives"] and h["ps_invocations"] == c["ps_invocations"]


def test_slivers_reach_many_tiles():
    """The scenes hold what the cases are for: slivers over more than 32 64x64 tiles and triangles
    crossing the near or far plane."""
    pos, _ = slivers(400, 901)
    t = pos.reshape(-1, 3, 4)
    x = np.clip((t[..., 0] + 1) * 0.5 * W, 0, W)
    y = np.clip((1 - (t[..., 1] + 1) * 0.5) * H, 0, H)
    span = (np.floor(x.max(1) / 64) - np.floor(x.min(1) / 64) + 1) * (np.floor(y.max(1) / 64) - np.floor(y.min(1) / 64) + 1)
    assert (span > 32).sum() > 100
    assert ((t[..., 2] < 0) | (t[..., 2] > 1)).any(1).sum() > 50


def _large_target_frame(lib):
    """60 plain triangles (most over more than 32 64x64 tiles) into a 4160x4128 target:
    67,080 16x16 GPU tiles, more than the multi-workgroup tile-range kernel takes (65,536), so the
    counting sort's ranges and scheduling order come from its one-workgroup form."""
    from salviarenderer_amd import scenes as S
    r = lib.renderer()
    f = S.setup_plain(r, 4160, 4128, 60, 1)
    f.render()
    r.flush()
    return f.color.read().copy(), f.ds.read().copy()


@pytest.mark.gpu
def test_target_beyond_65536_tiles(hip_lib, cpu_lib):
    hc, hd = _large_target_frame(hip_lib)
    cc, cd = _large_target_frame(cpu_lib)
    assert np.array_equal(hd, cd)
    assert np.array_equal(hc, cc), f"{np.count_nonzero(hc != cc)} colour bytes differ"
    assert np.count_nonzero(cd[..., 0] < 1.0) > 1_000_000  # the triangles cover a large part
