"""Request server: continuous batching of render requests over resident meshes.  Each result must
equal the plain pipeline's render of the same request, alone and in a mix of concurrent requests,
through retries (bins outgrowing their estimate), LRU evictions and footprint admission."""
import numpy as np
import pytest

from salviarenderer_amd import api as A
from salviarenderer_amd import scenes as S

SPECS = [
    S.JobSpec(0x5A11C001, 64, 64, 300, 64, 1),
    S.JobSpec(0x5A11C002, 200, 130, 2500, 128, 4),
    S.JobSpec(0x5A11C003, 256, 256, 1000, 256, 1),
    S.JobSpec(0x5A11C004, 300, 170, 5000, 64, 4),
    S.JobSpec(0x5A11C005, 16, 16, 50, 64, 1),
    S.JobSpec(0x5A11C006, 640, 360, 20000, 128, 4),
]
CLEAR = (0.1, 0.2, 0.3, 1.0)


def plain(lib, spec, samples):
    """The renderer API's output for one request: colour (resolved) and depth/stencil bytes."""
    r = lib.renderer()
    sp = S.JobSpec(spec.seed, spec.width, spec.height, spec.n_tris, spec.tex_size, samples)
    jobs = S.BatchedJobs(r, [sp], with_textures=False)
    jobs.bind_common()
    jb = jobs.jobs[0]
    r.set_render_targets([jb["color"]], jb["ds"])
    r.set_viewport(0, 0, sp.width, sp.height, 0.0, 1.0)
    r.set_vertex_buffers(0, [jb["vb"]], [32])
    r.clear_color(jb["color"], CLEAR)
    r.clear_depth_stencil(jb["ds"], A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
    r.draw(0, sp.n_tris)
    if jb["resolved"] is not None:
        r.resolve(jb["color"], jb["resolved"])
    r.flush()
    return jobs.output(0).ravel().copy(), jb["ds"].read().ravel().copy()


def submit(srv, j, samples, mesh=True, arrival_ns=None, key=None):
    sp = SPECS[j]
    return srv.submit(key if key is not None else sp.seed, sp.n_tris, sp.width, sp.height, samples,
                      mesh=S.batch_triangles(sp) if mesh else None, texture_key=sp.seed,
                      texture=S.noise_texture(sp.tex_size, sp.seed), clear=CLEAR, want_depth=True,
                      arrival_ns=arrival_ns)


def drain(srv, n, timeout_ms=60000):
    got = {}
    waited = 0
    while len(got) < n and waited < timeout_ms:
        for c in srv.poll(64, wait_ms=100):
            got[c["ticket"]] = c
        waited += 100
    assert len(got) == n, (len(got), n)
    return got


def test_oracle_server_matches_plain(cpu_lib):
    srv = cpu_lib.renderer().create_server()
    tk = {submit(srv, j, s): (j, s) for j in (0, 1, 4) for s in (1, 4)}
    got = drain(srv, len(tk))
    for t, (j, s) in tk.items():
        c, d = plain(cpu_lib, SPECS[j], s)
        assert got[t]["status"] == 0
        assert np.array_equal(got[t]["color"], c) and np.array_equal(got[t]["depth"], d)


@pytest.mark.gpu
def test_served_alone_and_in_mix_equal_plain(hip_lib, cpu_lib):
    r = hip_lib.renderer()
    srv = r.create_server()
    ref, alone = {}, {}
    for j in range(len(SPECS)):
        for s in (1, 4):
            ref[(j, s)] = plain(hip_lib, SPECS[j], s)
            t = submit(srv, j, s)
            alone[(j, s)] = drain(srv, 1)[t]
            assert alone[(j, s)]["status"] == 0 and alone[(j, s)]["batch_jobs"] == 1
            assert np.array_equal(alone[(j, s)]["color"], ref[(j, s)][0]), (j, s)
            assert np.array_equal(alone[(j, s)]["depth"], ref[(j, s)][1]), (j, s)
    for j in (0, 1, 4):  # the plain pipeline here is the oracle's too
        c, d = plain(cpu_lib, SPECS[j], 4)
        assert np.array_equal(ref[(j, 4)][0], c) and np.array_equal(ref[(j, 4)][1], d)
    # the mix: 60 requests arriving over ~20 ms (after all are submitted), resident meshes, several
    # batches in flight
    rng = np.random.default_rng(3)
    t0 = srv.now() + 500_000_000
    arrivals = t0 + np.cumsum(rng.exponential(300_000, 60)).astype(np.int64)
    tk = {}
    for a in arrivals:
        j, s = int(rng.integers(len(SPECS))), int(rng.choice([1, 4]))
        tk[submit(srv, j, s, mesh=False, arrival_ns=int(a))] = (j, s)
    got = drain(srv, len(tk))
    sizes = []
    for t, (j, s) in tk.items():
        g = got[t]
        assert g["status"] == 0
        assert g["admitted_ns"] >= g["arrival_ns"] and g["done_ns"] >= g["admitted_ns"]
        assert np.array_equal(g["color"], alone[(j, s)]["color"]), (j, s)
        assert np.array_equal(g["depth"], alone[(j, s)]["depth"]), (j, s)
        sizes.append(g["batch_jobs"])
    st = srv.stats()
    print(f"\n  batches {st.batches} jobs {st.jobs} largest batch {st.max_batch_jobs} mean batch "
          f"{np.mean(sizes):.1f} retries {st.retries}", flush=True)
    assert max(sizes) > 1  # requests were batched together


@pytest.mark.gpu
def test_server_retries_and_evicts(hip_lib):
    r = hip_lib.renderer()
    d = A.default_server_desc(hip_lib)
    d.pool_bytes = 2 << 20  # one or two meshes at most
    srv = r.create_server(d)
    # 40 triangles spanning the target: ~500 records per triangle, far above the first estimate
    big = S.JobSpec(0x5A11C101, 640, 360, 40, 64, 1)
    mesh = S.plain_triangles(40, big.seed)
    t = srv.submit(7, 40, 640, 360, 1, mesh=mesh, clear=CLEAR, want_depth=True)
    g = drain(srv, 1)[t]
    assert g["status"] == 0 and srv.stats().retries >= 1
    rr = hip_lib.renderer()
    jobs = S.BatchedJobs(rr, [big], with_textures=False)
    jobs.jobs[0]["vb"].write(mesh)
    jobs.bind_common()
    jb = jobs.jobs[0]
    rr.set_render_targets([jb["color"]], jb["ds"])
    rr.set_viewport(0, 0, 640, 360, 0.0, 1.0)
    rr.set_vertex_buffers(0, [jb["vb"]], [32])
    rr.clear_color(jb["color"], CLEAR)
    rr.clear_depth_stencil(jb["ds"], A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
    rr.draw(0, 40)
    rr.flush()
    assert np.array_equal(g["color"], jb["color"].read().ravel())
    assert np.array_equal(g["depth"], jb["ds"].read().ravel())
    # meshes and textures of 0.3-2 MB through a 2 MiB pool: older ones are evicted, results unchanged
    for j in (1, 3, 1, 5, 3):
        t = submit(srv, j, 1, mesh=True)
        g = drain(srv, 1)[t]
        assert g["status"] == 0
        assert np.array_equal(g["color"], plain(hip_lib, SPECS[j], 1)[0])
    st = srv.stats()
    assert st.evictions >= 1 and st.pool_bytes <= d.pool_bytes


@pytest.mark.gpu
def test_server_admission_by_footprint(hip_lib):
    r = hip_lib.renderer()
    d = A.default_server_desc(hip_lib)
    d.arena_bytes = 48 << 20
    srv = r.create_server(d)
    # ~31 MB each (640x360 4x targets, scratch and bins): fits alone, not twice, so requests that
    # arrive together still run one batch each
    arrive = srv.now() + 5_000_000
    tk = [submit(srv, 5, 4, arrival_ns=arrive) for _ in range(3)]
    got = drain(srv, 3)
    assert all(got[t]["status"] == 0 for t in tk)
    assert max(got[t]["batch_jobs"] for t in tk) == 1
    ref = plain(hip_lib, SPECS[5], 4)[0]
    assert all(np.array_equal(got[t]["color"], ref) for t in tk)
    # larger than the whole arena (1280x720 4x targets alone are 48 MB): refused with OUT_OF_MEMORY,
    # and the server keeps serving
    sp = SPECS[5]
    t = srv.submit(sp.seed, sp.n_tris, 1280, 720, 4, clear=CLEAR)
    assert drain(srv, 1)[t]["status"] == A.OUT_OF_MEMORY
    t = submit(srv, 0, 1)
    assert drain(srv, 1)[t]["status"] == 0


def _release_before_poll(lib):
    r = lib.renderer()
    srv = r.create_server(A.default_server_desc(lib))
    a = submit(srv, 4, 1)
    b = submit(srv, 0, 1)
    # wait (capacity 0 reports nothing) until both are done, then release one unreported ticket
    import time

    deadline = time.time() + 60
    while srv.stats().jobs < 2 and time.time() < deadline:
        srv.poll(0, wait_ms=50)
    srv.release(a)
    got = drain(srv, 1)
    assert set(got) == {b} and got[b]["status"] == 0 and "color" in got[b]
    assert srv.poll(16, wait_ms=200) == []  # the released ticket is never reported
    with pytest.raises(A.SrxError):
        srv.release(a)


def test_oracle_server_release_before_poll(cpu_lib):
    """A finished ticket released before poll reports it is gone: poll never reports it."""
    _release_before_poll(cpu_lib)


@pytest.mark.gpu
def test_server_release_before_poll(hip_lib):
    _release_before_poll(hip_lib)



@pytest.mark.gpu
def test_served_target_beyond_8192_tiles(hip_lib):
    """A 2048x2048 request (16,384 16x16 tiles: more than one LDS range of tile counts), 1x and 4x
    MSAA, alone and beside a small one, equals the plain pipeline."""
    big = S.JobSpec(0x5A11C101, 2048, 2048, 30000, 64, 1)
    r = hip_lib.renderer()
    srv = r.create_server()
    for s in (1, 4):
        ref = plain(hip_lib, big, s)
        t = srv.submit(big.seed, big.n_tris, big.width, big.height, s, mesh=S.batch_triangles(big),
                       texture_key=0, clear=CLEAR, want_depth=True)
        t2 = submit(srv, 2, s)
        got = drain(srv, 2)
        assert got[t]["status"] == 0 and got[t2]["status"] == 0
        assert np.array_equal(got[t]["color"], ref[0]) and np.array_equal(got[t]["depth"], ref[1]), s
