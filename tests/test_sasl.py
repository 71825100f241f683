"""SASL pixel shaders: compiler, quad execution and draws.

Known-answer tests restate the reference's pixel-shader JIT tests
(sasl/test/jit_test/general.cpp:1247-1800 with the shaders of sasl/test/repo/*.sps) in shaders
written here: the same inputs, the same computation and the test's own expected values, evaluated
pixel by pixel in float32 (the tests' inputs from glibc rand() are replaced by a seeded generator).
The CPU oracle must reproduce them bit for bit (the reference's tolerances are looser); the HIP
library must reproduce the oracle bit for bit.  Intrinsics of the host/device math libraries
(exp, pow, sin, ...) are held to 4 ulp instead.

The SIMD semantics the KATs leave open -- masked stores with unmasked temporaries, derivatives
across inactive pixels, loop masks with break -- are pinned by programs whose per-pixel scalar
meaning is known (a SIMD execution without derivatives equals four independent scalar runs) and
by one derivative-in-a-branch case derived from cgs_simd::store (cgs_simd.cpp:85-146).
"""
import numpy as np
import pytest

from salviarenderer_amd import api as A

f32 = np.float32


def quads(per_pixel_inputs):
    """(npixels, ninputs, 4) -> (nquads, 4, ninputs, 4), pixels in quad order 0 1 / 2 3."""
    x = np.asarray(per_pixel_inputs, f32)
    assert x.shape[0] % 4 == 0
    return x.reshape(-1, 4, x.shape[1], 4)


def run(lib, src, inputs, variables=None):
    r = lib.renderer()
    code = lib.compile_shader(src)
    for k, v in (variables or {}).items():
        r.set_ps_variable(k, v)
    out = r.ps_execute_quads(code, quads(inputs))
    return out.reshape(-1, out.shape[2], 4)


def vec_inputs(*cols):
    """Per-pixel inputs from arrays of shape (npixels, ncomp) (padded to 4 with zeros)."""
    n = len(cols[0])
    x = np.zeros((n, len(cols), 4), f32)
    for i, c in enumerate(cols):
        c = np.asarray(c, f32).reshape(n, -1)
        x[:, i, :c.shape[1]] = c
    return x


def rand_like_glibc(rng, n, div):
    """rand() / div as float: an integer in [0, 2^31) converted to float, then divided."""
    return (rng.integers(0, 2 ** 31, size=n).astype(f32) / f32(div)).astype(f32)


# ---------------------------------------------------------------- known-answer programs
# arithmetic.sps (general.cpp:1247-1281)
SRC_ARITH = """
struct In { float v : TEXCOORD(0); };
struct Out { float c : COLOR; };
Out fn(In i) {
  Out o;
  float a, b;
  a = i.v;
  b = 5.0f;
  o.c = a + b;
  o.c = o.c + (a - b);
  o.c = o.c + (a * b);
  return o;
}
"""


def kat_arith():
    src = np.array([(f32(i) + f32(3.77)) * (f32(0.76) * f32(i)) for i in range(16)], f32)
    want = np.zeros((16, 1, 4), f32)
    for i, s in enumerate(src):
        d = s + f32(5.0)
        d = d + (s - f32(5.0))
        d = d + (s * f32(5.0))
        want[i, 0, 0] = d
    return SRC_ARITH, vec_inputs(src), want


# swizzle_and_wm.sps (general.cpp:1309-1345): o.f3.yzx = x.xyz + x.wxy
SRC_SWZ = """
struct PIn { float4 v : Texcoord(0); };
struct POut { float3 f : SV_Target; };
POut fn(PIn p) {
  POut o;
  float3 a, b;
  a = (p.v).xyz;
  b = (p.v).wxy;
  o.f.yzx = a + b;
  return o;
}
"""


def kat_swizzle():
    v = np.array([(f32(i) + f32(3.77)) * (f32(0.76) * f32(i)) for i in range(64)], f32).reshape(16, 4)
    want = np.zeros((16, 1, 4), f32)
    s = v[:, [0, 1, 2]] + v[:, [3, 0, 1]]
    want[:, 0, 1], want[:, 0, 2], want[:, 0, 0] = s[:, 0], s[:, 1], s[:, 2]
    return SRC_SWZ, vec_inputs(v), want


# intrinsics.sps (general.cpp:1374-1428): cross (overwritten), dot, sqrt
SRC_INTRIN = """
struct PIn { float3 a : Texcoord(0); float3 b : Texcoord(1); };
struct POut { float3 p : COLOR(0); float2 q : COLOR(1); };
POut fn(PIn i) {
  POut o;
  float3 c = cross(i.a, i.b);
  o.p = c;
  float d = dot(i.a, i.b);
  o.p = sqrt(i.a);
  o.q.x = d;
  o.q.y = sqrt(i.a.x);
  return o;
}
"""


def kat_intrinsics():
    n = 16
    a = np.zeros((n, 3), f32)
    b = np.zeros((n, 3), f32)
    for i in range(n):
        for j in range(3):
            base = f32(i * 3 + j)
            a[i, j] = (base + f32(3.77)) * (f32(0.76) * base)
            b[i, j] = (base - f32(4.62)) * (f32(0.11) * base)
    want = np.zeros((n, 2, 4), f32)
    prod = a * b
    want[:, 0, :3] = np.sqrt(a)
    want[:, 1, 0] = ((f32(0) + prod[:, 0]) + prod[:, 1]) + prod[:, 2]
    want[:, 1, 1] = np.sqrt(a[:, 0])
    return SRC_INTRIN, vec_inputs(a, b), want


# branches.sps (general.cpp:1431-1493), over in0 = i * 0.34 - 1 for 16 pixels (every branch)
SRC_BRANCH = """
struct PIn { float x : TEXCOORD(0); float3 y : TEXCOORD(1); };
struct POut { float2 r : COLOR; };
POut fn(PIn i) {
  POut o;
  o.r.x = 88.3f;
  o.r.y = 75.4f;
  if (i.x > 0.0f) { o.r.x = i.x; }
  if (i.x > 1.0f) { o.r.y = i.y.x; } else { o.r.y = i.y.y; }
  if (i.x > 2.0f) {
    o.r.y = i.y.z;
    if (i.x > 3.0f) { o.r.y = o.r.y + 1.0f; }
    else if (i.x > 2.5f) { o.r.y = o.r.y + 2.0f; }
  }
  return o;
}
"""


def kat_branches():
    rng = np.random.default_rng(5)
    n = 16
    x = np.array([f32(i) * f32(0.34) - f32(1.0) for i in range(n)], f32)
    y = rand_like_glibc(rng, 3 * n, 35.0).reshape(n, 3)
    want = np.zeros((n, 1, 4), f32)
    for i in range(n):
        r0, r1 = f32(88.3), f32(75.4)
        if x[i] > 0:
            r0 = x[i]
        r1 = y[i, 0] if x[i] > 1 else y[i, 1]
        if x[i] > 2:
            r1 = y[i, 2]
            if x[i] > 3:
                r1 = f32(r1 + f32(1))
            elif x[i] > f32(2.5):
                r1 = f32(r1 + f32(2))
        want[i, 0, :2] = (r0, r1)
    return SRC_BRANCH, vec_inputs(x, y), want


# ddx_ddy.sps (general.cpp:1496-1598)
SRC_DDXY = """
struct PIn { float a : TEXCOORD(0); float2 b : TEXCOORD(1); float3 c : TEXCOORD(2); float4 d : TEXCOORD(3); };
struct POut { float a : COLOR(0); float2 b : COLOR(1); float3 c : COLOR(2); float4 d : COLOR(3); };
POut fn(PIn i) {
  POut o;
  o.a = ddx(i.a) + ddy(i.a);
  o.b = ddx(i.b).xy + ddy(i.b).yx;
  o.c = ddx(i.c).xyz + ddy(i.c).yzx;
  o.d = ddx(i.d).xwzy + ddy(i.d).yzxw;
  return o;
}
"""


def quad_dd(v):
    """get_ddx / get_ddy of the test (general.cpp:1496-1523) for (npixels, k) values."""
    q = v.reshape(-1, 4, v.shape[1])
    dx = np.empty_like(q)
    dy = np.empty_like(q)
    dx[:, 0] = dx[:, 1] = q[:, 1] - q[:, 0]
    dx[:, 2] = dx[:, 3] = q[:, 3] - q[:, 2]
    dy[:, 0] = dy[:, 2] = q[:, 2] - q[:, 0]
    dy[:, 1] = dy[:, 3] = q[:, 3] - q[:, 1]
    return dx.reshape(v.shape), dy.reshape(v.shape)


def kat_ddxy():
    rng = np.random.default_rng(6)
    n = 16
    cols = [rand_like_glibc(rng, n * k, 67.0).reshape(n, k) for k in (1, 2, 3, 4)]
    want = np.zeros((n, 4, 4), f32)
    swz = [([0], [0]), ([0, 1], [1, 0]), ([0, 1, 2], [1, 2, 0]), ([0, 3, 2, 1], [1, 2, 0, 3])]
    for k, (c, (sx, sy)) in enumerate(zip(cols, swz)):
        dx, dy = quad_dd(c)
        want[:, k, :k + 1] = dx[:, sx] + dy[:, sy]
    return SRC_DDXY, vec_inputs(*cols), want


# for_loop.sps (general.cpp:1664-1705)
SRC_FOR = """
struct PIn { float v : TEXCOORD(0); };
struct POut { float r : COLOR; };
POut fn(PIn i) {
  POut o;
  float x = i.v;
  for (int k = 0; k < 10; k = k + 1) {
    x = x * 2.0f;
    if (x > 5000.0f) { break; }
  }
  o.r = x;
  return o;
}
"""


def kat_for():
    rng = np.random.default_rng(7)
    v = np.concatenate([rand_like_glibc(rng, 12, 1000.0) / f32(4096), np.array([0, 1, 2.5, 4999], f32)])
    v = np.concatenate([v, [f32(0.001)]]).astype(f32)[:16]
    want = np.zeros((16, 1, 4), f32)
    for i, x in enumerate(v):
        for _ in range(10):
            x = f32(x * f32(2))
            if x > 5000:
                break
        want[i, 0, 0] = x
    return SRC_FOR, vec_inputs(v), want


# while.sps (general.cpp:1708-1750)
SRC_WHILE = """
struct PIn { float v : TEXCOORD(0); };
struct POut { float r : COLOR; };
POut fn(PIn i) {
  POut o;
  float x = i.v;
  while (x < 3000.0f) {
    if (x < 1.0f) { break; }
    x = x * x;
  }
  o.r = x;
  return o;
}
"""

# do_while.sps (general.cpp:1753-1795)
SRC_DO = """
struct PIn { float v : TEXCOORD(0); };
struct POut { float r : COLOR; };
POut fn(PIn i) {
  POut o;
  float x = i.v;
  do {
    if (x < 1.0f) { break; }
    x = x * x;
  } while (x < 3000.0f);
  o.r = x;
  return o;
}
"""


def kat_while(do):
    v = np.array([f32(0.79) * f32(i * i * i) for i in range(16)], f32)
    want = np.zeros((16, 1, 4), f32)
    for i, x in enumerate(v):
        if do:
            while True:
                if x < 1:
                    break
                x = f32(x * x)
                if not x < 3000:
                    break
        else:
            while x < 3000:
                if x < 1:
                    break
                x = f32(x * x)
        want[i, 0, 0] = x
    return (SRC_DO if do else SRC_WHILE), vec_inputs(v), want


KATS = {
    "arithmetic": kat_arith, "swizzle_and_wm": kat_swizzle, "intrinsics": kat_intrinsics, "branches": kat_branches,
    "ddx_ddy": kat_ddxy, "for_loop": kat_for, "while": lambda: kat_while(False), "do_while": lambda: kat_while(True),
}


@pytest.mark.parametrize("name", list(KATS))
def test_reference_kat_on_oracle(cpu_lib, name):
    src, inputs, want = KATS[name]()
    got = run(cpu_lib, src, inputs)
    np.testing.assert_array_equal(got, want)


# ---------------------------------------------------------------- SIMD semantics and intrinsics
# Loops with break at several depths and data-dependent trip counts: with no derivative or texture
# in divergent code, the masked SIMD execution equals a scalar run per pixel.
SRC_LOOPS = """
float4 fn(float a : TEXCOORD(0), float2 b : TEXCOORD(1)) : COLOR {
  float x = a;
  float n = 0.0;
  float acc = 0.0;
  while (n < 9.0) {
    n = n + 1.0;
    if (x > 40.0) {
      if (b.x > 0.5) { break; }
      x = x - 35.0;
    }
    for (int k = 0; k < 3; k = k + 1) {
      acc = acc + x;
      if (acc > b.y) { break; }
    }
    x = x * 1.75;
  }
  float y = b.x;
  do { y = y * 2.0; } while (y < 3.0);
  return float4(x, n, acc, y);
}
"""


def scalar_loops(a, b):
    x, n, acc = f32(a), f32(0), f32(0)
    while n < 9:
        n = f32(n + 1)
        if x > 40:
            if b[0] > f32(0.5):
                break
            x = f32(x - f32(35))
        for _ in range(3):
            acc = f32(acc + x)
            if acc > b[1]:
                break
        x = f32(x * f32(1.75))
    y = f32(b[0])
    while True:
        y = f32(y * f32(2))
        if not y < 3:
            break
    return x, n, acc, y


def test_loops_match_scalar_execution(cpu_lib):
    rng = np.random.default_rng(11)
    n = 64
    a = (rng.random(n) * 60).astype(f32)
    b = np.stack([rng.random(n), rng.random(n) * 300 + 1], 1).astype(f32)
    got = run(cpu_lib, SRC_LOOPS, vec_inputs(a, b))
    want = np.array([scalar_loops(a[i], b[i]) for i in range(n)], f32)
    np.testing.assert_array_equal(got[:, 0], want)


# A derivative inside a branch reads the variable in every pixel of the quad: pixels outside the
# branch keep their earlier value (masked store, unmasked load).
SRC_DD_BRANCH = """
float4 fn(float a : TEXCOORD(0)) : COLOR {
  float v = a;
  float d = -1.0;
  if (a > 1.5) {
    v = a * 10.0;
    d = ddx(v) + ddy(v) * 100.0;
  }
  return float4(v, d, 0.0, 1.0);
}
"""


def test_derivative_in_branch_sees_inactive_pixels(cpu_lib):
    a = np.array([1.0, 2.0, 3.0, 1.0, 2.0, 2.0, 2.0, 2.0, 0.0, 0.0, 0.0, 5.0], f32)
    got = run(cpu_lib, SRC_DD_BRANCH, vec_inputs(a))
    want = np.zeros((12, 1, 4), f32)
    for q in range(3):
        p = a[4 * q:4 * q + 4]
        act = p > f32(1.5)
        v = np.where(act, p * f32(10), p).astype(f32)
        dx = np.array([v[1] - v[0], v[1] - v[0], v[3] - v[2], v[3] - v[2]], f32)
        dy = np.array([v[2] - v[0], v[3] - v[1], v[2] - v[0], v[3] - v[1]], f32)
        d = np.where(act, dx + dy * f32(100), f32(-1)).astype(f32)
        want[4 * q:4 * q + 4, 0] = np.stack([v, d, np.zeros(4, f32), np.ones(4, f32)], 1)
    np.testing.assert_array_equal(got, want)


# The intrinsic formulas of cg_impl.cpp:1030-1195: each is a fixed chain of float ops.
SRC_FORMULAS = """
float4 k;
float4 fn(float4 a : TEXCOORD(0), float4 b : TEXCOORD(1)) : COLOR(0) {
  float3 n = normalize(a.xyz);
  float l = length(b.xyz);
  float ds = distance(a.xyz, b.xyz);
  float3 r = reflect(a.xyz, n);
  float3 le = lerp(a.xyz, b.xyz, k.x);
  float4 c = clamp(a, -k.y, k.y);
  float4 s = saturate(b);
  float4 m = min(a, b) + max(a, b) * 2.0;
  float4 f = frac(a * 3.0);
  float4 q = rcp(b + 2.0);
  float4 w = abs(-a) + sqrt(abs(b)) + floor(a) + ceil(b) + round(a * 0.5) + trunc(b * 3.0);
  float4 g = fmod(a * 7.0, 3.0) + ldexp(b, 3.0) + exp2(b * 4.0);
  return float4(dot(n, r) + l + ds, dot(le, float3(1.0, 2.0, 3.0)), dot(c + s + m, f + q), dot(w, g));
}
"""


def eflib_round(v):
    """eflib fast_round (math.h:111-124): add and subtract 2^23 with v's sign."""
    v = np.asarray(v, f32)
    bias = np.where(np.signbit(v), f32(-8388608.0), f32(8388608.0)).astype(f32)
    return ((v + bias).astype(f32) - bias).astype(f32)


def test_intrinsic_formulas(cpu_lib):
    rng = np.random.default_rng(12)
    n = 32
    a = (rng.random((n, 4)) * 8 - 4).astype(f32)
    b = (rng.random((n, 4)) * 2 - 1).astype(f32)
    kv = np.array([0.3, 1.5, 0, 0], f32)
    got = run(cpu_lib, SRC_FORMULAS, vec_inputs(a, b), {"k": kv})

    def dot(x, y):
        s = f32(0)
        for i in range(x.shape[-1]):
            s = f32(s + f32(x[i] * y[i]))
        return s

    def floor_(v):
        r = eflib_round(v)
        return np.where(r > v, r - f32(1), r).astype(f32)

    def ceil_(v):
        r = eflib_round(v)
        return np.where(r < v, r + f32(1), r).astype(f32)

    want = np.zeros((n, 4), f32)
    for i in range(n):
        A3, B3 = a[i, :3], b[i, :3]
        nn = (A3 / np.sqrt(dot(A3, A3))).astype(f32)
        l = np.sqrt(dot(B3, B3))
        d3 = (B3 - A3).astype(f32)
        ds = np.sqrt(dot(d3, d3))
        r = (A3 - (f32(2) * dot(A3, nn)) * nn).astype(f32)
        le = (A3 + (B3 - A3).astype(f32) * kv[0]).astype(f32)
        c = np.where(a[i] >= -kv[1], a[i], -kv[1]).astype(f32)
        c = np.where(c <= kv[1], c, kv[1]).astype(f32)
        s = np.where(b[i] >= 0, b[i], f32(0)).astype(f32)
        s = np.where(s <= 1, s, f32(1)).astype(f32)
        m = (np.where(a[i] < b[i], a[i], b[i]) + np.where(a[i] > b[i], a[i], b[i]) * f32(2)).astype(f32)
        a3 = np.abs((a[i] * f32(3)).astype(f32))
        f = (a3 - floor_(a3)).astype(f32)
        q = (f32(1) / (b[i] + f32(2)).astype(f32)).astype(f32)
        w = (np.abs(a[i]) + np.sqrt(np.abs(b[i]))).astype(f32)
        w = (w + floor_(a[i])).astype(f32)
        w = (w + ceil_(b[i])).astype(f32)
        w = (w + eflib_round((a[i] * f32(0.5)).astype(f32))).astype(f32)
        b3 = (b[i] * f32(3)).astype(f32)
        w = (w + np.where(b3 > 0, floor_(b3), ceil_(b3))).astype(f32)
        g = np.fmod((a[i] * f32(7)).astype(f32), f32(3)).astype(f32)
        g = (g + np.ldexp(b[i], 3).astype(f32)).astype(f32)
        e4 = (b[i] * f32(4)).astype(f32)
        g = (g + np.ldexp(f32(1), np.trunc(e4).astype(np.int32)).astype(f32)).astype(f32)
        want[i] = (f32(dot(nn, r) + l) + ds, dot(le, np.array([1, 2, 3], f32)),
                   dot((c + s + m).astype(f32), (f + q).astype(f32)), dot(w, g))
    np.testing.assert_array_equal(got[:, 0], want)


# Library intrinsics, one per output component: within 4 ulp of numpy's float32 results on the
# oracle (glibc); the HIP library's device math within 4 ulp of the oracle.
SRC_LIBM = ["""
float4 fn(float4 a : TEXCOORD(0)) : COLOR { return float4(exp(a.x), sin(a.y), pow(a.z, 2.5), cos(a.w)); }
""", """
float4 fn(float4 a : TEXCOORD(0)) : COLOR { return float4(tanh(a.x), atan(a.y), log10(a.z), tan(a.w * 0.25)); }
""", """
float4 fn(float4 a : TEXCOORD(0)) : COLOR { return float4(sinh(a.x), cosh(a.y), asin(a.y * 0.2), acos(a.w * 0.2)); }
""", """
float4 fn(float4 a : TEXCOORD(0)) : COLOR { return float4(rsqrt(a.z), exp(a.y * 3.0), pow(a.x * a.x + 0.5, a.w), sin(a.z * 9.0)); }
"""]


def libm_expected(a):
    x, y, z, w = (a[:, i].astype(np.float64) for i in range(4))
    f = lambda v: np.asarray(v, np.float64).astype(f32).astype(np.float64)  # noqa: E731
    return [np.stack([np.exp(x), np.sin(y), z ** 2.5, np.cos(w)], 1),
            np.stack([np.tanh(x), np.arctan(y), np.log10(z), np.tan(f(w * 0.25))], 1),
            np.stack([np.sinh(x), np.cosh(y), np.arcsin(f(y * 0.2)), np.arccos(f(w * 0.2))], 1),
            np.stack([1 / np.sqrt(z), np.exp(f(y * 3.0)), f(f(x * x) + 0.5) ** w, np.sin(f(z * 9.0))], 1)]


def libm_inputs():
    rng = np.random.default_rng(13)
    a = np.stack([rng.random(32) * 4 - 2, rng.random(32) * 4 - 2, rng.random(32) * 3 + 0.1, rng.random(32) * 4 - 2], 1)
    return a.astype(f32)


def ulp_diff(a, b):
    ai = np.asarray(a, f32).view(np.int32).astype(np.int64)
    bi = np.asarray(b, f32).view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


def test_library_intrinsics(cpu_lib):
    a = libm_inputs()
    for src, want in zip(SRC_LIBM, libm_expected(a)):
        got = run(cpu_lib, src, vec_inputs(a))[:, 0]
        d = ulp_diff(got, want.astype(f32))
        assert d.max() <= 4, (src, d.max(axis=0))


def test_fast_log_matches_eflib(cpu_lib):
    """log/log2 are eflib's fast_log/fast_log2 (compiler_impl.cpp:382-388), not libm: log2 equals
    the oracle's pinned fast_log2 (test_oracle_pins), log is it times 0.69314718, and both differ
    from the library logarithm between powers of two."""
    import ctypes
    d = cpu_lib.dll
    d.srxo_test_fast_log2.restype = ctypes.c_float
    d.srxo_test_fast_log2.argtypes = [ctypes.c_float]
    src = """float4 fn(float a : TEXCOORD(0)) : COLOR { return float4(log2(a), log(a), 0.0, 0.0); }"""
    a = np.array([0.25, 0.5, 1.0, 2.0, 3.0, 10.0, 1024.0, 2.0 ** -10], f32)
    got = run(cpu_lib, src, vec_inputs(a))[:, 0]
    fl = np.array([d.srxo_test_fast_log2(float(x)) for x in a], f32)
    np.testing.assert_array_equal(got[:, 0], fl)
    np.testing.assert_array_equal(got[:, 1], (fl * f32(0.69314718)).astype(f32))
    assert abs(float(got[5, 0]) - np.log2(10.0)) > 1e-4  # the approximation, not log2f


def test_uniforms_and_reflection(cpu_lib):
    src = """
    float4 tint;
    int steps;
    float scale;
    float4 fn(float2 uv : TEXCOORD(3), float s : TEXCOORD(1)) : COLOR(0) {
      float x = 0.0;
      for (int i = 0; i < steps; i = i + 1) { x = x + scale; }
      return tint * uv.x + float4(x, s, uv.y, 0.0);
    }
    """
    code = cpu_lib.compile_shader(src)
    info = code.reflect()
    assert (info.input_count, info.output_count, info.uniform_count, info.sampler_count) == (2, 1, 3, 0)
    # inputs in TEXCOORD index order: s (TEXCOORD1) is attribute 0, uv (TEXCOORD3) attribute 1
    assert list(info.input_components[:2]) == [1, 2]
    r = cpu_lib.renderer()
    r.set_ps_variable("tint", [1.0, 2.0, 3.0, 4.0])
    r.set_ps_variable("steps", [3], dtype=np.int32)
    r.set_ps_variable("scale", 0.5)
    r.set_ps_variable("unused_name", 7.0)  # ignored
    x = np.zeros((4, 2, 4), f32)
    x[:, 0, 0] = [10, 11, 12, 13]      # s
    x[:, 1, :2] = [[1, 2], [3, 4], [5, 6], [7, 8]]  # uv
    out = r.ps_execute_quads(code, quads(x)).reshape(4, 4)
    for p in range(4):
        u, v = x[p, 1, 0], x[p, 1, 1]
        np.testing.assert_array_equal(out[p], np.array([1 * u + 1.5, 2 * u + x[p, 0, 0], 3 * u + v, 4 * u], f32))


@pytest.mark.parametrize("src,code", [
    ("float4 f(float a : TEXCOORD0) : COLOR { if (a > 0.0 && a < 1.0) { a = 1.0; } return float4(a, 0, 0, 0); }", A.E_UNSUPPORTED),
    ("float4 f(float a : TEXCOORD0) : COLOR { return float4(a > 0.0 ? 1.0 : 0.0, 0, 0, 0); }", A.E_UNSUPPORTED),
    ("float4 f(float a : TEXCOORD0) : COLOR { while (a < 5.0) { a = a + 1.0; continue; } return float4(a, 0, 0, 0); }", A.E_UNSUPPORTED),
    ("float4 f(float a : TEXCOORD0) : COLOR { return float4((int)a, 0, 0, 0); }", A.E_UNSUPPORTED),
    ("sampler s; float4 f(float4 a : TEXCOORD0) : COLOR { return tex2Dlod(s, a); }", A.E_UNSUPPORTED),
    ("float4x4 m; float4 f(float4 a : TEXCOORD0) : COLOR { return a; }", A.E_UNSUPPORTED),
    ("float4 f(float4 a : POSITION) : COLOR { return a; }", A.E_UNSUPPORTED),
    ("float4 f(float4 a : TEXCOORD0) : DEPTH { return a; }", A.E_UNSUPPORTED),
    ("float4 g(float4 a) { return a; } float4 f(float4 a : TEXCOORD0) : COLOR { return g(a); }", A.E_UNSUPPORTED),
    ("float4 f(float4 a : TEXCOORD0) : COLOR { return a + ; }", A.INVALID_PARAMETER),
    ("float4 f(float4 a : TEXCOORD0) : COLOR { return b; }", A.INVALID_PARAMETER),
    ("float4 f(float4 a : TEXCOORD0, float4 b : TEXCOORD0) : COLOR { return a; }", A.INVALID_PARAMETER),
    ("float4 f(float4 a : TEXCOORD0) : COLOR { a.xx = a.yz; return a; }", A.INVALID_PARAMETER),
    ("float4 f(float2 a : TEXCOORD0) : COLOR { return a.xyz; }", A.INVALID_PARAMETER),
    ("float4 f(float4 a : TEXCOORD0) : COLOR { break; return a; }", A.INVALID_PARAMETER),
])
def test_compile_errors(cpu_lib, src, code):
    with pytest.raises(A.SrxError) as e:
        cpu_lib.compile_shader(src)
    assert e.value.code == code, str(e.value)


def test_step_limit_ends_a_runaway_loop(cpu_lib):
    """A quad whose loop never ends stops after SASL_STEP_LIMIT instructions (the reference would
    not return); other quads of the call are unaffected."""
    src = """float4 f(float a : TEXCOORD0) : COLOR { float x = a; while (x < 10.0) { x = x * 1.0; } return float4(x, 1, 0, 0); }"""
    got = run(cpu_lib, src, vec_inputs(np.array([0.5, 20, 20, 20, 20, 30, 40, 50], f32)))[:, 0]
    np.testing.assert_array_equal(got[4:], [[20, 1, 0, 0], [30, 1, 0, 0], [40, 1, 0, 0], [50, 1, 0, 0]])


# ---------------------------------------------------------------- HIP against the oracle
def compare_hip(hip_lib, cpu_lib, src, inputs, variables=None, ulp=0):
    want = run(cpu_lib, src, inputs, variables)
    got = run(hip_lib, src, inputs, variables)
    if ulp == 0:
        np.testing.assert_array_equal(got, want)
    else:
        d = ulp_diff(got, want)
        print(f"\n  max ulp per component {d.reshape(-1, 4).max(axis=0)}", flush=True)
        assert d.max() <= ulp


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(KATS))
def test_reference_kat_hip(hip_lib, cpu_lib, name):
    src, inputs, want = KATS[name]()
    np.testing.assert_array_equal(run(hip_lib, src, inputs), want)


@pytest.mark.gpu
def test_simd_semantics_hip(hip_lib, cpu_lib):
    rng = np.random.default_rng(21)
    a = (rng.random(256) * 60).astype(f32)
    b = np.stack([rng.random(256), rng.random(256) * 300 + 1], 1).astype(f32)
    compare_hip(hip_lib, cpu_lib, SRC_LOOPS, vec_inputs(a, b))
    compare_hip(hip_lib, cpu_lib, SRC_DD_BRANCH, vec_inputs((rng.random(64) * 3).astype(f32)))
    a4 = (rng.random((64, 4)) * 8 - 4).astype(f32)
    b4 = (rng.random((64, 4)) * 2 - 1).astype(f32)
    compare_hip(hip_lib, cpu_lib, SRC_FORMULAS, vec_inputs(a4, b4), {"k": np.array([0.3, 1.5, 0, 0], f32)})
    for src in SRC_LIBM:
        compare_hip(hip_lib, cpu_lib, src, vec_inputs(libm_inputs()), ulp=4)


@pytest.mark.gpu
def test_step_limit_hip(hip_lib, cpu_lib):
    src = """float4 f(float a : TEXCOORD0) : COLOR { float x = a; while (x < 10.0) { x = x * 1.0; } return float4(x, 1, 0, 0); }"""
    compare_hip(hip_lib, cpu_lib, src, vec_inputs(np.array([0.5, 20, 20, 20, 30, 30, 30, 30], f32)))


# ---------------------------------------------------------------- draws
from salviarenderer_amd import scenes as S  # noqa: E402


def render_sasl(lib, name, stats=False, **kw):
    r = lib.renderer()
    f = S.SaslFrame(r, name, **kw)
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
    return out


def test_sasl_draws_on_the_oracle(cpu_lib):
    for name in S.SASL_SHADERS:
        out = render_sasl(cpu_lib, name)
        covered = (out["ds"].view(np.float32).reshape(-1, 2)[:, 0] < 1.0).mean()
        assert covered > 0.5, name
        c = out["color"].reshape(-1, 4)
        assert len(np.unique(c, axis=0)) > (1 if name != "constant" else 0), name
    # the constant shader reads its uniform (0.25, 0.5, 0.75, 1) in UNORM8
    out = render_sasl(cpu_lib, "constant")
    c = out["color"].reshape(-1, 4)
    covered = out["ds"].view(np.float32).reshape(-1, 2)[:, 0] < 1.0
    assert (c[covered] == np.array([64, 128, 191, 255], np.uint8)).all()


def test_sasl_code_takes_precedence_over_the_shader_id(cpu_lib):
    r = cpu_lib.renderer()
    f = S.SaslFrame(r, "constant")
    f.render()
    r.flush()
    with_code = f.color.read().copy()
    assert r.get_pixel_shader_code() is not None
    r.set_pixel_shader_code(None)
    assert r.get_pixel_shader_code() is None
    r.set_render_targets([f.color], f.ds)
    r.clear_color(f.color, f.clear_rgba)
    r.clear_depth_stencil(f.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
    r.draw_index(0, f.n_tris, 0)  # now PS_CONSTANT (1, 0, 1, 1)
    r.flush()
    without = f.color.read().reshape(-1, 4)
    covered = f.ds.read().view(np.float32).reshape(-1, 2)[:, 0] < 1.0
    assert (without[covered] == np.array([255, 0, 255, 255], np.uint8)).all()
    assert (with_code.reshape(-1, 4)[covered] == np.array([64, 128, 191, 255], np.uint8)).all()


def test_sasl_reads_more_attributes_than_written(cpu_lib):
    r = cpu_lib.renderer()
    f = S.SaslFrame(r, "box")
    f.code = cpu_lib.compile_shader(
        "float4 f(float4 a : TEXCOORD0, float4 b : TEXCOORD1, float4 c : TEXCOORD2) : COLOR { return a + b + c; }")
    with pytest.raises(A.SrxError) as e:
        f.render()
    assert e.value.code == A.INVALID_PARAMETER


LIBM_SHADERS = {"procedural"}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(S.SASL_SHADERS))
@pytest.mark.parametrize("samples", [1, 4])
def test_sasl_draw_hip(hip_lib, cpu_lib, name, samples):
    kw = dict(samples=samples)
    if name == "derivatives":
        kw["fmt"] = A.PF_RGBA32F
    if name == "box":
        kw["blend"] = A.BLEND_ON
    want = render_sasl(cpu_lib, name, stats=True, **kw)
    for stats in (False, True):
        got = render_sasl(hip_lib, name, stats=stats, **kw)
        assert np.array_equal(got["ds"], want["ds"]), (name, "depth/stencil")
        keys = ["color"] + (["resolved"] if samples > 1 else [])
        for k in keys:
            if name in LIBM_SHADERS:
                d = np.abs(got[k].astype(np.int32) - want[k].astype(np.int32))
                print(f"\n  {name} {k}: UNORM8 bytes differing {int((d > 0).sum())} of {d.size}", flush=True)
                assert d.max() <= 1, name
            else:
                assert np.array_equal(got[k], want[k]), (name, k)
        if stats:
            assert got["ps_invocations"] == want["ps_invocations"]


@pytest.mark.gpu
def test_sasl_lambert_aniso_hip(hip_lib, cpu_lib):
    """The textured Lambert shader with a 16x anisotropic sampler (the fast RGBA8 sampling path)."""
    want = render_sasl(cpu_lib, "lambert", aniso=True, width=480, height=270, grid=96, tex_size=256)
    got = render_sasl(hip_lib, "lambert", aniso=True, width=480, height=270, grid=96, tex_size=256)
    assert np.array_equal(got["ds"], want["ds"])
    assert np.array_equal(got["color"], want["color"])


@pytest.mark.gpu
def test_sasl_textured_mesh_small_hip(hip_lib, cpu_lib):
    """The textured-mesh frame (320x176, 4x MSAA, 16x anisotropic) shaded by the SASL Lambert shader."""
    outs = []
    for lib in (cpu_lib, hip_lib):
        r = lib.renderer()
        f = S.setup_mesh(r, 320, 176, 4, grid=128, tex_size=256, sasl=True)
        f.render()
        r.flush()
        outs.append((f.color.read().copy(), f.ds.read().copy(), f.resolved.read().copy()))
    for w, g in zip(*outs):
        assert np.array_equal(w, g)


@pytest.mark.gpu
def test_sasl_draw_refused_in_a_batch(hip_lib):
    r = hip_lib.renderer()
    f = S.SaslFrame(r, "lambert")
    r.batch_begin()
    with pytest.raises(A.SrxError) as e:
        f.render(resolve=False)
    assert e.value.code == A.E_UNSUPPORTED
    with pytest.raises(A.SrxError):
        r.batch_end()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["lambert", "procedural"])
def test_sasl_in_order_paths_hip(hip_lib, cpu_lib, name):
    """Blended SASL draws and SASL draws under a statistics query run the in-order kernel's quad
    shading; both against the oracle."""
    for blend, stats in ((A.BLEND_ON, False), (A.BLEND_OFF, True)):
        want = render_sasl(cpu_lib, name, stats=stats, samples=4, blend=blend)
        got = render_sasl(hip_lib, name, stats=stats, samples=4, blend=blend)
        assert np.array_equal(got["ds"], want["ds"])
        d = np.abs(got["color"].astype(np.int32) - want["color"].astype(np.int32))
        assert d.max() <= (1 if name in LIBM_SHADERS else 0), (name, blend, stats)
        if stats:
            assert got["ps_invocations"] == want["ps_invocations"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["lambert", "procedural", "derivatives", "tex_derivatives"])
def test_sasl_interpreter_and_native_agree_hip(hip_lib, cpu_lib, name, monkeypatch):
    """Opaque SASL draws shade with the program compiled to native code at run time; with
    SRX_SASL_JIT=0 the quad interpreter shades them instead.  Both against the oracle.
    "tex_derivatives" takes derivatives of texture samples and branches on one, so the samples of
    pixels that store nothing reach the stored colours."""
    kw = dict(samples=4, fmt=A.PF_RGBA32F if name in ("derivatives", "tex_derivatives") else A.PF_RGBA8)
    want = render_sasl(cpu_lib, name, **kw)
    got = {}
    for jit in ("1", "0"):
        monkeypatch.setenv("SRX_SASL_JIT", jit)
        got[jit] = render_sasl(hip_lib, name, **kw)
    for k in ("ds", "color", "resolved"):
        if name in LIBM_SHADERS and k != "ds":
            for jit in ("1", "0"):
                assert np.abs(got[jit][k].astype(np.int32) - want[k].astype(np.int32)).max() <= 1
            assert np.array_equal(got["1"][k], got["0"][k])  # the same device math either way
        else:
            assert np.array_equal(got["1"][k], want[k]), k
            assert np.array_equal(got["0"][k], want[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["lambert", "procedural"])
def test_sasl_noalloc_draw_hip(hip_lib, cpu_lib, name, monkeypatch):
    """srx_draw_noalloc with a SASL shader: a program not yet compiled shades with the interpreter
    (compiling would allocate), a compiled one with its native code; both as srx_draw_index."""
    want = render_sasl(cpu_lib, name, samples=2)
    monkeypatch.setenv("SRX_SASL_JIT", "1")
    for warm in (False, True):
        r = hip_lib.renderer()
        f = S.SaslFrame(r, name, samples=2)
        # binds the frame's state; with warm, an allocating draw compiles the program first (no
        # other test draws these programs at 2 samples, so the first pass finds them uncompiled)
        monkeypatch.setenv("SRX_SASL_JIT", "1" if warm else "0")
        f.render()
        r.flush()
        monkeypatch.setenv("SRX_SASL_JIT", "1")
        n = f.n_tris
        r2 = r
        r2.set_render_targets([f.color], f.ds)
        r2.set_viewport(0, 0, f.width, f.height, 0.0, 1.0)
        r2.clear_color(f.color, f.clear_rgba)
        r2.clear_depth_stencil(f.ds, A.CLEAR_DEPTH | A.CLEAR_STENCIL, 1.0, 0)
        # a non-indexed draw of the same triangles: expand the index buffer into a vertex buffer
        verts, idx = S.grid_mesh(48)
        vb = r2.create_buffer_from(np.ascontiguousarray(verts[idx.ravel()]))
        r2.set_vertex_buffers(0, [vb], [32])
        r2.draw_noalloc(0, n)
        r2.resolve(f.color, f.resolved)
        r2.flush()
        got_ds = f.ds.read()
        assert np.array_equal(got_ds, want["ds"]), (name, warm)
        d = np.abs(f.color.read().astype(np.int32) - want["color"].astype(np.int32))
        assert d.max() <= (1 if name in LIBM_SHADERS else 0), (name, warm)
