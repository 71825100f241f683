// HIP execution of a compiled SASL pixel shader (srx_sasl_program.h): one thread runs the four
// pixels of a 2x2 quad together, as the reference's SIMD code runs its 4-pixel package
// (sasl/src/codegen/cgs_simd.cpp).  The register file (32 registers x 4 pixels x float4) is the
// thread's private memory; the program and its constants are read from device memory.
#pragma once

#include "srx_sasl_program.h"
#include "srx_shade_common.h"

namespace srx {

struct sasl_regs {
  float4 r[SASL_MAX_REGS][4];  // register, pixel
};

__device__ __forceinline__ float sasl_comp(const float4& v, int c) {
  return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}
__device__ __forceinline__ float4 sasl_swizzle(const float4& v, uint32_t s) {
  return make_float4(sasl_comp(v, s & 3), sasl_comp(v, (s >> 2) & 3), sasl_comp(v, (s >> 4) & 3), sasl_comp(v, (s >> 6) & 3));
}

// eflib fast_ceil / fast_log (math.h:107-109, 126-129)
__device__ __forceinline__ float sasl_fast_ceil(float v) {
  const float f = fast_round(v);
  return (f < v) ? f + 1.0f : f;
}

// The reference's scalar externals (sasl/src/drivers/compiler_impl.cpp:340-416).  expf, powf and
// the trigonometric functions are the device library's (within a few ulp of the host's libm);
// the eflib approximations, ldexpf and fmodf are exact restatements.
__device__ __forceinline__ float sasl_fn1(int fn, float v) {
  switch (fn) {
  case SFN_EXP: return expf(v);
  case SFN_EXP2: return ldexpf(1.0f, cvtt(v));
  case SFN_LOG: return fast_log2(v) * 0.69314718f;
  case SFN_LOG2: return fast_log2(v);
  case SFN_LOG10: return log10f(v);
  case SFN_SIN: return sinf(v);
  case SFN_COS: return cosf(v);
  case SFN_TAN: return tanf(v);
  case SFN_ASIN: return asinf(v);
  case SFN_ACOS: return acosf(v);
  case SFN_ATAN: return atanf(v);
  case SFN_SINH: return sinhf(v);
  case SFN_COSH: return coshf(v);
  case SFN_TANH: return tanhf(v);
  case SFN_CEIL: return sasl_fast_ceil(v);
  case SFN_FLOOR: return fast_floor(v);
  case SFN_ROUND: return fast_round(v);
  case SFN_TRUNC: return v > 0.0f ? fast_floor(v) : sasl_fast_ceil(v);
  case SFN_RSQRT: return 1.0f / sqrtf(v);
  }
  return 0.0f;
}
__device__ __forceinline__ float sasl_fn2(int fn, float a, float b) {
  switch (fn) {
  case SFN_POW: return powf(a, b);
  case SFN_FMOD: return fmodf(a, b);
  case SFN_LDEXP: return ldexpf(a, cvtt(b));
  }
  return 0.0f;
}

__device__ __forceinline__ float4 sasl_map(const float4& a, const float4& b, int op, int imm) {
  auto f = [&](float x, float y) -> float {
    switch (op) {
    case SOP_MOV: return x;
    case SOP_ADD: return x + y;
    case SOP_SUB: return x - y;
    case SOP_MUL: return x * y;
    case SOP_DIV: return x / y;
    case SOP_NEG: return -x;
    case SOP_SLT: return x < y ? x : y;
    case SOP_SGT: return x > y ? x : y;
    case SOP_SGE: return x >= y ? x : y;
    case SOP_SLE: return x <= y ? x : y;
    case SOP_CLT: return x < y ? 1.0f : 0.0f;
    case SOP_CGT: return x > y ? 1.0f : 0.0f;
    case SOP_CLE: return x <= y ? 1.0f : 0.0f;
    case SOP_CGE: return x >= y ? 1.0f : 0.0f;
    case SOP_CEQ: return x == y ? 1.0f : 0.0f;
    case SOP_CNE: return x != y ? 1.0f : 0.0f;
    case SOP_ABS: return fabsf(x);
    case SOP_SQRT: return sqrtf(x);
    case SOP_FN1: return sasl_fn1(imm, x);
    case SOP_FN2: return sasl_fn2(imm, x, y);
    }
    return 0.0f;
  };
  return make_float4(f(a.x, b.x), f(a.y, b.y), f(a.z, b.z), f(a.w, b.w));
}

struct sasl_frame_dev {
  uint8_t loop, saved, cond, mv, brk;
};

// Runs p on a quad whose input registers the caller has loaded into R; outputs are read from R
// afterwards (p->outputs).  tex(slot, px, u, v, dxu, dxv, dyu, dyv) samples for one pixel.
template <class TEX>
__device__ void sasl_exec(const sasl_program* __restrict__ p, sasl_regs& R, TEX&& tex) {
  const float4* __restrict__ K = reinterpret_cast<const float4*>(p->consts);
  auto read = [&](uint32_t o, uint32_t s, int px) -> float4 {
    const float4 v = o >= SASL_CONST ? K[o - SASL_CONST] : R.r[o][px];
    return sasl_swizzle(v, s);
  };
  sasl_frame_dev st[SASL_MAX_DEPTH];
  int depth = 0;
  uint32_t exec = 15u;
  auto loop_brk = [&]() -> uint32_t {
    for (int d = depth - 1; d >= 0; --d)
      if (st[d].loop) return st[d].brk;
    return 0u;
  };
  auto cond_bits = [&](uint32_t o, uint32_t s) -> uint32_t {
    uint32_t m = 0;
    for (int px = 0; px < 4; ++px)
      if (read(o, s, px).x != 0.0f) m |= 1u << px;
    return m;
  };
  const uint32_t n = p->ninst;
  uint32_t pc = 0, steps = 0;
  while (pc < n) {
    const sasl_inst in = p->code[pc];
    uint32_t next = pc + 1;
    float4 res[4];
    bool writes = false;
    switch (in.op) {
    case SOP_IF: {
      const uint32_t c = cond_bits(in.a, in.sa);
      st[depth++] = sasl_frame_dev{0, static_cast<uint8_t>(exec), static_cast<uint8_t>(c), 0, 0};
      exec &= c;
      if (!exec) next = in.target;
      break;
    }
    case SOP_ELSE: {
      sasl_frame_dev& f = st[depth - 1];
      f.saved = static_cast<uint8_t>(f.saved & ~loop_brk());
      exec = f.saved & ~static_cast<uint32_t>(f.cond) & 15u;
      if (!exec) next = in.target;
      break;
    }
    case SOP_ENDIF:
      --depth;
      exec = st[depth].saved & ~loop_brk() & 15u;
      break;
    case SOP_LOOP:
      st[depth++] = sasl_frame_dev{1, static_cast<uint8_t>(exec), 0, static_cast<uint8_t>(exec), 0};
      break;
    case SOP_LTEST: {
      sasl_frame_dev& f = st[depth - 1];
      exec = f.mv & (in.imm ? 15u : cond_bits(in.a, in.sa));
      f.brk = 0;
      if (!exec) next = in.target;
      break;
    }
    case SOP_DTEST: {
      sasl_frame_dev& f = st[depth - 1];
      exec = f.mv & cond_bits(in.a, in.sa);
      f.brk = 0;
      f.mv = static_cast<uint8_t>(exec);
      if (exec) {
        if (++steps > SASL_STEP_LIMIT) return;
        next = in.target;
      }
      break;
    }
    case SOP_DBODY: exec = st[depth - 1].mv; break;
    case SOP_LSAVE: st[depth - 1].mv = static_cast<uint8_t>(exec); break;
    case SOP_JMP:
      if (++steps > SASL_STEP_LIMIT) return;
      next = in.target;
      break;
    case SOP_LEND:
      --depth;
      exec = st[depth].saved;
      break;
    case SOP_BREAK:
      for (int d = depth - 1; d >= 0; --d)
        if (st[d].loop) {
          st[d].brk = static_cast<uint8_t>(st[d].brk | exec);
          break;
        }
      exec = 0;
      break;
    case SOP_NOP: break;
    case SOP_DDX:
    case SOP_DDY: {
      const float4 v0 = read(in.a, in.sa, 0), v1 = read(in.a, in.sa, 1), v2 = read(in.a, in.sa, 2), v3 = read(in.a, in.sa, 3);
      if (in.op == SOP_DDX) {  // p1 - p0 for the top row, p3 - p2 for the bottom
        res[0] = res[1] = make_float4(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z, v1.w - v0.w);
        res[2] = res[3] = make_float4(v3.x - v2.x, v3.y - v2.y, v3.z - v2.z, v3.w - v2.w);
      } else {  // p2 - p0 for the left column, p3 - p1 for the right
        res[0] = res[2] = make_float4(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z, v2.w - v0.w);
        res[1] = res[3] = make_float4(v3.x - v1.x, v3.y - v1.y, v3.z - v1.z, v3.w - v1.w);
      }
      writes = true;
      break;
    }
    case SOP_TEX:
      for (int px = 0; px < 4; ++px) {
        res[px] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (!((exec >> px) & 1u)) continue;  // tex2Dgrad_ps samples only executing pixels
        const float4 c = read(in.a, in.sa, px), dx = read(in.b, in.sb, px), dy = read(in.c, in.sc, px);
        res[px] = tex(in.imm, px, c.x, c.y, dx.x, dx.y, dy.x, dy.y);
      }
      writes = true;
      break;
    case SOP_DOT:
      for (int px = 0; px < 4; ++px) {
        const float4 a = read(in.a, in.sa, px), b = read(in.b, in.sb, px);
        float s = 0.0f;
        s = s + a.x * b.x;
        if (in.imm > 1) s = s + a.y * b.y;
        if (in.imm > 2) s = s + a.z * b.z;
        if (in.imm > 3) s = s + a.w * b.w;
        res[px] = make_float4(s, s, s, s);
      }
      writes = true;
      break;
    case SOP_CROSS:
      for (int px = 0; px < 4; ++px) {
        const float4 a = read(in.a, in.sa, px), b = read(in.b, in.sb, px);
        res[px] = make_float4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f);
      }
      writes = true;
      break;
    default:
      for (int px = 0; px < 4; ++px) res[px] = sasl_map(read(in.a, in.sa, px), read(in.b, in.sb, px), in.op, in.imm);
      writes = true;
      break;
    }
    if (writes) {
      const uint32_t pm = (in.wmask & SASL_WM_MASKED) ? exec : 15u;
      for (int px = 0; px < 4; ++px) {
        if (!((pm >> px) & 1u)) continue;
        float4& d = R.r[in.dst][px];
        if (in.wmask & 1u) d.x = res[px].x;
        if (in.wmask & 2u) d.y = res[px].y;
        if (in.wmask & 4u) d.z = res[px].z;
        if (in.wmask & 8u) d.w = res[px].w;
      }
    }
    pc = next;
  }
}

// Colour target 0's value for each pixel of the quad after sasl_exec (the reference leaves a
// target the shader does not write undefined; here it is zero).
__device__ __forceinline__ float4 sasl_color0(const sasl_program* __restrict__ p, const sasl_regs& R, int px) {
  for (uint32_t o = 0; o < p->noutputs; ++o) {
    if (p->outputs[o].target != 0) continue;
    const float4 v = R.r[p->outputs[o].reg][px];
    const int nc = p->outputs[o].ncomp;
    return make_float4(v.x, nc > 1 ? v.y : 0.0f, nc > 2 ? v.z : 0.0f, nc > 3 ? v.w : 0.0f);
  }
  return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace srx
