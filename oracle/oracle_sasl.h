// ORACLE -- test infrastructure only. Nothing in the product links, includes or calls this.
//
// CPU execution of a compiled SASL pixel shader (srx_sasl_program.h) on one 2x2 quad, restating
// the reference's 4-pixel SIMD code (sasl/src/codegen/cgs_simd.cpp) pixel by pixel: each
// instruction computes its result for the four pixels before writing any of them; temporaries are
// written in all four, variables only in the pixels of the execution mask (cgs_simd::store,
// cgs_simd.cpp:85-146); if/loop masks follow then_beg/else_beg/apply_break (:246-265, :340-377) and
// enter_loop/apply_loop_condition/save_next_iteration_exec_mask (:414-476).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#include "../salviarenderer_amd/csrc/srx_sasl_program.h"
#include "oracle_math.h"

namespace srxo {

// math.h:107-109, 126-129: fast_log = fast_log2 * ln 2; fast_ceil from fast_round
inline float sasl_fast_log(float v) { return fast_log2(v) * 0.69314718f; }
inline float sasl_fast_ceil(float v) {
  float f = fast_round(v);
  return (f < v) ? f + 1 : f;
}

// The scalar externals the reference injects (sasl/src/drivers/compiler_impl.cpp:340-416)
inline float sasl_fn1(int fn, float v) {
  using namespace srx;
  switch (fn) {
  case SFN_EXP: return expf(v);
  case SFN_EXP2: return ldexpf(1.0f, cvtt(v));
  case SFN_LOG: return sasl_fast_log(v);
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
  case SFN_TRUNC: return v > 0.0f ? fast_floor(v) : sasl_fast_ceil(v);  // eflib::trunc, math.h:56-60
  case SFN_RSQRT: return 1.0f / std::sqrt(v);
  }
  return 0.0f;
}
inline float sasl_fn2(int fn, float a, float b) {
  using namespace srx;
  switch (fn) {
  case SFN_POW: return powf(a, b);
  case SFN_FMOD: return fmodf(a, b);
  case SFN_LDEXP: return ldexpf(a, cvtt(b));
  }
  return 0.0f;
}

struct sasl_frame {
  bool loop;
  unsigned saved, cond, mv, brk;
};

// Runs `p` on a quad.  in[px][i] is input i of pixel px (pixels 0 1 / 2 3); out[px][o] receives
// output o.  tex(slot, px, u, v, dxu, dxv, dyu, dyv, res) samples for one pixel.
template <class TEX>
void sasl_run_quad(const srx::sasl_program& p, const float (*in)[srx::SASL_MAX_INPUTS][4],
                   float (*out)[srx::SASL_MAX_OUTPUTS][4], TEX&& tex) {
  using namespace srx;
  float R[SASL_MAX_REGS][4][4];  // register, pixel, component
  std::memset(R, 0, sizeof R);
  for (uint32_t i = 0; i < p.ninputs; ++i)
    for (int px = 0; px < 4; ++px)
      for (int k = 0; k < 4; ++k) R[p.inputs[i].reg][px][k] = k < p.inputs[i].ncomp ? in[px][i][k] : 0.0f;

  auto read = [&](uint8_t o, uint8_t swz, int px, float* v) {
    for (int k = 0; k < 4; ++k) {
      const int c = (swz >> (2 * k)) & 3;
      v[k] = o >= SASL_CONST ? p.consts[o - SASL_CONST][c] : R[o][px][c];
    }
  };
  sasl_frame st[SASL_MAX_DEPTH];
  int depth = 0;
  unsigned exec = 15;
  auto loop_brk = [&]() -> unsigned {
    for (int d = depth - 1; d >= 0; --d)
      if (st[d].loop) return st[d].brk;
    return 0u;
  };
  auto cond_bits = [&](const sasl_inst& in) {
    unsigned m = 0;
    for (int px = 0; px < 4; ++px) {
      float v[4];
      read(in.a, in.sa, px, v);
      if (v[0] != 0.0f) m |= 1u << px;
    }
    return m;
  };

  uint32_t steps = 0;  // loop iterations: taken backward jumps
  uint32_t pc = 0;
  bool stopped = false;
  while (pc < p.ninst && !stopped) {
    const sasl_inst& in = p.code[pc];
    uint32_t next = pc + 1;
    float res[4][4] = {};
    bool writes = true;
    switch (in.op) {
    case SOP_IF: {
      const unsigned c = cond_bits(in);
      st[depth++] = sasl_frame{false, exec, c, 0, 0};
      exec &= c;
      if (!exec) next = in.target;
      writes = false;
      break;
    }
    case SOP_ELSE: {
      sasl_frame& f = st[depth - 1];
      f.saved &= ~loop_brk();
      exec = f.saved & ~f.cond;
      if (!exec) next = in.target;
      writes = false;
      break;
    }
    case SOP_ENDIF:
      --depth;
      exec = st[depth].saved & ~loop_brk();
      writes = false;
      break;
    case SOP_LOOP:
      st[depth++] = sasl_frame{true, exec, 0, exec, 0};
      writes = false;
      break;
    case SOP_LTEST: {
      sasl_frame& f = st[depth - 1];
      exec = f.mv & (in.imm ? 15u : cond_bits(in));
      f.brk = 0;
      if (!exec) next = in.target;
      writes = false;
      break;
    }
    case SOP_DTEST: {
      sasl_frame& f = st[depth - 1];
      exec = f.mv & cond_bits(in);
      f.brk = 0;
      f.mv = exec;
      if (exec) {
        if (++steps > SASL_STEP_LIMIT) stopped = true;
        next = in.target;
      }
      writes = false;
      break;
    }
    case SOP_DBODY:
      exec = st[depth - 1].mv;
      writes = false;
      break;
    case SOP_LSAVE:
      st[depth - 1].mv = exec;
      writes = false;
      break;
    case SOP_JMP:
      if (++steps > SASL_STEP_LIMIT) stopped = true;
      next = in.target;
      writes = false;
      break;
    case SOP_LEND:
      --depth;
      exec = st[depth].saved;
      writes = false;
      break;
    case SOP_BREAK:
      for (int d = depth - 1; d >= 0; --d)
        if (st[d].loop) {
          st[d].brk |= exec;
          break;
        }
      exec = 0;
      writes = false;
      break;
    case SOP_NOP: writes = false; break;
    case SOP_DDX:
    case SOP_DDY: {
      float v[4][4];
      for (int px = 0; px < 4; ++px) read(in.a, in.sa, px, v[px]);
      for (int px = 0; px < 4; ++px) {
        // ddx: p1 - p0 for the top row, p3 - p2 for the bottom; ddy: p2 - p0, p3 - p1 per column
        const int lo = in.op == SOP_DDX ? (px & 2) : (px & 1);
        const int hi = in.op == SOP_DDX ? lo + 1 : lo + 2;
        for (int k = 0; k < 4; ++k) res[px][k] = v[hi][k] - v[lo][k];
      }
      break;
    }
    case SOP_TEX:
      for (int px = 0; px < 4; ++px) {
        if (!(exec >> px & 1)) continue;  // tex2Dgrad_ps samples only executing pixels
        float c[4], dx[4], dy[4];
        read(in.a, in.sa, px, c);
        read(in.b, in.sb, px, dx);
        read(in.c, in.sc, px, dy);
        tex(in.imm, px, c[0], c[1], dx[0], dx[1], dy[0], dy[1], res[px]);
      }
      break;
    default:
      for (int px = 0; px < 4; ++px) {
        float a[4], b[4], c[4];
        read(in.a, in.sa, px, a);
        read(in.b, in.sb, px, b);
        read(in.c, in.sc, px, c);
        float* r = res[px];
        switch (in.op) {
        case SOP_MOV: for (int k = 0; k < 4; ++k) r[k] = a[k]; break;
        case SOP_ADD: for (int k = 0; k < 4; ++k) r[k] = a[k] + b[k]; break;
        case SOP_SUB: for (int k = 0; k < 4; ++k) r[k] = a[k] - b[k]; break;
        case SOP_MUL: for (int k = 0; k < 4; ++k) r[k] = a[k] * b[k]; break;
        case SOP_DIV: for (int k = 0; k < 4; ++k) r[k] = a[k] / b[k]; break;
        case SOP_NEG: for (int k = 0; k < 4; ++k) r[k] = -a[k]; break;
        case SOP_SLT: for (int k = 0; k < 4; ++k) r[k] = a[k] < b[k] ? a[k] : b[k]; break;
        case SOP_SGT: for (int k = 0; k < 4; ++k) r[k] = a[k] > b[k] ? a[k] : b[k]; break;
        case SOP_SGE: for (int k = 0; k < 4; ++k) r[k] = a[k] >= b[k] ? a[k] : b[k]; break;
        case SOP_SLE: for (int k = 0; k < 4; ++k) r[k] = a[k] <= b[k] ? a[k] : b[k]; break;
        case SOP_CLT: for (int k = 0; k < 4; ++k) r[k] = a[k] < b[k] ? 1.0f : 0.0f; break;
        case SOP_CGT: for (int k = 0; k < 4; ++k) r[k] = a[k] > b[k] ? 1.0f : 0.0f; break;
        case SOP_CLE: for (int k = 0; k < 4; ++k) r[k] = a[k] <= b[k] ? 1.0f : 0.0f; break;
        case SOP_CGE: for (int k = 0; k < 4; ++k) r[k] = a[k] >= b[k] ? 1.0f : 0.0f; break;
        case SOP_CEQ: for (int k = 0; k < 4; ++k) r[k] = a[k] == b[k] ? 1.0f : 0.0f; break;
        case SOP_CNE: for (int k = 0; k < 4; ++k) r[k] = a[k] != b[k] ? 1.0f : 0.0f; break;
        case SOP_DOT: {
          float s = 0.0f;
          for (int k = 0; k < in.imm; ++k) {
            const float prod = a[k] * b[k];
            s = s + prod;
          }
          for (int k = 0; k < 4; ++k) r[k] = s;
          break;
        }
        case SOP_CROSS:
          r[0] = a[1] * b[2] - a[2] * b[1];
          r[1] = a[2] * b[0] - a[0] * b[2];
          r[2] = a[0] * b[1] - a[1] * b[0];
          r[3] = 0.0f;
          break;
        case SOP_ABS: for (int k = 0; k < 4; ++k) r[k] = std::fabs(a[k]); break;
        case SOP_SQRT: for (int k = 0; k < 4; ++k) r[k] = std::sqrt(a[k]); break;
        case SOP_FN1: for (int k = 0; k < 4; ++k) r[k] = sasl_fn1(in.imm, a[k]); break;
        case SOP_FN2: for (int k = 0; k < 4; ++k) r[k] = sasl_fn2(in.imm, a[k], b[k]); break;
        default: break;
        }
      }
      break;
    }
    if (writes) {
      const bool masked = (in.wmask & SASL_WM_MASKED) != 0;
      for (int px = 0; px < 4; ++px) {
        if (masked && !(exec >> px & 1)) continue;
        for (int k = 0; k < 4; ++k)
          if (in.wmask >> k & 1) R[in.dst][px][k] = res[px][k];
      }
    }
    pc = next;
  }
  for (uint32_t o = 0; o < p.noutputs; ++o)
    for (int px = 0; px < 4; ++px)
      for (int k = 0; k < 4; ++k) out[px][o][k] = k < p.outputs[o].ncomp ? R[p.outputs[o].reg][px][k] : 0.0f;
}

}  // namespace srxo
