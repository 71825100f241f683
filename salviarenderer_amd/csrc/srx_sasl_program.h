// Compiled form of a SASL pixel shader (renderer::set_pixel_shader_code, renderer.h:86): a register
// program that runs the four pixels of a 2x2 quad together, as the reference's SIMD code generator
// runs its 4-pixel package (sasl/src/codegen/cg_simd.cpp, cgs_simd.cpp), so ddx/ddy and tex2D see
// the whole quad.  Plain data shared by the compiler (srx_sasl.cpp), the HIP interpreter
// (srx_sasl_exec.h) and the CPU oracle.
//
// Execution model, restating cgs_simd.cpp:
//  * Every instruction computes all four pixels.  Writes to temporaries are unmasked (the
//    reference's SSA values); writes to variables and outputs (wmask bit SASL_WM_MASKED) keep the
//    old value in pixels outside the execution mask (cgs_simd::store, cgs_simd.cpp:85-146).
//  * if: then-mask = exec & cond, else-mask = exec' & ~cond, where exec' is the enclosing mask
//    with the innermost loop's broken pixels removed (then_beg/else_beg/then_end,
//    cgs_simd.cpp:246-265).
//  * loops keep a mask variable: the condition sets exec = mask_var & cond and clears the break
//    mask; the body's end stores exec into mask_var; the loop exits when exec is empty
//    (enter_loop/apply_loop_condition/save_next_iteration_exec_mask, cgs_simd.cpp:414-476); after
//    the loop the enclosing mask is the one the loop was entered with.
//  * break: break mask |= exec, exec &= ~break mask (break_/apply_break, cgs_simd.cpp:352-377).
//  * ddx pairs the columns of each row, ddy the rows of each column (derivation,
//    cgs_simd.cpp:267-312); tex2D(s, c) is tex2Dgrad(s, c, ddx(c), ddy(c)) (cg_impl.cpp:902-909),
//    sampled only for pixels in the execution mask (tex2Dgrad_ps, sampler_api.cpp:13-18).
#pragma once

#ifndef __HIPCC_RTC__
#include <cstdint>
#endif

namespace srx {

constexpr int SASL_MAX_INST = 256;
constexpr int SASL_MAX_REGS = 32;     // inputs, variables and temporaries
constexpr int SASL_MAX_CONSTS = 32;   // literals and uniform variables
constexpr int SASL_MAX_INPUTS = 8;
constexpr int SASL_MAX_OUTPUTS = 8;
constexpr int SASL_MAX_SAMPLERS = 4;
constexpr int SASL_MAX_DEPTH = 8;     // nesting of if and loops
// Loop iterations (taken backward jumps) one quad may run; a quad that reaches the limit stops with
// its registers as they are (a loop that never ends in the reference never returns).
constexpr uint32_t SASL_STEP_LIMIT = 1u << 16;

constexpr uint8_t SASL_CONST = 32;    // operand >= 32: constant operand - 32
constexpr uint8_t SASL_WM_MASKED = 16;

enum sasl_op : uint8_t {
  SOP_NOP = 0,
  SOP_MOV,    // d = a
  SOP_ADD,    // d = a + b
  SOP_SUB,
  SOP_MUL,
  SOP_DIV,
  SOP_NEG,    // d = -a
  SOP_SLT,    // d = a < b ? a : b   (min, cg_impl.cpp:1115-1126)
  SOP_SGT,    // d = a > b ? a : b   (max)
  SOP_SGE,    // d = a >= b ? a : b  (clamp's lower bound, cgs.cpp:1860-1864)
  SOP_SLE,    // d = a <= b ? a : b  (clamp's upper bound)
  SOP_CLT,    // d = a < b ? 1 : 0
  SOP_CGT,
  SOP_CLE,
  SOP_CGE,
  SOP_CEQ,
  SOP_CNE,
  SOP_DOT,    // d = ((0 + a0*b0) + a1*b1) + ... over imm components (emit_dot_vv, cgs.cpp:1027-1040)
  SOP_CROSS,  // d = a.yzx * b.zxy - a.zxy * b.yzx (emit_cross, cgs.cpp:755-768)
  SOP_ABS,
  SOP_SQRT,
  SOP_FN1,    // d = f(a), f = imm (sasl_fn), per component
  SOP_FN2,    // d = f(a, b)
  SOP_DDX,
  SOP_DDY,
  SOP_TEX,    // d = sample_2d_grad(sampler imm, a.xy, b.xy, c.xy, 0) in executing pixels, else 0
  SOP_IF,     // push; cond = a.x != 0; exec &= cond; if exec == 0 jump to target (its ELSE/ENDIF)
  SOP_ELSE,   // exec' = saved & ~break; exec = exec' & ~cond; if exec == 0 jump to target (ENDIF)
  SOP_ENDIF,  // exec = saved & ~break; pop
  SOP_LOOP,   // push: saved = mask_var = exec, break = 0
  SOP_LTEST,  // exec = mask_var & (a.x != 0 unless imm); break = 0; if exec == 0 jump to target (LEND)
  SOP_DTEST,  // exec = mask_var & (a.x != 0); break = 0; mask_var = exec; if exec != 0 jump to target
  SOP_DBODY,  // exec = mask_var (start of a do-while body)
  SOP_LSAVE,  // mask_var = exec
  SOP_JMP,    // jump to target
  SOP_LEND,   // exec = saved; pop
  SOP_BREAK,  // break |= exec; exec = 0
  SOP_COUNT
};

// Scalar functions of SOP_FN1/SOP_FN2: the reference's externals (compiler_impl.cpp:340-430).
enum sasl_fn : uint8_t {
  SFN_EXP = 0,   // expf
  SFN_EXP2,      // ldexpf(1, (int)v)
  SFN_LOG,       // eflib::fast_log
  SFN_LOG2,      // eflib::fast_log2
  SFN_LOG10,     // log10f
  SFN_SIN, SFN_COS, SFN_TAN, SFN_ASIN, SFN_ACOS, SFN_ATAN, SFN_SINH, SFN_COSH, SFN_TANH,
  SFN_CEIL,      // eflib::fast_ceil
  SFN_FLOOR,     // eflib::fast_floor
  SFN_ROUND,     // eflib::fast_round
  SFN_TRUNC,     // eflib::trunc: v > 0 ? fast_floor(v) : fast_ceil(v)
  SFN_RSQRT,     // 1 / sqrtf(v)
  SFN_POW,       // powf(a, b)
  SFN_FMOD,      // fmodf(a, b)
  SFN_LDEXP,     // ldexpf(a, (int)b)
  SFN_COUNT
};

// One instruction: for each destination component k in wmask, the operands' components are
// a[(sa >> 2k) & 3], b[(sb >> 2k) & 3], c[(sc >> 2k) & 3].
struct sasl_inst {
  uint8_t op, dst, wmask, imm;
  uint8_t a, b, c, sa;
  uint8_t sb, sc;
  uint16_t target;
};

// Input register i holds vertex-shader output attribute `attr` (interpolated and divided by w as the
// quad's attributes are), its first ncomp components and zeros after (pixel_shader_unit::update,
// shader_unit.cpp:106-145).
struct sasl_input {
  uint8_t reg, attr, ncomp, pad;
};
// Output: register reg's first ncomp components go to colour target `target` (COLOR(target)).
struct sasl_output {
  uint8_t reg, target, ncomp, pad;
};

struct sasl_program {
  uint32_t ninst, ninputs, noutputs, nregs;
  sasl_input inputs[SASL_MAX_INPUTS];
  sasl_output outputs[SASL_MAX_OUTPUTS];
  float consts[SASL_MAX_CONSTS][4];
  sasl_inst code[SASL_MAX_INST];
};

}  // namespace srx
