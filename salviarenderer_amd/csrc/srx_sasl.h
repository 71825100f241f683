// SASL pixel-shader compiler: source text -> sasl_program (srx_sasl_program.h).
//
// The reference compiles SASL through its own front end and LLVM into a 4-pixel SIMD function
// (salvia/src/core/renderer.cpp compile(), sasl/src/codegen/cg_simd.cpp).  This compiler accepts
// the pixel-shader subset that code generator implements and emits the register program that the
// HIP interpreter and the CPU oracle execute:
//   * struct definitions, global uniforms (float/int scalars and vectors) and samplers, one entry
//     function whose parameters and result carry semantics: inputs TEXCOORD(n) only, outputs
//     COLOR(n)/SV_Target(n) only (psinput/psoutput_semantic_usage, reflector.cpp:43-57);
//   * local declarations, assignment (=, +=, -=, *=, /=) to variables, members and swizzles,
//     if/else, while, do-while, for, break, one return as the entry's last statement;
//   * + - * / % on floats and + - * on ints, comparisons, unary minus, member access and
//     swizzles, vector constructors, and the intrinsics listed in sasl_compile_ps.
// Constructs the reference's SIMD generator leaves unimplemented (&&, ||, ?:, casts, continue,
// tex2Dlod/tex2Dbias/tex2Dproj, cg_simd.cpp:62-66,522-541, sampler_api.cpp:20-50) and ones not
// built here (matrices, helper functions, arrays) are refused with SASL_UNSUPPORTED.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "srx_sasl_program.h"

namespace srx {

enum sasl_status { SASL_OK = 0, SASL_ERROR = 1, SASL_UNSUPPORTED = 2 };

struct sasl_uniform {
  std::string name;
  uint8_t cslot, ncomp, is_int, pad;
};

struct sasl_module {
  sasl_program prog;                   // uniform constant slots hold zeros
  std::vector<sasl_uniform> uniforms;  // global variables (shader_reflection buffer_in layouts)
  std::vector<std::string> samplers;   // sampler slot -> global sampler name
};

// Compiles a pixel shader.  Returns SASL_OK, or SASL_ERROR / SASL_UNSUPPORTED with a message
// naming the line.
int sasl_compile_ps(const char* src, size_t len, sasl_module& out, std::string& msg);

// pixel_shader_unit::set_variable (shader_unit.cpp:97-100): copies a uniform's bytes (at most the
// variable's size; ints converted to the float registers) into its constant slot of `prog`.
// Returns false when the module has no such variable.
bool sasl_set_uniform(const sasl_module& m, sasl_program& prog, const std::string& name, const void* data, size_t size);

}  // namespace srx
