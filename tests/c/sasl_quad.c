/* A C program compiled against include/srx.h: compiles a SASL pixel shader (compile(),
   set_ps_variable), runs it on one 2x2 quad (pixel_shader_unit::update + execute) and checks the
   outputs against values worked out by hand; also checks the error codes of a syntax error and of
   a construct the reference's SIMD code generator leaves unimplemented. */
#include <stdio.h>
#include <string.h>

#include "srx.h"

static const char* kShader =
    "struct In { float4 v : TEXCOORD(0); float2 w : TEXCOORD(1); };\n"
    "float4 tint;\n"
    "float4 quad_main(In i) : COLOR {\n"
    "  float4 c = i.v * tint + float4(ddx(i.w.x), ddy(i.w.y), 0.0, 1.0);\n"
    "  if (c.x > 2.0) { c.y = c.y + 10.0; }\n"
    "  return c;\n"
    "}\n";

int main(void) {
  srx_renderer r;
  srx_shader_code code, bad;
  srx_shader_info info;
  float tint[4] = {1.0f, 2.0f, 3.0f, 4.0f};
  float in[4][2][4], out[4][1][4];
  /* ddx of w.x = p*p pairs the row's pixels, ddy of w.y = 3p the column's */
  const float ddx[4] = {1.0f, 1.0f, 5.0f, 5.0f}, ddy[4] = {6.0f, 6.0f, 6.0f, 6.0f};
  int p, k, errors = 0;
  if (srx_create_renderer(SRX_RENDERER_BENCHMARK, &r) != SRX_OK) return 1;
  if (srx_compile_shader(kShader, strlen(kShader), SRX_LANG_PIXEL_SHADER, &code) != SRX_OK) {
    fprintf(stderr, "compile: %s\n", srx_last_error_string());
    return 2;
  }
  if (srx_shader_reflect(code, &info) != SRX_OK || info.input_count != 2 || info.output_count != 1 ||
      info.input_components[0] != 4 || info.input_components[1] != 2 || info.output_targets[0] != 0 ||
      info.uniform_count != 1) {
    fprintf(stderr, "reflection\n");
    return 3;
  }
  if (srx_set_ps_variable(r, "tint", tint, sizeof tint) != SRX_OK) return 4;
  for (p = 0; p < 4; ++p) {
    for (k = 0; k < 4; ++k) in[p][0][k] = (float)(p + k);
    in[p][1][0] = (float)(p * p);
    in[p][1][1] = (float)(3 * p);
    in[p][1][2] = in[p][1][3] = 0.0f;
  }
  if (srx_ps_execute_quads(r, code, &in[0][0][0], &out[0][0][0], 1) != SRX_OK) {
    fprintf(stderr, "execute: %s\n", srx_last_error_string());
    return 5;
  }
  for (p = 0; p < 4; ++p) {
    float want[4];
    want[0] = in[p][0][0] * tint[0] + ddx[p];
    want[1] = in[p][0][1] * tint[1] + ddy[p];
    want[2] = in[p][0][2] * tint[2] + 0.0f;
    want[3] = in[p][0][3] * tint[3] + 1.0f;
    if (want[0] > 2.0f) want[1] = want[1] + 10.0f;
    for (k = 0; k < 4; ++k)
      if (out[p][0][k] != want[k]) {
        fprintf(stderr, "pixel %d component %d: %g, want %g\n", p, k, out[p][0][k], want[k]);
        ++errors;
      }
    printf("pixel %d: %g %g %g %g\n", p, out[p][0][0], out[p][0][1], out[p][0][2], out[p][0][3]);
  }
  if (srx_compile_shader("float4 f(float4 a : TEXCOORD0) : COLOR { return a + ; }", 54, SRX_LANG_PIXEL_SHADER, &bad) !=
      SRX_INVALID_PARAMETER)
    ++errors;
  {
    const char* s = "float4 f(float a : TEXCOORD0) : COLOR { return float4(a > 0.0 ? 1.0 : 0.0, 0, 0, 0); }";
    if (srx_compile_shader(s, strlen(s), SRX_LANG_PIXEL_SHADER, &bad) != SRX_E_UNSUPPORTED) ++errors;
  }
  srx_release(code);
  srx_release(r);
  return errors ? 6 : 0;
}
