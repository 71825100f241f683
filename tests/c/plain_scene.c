/* The plain small scene (512x512 RGBA8 + RG32F depth/stencil, 1 sample, 2,000 non-indexed
 * Gouraud triangles from splitmix64 seed 0x5A11A001, cull none, depth LESS) rendered through the
 * C ABI of include/srx.h alone, as a C caller linked against libsrx_hip.so (or the oracle's
 * libsrx_cpu.so) would.  Writes the colour bytes then the depth/stencil bytes to argv[1].
 * Also exercises map/unmap, the state getters and a query polled with do_not_wait. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "srx.h"

#define W 512
#define H 512
#define NTRI 2000

static int fails = 0;
#define CHECK(expr)                                                                          \
  do {                                                                                       \
    int32_t rc_ = (expr);                                                                    \
    if (rc_ != SRX_OK) {                                                                     \
      fprintf(stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #expr, rc_, srx_last_error_string()); \
      exit(1);                                                                               \
    }                                                                                        \
  } while (0)
#define EXPECT(cond)                                                        \
  do {                                                                      \
    if (!(cond)) {                                                          \
      fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #cond);   \
      ++fails;                                                              \
    }                                                                       \
  } while (0)

/* splitmix64 output k (k = 1, 2, ...) of `seed`; a float in [0,1) is (u >> 40) * 2^-24 */
static double uniform01(uint64_t seed, uint64_t k) {
  uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 40) * (1.0 / 16777216.0);
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s out.bin\n", argv[0]);
    return 2;
  }
  printf("backend %s, abi %u\n", srx_backend_name(), srx_abi_version());
  srx_renderer r;
  CHECK(srx_create_renderer(SRX_RENDERER_BENCHMARK, &r));

  /* defaults of renderer_impl::renderer_impl */
  EXPECT(srx_get_index_format(r) == SRX_FMT_R16_UINT);
  EXPECT(srx_get_primitive_topology(r) == SRX_TOPOLOGY_TRIANGLE_LIST);
  srx_raster_state rs0;
  srx_raster_desc rd0;
  CHECK(srx_get_rasterizer_state(r, &rs0));
  CHECK(srx_raster_state_desc(rs0, &rd0));
  EXPECT(rd0.cull_mode == SRX_CULL_BACK && rd0.fill_mode == SRX_FILL_SOLID && rd0.front_ccw == 0);
  srx_release(rs0);

  srx_texture ct, dt;
  srx_surface color, ds;
  CHECK(srx_create_tex2d(r, W, H, 1, SRX_PF_RGBA8, &ct));
  CHECK(srx_create_tex2d(r, W, H, 1, SRX_PF_RG32F, &dt));
  CHECK(srx_texture_surface(ct, 0, &color));
  CHECK(srx_texture_surface(dt, 0, &ds));
  CHECK(srx_set_render_targets(r, 1, &color, ds));
  srx_viewport vp = {0, 0, W, H, 0.0f, 1.0f};
  CHECK(srx_set_viewport(r, &vp));

  /* vertices x y z w r g b a, uploaded through map(write_discard) */
  const size_t bytes = (size_t)NTRI * 3 * 8 * sizeof(float);
  srx_buffer vb;
  CHECK(srx_create_buffer(r, bytes, &vb));
  srx_mapped_resource m;
  CHECK(srx_map_buffer(r, &m, vb, SRX_MAP_WRITE_DISCARD));
  EXPECT(m.row_pitch == bytes && m.depth_pitch == bytes);
  EXPECT(srx_map_buffer(r, &m, vb, SRX_MAP_READ) == SRX_FAILED); /* one mapping at a time */
  float* v = (float*)m.data;
  for (uint64_t i = 0; i < (uint64_t)NTRI * 3; ++i) {
    double u[7];
    for (int k = 0; k < 7; ++k) u[k] = uniform01(0x5A11A001ull, i * 7 + k + 1);
    v[i * 8 + 0] = (float)(-1.0 + 2.0 * u[0]);
    v[i * 8 + 1] = (float)(-1.0 + 2.0 * u[1]);
    v[i * 8 + 2] = (float)(0.1 + 0.8 * u[2]);
    v[i * 8 + 3] = 1.0f;
    for (int k = 0; k < 4; ++k) v[i * 8 + 4 + k] = (float)u[3 + k];
  }
  CHECK(srx_unmap(r));
  EXPECT(srx_unmap(r) == SRX_FAILED);
  CHECK(srx_map_buffer(r, &m, vb, SRX_MAP_READ)); /* reads back what was written */
  EXPECT(((const float*)m.data)[3] == 1.0f && ((const float*)m.data)[11] == 1.0f);
  CHECK(srx_unmap(r));

  srx_input_element_desc el[2] = {{"POSITION", 0, SRX_FMT_R32G32B32A32_FLOAT, 0, 0, 0, 0},
                                  {"COLOR", 0, SRX_FMT_R32G32B32A32_FLOAT, 0, 16, 0, 0}};
  srx_input_layout layout;
  CHECK(srx_create_input_layout(r, el, 2, SRX_VS_PASSTHROUGH_COLOR, &layout));
  CHECK(srx_set_input_layout(r, layout));
  size_t stride = 32, offset = 0;
  CHECK(srx_set_vertex_buffers(r, 0, 1, &vb, &stride, &offset));
  CHECK(srx_set_primitive_topology(r, SRX_TOPOLOGY_TRIANGLE_LIST));
  CHECK(srx_set_vertex_shader(r, SRX_VS_PASSTHROUGH_COLOR, NULL, 0));
  CHECK(srx_set_pixel_shader(r, SRX_PS_GOURAUD, NULL, 0));
  CHECK(srx_set_blend_shader(r, SRX_BLEND_OFF));
  srx_raster_desc rd;
  srx_default_raster_desc(&rd);
  rd.cull_mode = SRX_CULL_NONE;
  srx_raster_state rs;
  CHECK(srx_create_raster_state(r, &rd, &rs));
  CHECK(srx_set_rasterizer_state(r, rs));
  srx_depth_stencil_desc dd;
  srx_default_depth_stencil_desc(&dd);
  srx_ds_state dss;
  CHECK(srx_create_depth_stencil_state(r, &dd, &dss));
  CHECK(srx_set_depth_stencil_state(r, dss, 0));

  EXPECT(srx_get_vertex_shader(r) == SRX_VS_PASSTHROUGH_COLOR);
  EXPECT(srx_get_pixel_shader(r) == SRX_PS_GOURAUD);
  EXPECT(srx_get_blend_shader(r) == SRX_BLEND_OFF);
  srx_raster_state got_rs;
  CHECK(srx_get_rasterizer_state(r, &got_rs));
  EXPECT(got_rs == rs);
  srx_release(got_rs);
  srx_viewport gvp;
  CHECK(srx_get_viewport(r, &gvp));
  EXPECT(gvp.w == W && gvp.h == H && gvp.maxz == 1.0f);

  srx_query q;
  CHECK(srx_create_query(r, SRX_QUERY_PIPELINE_STATISTICS, &q));
  srx_pipeline_statistics st;
  EXPECT(srx_get_data(r, q, &st, sizeof st, 1) == SRX_ASYNC_ERROR); /* never ended */
  const float black[4] = {0, 0, 0, 0};
  CHECK(srx_clear_color(r, color, black));
  CHECK(srx_clear_depth_stencil(r, ds, SRX_CLEAR_DEPTH | SRX_CLEAR_STENCIL, 1.0f, 0));
  CHECK(srx_begin(r, q));
  EXPECT(srx_get_data(r, q, &st, sizeof st, 1) == SRX_ASYNC_ERROR); /* running */
  CHECK(srx_draw(r, 0, NTRI));
  CHECK(srx_end(r, q));
  int polls = 0;
  int32_t as;
  while ((as = srx_get_data(r, q, &st, sizeof st, 1)) == SRX_ASYNC_TIMEOUT) ++polls;
  EXPECT(as == SRX_ASYNC_READY);
  EXPECT(st.ia_primitives == NTRI && st.ia_vertices == 3 * NTRI && st.cinvocations == NTRI);
  EXPECT(st.ps_invocations > 0 && st.ps_invocations % 4 == 0);
  printf("pipeline statistics after %d polls: ia_primitives %llu cprimitives %llu ps_invocations %llu\n", polls,
         (unsigned long long)st.ia_primitives, (unsigned long long)st.cprimitives,
         (unsigned long long)st.ps_invocations);
  CHECK(srx_flush(r));

  /* read the targets through map(read); the pitches follow the linear layout */
  FILE* f = fopen(argv[1], "wb");
  if (!f) return 1;
  CHECK(srx_map_surface(r, &m, color, SRX_MAP_READ));
  EXPECT(m.row_pitch == W * 4 && m.depth_pitch == W * 4 * H);
  fwrite(m.data, 1, srx_surface_size_bytes(color), f);
  CHECK(srx_unmap(r));
  CHECK(srx_map_surface(r, &m, ds, SRX_MAP_READ));
  EXPECT(m.row_pitch == W * 8);
  fwrite(m.data, 1, srx_surface_size_bytes(ds), f);
  CHECK(srx_unmap(r));
  fclose(f);

  /* a read_write mapping persists its writes, a read mapping does not */
  CHECK(srx_map_surface(r, &m, color, SRX_MAP_READ_WRITE));
  unsigned char first = ((unsigned char*)m.data)[0];
  ((unsigned char*)m.data)[0] = (unsigned char)(first ^ 0xFF);
  CHECK(srx_unmap(r));
  CHECK(srx_map_surface(r, &m, color, SRX_MAP_READ));
  EXPECT(((unsigned char*)m.data)[0] == (unsigned char)(first ^ 0xFF));
  ((unsigned char*)m.data)[0] = first;
  CHECK(srx_unmap(r));
  unsigned char back[4];
  CHECK(srx_map_surface(r, &m, color, SRX_MAP_WRITE)); /* write keeps the bytes it does not touch */
  memcpy(back, m.data, 4);
  CHECK(srx_unmap(r));
  EXPECT(back[0] == (unsigned char)(first ^ 0xFF));

  srx_release(q);
  srx_release(dss);
  srx_release(rs);
  srx_release(layout);
  srx_release(vb);
  srx_release(color);
  srx_release(ds);
  srx_release(ct);
  srx_release(dt);
  srx_release(r);
  if (fails) {
    fprintf(stderr, "%d checks failed\n", fails);
    return 1;
  }
  printf("ok\n");
  return 0;
}
