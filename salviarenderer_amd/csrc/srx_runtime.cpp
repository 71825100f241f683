// Host runtime of the HIP backend: the srx C ABI (include/srx.h) over device resources.
//
// One HIP stream per renderer carries every command in submission order (the reference's
// async renderer queue, async_renderer.cpp:17-110); srx_flush() waits for it.  Resources live
// in HBM in the reference's linear layout ((y*w + x)*samples + s)*elem_size (surface.cpp:277-295),
// so map/read-back bytes are identical.  State validation follows renderer_impl.cpp.
#include <cstring>

#include "srx_runtime_internal.h"

namespace srx_rt {

// A swap chain without a window (swap_chain_impl.cpp): back buffer and presented surface.
struct swap_obj : object {
  ref<renderer_obj> r;
  ref<texture_obj> back, resolved;
  swap_obj() : object(K_SWAP_CHAIN) {}
};


// The device descriptor of a sampler and its texture (sampler.cpp's state); queues pending clears
// of the texture's levels.
int32_t fill_sampler(renderer_obj* r, sampler_obj* so, sampler_desc_dev& sd) {
  texture_obj* t = so->tex.get();
  sd.min_filter = so->desc.min_filter;
  sd.mag_filter = so->desc.mag_filter;
  sd.mip_filter = so->desc.mip_filter;
  sd.mip_qual = so->desc.mip_qual;
  sd.addr_u = so->desc.addr_mode_u;
  sd.addr_v = so->desc.addr_mode_v;
  sd.lod_bias = so->desc.mip_lod_bias;
  sd.max_aniso = so->desc.max_anisotropy;
  for (int k = 0; k < 4; ++k) sd.border[k] = so->desc.border_color[k];
  sd.nlevels = static_cast<int>(std::min<size_t>(t->surfs.size(), MAX_LEVELS));
  sd.max_lod = t->max_lod;
  sd.min_lod = std::min(t->min_lod, sd.nlevels - 1);
  sd.size[0] = t->size[0];
  sd.size[1] = t->size[1];
  sd.size[2] = t->size[2];
  for (int l = 0; l < sd.nlevels; ++l) {
    RCHK(run_pending_clear(t->surfs[l].get(), r->stream));
    sd.levels[l] = t->surfs[l]->desc();
  }
  bool rgba8 = sd.nlevels > 0;
  for (int l = 0; l < sd.nlevels; ++l) rgba8 &= sd.levels[l].pf == SRX_PF_RGBA8 && sd.levels[l].samples == 1;
  sd.fast_aniso = rgba8 && sd.mip_filter == SRX_FILTER_ANISOTROPIC && sd.max_aniso > 1 &&
                  sd.min_filter == SRX_FILTER_LINEAR && sd.mag_filter == SRX_FILTER_LINEAR &&
                  sd.addr_u == SRX_ADDRESS_WRAP && sd.addr_v == SRX_ADDRESS_WRAP;
  return SRX_OK;
}

// The draw's parameters and sampler table from the renderer's state (renderer_impl.cpp:159-238
// validation); skip = nothing to draw into.
int32_t build_draw(renderer_obj* r, bool indexed, uint32_t start, uint32_t primcnt, int32_t basevert, draw_params& P,
                   ps_samplers& SM, bool& skip) {
  render_state& st = r->st;
  skip = st.color_targets.empty() && !st.ds_target;  // render_core.cpp:44-46
  if (skip) return SRX_OK;
  if (st.topo != SRX_TOPOLOGY_TRIANGLE_LIST && st.topo != SRX_TOPOLOGY_TRIANGLE_STRIP)
    return fail(SRX_E_UNSUPPORTED, "only triangle lists and strips are rasterised (lines are unimplemented upstream)");
  if (st.raster.fill_mode != SRX_FILL_SOLID) return fail(SRX_E_UNSUPPORTED, "wireframe fill is unimplemented upstream");
  int nout = vs_output_count(st.vs_id);
  if (nout <= 0) return fail(SRX_INVALID_PARAMETER, "no vertex shader bound");
  if (!st.layout) return fail(SRX_INVALID_PARAMETER, "no input layout");
  if (indexed && !st.ib) return fail(SRX_INVALID_PARAMETER, "no index buffer");
  int samples = static_cast<int>(st.target_samples);
  if (samples != 1 && samples != 2 && samples != 4) return fail(SRX_E_UNSUPPORTED, "sample count must be 1, 2 or 4");

  std::memset(&P, 0, sizeof P);
  P.vs_id = st.vs_id;
  P.nattr = nout;
  for (int i = 0; i < MAX_ATTRS; ++i) P.attr_mod[i] = SRX_AM_LINEAR;
  std::memcpy(P.vs_consts, st.vs_consts, sizeof P.vs_consts);

  // stream assembler register map (stream_assembler.cpp:57-93)
  vs_binding binds[8];
  int nb = vs_bindings(st.vs_id, binds);
  uint64_t count_max = 0;
  for (int i = 0; i < nb; ++i) {
    const layout_elem* le = nullptr;
    for (auto& e : st.layout->elems)
      if (e.sem == binds[i].sem && e.index == binds[i].index) le = &e;
    if (!le) return fail(SRX_INVALID_PARAMETER, "input layout lacks a semantic the vertex shader reads");
    if (le->slot >= SRX_MAX_INPUT_SLOTS || !st.vbs[le->slot].buf) return fail(SRX_INVALID_PARAMETER, "vertex buffer slot unbound");
    const vb_slot& vb = st.vbs[le->slot];
    uint32_t need = fmt_bytes(le->fmt);
    if (!need) return fail(SRX_E_UNSUPPORTED, "vertex element format");
    if (vb.stride == 0) return fail(SRX_E_UNSUPPORTED, "zero vertex stride");
    size_t off = vb.offset + le->offset;
    input_elem_dev& d = P.in[P.nin++];
    d.base = vb.buf->d + std::min(off, vb.buf->size);
    d.avail = off < vb.buf->size ? vb.buf->size - off : 0;
    d.stride = static_cast<uint32_t>(vb.stride);
    d.fmt = le->fmt;
    d.reg = binds[i].reg;
    d.default_w = (std::string(binds[i].sem) == "position") ? 1.0f : 0.0f;
    uint64_t c = d.avail >= need ? (d.avail - need) / d.stride + 1 : 0;
    count_max = std::max(count_max, c);
  }
  uint64_t needed = count_max;
  if (!indexed) needed = std::min<uint64_t>(count_max, static_cast<uint64_t>(primcnt) * 3 + 2);
  if (needed > 0x7FFFFFFFull) return fail(SRX_E_UNSUPPORTED, "vertex count");
  P.nverts = static_cast<uint32_t>(needed);

  P.ib = indexed ? st.ib->d : nullptr;
  P.ib_size = indexed ? st.ib->size : 0;
  P.istride = st.index_fmt == SRX_FMT_R16_UINT ? 2 : 4;
  P.istart_bytes = start * P.istride;
  P.indexed = indexed ? 1 : 0;
  P.topo = st.topo;
  P.basevert = basevert;
  P.prim_count = primcnt;

  P.cull_mode = st.raster.cull_mode;
  P.front_ccw = st.raster.front_ccw;
  P.vp[0] = st.vp.x; P.vp[1] = st.vp.y; P.vp[2] = st.vp.w; P.vp[3] = st.vp.h; P.vp[4] = st.vp.minz; P.vp[5] = st.vp.maxz;
  P.target_w = st.target_vp.w;
  P.target_h = st.target_vp.h;
  P.samples = samples;
  static const float p1[1][2] = {{0.5f, 0.5f}};
  static const float p2[2][2] = {{0.25f, 0.25f}, {0.75f, 0.75f}};
  static const float p4[4][2] = {{0.375f, 0.125f}, {0.875f, 0.375f}, {0.125f, 0.625f}, {0.625f, 0.875f}};
  const float(*pat)[2] = samples == 1 ? p1 : samples == 2 ? p2 : p4;
  for (int s = 0; s < samples; ++s) {
    P.sp[s][0] = pat[s][0];
    P.sp[s][1] = pat[s][1];
  }
  // rasterizer.cpp:1106-1108
  P.tiles_x = static_cast<int32_t>(static_cast<size_t>(st.vp.w + 64 - 1) / 64);
  P.tiles_y = static_cast<int32_t>(static_cast<size_t>(st.vp.h + 64 - 1) / 64);
  P.gt = gpu_tile_size();
  float cw = std::min(st.target_vp.w, static_cast<float>(P.tiles_x * 64));
  float ch = std::min(st.target_vp.h, static_cast<float>(P.tiles_y * 64));
  P.gtiles_x = static_cast<int32_t>(std::ceil(std::max(cw, 0.0f) / P.gt));
  P.gtiles_y = static_cast<int32_t>(std::ceil(std::max(ch, 0.0f) / P.gt));

  surface_obj* ct = nullptr;
  for (auto& c : st.color_targets)
    if (c) {
      ct = c.get();
      break;
    }
  if (!st.color_targets.empty()) ct = st.color_targets[0].get();
  if (ct) P.color = ct->desc();
  surface_obj* ds = st.ds_target.get();
  if (ds) P.ds = ds->desc();
  const srx_depth_stencil_desc& d = st.ds;
  P.depth_enable = d.depth_enable;
  P.depth_func = d.depth_func;
  P.stencil_enable = ds && d.stencil_enable;
  P.read_depth = ds && d.depth_enable && d.depth_func != SRX_CMP_NEVER && d.depth_func != SRX_CMP_ALWAYS;
  P.write_depth = ds && d.depth_enable && d.depth_write_mask && d.depth_func != SRX_CMP_NEVER;
  P.read_mask = d.stencil_read_mask & 0xFF;
  P.write_mask = d.stencil_write_mask & 0xFF;
  P.stencil_ref = P.stencil_enable ? (static_cast<uint32_t>(st.stencil_ref) & P.read_mask) : 0;
  P.st_func[0] = d.front_face.stencil_func;
  P.st_func[1] = d.back_face.stencil_func;
  P.st_pass_op[0] = d.front_face.stencil_pass_op;
  P.st_pass_op[1] = d.back_face.stencil_pass_op;
  // early-Z: stencil off and no depth-writing pixel shader (framebuffer.cpp:424), per draw
  P.early_z = ds && !d.stencil_enable;
  P.count_stats = 0;
  for (auto& q : r->active)
    if (q->id != SRX_QUERY_PIPELINE_PROFILES) P.count_stats = 1;
  P.blend_id = st.blend_id;
  P.ps_id = st.ps_id;
  std::memcpy(P.ps_consts, st.ps_consts, sizeof P.ps_consts);

  std::memset(&SM, 0, sizeof SM);
  // SASL pixel shader: its samplers are bound by name, slot k being its k-th sampler variable
  sampler_obj* bound[MAX_SAMPLERS] = {};
  if (st.ps_code) {
    const sasl_module& m = st.ps_code->mod;
    for (uint32_t i = 0; i < m.prog.ninputs; ++i)
      if (m.prog.inputs[i].attr >= nout)
        return fail(SRX_INVALID_PARAMETER, "the pixel shader reads more attributes than the vertex shader writes");
    P.ps_id = PS_SASL;
    r->sasl_cur = m.prog;
    for (auto& v : st.ps_vars) sasl_set_uniform(m, r->sasl_cur, v.first, v.second.data(), v.second.size());
    for (size_t k = 0; k < m.samplers.size() && k < MAX_SAMPLERS; ++k)
      for (auto& ns : st.ps_named_samplers)
        if (ns.first == m.samplers[k]) bound[k] = ns.second.get();
  } else {
    for (int i = 0; i < MAX_SAMPLERS; ++i) bound[i] = st.samplers[i].get();
  }
  for (int i = 0; i < MAX_SAMPLERS; ++i) {
    sampler_obj* so = bound[i];
    if (!so || !so->tex) continue;
    RCHK(fill_sampler(r, so, SM.s[i]));
    P.has_sampler[i] = 1;
  }
  // Opaque draws whose samples keep or lose a colour in one depth reduction: early-Z draws, and
  // draws whose stencil never changes (KEEP on pass; fail ops never run), whose stencil test is
  // then a fixed per-sample mask.
  P.vis_stencil = P.stencil_enable && P.st_pass_op[0] == SRX_SOP_KEEP && P.st_pass_op[1] == SRX_SOP_KEEP &&
                  P.st_func[0] == P.st_func[1];
  return SRX_OK;
}

// Device scratch of one draw: the renderer's grow-only arrays, or (fixed != null) one caller-free
// region carved front to back at 256-byte alignment, which refuses what does not fit.
struct draw_scratch {
  renderer_obj* r;
  const fixed_scratch* fixed;
  size_t used = 0;
  template <class T>
  int32_t take(dev_array<T>& a, size_t n, T*& p) {
    if (!fixed) {
      RCHK(ensure(r, a, n));
      p = a.p;
      return SRX_OK;
    }
    return carve(n, p);
  }
  // n elements of the fixed region
  template <class T>
  int32_t carve(size_t n, T*& p) {
    const size_t off = (used + 255) & ~static_cast<size_t>(255);
    if (off > fixed->bytes || n > (fixed->bytes - off) / sizeof(T))
      return fail(SRX_OUT_OF_MEMORY, "the draw's scratch exceeds the no-allocation region");
    p = reinterpret_cast<T*>(fixed->base + off);
    used = off + n * sizeof(T);
    return SRX_OK;
  }
};

static size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Bytes the record-sized buffers take for `records` binned records (keys/values twice, the sort
// histogram and its scan space, the tile order and the visibility work table), in take() order.
static size_t record_scratch_bytes(uint32_t P3, uint32_t T, int samples, bool vis_tables, uint64_t records) {
  const uint32_t n = static_cast<uint32_t>(std::min<uint64_t>(records, 0xFFFFFFFFu));
  const size_t rn = std::max<uint32_t>(n, 1);
  const uint32_t nblocks = (std::max(n, T) + 4095) / 4096;
  const size_t hist_n = std::max<size_t>(static_cast<size_t>(256) * std::max<uint32_t>(nblocks, 1), bucket_hist_size(T, n));
  size_t b = 2 * align256(rn * 4) + 2 * align256(rn * 4) + align256(hist_n * 4) +
             align256(std::max(scan_tmp_size(P3), scan_tmp_size(static_cast<uint32_t>(hist_n))) * 4) +
             align256(std::max<uint32_t>(T, 1) * 4);
  if (vis_tables) {
    uint32_t ne, nm, nsl;
    vis_table_bounds(T, n, &ne, &nm, &nsl);
    b += align256(static_cast<size_t>(ne) * 16) + align256(static_cast<size_t>(nm) * 16) + align256(16) +
         align256(static_cast<size_t>(nsl) * 256 * samples * 8);
  }
  return b + 256;
}

int32_t execute_draw(renderer_obj* r, bool indexed, uint32_t start, uint32_t primcnt, int32_t basevert,
                     const fixed_scratch* fixed) {
  draw_params P;
  ps_samplers SM;
  bool skip = false;
  RCHK(build_draw(r, indexed, start, primcnt, basevert, P, SM, skip));
  if (skip) return SRX_OK;
  if (r->recording) {
    if (fixed) return batch_refuse(r, SRX_E_UNSUPPORTED, "a batch records allocating draws only");
    for (auto& ct : r->st.color_targets) RCHK(run_pending_clear(ct.get(), r->stream));
    RCHK(run_pending_clear(r->st.ds_target.get(), r->stream));
    return record_draw(r, P, SM);
  }
  const int nout = P.nattr;
  const int samples = P.samples;
  hipStream_t s = r->stream;
  uint32_t P3 = 3 * primcnt;
  uint64_t& hs_vs = r->host_stats[ST_VS_INVOCATIONS];
  r->host_stats[ST_IA_VERTICES] += static_cast<uint64_t>(primcnt) * 3;
  r->host_stats[ST_IA_PRIMITIVES] += primcnt;
  r->host_stats[ST_CINVOCATIONS] += primcnt;
  hs_vs += P.nverts;
  if (primcnt == 0) return SRX_OK;

  draw_scratch X{r, fixed};
  draw_buffers B{};
  uint32_t T = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
  bool vis_pass = (P.early_z || P.vis_stencil) && (P.blend_id == SRX_BLEND_OFF || P.blend_id == SRX_BLEND_DISABLED);
  RCHK(X.take(r->vbuf, (static_cast<size_t>(P.nverts) + 1) * (nout + 1), B.vbuf));
  RCHK(X.take(r->tri_rec, static_cast<size_t>(P3) * 5, B.tri_rec));
  RCHK(X.take(r->tri_attr, static_cast<size_t>(P3) * 3 * (nout + 1), B.tri_attr));
  if (vis_pass)
    RCHK(X.take(r->vis, static_cast<size_t>(std::max(P.target_w, 1.0f)) * static_cast<size_t>(std::max(P.target_h, 1.0f)) * samples,
                B.vis));
  ps_samplers* samplers_d = nullptr;
  RCHK(X.take(r->samplers_d, 1, samplers_d));
  B.samplers = samplers_d;
  if (P.ps_id == PS_SASL) {
    sasl_program* prog_d = nullptr;
    if (fixed) {
      RCHK(X.carve(1, prog_d));
    } else {
      if (!r->sasl_d) HIPCHK(hipMalloc(&r->sasl_d, sizeof(sasl_program)));
      prog_d = r->sasl_d;
    }
    P.sasl = prog_d;
  }
  RCHK(X.take(r->cnt, P3, B.cnt));
  RCHK(X.take(r->clip_flag, primcnt, B.clip_flag));
  RCHK(X.take(r->clip_slot, primcnt, B.clip_slot));
  RCHK(X.take(r->wide, primcnt, B.wide_list));
  RCHK(X.take(r->bsum, (static_cast<size_t>(primcnt) + 255) / 256, B.bsum));
  RCHK(X.take(r->pairs, primcnt, B.pairs));
  RCHK(X.take(r->scan_tmp, scan_tmp_size(P3), B.scan_tmp));
  RCHK(X.take(r->partial, std::max<size_t>(T, (primcnt + 255) / 256) + 1, B.partial));
  RCHK(X.take(r->tstart, std::max<uint32_t>(T, 1), B.tstart));
  RCHK(X.take(r->tend, std::max<uint32_t>(T, 1), B.tend));
  RCHK(X.take(r->tile_count, 2 * static_cast<size_t>(std::max<uint32_t>(T, 1)), B.tile_count));
  // clipping capacity for every primitive, so no read-back decides whether clipping runs
  RCHK(X.take(r->clip_verts, static_cast<size_t>(primcnt) * 5 * (nout + 1), B.clip_verts));
  RCHK(X.take(r->clip_hdr, primcnt, B.clip_hdr));
  B.clip_capacity = primcnt;
  B.counters = r->counters.p;
  B.stats = r->stats_d;

  // statistics queries need the in-order walk (ps_invocations counts early-Z passes as they happen)
  const bool counting = P.count_stats != 0;
  const bool vis_tables = vis_pass && !counting;
  // Records are emitted before their count is read back, so the read-back overlaps the emission.
  // Allocating draws emit into a grow-only buffer sized from the previous draw and repeat the
  // emission when it was too small; a fixed region gets the largest record capacity whose
  // record-sized buffers all fit, and a draw with more records is refused.
  uint32_t capacity, sort_cap;  // records the buffer holds; records the early sort is sized for
  if (!fixed) {
    uint32_t recent = 0;
    for (uint32_t v : r->recent_records) recent = std::max(recent, v);
    const uint64_t want = std::max<uint64_t>(std::max<uint64_t>(recent + static_cast<uint64_t>(recent) / 4, static_cast<uint64_t>(P3) + T),
                                             1u << 16);
    const uint32_t guess = static_cast<uint32_t>(std::min<uint64_t>(want, 0xFFFFFFF0u));
    if (r->keys0.cap < guess) {
      RCHK(ensure(r, r->keys0, guess));
      RCHK(ensure(r, r->vals0, guess));
    }
    B.keys[0] = r->keys0.p;
    B.vals[0] = r->vals0.p;
    capacity = static_cast<uint32_t>(std::min<size_t>(r->keys0.cap, r->vals0.cap));
    sort_cap = std::min(capacity, guess);
  } else {
    const size_t room = fixed->bytes - std::min(fixed->bytes, align256(X.used));
    uint64_t lo = 0, hi = 0xFFFFFFFEu;  // largest capacity whose buffers fit
    while (lo < hi) {
      const uint64_t mid = (lo + hi + 1) / 2;
      if (record_scratch_bytes(P3, T, samples, vis_tables, mid) <= room) lo = mid;
      else hi = mid - 1;
    }
    if (lo == 0) return fail(SRX_OUT_OF_MEMORY, "the draw's scratch exceeds the no-allocation region");
    capacity = static_cast<uint32_t>(lo);
    sort_cap = capacity;
    RCHK(X.take(r->keys0, capacity, B.keys[0]));
    RCHK(X.take(r->vals0, capacity, B.vals[0]));
  }

  bool prof = r->profiling() && !fixed;  // event creation may allocate: no-allocation draws are not profiled
  draw_events de{};
  if (prof)
    for (auto& e : de.ev) {
      e = r->take_event();
      if (!e) return fail(SRX_E_HIP, "hipEventCreate");
    }
  auto mark = [&](int i) {
    if (prof) (void)hipEventRecord(de.ev[i], s);
  };
  // sampler table: staged in pinned memory (the counter read-back below waits for the copy before the
  // staging is reused); a renderer's own table is uploaded only when it changed
  const bool same_table = !fixed && r->samplers_valid && samplers_d == r->samplers_last_d &&
                          std::memcmp(&r->samplers_last, &SM, sizeof(ps_samplers)) == 0;
  if (!same_table) {
    *r->samplers_h = SM;
    HIPCHK(hipMemcpyAsync(samplers_d, r->samplers_h, sizeof(ps_samplers), hipMemcpyHostToDevice, s));
    r->samplers_last = SM;
    r->samplers_last_d = samplers_d;
    r->samplers_valid = !fixed;
  }
  // the SASL program with its uniforms, staged like the sampler table
  if (P.ps_id == PS_SASL) {
    const bool same_prog = !fixed && r->sasl_valid && std::memcmp(&r->sasl_last, &r->sasl_cur, sizeof(sasl_program)) == 0;
    if (!same_prog) {
      *r->sasl_h = r->sasl_cur;
      HIPCHK(hipMemcpyAsync(const_cast<sasl_program*>(P.sasl), r->sasl_h, sizeof(sasl_program), hipMemcpyHostToDevice, s));
      r->sasl_last = r->sasl_cur;
      r->sasl_valid = !fixed;
    }
  }
  mark(0);
  launch_vs(P, B.vbuf, B.counters, B.tile_count, 2 * static_cast<size_t>(T), s);
  mark(1);
  // setup leaves each 256-primitive block's record total in bsum; the clip launch (which returns at
  // once on the device when setup flagged no primitive) adds the counts it makes and writes the
  // draw's total, from which the emission takes every triangle's first record: no scan runs
  launch_setup(P, B, s);
  launch_clip(P, B, s);
  mark(2);
  mark(3);
  // the counters are final after the clip launch: their read-back overlaps the emission
  HIPCHK(hipEventRecord(r->ev_counted, s));
  launch_emit(P, B, capacity, s);
  // A target's pending full clear is applied by the order-independent visibility pass as it writes
  // the target, when that pass covers the whole target (it then reads nothing back); the others run
  // here, while the record count travels to the host.
  const bool will_vis = vis_tables && vis_applicable(P);
  surface_obj* const ct0 = r->st.color_targets.empty() ? nullptr : r->st.color_targets[0].get();
  surface_obj* const dst = r->st.ds_target.get();
  auto covers = [&](const surface_obj* o) {
    return o && P.target_w == static_cast<float>(o->w) && P.target_h == static_cast<float>(o->h) &&
           P.gtiles_x * P.gt >= o->w && P.gtiles_y * P.gt >= o->h;
  };
  const bool eat_color = will_vis && needs_shading(P) && P.ps_id != PS_SASL && ct0 && ct0->pend.kind == 1 && covers(ct0);
  const bool eat_ds = will_vis && dst && dst->pend.kind == 2 && dst->pend.flags == (SRX_CLEAR_DEPTH | SRX_CLEAR_STENCIL) &&
                      covers(dst);
  for (auto& ct : r->st.color_targets)
    if (!(eat_color && ct.get() == ct0)) RCHK(run_pending_clear(ct.get(), s));
  if (!eat_ds) RCHK(run_pending_clear(dst, s));
  RCHK(start_counter_read(r));
  const bool order_free = vis_tables && vis_applicable(P);  // tile lists need no primitive order
  // record-sized buffers of the sort, for `sized` records
  uint32_t sized = 0;
  auto take_sort_buffers = [&](uint32_t n) -> int32_t {
    sized = n;
    const uint32_t nblocks = (std::max(n, T) + 4095) / 4096;
    const size_t hist_n = std::max<size_t>(static_cast<size_t>(256) * std::max<uint32_t>(nblocks, 1), bucket_hist_size(T, n));
    RCHK(X.take(r->keys1, std::max<uint32_t>(n, 1), B.keys[1]));
    RCHK(X.take(r->vals1, std::max<uint32_t>(n, 1), B.vals[1]));
    RCHK(X.take(r->hist, hist_n, B.hist));
    RCHK(X.take(r->scan_tmp, std::max(scan_tmp_size(P3), scan_tmp_size(static_cast<uint32_t>(hist_n))), B.scan_tmp));
    return SRX_OK;
  };
  uint32_t* order = nullptr;
  // Order-free draws sort their records before the count reaches the host: the sort is sized for the
  // record buffer's capacity and reads the count on the device, and only a count beyond the
  // capacity makes the draw emit and sort again.
  bool bucketed = false, early = false;
  if (order_free && bucket_sort_applicable(P)) {
    early = true;
    RCHK(take_sort_buffers(sort_cap));
    RCHK(X.take(r->order, std::max<uint32_t>(T, 1), order));
    bucketed = launch_bucket_sort(P, B, sort_cap, order, s);
  }
  uint32_t ctr[CT_COUNT];
  RCHK(finish_counter_read(r, ctr));
  if (ctr[CT_CLIP_OVERFLOW]) return fail(SRX_OUT_OF_MEMORY, "clip buffer overflow");
  const uint32_t total = ctr[CT_TOTAL];
  if (fixed && total > capacity)
    return fail(SRX_OUT_OF_MEMORY, "the draw's binned records exceed the no-allocation region");
  if (!fixed) {
    r->recent_records[r->recent_next++ % 8] = total;
    if (total > capacity) {
      RCHK(ensure(r, r->keys0, total));
      RCHK(ensure(r, r->vals0, total));
      B.keys[0] = r->keys0.p;
      B.vals[0] = r->vals0.p;
      launch_emit(P, B, total, s);
    }
  }
  if (total > sort_cap) bucketed = false;  // the early sort saw only part of the records
  int which = 1;
  if (!bucketed || !order) {
    // the early sort's tile counts and classes are counted again below
    if (early) {
      launch_fill_u32(B.counters + CT_CLASS, CT_COUNT - CT_CLASS, 0u, s);
      launch_fill_u32(B.tile_count, 2 * T, 0u, s);
    }
    // sized for the capacity on a fixed region (the layout the capacity was chosen for), else for the count
    RCHK(take_sort_buffers(fixed ? capacity : total));
    RCHK(X.take(r->order, std::max<uint32_t>(T, 1), order));
    bucketed = order_free && launch_bucket_sort(P, B, std::max<uint32_t>(total, 1), order, s);
    if (!bucketed) {
      int bits = 1;
      while ((1u << bits) < T) ++bits;
      which = launch_sort(B, total, bits, s);
      launch_tile_ranges(P, B, which, total, s);
      launch_tile_order(P, B, order, false, s);
    }
  }
  B.order = order;
  mark(4);
  if (vis_tables) {
    uint32_t ne, nm, nsl;
    vis_table_bounds(T, sized, &ne, &nm, &nsl);
    RCHK(X.take(r->vis_entries, ne, B.vis_entries));
    RCHK(X.take(r->vis_merges, nm, B.vis_merges));
    RCHK(X.take(r->vis_counts, 4, B.vis_counts));
    RCHK(X.take(r->vis_partial, static_cast<size_t>(nsl) * 256 * samples, B.vis_partial));
  }
  if (eat_color) {
    P.clear_color = 1;
    for (int k = 0; k < 4; ++k) P.clear_rgba[k] = (&ct0->pend.color.x)[k];
    RCHK(run_pending_clear(ct0, s, true));
  }
  if (eat_ds) {
    P.clear_ds = 1;
    P.clear_depth = dst->pend.depth;
    P.clear_stencil = dst->pend.stencil;
    RCHK(run_pending_clear(dst, s, true));
  }
  const bool fused = vis_tables && launch_vis(P, B, which, needs_shading(P), total, s);
  if (!fused) launch_raster(P, B, which, s);
  mark(5);
  // the in-order kernel defers shading (to shade_kernel) only for early-Z draws
  // deferred shading: the in-order kernel's early-Z winners, or a SASL shader over vis_kernel's winners
  if (vis_pass && needs_shading(P) && ((!fused && P.early_z) || (fused && P.ps_id == PS_SASL))) {
    bool jit = false;
    if (P.ps_id == PS_SASL) RCHK(sasl_jit_launch(r->sasl_cur, P, B, r->device, s, !fixed, &jit));
    if (!jit) launch_shade(P, B, s);
  }
  mark(6);
  if (prof) r->pending_prof.push_back(de);
  HIPCHK(hipGetLastError());
  return SRX_OK;
}

renderer_obj::~renderer_obj() {
  drop_recording(this);
  if (stream) (void)run_pending_clears(this);
  if (stream) (void)hipStreamSynchronize(stream);
  if (side) (void)hipStreamSynchronize(side);

  auto fr = [](auto& a) {
    if (a.p) (void)hipFree(a.p);
  };
  fr(vbuf); fr(tri_rec); fr(tri_attr); fr(clip_verts); fr(samplers_d); fr(vis);
  if (samplers_h) (void)hipHostFree(samplers_h);
  if (sasl_h) (void)hipHostFree(sasl_h);
  if (sasl_d) (void)hipFree(sasl_d); fr(cnt); fr(keys0); fr(keys1); fr(vals0); fr(vals1);
  fr(tstart); fr(tend); fr(hist); fr(scan_tmp); fr(clip_hdr); fr(counters); fr(clip_flag); fr(clip_slot); fr(partial); fr(order); fr(tile_count); fr(wide); fr(bsum); fr(pairs); fr(vis_entries); fr(vis_merges); fr(vis_counts); fr(vis_partial);
  if (stats_d) (void)hipFree(stats_d);
  if (pinned) (void)hipHostFree(pinned);
  if (map_host) (void)hipHostFree(map_host);
  for (auto& d : pending_prof)
    for (auto e : d.ev)
      if (e) (void)hipEventDestroy(e);
  for (auto e : event_pool) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
  if (side) (void)hipStreamDestroy(side);
  if (ev_counted) (void)hipEventDestroy(ev_counted);
  if (ev_read) (void)hipEventDestroy(ev_read);
}


// A finished query's result (its `done` event has completed): counters as end - begin, and the
// stage times of its profiled draws in ns (their events go back to the renderer's pool).
int32_t resolve_query(renderer_obj* r, query_obj* q) {
  if (q->resolved) return SRX_OK;
  uint64_t d[ST_COUNT];
  for (int i = 0; i < ST_COUNT; ++i)
    d[i] = (q->host_end[i] + q->dev_snap[ST_COUNT + i]) - (q->host_begin[i] + q->dev_snap[i]);
  q->stats = srx_pipeline_statistics{};
  q->stats.ia_vertices = d[ST_IA_VERTICES];
  q->stats.ia_primitives = d[ST_IA_PRIMITIVES];
  q->stats.vs_invocations = d[ST_VS_INVOCATIONS];
  q->stats.cinvocations = d[ST_CINVOCATIONS];
  q->stats.cprimitives = d[ST_CPRIMITIVES];
  q->stats.ps_invocations = d[ST_PS_INVOCATIONS];
  q->istats.backend_input_pixels = d[ST_BACKEND_PIXELS];
  for (auto& v : q->prof) v = 0;
  for (auto& de : q->draws) {
    float ms[6];
    for (int i = 0; i < 6; ++i) HIPCHK(hipEventElapsedTime(&ms[i], de.ev[i], de.ev[i + 1]));
    auto ns = [](float m) { return static_cast<uint64_t>(static_cast<double>(m) * 1.0e6); };
    q->prof[PROF_VTX] += ns(ms[0]);
    q->prof[PROF_CLIP] += ns(ms[1]);
    q->prof[PROF_COMPACT] += ns(ms[2]);
    q->prof[PROF_DISPATCH] += ns(ms[3]);
    q->prof[PROF_RAS] += ns(ms[4]) + ns(ms[5]);
    q->prof[PROF_RAS_VIS] += ns(ms[4]);
    q->prof[PROF_RAS_SHADE] += ns(ms[5]);
    q->prof[PROF_DRAWS] += 1;
    for (auto e : de.ev) r->event_pool.push_back(e);
  }
  q->draws.clear();
  q->resolved = true;
  return SRX_OK;
}

// ---- pending clears
static void forget_pending(renderer_obj* r, surface_obj* o) {
  if (!r) return;
  auto& v = r->pending;
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i].get() == o) {
      v.erase(v.begin() + static_cast<std::ptrdiff_t>(i));
      return;
    }
}
static void set_pending(renderer_obj* r, surface_obj* o, int kind) {
  if (o->pend.owner != r) {
    forget_pending(o->pend.owner, o);
    r->pending.emplace_back(o);
  }
  o->pend.kind = kind;
  o->pend.owner = r;
}
int32_t run_pending_clear(surface_obj* o, hipStream_t s, bool drop) {
  if (!o || !o->pend.kind) return SRX_OK;
  renderer_obj* owner = o->pend.owner;
  const pending_clear c = o->pend;
  o->pend = pending_clear{};
  ref<surface_obj> keep(o);  // the owner's list may hold the last reference
  forget_pending(owner, o);
  if (drop) return SRX_OK;
  if (!s) s = owner->stream;
  if (c.kind == 1) launch_clear(o->desc(), c.color, s);
  else launch_clear_ds(o->desc(), c.flags, c.depth, c.stencil, s);
  HIPCHK(hipGetLastError());
  return SRX_OK;
}
int32_t run_pending_clears(renderer_obj* r) {
  while (!r->pending.empty()) RCHK(run_pending_clear(r->pending.back().get(), r->stream));
  return SRX_OK;
}

}  // namespace srx_rt

using namespace srx_rt;

// ==================================================================== C ABI
extern "C" {

uint32_t srx_abi_version(void) { return SRX_ABI_VERSION; }
const char* srx_backend_name(void) { return "hip"; }
int32_t srx_last_error(void) { return g_last_error; }
const char* srx_last_error_string(void) { return g_last_error_str.c_str(); }
void srx_retain(void* h) { retain(static_cast<object*>(h)); }
void srx_release(void* h) { release(static_cast<object*>(h)); }

void srx_default_raster_desc(srx_raster_desc* d) {
  *d = srx_raster_desc{SRX_FILL_SOLID, SRX_CULL_BACK, 0, 0, 0.0f, 0.0f, 1, 0, 1, 0};
}
void srx_default_depth_stencil_desc(srx_depth_stencil_desc* d) {
  srx_depth_stencil_op_desc op{SRX_SOP_KEEP, SRX_SOP_KEEP, SRX_SOP_KEEP, SRX_CMP_ALWAYS};
  *d = srx_depth_stencil_desc{1, 1, SRX_CMP_LESS, 0, 0xFF, 0xFF, op, op};
}
void srx_default_sampler_desc(srx_sampler_desc* d) {
  *d = srx_sampler_desc{SRX_FILTER_POINT, SRX_FILTER_POINT, SRX_FILTER_POINT, SRX_MIP_MI, SRX_ADDRESS_WRAP,
                        SRX_ADDRESS_WRAP, SRX_ADDRESS_WRAP, 0.0f, 0, SRX_CMP_ALWAYS, {0, 0, 0, 0}, -1e20f, 1e20f};
}

int32_t srx_create_renderer(uint32_t kind, srx_renderer* out) {
  (void)kind;
  if (!out) return SRX_INVALID_PARAMETER;
  std::call_once(g_ewa_once, [] {
    float t[256];
    make_ewa(t);
    hipError_t e = upload_ewa_table(t, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) g_ewa_status = hip_fail(e, "EWA table upload");
  });
  if (g_ewa_status != SRX_OK) return g_ewa_status;
  auto* r = new renderer_obj();
  {  // renderer_impl::renderer_impl (renderer_impl.cpp:269-284): default raster and depth/stencil states
    auto* rs = new raster_obj();
    rs->desc = r->st.raster;
    r->st.raster_state = ref<raster_obj>(rs);
    release(rs);
    auto* ds = new ds_obj();
    ds->desc = r->st.ds;
    r->st.ds_state = ref<ds_obj>(ds);
    release(ds);
  }
  hipError_t e = hipGetDevice(&r->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->stream, hipStreamDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_counted, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_read, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&r->counters.p, CT_COUNT * sizeof(uint32_t));
  if (e == hipSuccess) r->counters.cap = CT_COUNT;
  if (e == hipSuccess) e = hipMalloc(&r->stats_d, ST_COUNT * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(r->stats_d, 0, ST_COUNT * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipHostMalloc(&r->pinned, CT_COUNT * sizeof(uint32_t), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&r->samplers_h, sizeof(ps_samplers), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&r->sasl_h, sizeof(sasl_program), hipHostMallocDefault);
  if (e != hipSuccess) {
    delete r;
    return hip_fail(e, "srx_create_renderer");
  }
  *out = reinterpret_cast<srx_renderer>(static_cast<object*>(r));
  return SRX_OK;
}

int32_t srx_create_buffer(srx_renderer r, size_t size, srx_buffer* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !out) return SRX_INVALID_PARAMETER;
  auto* b = new buffer_obj();
  b->size = size;
  hipError_t e = hipMalloc(&b->d, size ? size : 1);
  if (e == hipSuccess) e = hipMemset(b->d, 0, size ? size : 1);
  if (e != hipSuccess) {
    delete b;
    return hip_fail(e, "srx_create_buffer");
  }
  *out = reinterpret_cast<srx_buffer>(static_cast<object*>(b));
  return SRX_OK;
}
size_t srx_buffer_size(srx_buffer b) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  return o ? o->size : 0;
}
// Host copies go through the null stream, which waits for every renderer stream.
int32_t srx_buffer_write(srx_buffer b, size_t off, const void* data, size_t size) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  if (!o || off + size > o->size || (!data && size)) return SRX_INVALID_PARAMETER;
  if (size) HIPCHK(hipMemcpy(o->d + off, data, size, hipMemcpyHostToDevice));
  return SRX_OK;
}
// Rows land with one 2-D copy (host pitch stride_src, device pitch stride_dest).
int32_t srx_buffer_transfer(srx_buffer b, size_t off, const void* data, size_t stride_dest, size_t stride_src, size_t size,
                            size_t count) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  if (!o || (!data && size && count)) return SRX_INVALID_PARAMETER;
  if (count == 0 || size == 0) return SRX_OK;
  if (stride_dest < size || stride_src < size || off > o->size || (o->size - off - size) / stride_dest < count - 1)
    return SRX_INVALID_PARAMETER;
  HIPCHK(hipMemcpy2D(o->d + off, stride_dest, data, stride_src, size, count, hipMemcpyHostToDevice));
  return SRX_OK;
}
int32_t srx_buffer_read(srx_buffer b, size_t off, void* data, size_t size) {
  auto* o = as<buffer_obj>(b, K_BUFFER);
  if (!o || off + size > o->size || (!data && size)) return SRX_INVALID_PARAMETER;
  if (size) HIPCHK(hipMemcpy(data, o->d + off, size, hipMemcpyDeviceToHost));
  return SRX_OK;
}

int32_t srx_create_tex2d(srx_renderer r, size_t w, size_t h, size_t ns, int32_t fmt, srx_texture* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !out || w == 0 || h == 0 || ns == 0 || fmt < 0 || fmt > 6 ||
      w > SRX_MAX_RENDER_TARGET_SIZE * 2 || h > SRX_MAX_RENDER_TARGET_SIZE * 2 || ns > SRX_MAX_SAMPLE_COUNT)
    return SRX_INVALID_PARAMETER;
  auto* t = new texture_obj();
  t->pf = fmt;
  t->samples = static_cast<int>(ns);
  t->size[0] = static_cast<uint32_t>(w);
  t->size[1] = static_cast<uint32_t>(h);
  t->size[2] = 1;
  surface_obj* s = nullptr;
  int32_t rc = make_surface(static_cast<int>(w), static_cast<int>(h), static_cast<int>(ns), fmt, &s);
  if (rc != SRX_OK) {
    delete t;
    return rc;
  }
  t->surfs.emplace_back(s);
  release(s);
  *out = reinterpret_cast<srx_texture>(static_cast<object*>(t));
  return SRX_OK;
}
int32_t srx_texture_surface(srx_texture t, size_t level, srx_surface* out) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  if (!o || !out || level >= o->surfs.size()) return SRX_INVALID_PARAMETER;
  if (static_cast<int>(level) < o->max_lod || static_cast<int>(level) > o->min_lod) return SRX_INVALID_PARAMETER;
  retain(o->surfs[level].get());
  *out = reinterpret_cast<srx_surface>(static_cast<object*>(o->surfs[level].get()));
  return SRX_OK;
}
// texture_2d::gen_mipmap (texture2d.cpp:25-36) with the box filter on the device
int32_t srx_texture_gen_mipmap(srx_texture t, int32_t filter, int32_t auto_gen) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  if (!o || (filter != SRX_FILTER_POINT && filter != SRX_FILTER_LINEAR)) return SRX_INVALID_PARAMETER;
  if (auto_gen) {
    uint32_t m = std::max(o->size[0], std::max(o->size[1], o->size[2]));
    int n = 0;
    while (m > 0) {
      m >>= 1;
      ++n;
    }
    o->max_lod = 0;
    o->min_lod = n - 1;
  }
  o->surfs.resize(1);
  RCHK(run_pending_clear(o->surfs[0].get(), nullptr));  // the mip kernels wait for the issuing stream
  for (int l = o->max_lod; l < o->min_lod; ++l) {
    surface_obj* src = o->surfs.back().get();
    surface_obj* dst = nullptr;
    RCHK(make_surface((src->w + 1) / 2, (src->h + 1) / 2, src->samples, src->pf, &dst));
    launch_mip(src->desc(), dst->desc(), filter, nullptr);
    o->surfs.emplace_back(dst);
    release(dst);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return SRX_OK;
}
size_t srx_texture_max_lod(srx_texture t) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  return o ? o->max_lod : 0;
}
size_t srx_texture_min_lod(srx_texture t) {
  auto* o = as<texture_obj>(t, K_TEXTURE);
  return o ? o->min_lod : 0;
}
size_t srx_surface_width(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->w : 0; }
size_t srx_surface_height(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->h : 0; }
size_t srx_surface_sample_count(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->samples : 0; }
int32_t srx_surface_format(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->pf : -1; }
size_t srx_surface_size_bytes(srx_surface s) { auto* o = as<surface_obj>(s, K_SURFACE); return o ? o->bytes : 0; }
int32_t srx_surface_write(srx_surface s, const void* d, size_t n) {
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!o || n != o->bytes || !d) return SRX_INVALID_PARAMETER;
  RCHK(run_pending_clear(o, nullptr, true));  // overwritten whole
  HIPCHK(hipMemcpy(o->d, d, n, hipMemcpyHostToDevice));
  return SRX_OK;
}
int32_t srx_surface_read(srx_surface s, void* d, size_t n) {
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!o || n != o->bytes || !d) return SRX_INVALID_PARAMETER;
  RCHK(run_pending_clear(o, nullptr));  // on the issuing renderer's stream, which the copy waits for
  HIPCHK(hipMemcpy(d, o->d, n, hipMemcpyDeviceToHost));
  return SRX_OK;
}
int32_t srx_surface_resolve(srx_renderer rr, srx_surface src, srx_surface dst) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* a = as<surface_obj>(src, K_SURFACE);
  auto* b = as<surface_obj>(dst, K_SURFACE);
  if (!r || !a || !b || b->samples != 1 || a->w != b->w || a->h != b->h) return SRX_INVALID_PARAMETER;
  RCHK(run_pending_clear(a, r->stream));
  RCHK(run_pending_clear(b, r->stream));
  if (r->recording) return record_resolve(r, a, b);
  launch_resolve(a->desc(), b->desc(), r->stream);
  HIPCHK(hipGetLastError());
  return SRX_OK;
}

int32_t srx_create_sampler(srx_renderer r, const srx_sampler_desc* d, srx_texture t, srx_sampler* out) {
  auto* tex = as<texture_obj>(t, K_TEXTURE);
  if (!as<renderer_obj>(r, K_RENDERER) || !d || !tex || !out) return SRX_INVALID_PARAMETER;
  auto* s = new sampler_obj();
  s->desc = *d;
  s->tex = ref<texture_obj>(tex);
  *out = reinterpret_cast<srx_sampler>(static_cast<object*>(s));
  return SRX_OK;
}
int32_t srx_create_query(srx_renderer r, uint32_t id, srx_query* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !out) return SRX_INVALID_PARAMETER;
  if (id != SRX_QUERY_PIPELINE_STATISTICS && id != SRX_QUERY_INTERNAL_STATISTICS && id != SRX_QUERY_PIPELINE_PROFILES)
    return SRX_INVALID_PARAMETER;
  auto* q = new query_obj();
  q->id = id;
  *out = reinterpret_cast<srx_query>(static_cast<object*>(q));
  return SRX_OK;
}
int32_t srx_create_input_layout(srx_renderer r, const srx_input_element_desc* d, size_t n, int32_t vs_id,
                                srx_input_layout* out) {
  (void)vs_id;
  if (!as<renderer_obj>(r, K_RENDERER) || (!d && n) || !out) return SRX_INVALID_PARAMETER;
  auto* l = new layout_obj();
  for (size_t i = 0; i < n; ++i)
    l->elems.push_back(layout_elem{canon_sem(lower(d[i].semantic_name)), d[i].semantic_index, d[i].data_format,
                                   d[i].input_slot, d[i].aligned_byte_offset});
  *out = reinterpret_cast<srx_input_layout>(static_cast<object*>(l));
  return SRX_OK;
}
int32_t srx_create_raster_state(srx_renderer r, const srx_raster_desc* d, srx_raster_state* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !d || !out) return SRX_INVALID_PARAMETER;
  if (d->cull_mode < 0 || d->cull_mode > 2) return SRX_INVALID_PARAMETER;
  auto* o = new raster_obj();
  o->desc = *d;
  *out = reinterpret_cast<srx_raster_state>(static_cast<object*>(o));
  return SRX_OK;
}
int32_t srx_create_depth_stencil_state(srx_renderer r, const srx_depth_stencil_desc* d, srx_ds_state* out) {
  if (!as<renderer_obj>(r, K_RENDERER) || !d || !out) return SRX_INVALID_PARAMETER;
  auto* o = new ds_obj();
  o->desc = *d;
  *out = reinterpret_cast<srx_ds_state>(static_cast<object*>(o));
  return SRX_OK;
}

int32_t srx_set_vertex_buffers(srx_renderer rr, size_t start, size_t n, const srx_buffer* bufs, const size_t* strides,
                               const size_t* offsets) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || start + n > SRX_MAX_INPUT_SLOTS || (n && (!bufs || !strides))) return SRX_INVALID_PARAMETER;
  for (size_t i = 0; i < n; ++i) {
    r->st.vbs[start + i].buf = ref<buffer_obj>(as<buffer_obj>(bufs[i], K_BUFFER));
    r->st.vbs[start + i].stride = strides[i];
    r->st.vbs[start + i].offset = offsets ? offsets[i] : 0;
  }
  return SRX_OK;
}
int32_t srx_set_index_buffer(srx_renderer rr, srx_buffer b, uint32_t fmt) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  if (fmt != SRX_FMT_R16_UINT && fmt != SRX_FMT_R32_UINT) return SRX_FAILED;  // renderer_impl.cpp:49-53
  r->st.ib = ref<buffer_obj>(as<buffer_obj>(b, K_BUFFER));
  r->st.index_fmt = fmt;
  return SRX_OK;
}
int32_t srx_set_input_layout(srx_renderer rr, srx_input_layout l) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  r->st.layout = ref<layout_obj>(as<layout_obj>(l, K_LAYOUT));
  return SRX_OK;
}
int32_t srx_set_primitive_topology(srx_renderer rr, uint32_t t) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  switch (t) {  // renderer_impl.cpp:70-82
  case SRX_TOPOLOGY_LINE_LIST:
  case SRX_TOPOLOGY_LINE_STRIP:
  case SRX_TOPOLOGY_TRIANGLE_LIST:
  case SRX_TOPOLOGY_TRIANGLE_STRIP: r->st.topo = t; return SRX_OK;
  }
  return SRX_FAILED;
}
int32_t srx_set_vertex_shader(srx_renderer rr, int32_t id, const void* consts, size_t size) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || vs_output_count(id) <= 0) return SRX_INVALID_PARAMETER;
  if (size < vs_const_bytes(id) || (vs_const_bytes(id) && !consts)) return SRX_INVALID_PARAMETER;
  r->st.vs_id = id;
  std::memset(r->st.vs_consts, 0, sizeof r->st.vs_consts);
  if (consts && size) std::memcpy(r->st.vs_consts, consts, std::min<size_t>(size, sizeof r->st.vs_consts));
  return SRX_OK;
}
int32_t srx_set_pixel_shader(srx_renderer rr, int32_t id, const void* consts, size_t size) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || id < SRX_PS_GOURAUD || id > SRX_PS_BOX) return SRX_INVALID_PARAMETER;
  if (size < ps_const_bytes(id) || (ps_const_bytes(id) && !consts)) return SRX_INVALID_PARAMETER;
  r->st.ps_id = id;
  std::memset(r->st.ps_consts, 0, sizeof r->st.ps_consts);
  if (consts && size) std::memcpy(r->st.ps_consts, consts, std::min<size_t>(size, 64));
  return SRX_OK;
}
int32_t srx_set_ps_sampler(srx_renderer rr, uint32_t slot, srx_sampler s) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || slot >= SRX_MAX_SAMPLER_SLOTS) return SRX_INVALID_PARAMETER;
  r->st.samplers[slot] = ref<sampler_obj>(as<sampler_obj>(s, K_SAMPLER));
  return SRX_OK;
}
int32_t srx_compile_shader(const char* code, size_t length, int32_t lang, srx_shader_code* out) {
  if (!code || !out) return SRX_INVALID_PARAMETER;
  *out = nullptr;
  if (lang != SRX_LANG_PIXEL_SHADER)
    return fail(SRX_E_UNSUPPORTED, "only pixel shaders are compiled; vertex shaders are the built-in ids");
  auto* o = new shader_obj();
  std::string msg;
  const int st = sasl_compile_ps(code, length, o->mod, msg);
  if (st != SASL_OK) {
    delete o;
    return fail(st == SASL_UNSUPPORTED ? SRX_E_UNSUPPORTED : SRX_INVALID_PARAMETER, msg.c_str());
  }
  *out = reinterpret_cast<srx_shader_code>(static_cast<object*>(o));
  return SRX_OK;
}
int32_t srx_shader_reflect(srx_shader_code c, srx_shader_info* info) {
  auto* o = as<shader_obj>(c, K_SHADER);
  if (!o || !info) return SRX_INVALID_PARAMETER;
  std::memset(info, 0, sizeof *info);
  const sasl_program& p = o->mod.prog;
  info->input_count = p.ninputs;
  info->output_count = p.noutputs;
  for (uint32_t i = 0; i < p.ninputs; ++i) info->input_components[i] = p.inputs[i].ncomp;
  for (uint32_t i = 0; i < p.noutputs; ++i) {
    info->output_targets[i] = p.outputs[i].target;
    info->output_components[i] = p.outputs[i].ncomp;
  }
  info->uniform_count = static_cast<uint32_t>(o->mod.uniforms.size());
  info->sampler_count = static_cast<uint32_t>(o->mod.samplers.size());
  info->instruction_count = p.ninst;
  info->register_count = p.nregs;
  return SRX_OK;
}
int32_t srx_set_pixel_shader_code(srx_renderer rr, srx_shader_code c) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  auto* o = c ? as<shader_obj>(c, K_SHADER) : nullptr;
  if (c && !o) return SRX_INVALID_PARAMETER;
  r->st.ps_code = ref<shader_obj>(o);
  return SRX_OK;
}
int32_t srx_get_pixel_shader_code(srx_renderer rr, srx_shader_code* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  retain(r->st.ps_code.get());
  *out = reinterpret_cast<srx_shader_code>(static_cast<object*>(r->st.ps_code.get()));
  return SRX_OK;
}
int32_t srx_set_ps_variable(srx_renderer rr, const char* name, const void* data, size_t size) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !name || (size && !data)) return SRX_INVALID_PARAMETER;
  std::vector<uint8_t> bytes(static_cast<const uint8_t*>(data), static_cast<const uint8_t*>(data) + size);
  for (auto& v : r->st.ps_vars)
    if (v.first == name) {
      v.second = std::move(bytes);
      return SRX_OK;
    }
  r->st.ps_vars.emplace_back(name, std::move(bytes));
  return SRX_OK;
}
int32_t srx_set_ps_sampler_named(srx_renderer rr, const char* name, srx_sampler s) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !name) return SRX_INVALID_PARAMETER;
  auto* so = s ? as<sampler_obj>(s, K_SAMPLER) : nullptr;
  if (s && !so) return SRX_INVALID_PARAMETER;
  for (auto& v : r->st.ps_named_samplers)
    if (v.first == name) {
      v.second = ref<sampler_obj>(so);
      return SRX_OK;
    }
  r->st.ps_named_samplers.emplace_back(name, ref<sampler_obj>(so));
  return SRX_OK;
}
int32_t srx_ps_execute_quads(srx_renderer rr, srx_shader_code c, const float* in, float* out, size_t nquads) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<shader_obj>(c, K_SHADER);
  if (!r || !o || (nquads && (!out || (!in && o->mod.prog.ninputs)))) return SRX_INVALID_PARAMETER;
  if (nquads > (1u << 24)) return fail(SRX_E_UNSUPPORTED, "more than 2^24 quads in one call");
  if (!nquads) return SRX_OK;
  if (r->recording) return batch_refuse(r, SRX_E_UNSUPPORTED, "a batch does not record shader evaluations");
  const sasl_module& m = o->mod;
  sasl_program prog = m.prog;
  for (auto& v : r->st.ps_vars) sasl_set_uniform(m, prog, v.first, v.second.data(), v.second.size());
  // sampler table of the named samplers, as a draw builds it
  ps_samplers SM;
  std::memset(&SM, 0, sizeof SM);
  uint32_t has = 0;
  for (size_t k = 0; k < m.samplers.size() && k < MAX_SAMPLERS; ++k)
    for (auto& ns : r->st.ps_named_samplers)
      if (ns.first == m.samplers[k] && ns.second && ns.second->tex) {
        RCHK(fill_sampler(r, ns.second.get(), SM.s[k]));
        has |= 1u << k;
      }
  const size_t in_n = nquads * 4 * m.prog.ninputs, out_n = nquads * 4 * m.prog.noutputs;
  void* dev = nullptr;
  const size_t bytes = sizeof(sasl_program) + sizeof(ps_samplers) + (in_n + out_n) * sizeof(float4);
  HIPCHK(hipStreamSynchronize(r->stream));
  HIPCHK(hipMalloc(&dev, bytes));
  auto* prog_d = static_cast<sasl_program*>(dev);
  auto* sm_d = reinterpret_cast<ps_samplers*>(prog_d + 1);
  auto* in_d = reinterpret_cast<float4*>(sm_d + 1);
  float4* out_d = in_d + in_n;
  hipError_t he = hipMemcpyAsync(prog_d, &prog, sizeof prog, hipMemcpyHostToDevice, r->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(sm_d, &SM, sizeof SM, hipMemcpyHostToDevice, r->stream);
  if (he == hipSuccess && in_n) he = hipMemcpyAsync(in_d, in, in_n * sizeof(float4), hipMemcpyHostToDevice, r->stream);
  if (he == hipSuccess) {
    launch_sasl_quads(prog_d, sm_d, has, in_d, out_d, static_cast<uint32_t>(nquads), r->stream);
    he = hipGetLastError();
  }
  if (he == hipSuccess && out_n) he = hipMemcpyAsync(out, out_d, out_n * sizeof(float4), hipMemcpyDeviceToHost, r->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(r->stream);
  (void)hipFree(dev);
  if (he != hipSuccess) return hip_fail(he, "srx_ps_execute_quads");
  return SRX_OK;
}
int32_t srx_set_blend_shader(srx_renderer rr, int32_t id) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || id < 0 || id > 2) return SRX_INVALID_PARAMETER;
  r->st.blend_id = id;
  return SRX_OK;
}
int32_t srx_set_rasterizer_state(srx_renderer rr, srx_raster_state rs) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<raster_obj>(rs, K_RASTER);
  if (!r || !o) return SRX_INVALID_PARAMETER;
  r->st.raster = o->desc;
  r->st.raster_state = ref<raster_obj>(o);
  return SRX_OK;
}
int32_t srx_set_depth_stencil_state(srx_renderer rr, srx_ds_state d, int32_t ref_) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<ds_obj>(d, K_DS);
  if (!r || !o) return SRX_INVALID_PARAMETER;
  r->st.ds = o->desc;
  r->st.ds_state = ref<ds_obj>(o);
  r->st.stencil_ref = ref_;
  return SRX_OK;
}
// renderer_impl::set_render_targets (renderer_impl.cpp:159-238)
int32_t srx_set_render_targets(srx_renderer rr, size_t n, const srx_surface* cts, srx_surface dst) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  render_state& st = r->st;
  st.target_vp = srx_viewport{0, 0, 0, 0, 0, 0};
  st.target_samples = 0;
  st.color_targets.clear();
  st.ds_target = ref<surface_obj>();
  if (n >= 8) return SRX_FAILED;
  auto* ds = as<surface_obj>(dst, K_SURFACE);
  if ((n == 0 || !cts) && !ds) return SRX_FAILED;
  srx_viewport tvp{0, 0, 3.4028234663852886e38f, 3.4028234663852886e38f, 0, 0};
  size_t samples = 0;
  for (size_t i = 0; i < n; ++i) {
    auto* c = as<surface_obj>(cts[i], K_SURFACE);
    if (!c) continue;
    tvp.w = std::min(static_cast<float>(c->w), tvp.w);
    tvp.h = std::min(static_cast<float>(c->h), tvp.h);
    if (samples == 0) samples = c->samples;
    else if (static_cast<size_t>(c->samples) != samples) return SRX_FAILED;
  }
  if (ds) {
    if (ds->pf != PF_RG32F) return SRX_FAILED;
    if (n == 0) {
      samples = ds->samples;
      tvp.w = static_cast<float>(ds->w);
      tvp.h = static_cast<float>(ds->h);
    }
    if (static_cast<float>(ds->w) < tvp.w || static_cast<float>(ds->h) < tvp.h) return SRX_FAILED;
    if (static_cast<size_t>(ds->samples) != samples) return SRX_FAILED;
  }
  st.target_vp = tvp;
  st.target_samples = samples;
  for (size_t i = 0; i < n; ++i) st.color_targets.emplace_back(as<surface_obj>(cts[i], K_SURFACE));
  st.ds_target = ref<surface_obj>(ds);
  return SRX_OK;
}
int32_t srx_set_viewport(srx_renderer rr, const srx_viewport* vp) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !vp) return SRX_INVALID_PARAMETER;
  if (vp->x < 0 || vp->y < 0 || vp->w >= SRX_MAX_RENDER_TARGET_SIZE || vp->h >= SRX_MAX_RENDER_TARGET_SIZE) return SRX_FAILED;
  r->st.vp = *vp;
  return SRX_OK;
}
int32_t srx_get_viewport(srx_renderer rr, srx_viewport* vp) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !vp) return SRX_INVALID_PARAMETER;
  *vp = r->st.vp;
  return SRX_OK;
}

int32_t srx_begin(srx_renderer rr, srx_query qq) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* q = as<query_obj>(qq, K_QUERY);
  if (!r || !q) return SRX_INVALID_PARAMETER;
  if (q->running) return SRX_FAILED;
  for (auto& a : r->active)
    if (a->id == q->id) return SRX_FAILED;
  if (r->recording) return batch_refuse(r, SRX_E_UNSUPPORTED, "queries cannot run inside a batch");
  if (!q->done) HIPCHK(hipEventCreateWithFlags(&q->done, hipEventDefault));
  if (!q->dev_snap) HIPCHK(hipHostMalloc(&q->dev_snap, 2 * ST_COUNT * sizeof(unsigned long long), hipHostMallocDefault));
  if (q->has_data && !q->resolved) HIPCHK(hipEventSynchronize(q->done));  // the previous end's copies
  for (auto& d : q->draws)
    for (auto e : d.ev) r->event_pool.push_back(e);
  q->draws.clear();
  q->running = true;
  q->has_data = false;
  q->resolved = false;
  std::memcpy(q->host_begin, r->host_stats, sizeof q->host_begin);
  HIPCHK(hipMemcpyAsync(q->dev_snap, r->stats_d, ST_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost, r->stream));
  r->active.emplace_back(q);
  return SRX_OK;
}
// Queues the counters' read-back and the completion event; does not wait (async_renderer.cpp:17-110).
int32_t srx_end(srx_renderer rr, srx_query qq) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* q = as<query_obj>(qq, K_QUERY);
  if (!r || !q || !q->running) return SRX_FAILED;
  std::memcpy(q->host_end, r->host_stats, sizeof q->host_end);
  HIPCHK(hipMemcpyAsync(q->dev_snap + ST_COUNT, r->stats_d, ST_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        r->stream));
  HIPCHK(hipEventRecord(q->done, r->stream));
  if (q->id == SRX_QUERY_PIPELINE_PROFILES) q->draws = std::move(r->pending_prof);
  r->pending_prof.clear();
  q->running = false;
  q->has_data = true;
  for (size_t i = 0; i < r->active.size(); ++i)
    if (r->active[i].get() == q) {
      r->active.erase(r->active.begin() + static_cast<long>(i));
      break;
    }
  return SRX_OK;
}
// async_object::get (async_object.h:62-81): error while running or before the first end; timeout
// while the GPU has not reached the end (do_not_wait), else waits; then ready.
int32_t srx_get_data(srx_renderer rr, srx_query qq, void* data, size_t size, int32_t do_not_wait) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* q = as<query_obj>(qq, K_QUERY);
  if (!r || !q || !data || !q->has_data || q->running) return SRX_ASYNC_ERROR;
  if (!q->resolved) {
    if (do_not_wait) {
      const hipError_t e = hipEventQuery(q->done);
      if (e == hipErrorNotReady) return SRX_ASYNC_TIMEOUT;
      if (e != hipSuccess) {
        hip_fail(e, "hipEventQuery");
        return SRX_ASYNC_ERROR;
      }
    } else if (hipEventSynchronize(q->done) != hipSuccess) {
      return SRX_ASYNC_ERROR;
    }
    if (resolve_query(r, q) != SRX_OK) return SRX_ASYNC_ERROR;
  }
  if (q->id == SRX_QUERY_PIPELINE_STATISTICS && size >= sizeof(q->stats)) std::memcpy(data, &q->stats, sizeof(q->stats));
  else if (q->id == SRX_QUERY_INTERNAL_STATISTICS && size >= sizeof(q->istats)) std::memcpy(data, &q->istats, sizeof(q->istats));
  else if (q->id == SRX_QUERY_PIPELINE_PROFILES && size >= 8 * sizeof(uint64_t))
    std::memcpy(data, q->prof, std::min(size, sizeof(q->prof)));
  else return SRX_ASYNC_ERROR;
  return SRX_ASYNC_READY;
}

// ---- map / unmap (renderer.h:61-63; resource_manager.cpp:7-113, buffer.cpp:9-29, surface.cpp:94-121)
// The mapped bytes live in pinned host memory.  Every mode first finishes the renderer's queued
// work and copies the resource down, so the host sees its current bytes whatever it reads or leaves
// unwritten (the reference hands out the resource's own memory for every mode but map_read); unmap
// copies them back for every mode that may write.
static int32_t map_common(renderer_obj* r, srx_mapped_resource* m, uint8_t* dev, size_t bytes, int32_t mode,
                          uint32_t row_pitch, uint32_t depth_pitch) {
  if (mode < SRX_MAP_READ || mode > SRX_MAP_WRITE_NO_OVERWRITE) return SRX_FAILED;  // map_mode_none or unknown
  if (r->map_mode != 0) return SRX_FAILED;                                          // one mapping at a time
  if (r->map_cap < bytes) {
    if (r->map_host) HIPCHK(hipHostFree(r->map_host));
    r->map_host = nullptr;
    r->map_cap = 0;
    HIPCHK(hipHostMalloc(&r->map_host, std::max<size_t>(bytes, 4096), hipHostMallocDefault));
    r->map_cap = std::max<size_t>(bytes, 4096);
  }
  HIPCHK(hipStreamSynchronize(r->stream));
  if (bytes) {
    HIPCHK(hipMemcpyAsync(r->map_host, dev, bytes, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
  }
  m->data = r->map_host;
  m->row_pitch = row_pitch;
  m->depth_pitch = depth_pitch;
  r->map_mode = mode;
  return SRX_OK;
}
int32_t srx_map_buffer(srx_renderer rr, srx_mapped_resource* m, srx_buffer bb, int32_t mode) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* b = as<buffer_obj>(bb, K_BUFFER);
  if (!r || !m || !b) return SRX_INVALID_PARAMETER;
  const uint32_t pitch = static_cast<uint32_t>(b->size);
  RCHK(map_common(r, m, b->d, b->size, mode, pitch, pitch));
  r->mapped_buf = ref<buffer_obj>(b);
  return SRX_OK;
}
int32_t srx_map_surface(srx_renderer rr, srx_mapped_resource* m, srx_surface ss, int32_t mode) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<surface_obj>(ss, K_SURFACE);
  if (!r || !m || !o) return SRX_INVALID_PARAMETER;
  RCHK(run_pending_clear(o, r->stream));
  const uint32_t pitch = static_cast<uint32_t>(static_cast<size_t>(o->w) * o->samples * pf_bytes(o->pf));
  RCHK(map_common(r, m, o->d, o->bytes, mode, pitch, pitch * static_cast<uint32_t>(o->h)));
  r->mapped_surf = ref<surface_obj>(o);
  return SRX_OK;
}
int32_t srx_unmap(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  if (r->map_mode == 0) return SRX_FAILED;
  uint8_t* dev = r->mapped_buf ? r->mapped_buf->d : r->mapped_surf->d;
  const size_t bytes = r->mapped_buf ? r->mapped_buf->size : r->mapped_surf->bytes;
  const int32_t mode = r->map_mode;
  r->map_mode = 0;
  r->mapped_buf = ref<buffer_obj>();
  r->mapped_surf = ref<surface_obj>();
  if (mode != SRX_MAP_READ && bytes) {
    HIPCHK(hipMemcpyAsync(dev, r->map_host, bytes, hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));  // the staging may be mapped again at once
  }
  return SRX_OK;
}

// ---- state get (renderer.h:104-113, renderer_impl.cpp:61-157)
int32_t srx_get_index_buffer(srx_renderer rr, srx_buffer* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  retain(r->st.ib.get());
  *out = reinterpret_cast<srx_buffer>(static_cast<object*>(r->st.ib.get()));
  return SRX_OK;
}
uint32_t srx_get_index_format(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.index_fmt : 0;
}
uint32_t srx_get_primitive_topology(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.topo : 0;
}
int32_t srx_get_vertex_shader(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.vs_id : 0;
}
int32_t srx_get_pixel_shader(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.ps_id : 0;
}
int32_t srx_get_blend_shader(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  return r ? r->st.blend_id : -1;
}
int32_t srx_get_rasterizer_state(srx_renderer rr, srx_raster_state* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  retain(r->st.raster_state.get());
  *out = reinterpret_cast<srx_raster_state>(static_cast<object*>(r->st.raster_state.get()));
  return SRX_OK;
}
int32_t srx_get_depth_stencil_state(srx_renderer rr, srx_ds_state* out, int32_t* stencil_ref) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  retain(r->st.ds_state.get());
  *out = reinterpret_cast<srx_ds_state>(static_cast<object*>(r->st.ds_state.get()));
  if (stencil_ref) *stencil_ref = r->st.stencil_ref;
  return SRX_OK;
}
int32_t srx_raster_state_desc(srx_raster_state rs, srx_raster_desc* d) {
  auto* o = as<raster_obj>(rs, K_RASTER);
  if (!o || !d) return SRX_INVALID_PARAMETER;
  *d = o->desc;
  return SRX_OK;
}
int32_t srx_depth_stencil_state_desc(srx_ds_state ds, srx_depth_stencil_desc* d) {
  auto* o = as<ds_obj>(ds, K_DS);
  if (!o || !d) return SRX_INVALID_PARAMETER;
  *d = o->desc;
  return SRX_OK;
}
// set_vs_sampler (renderer_impl.cpp:336-339): bound by slot; kept for vertex shaders that sample
int32_t srx_set_vs_sampler(srx_renderer rr, uint32_t slot, srx_sampler s) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || slot >= SRX_MAX_SAMPLER_SLOTS) return SRX_INVALID_PARAMETER;
  r->st.vs_samplers[slot] = ref<sampler_obj>(as<sampler_obj>(s, K_SAMPLER));
  return SRX_OK;
}
int32_t srx_draw(srx_renderer rr, size_t start, size_t primcnt) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  if (primcnt > MAX_DRAW_PRIMITIVES) return fail(SRX_E_UNSUPPORTED, "more than 1,431,650,304 primitives in one draw");
  (void)start;  // the index fetcher ignores the start of non-indexed draws (index_fetcher.cpp:103-107)
  return execute_draw(r, false, 0, static_cast<uint32_t>(primcnt), 0, nullptr);
}
int32_t srx_draw_index(srx_renderer rr, size_t start, size_t primcnt, int32_t basevert) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  if (primcnt > MAX_DRAW_PRIMITIVES) return fail(SRX_E_UNSUPPORTED, "more than 1,431,650,304 primitives in one draw");
  if (start > 0xFFFFFFFFu) return fail(SRX_E_UNSUPPORTED, "start index beyond 2^32");
  return execute_draw(r, true, static_cast<uint32_t>(start), static_cast<uint32_t>(primcnt), basevert, nullptr);
}
int32_t srx_clear_color(srx_renderer rr, srx_surface s, const float* c) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!r || !o || !c) return SRX_INVALID_PARAMETER;
  if (r->recording) {
    RCHK(run_pending_clear(o, r->stream));
    return record_fill(r, o, make_float4(c[0], c[1], c[2], c[3]));
  }
  // a full colour clear supersedes whatever clear is still pending
  if (o->pend.kind == 2) RCHK(run_pending_clear(o, r->stream));
  set_pending(r, o, 1);
  o->pend.color = make_float4(c[0], c[1], c[2], c[3]);
  return SRX_OK;
}
int32_t srx_clear_depth_stencil(srx_renderer rr, srx_surface s, uint32_t flags, float d, uint32_t st) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  auto* o = as<surface_obj>(s, K_SURFACE);
  if (!r || !o || o->pf != PF_RG32F) return SRX_INVALID_PARAMETER;
  if (flags != (SRX_CLEAR_DEPTH | SRX_CLEAR_STENCIL) && flags != SRX_CLEAR_DEPTH && flags != SRX_CLEAR_STENCIL)
    return SRX_E_UNSUPPORTED;
  if (r->recording) {
    if (flags != (SRX_CLEAR_DEPTH | SRX_CLEAR_STENCIL))
      return batch_refuse(r, SRX_E_UNSUPPORTED, "batched depth/stencil clears set both");
    float sf;
    std::memcpy(&sf, &st, sizeof sf);
    RCHK(run_pending_clear(o, r->stream));
    return record_fill(r, o, make_float4(d, sf, 0.0f, 0.0f));
  }
  // merged with a pending depth/stencil clear: each channel keeps its latest clear value
  if (o->pend.kind == 1) RCHK(run_pending_clear(o, r->stream));
  const uint32_t had = o->pend.kind == 2 ? o->pend.flags : 0u;
  set_pending(r, o, 2);
  o->pend.flags = had | flags;
  if (flags & SRX_CLEAR_DEPTH) o->pend.depth = d;
  if (flags & SRX_CLEAR_STENCIL) o->pend.stencil = st;
  return SRX_OK;
}
int32_t srx_flush(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  RCHK(run_pending_clears(r));
  HIPCHK(hipStreamSynchronize(r->stream));
  return SRX_OK;
}

int32_t srx_create_swap_chain(srx_renderer rr, size_t w, size_t h, size_t ns, int32_t fmt, srx_swap_chain* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out) return SRX_INVALID_PARAMETER;
  srx_texture back = nullptr, res = nullptr;
  RCHK(srx_create_tex2d(rr, w, h, ns, fmt, &back));
  if (ns > 1) {
    const int32_t e = srx_create_tex2d(rr, w, h, 1, fmt, &res);
    if (e != SRX_OK) {
      srx_release(back);
      return e;
    }
  }
  auto* sc = new swap_obj();
  sc->r = ref<renderer_obj>(r);
  sc->back = ref<texture_obj>(as<texture_obj>(back, K_TEXTURE));
  sc->resolved = ref<texture_obj>(as<texture_obj>(res ? res : back, K_TEXTURE));
  srx_release(back);
  if (res) srx_release(res);
  *out = reinterpret_cast<srx_swap_chain>(static_cast<object*>(sc));
  return SRX_OK;
}
int32_t srx_swap_chain_surface(srx_swap_chain s, srx_surface* out) {
  auto* sc = as<swap_obj>(s, K_SWAP_CHAIN);
  if (!sc || !out) return SRX_INVALID_PARAMETER;
  return srx_texture_surface(reinterpret_cast<srx_texture>(static_cast<object*>(sc->back.get())), 0, out);
}
int32_t srx_swap_chain_presented(srx_swap_chain s, srx_surface* out) {
  auto* sc = as<swap_obj>(s, K_SWAP_CHAIN);
  if (!sc || !out) return SRX_INVALID_PARAMETER;
  return srx_texture_surface(reinterpret_cast<srx_texture>(static_cast<object*>(sc->resolved.get())), 0, out);
}
int32_t srx_swap_chain_present(srx_swap_chain s) {
  auto* sc = as<swap_obj>(s, K_SWAP_CHAIN);
  if (!sc) return SRX_INVALID_PARAMETER;
  auto* rr = reinterpret_cast<srx_renderer>(static_cast<object*>(sc->r.get()));
  RCHK(srx_flush(rr));
  if (sc->resolved.get() != sc->back.get()) {
    srx_surface a = nullptr, b = nullptr;
    RCHK(srx_swap_chain_surface(s, &a));
    int32_t e = srx_swap_chain_presented(s, &b);
    if (e == SRX_OK) e = srx_surface_resolve(rr, a, b);
    srx_release(a);
    if (b) srx_release(b);
    RCHK(e);
    RCHK(srx_flush(rr));
  }
  return SRX_OK;
}

}  // extern "C"
