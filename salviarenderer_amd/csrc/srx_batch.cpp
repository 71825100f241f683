// Batches of independent render jobs: commands recorded between srx_batch_begin and srx_batch_end,
// replayed as one launch per pipeline stage over all jobs from a HIP graph.
#include "srx_runtime_internal.h"

#include <map>

namespace srx_rt {

// ---------------------------------------------------------------- batches
// Recorded clears, draws and resolves of independent render jobs, replayed as one launch per
// pipeline stage (all jobs' workgroups in one grid), captured in a HIP graph.  The draws take the
// order-independent visibility path, so every job's output equals its one-by-one render.
enum { BS_DRAWN = 1, BS_FILLED = 2, BS_RESOLVE_SRC = 4, BS_RESOLVE_DST = 8, BS_SAMPLED = 16 };

struct batch_draw {
  draw_params P;
  ps_samplers SM;
  render_state keep;  // holds the draw's buffers, targets and textures
};
struct batch_fill {
  fill_item it;
  ref<surface_obj> s;
};
struct batch_resolve {
  resolve_item it;
  ref<surface_obj> a, b;
};

struct batch_obj : object {
  renderer_obj* r = nullptr;  // retained from srx_batch_end on
  bool owns_renderer = false;
  std::vector<batch_draw> draws;
  std::vector<batch_fill> fills;
  std::vector<batch_resolve> resolves;
  std::vector<std::pair<surface_obj*, uint32_t>> marks;  // recording hazards per surface
  int32_t error = SRX_OK;  // first recording error, reported by srx_batch_end
  // tuning
  int tile = 16, bin_pt = 16, vis_threads = 256;
  std::map<int, std::vector<uint32_t>> counted;  // records per draw, by GPU tile size
  // sizes
  std::vector<uint32_t> records;  // per draw, from the counting pass
  uint32_t capacity = 0;
  uint64_t primitives = 0, device_bytes = 0;
  // device memory: `fixed` lives as long as the batch, `tuned` is rebuilt by srx_batch_configure
  std::vector<void*> fixed, tuned;
  batch_job* jobs_d = nullptr;
  std::vector<batch_job> jobs_h;
  uint32_t *cnt_all = nullptr, *status_d = nullptr, *counts_d = nullptr, *scan_total = nullptr, *status_h = nullptr;
  uint32_t cnt_n = 0;  // per-triangle count slots + 1 trailing
  uint32_t* counters_all = nullptr;
  uint32_t* hist_all = nullptr;
  uint32_t hist_n = 0;
  uint32_t* scan_tmp = nullptr;
  uint32_t *maps_fixed = nullptr, *maps_tuned = nullptr;
  fill_item* fills_d = nullptr;
  resolve_item* resolves_d = nullptr;
  map_image::entry m_vs, m_emit, m_clip_all;
  std::vector<std::pair<int, map_image::entry>> m_setup, m_clip;  // per attribute count
  map_image::entry m_fill, m_resolve, m_bucket, m_range;
  struct vis_class {
    int na, samples;
    bool shade;
    map_image::entry m;
  };
  std::vector<vis_class> m_vis;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  batch_obj() : object(K_BATCH) {}
  uint64_t tuned_bytes = 0;
  void free_tuned() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    for (void* p : tuned) (void)hipFree(p);
    tuned.clear();
    device_bytes -= tuned_bytes;
    tuned_bytes = 0;
  }
  ~batch_obj() override {
    if (r && r->stream) (void)hipStreamSynchronize(r->stream);
    free_tuned();
    for (void* p : fixed) (void)hipFree(p);
    if (status_h) (void)hipHostFree(status_h);
    if (owns_renderer) release(r);
  }
  uint32_t& mark(surface_obj* s) {
    for (auto& m : marks)
      if (m.first == s) return m.second;
    marks.emplace_back(s, 0u);
    return marks.back().second;
  }
  int32_t refuse(int32_t code, const char* msg) {
    if (error == SRX_OK) error = fail(code, msg);
    return code;
  }
};

void drop_recording(renderer_obj* r) {
  if (r->recording) release(r->recording);
  r->recording = nullptr;
}

template <class T>
int32_t batch_alloc(batch_obj* b, std::vector<void*>& owner, T** p, size_t n) {
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, bytes));
  owner.push_back(q);
  b->device_bytes += bytes;
  if (&owner == &b->tuned) b->tuned_bytes += bytes;
  *p = static_cast<T*>(q);
  return SRX_OK;
}

int32_t record_draw(renderer_obj* r, const draw_params& P, const ps_samplers& SM) {
  batch_obj* b = r->recording;
  if (P.prim_count == 0) return SRX_OK;
  const uint32_t T = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
  if (T == 0) return SRX_OK;
  if (!r->active.empty()) return b->refuse(SRX_E_UNSUPPORTED, "queries do not observe batched draws");
  if (sasl_shades(P)) return b->refuse(SRX_E_UNSUPPORTED, "batched draws do not run SASL pixel shaders");
  if (!(P.early_z || P.vis_stencil) || !(P.blend_id == SRX_BLEND_OFF || P.blend_id == SRX_BLEND_DISABLED) ||
      !vis_applicable(P))
    return b->refuse(SRX_E_UNSUPPORTED, "a batched draw must be opaque with an order-independent depth test");
  if (P.nattr > 2) return b->refuse(SRX_E_UNSUPPORTED, "batched draws use vertex shaders with at most 2 attributes");
  if (3ull * P.prim_count >= 0x40000000ull) return b->refuse(SRX_E_UNSUPPORTED, "batched draw primitive count");
  surface_obj* ts[2] = {r->st.color_targets.empty() ? nullptr : r->st.color_targets[0].get(), r->st.ds_target.get()};
  for (surface_obj* t : ts) {
    if (!t) continue;
    uint32_t& m = b->mark(t);
    if (m & (BS_DRAWN | BS_RESOLVE_SRC | BS_RESOLVE_DST | BS_SAMPLED))
      return b->refuse(SRX_E_UNSUPPORTED, "batched jobs must not share or reuse targets");
    m |= BS_DRAWN;
  }
  for (int i = 0; i < MAX_SAMPLERS; ++i) {
    sampler_obj* so = r->st.samplers[i].get();
    if (!so || !so->tex) continue;
    for (auto& sf : so->tex->surfs) {
      uint32_t& m = b->mark(sf.get());
      if (m & (BS_DRAWN | BS_FILLED | BS_RESOLVE_DST))
        return b->refuse(SRX_E_UNSUPPORTED, "a batched draw samples a surface the batch writes");
      m |= BS_SAMPLED;
    }
  }
  b->draws.push_back(batch_draw{P, SM, r->st});
  b->primitives += P.prim_count;
  return SRX_OK;
}

int32_t record_fill(renderer_obj* r, surface_obj* o, float4 v) {
  batch_obj* b = r->recording;
  const int es = pf_bytes(o->pf);
  if (es == 12) return b->refuse(SRX_E_UNSUPPORTED, "batched clears of 12-byte texels");
  uint32_t& m = b->mark(o);
  if (m & (BS_DRAWN | BS_RESOLVE_SRC | BS_RESOLVE_DST | BS_SAMPLED))
    return b->refuse(SRX_E_UNSUPPORTED, "a batched clear follows a use of its surface");
  m |= BS_FILLED;
  const size_t bytes = static_cast<size_t>(o->w) * o->h * o->samples * es;
  fill_item it{o->d, bytes / 16, static_cast<uint32_t>(bytes % 16), o->pf, v};
  for (auto& f : b->fills)
    if (f.s.get() == o) {  // a later clear of the same surface replaces the earlier one
      f.it = it;
      return SRX_OK;
    }
  b->fills.push_back(batch_fill{it, ref<surface_obj>(o)});
  return SRX_OK;
}

int32_t record_resolve(renderer_obj* r, surface_obj* a, surface_obj* d) {
  batch_obj* b = r->recording;
  uint32_t& ma = b->mark(a);
  if (ma & BS_RESOLVE_DST) return b->refuse(SRX_E_UNSUPPORTED, "batched resolve chains");
  ma |= BS_RESOLVE_SRC;
  uint32_t& md = b->mark(d);
  if (md & (BS_DRAWN | BS_RESOLVE_SRC | BS_RESOLVE_DST | BS_SAMPLED))
    return b->refuse(SRX_E_UNSUPPORTED, "a batched resolve target is used elsewhere in the batch");
  md |= BS_RESOLVE_DST;
  b->resolves.push_back(batch_resolve{resolve_item{a->desc(), d->desc()}, ref<surface_obj>(a), ref<surface_obj>(d)});
  return SRX_OK;
}

// GPU tiles of a draw's viewport for tiles of side gt (as build_draw derives them for 16).
void gpu_tiles(const draw_params& P, int gt, int32_t* gx, int32_t* gy) {
  const float cw = std::min(P.target_w, static_cast<float>(P.tiles_x * 64));
  const float ch = std::min(P.target_h, static_cast<float>(P.tiles_y * 64));
  *gx = static_cast<int32_t>(std::ceil(std::max(cw, 0.0f) / gt));
  *gy = static_cast<int32_t>(std::ceil(std::max(ch, 0.0f) / gt));
}

// Sets every job to GPU tiles of side b->tile and, unless done before for that size, runs the
// counting pass (vertex shader, setup, clipping, scan of the per-triangle record counts; one
// host synchronisation) whose per-draw record counts size the binning.
int32_t batch_count(batch_obj* b) {
  hipStream_t s = b->r->stream;
  const uint32_t J = static_cast<uint32_t>(b->draws.size());
  for (uint32_t j = 0; j < J; ++j) {
    batch_job& jb = b->jobs_h[j];
    jb.P.gt = b->tile;
    gpu_tiles(jb.P, b->tile, &jb.P.gtiles_x, &jb.P.gtiles_y);
    jb.tiles = static_cast<uint32_t>(jb.P.gtiles_x * jb.P.gtiles_y);
  }
  if (b->counted.count(b->tile)) {
    b->records = b->counted[b->tile];
  } else {
    RCHK(batch_alloc(b, b->tuned, &b->scan_tmp, scan_tmp_size(b->cnt_n)));
    RCHK(batch_alloc(b, b->tuned, &b->hist_all, 1));
    HIPCHK(hipMemcpyAsync(b->jobs_d, b->jobs_h.data(), sizeof(batch_job) * J, hipMemcpyHostToDevice, s));
    const uint32_t* mf = b->maps_fixed;
    launch_batch_reset(b->counters_all, J * CT_COUNT, b->cnt_all + (b->cnt_n - 1), b->hist_all, b->status_d, s);
    launch_vs_batch(b->jobs_d, map_image::at(mf, b->m_vs), b->m_vs.blocks, s);
    for (auto& e : b->m_setup) launch_setup_batch(e.first, b->jobs_d, map_image::at(mf, e.second), e.second.blocks, s);
    for (auto& e : b->m_clip) launch_clip_batch(e.first, b->jobs_d, map_image::at(mf, e.second), e.second.blocks, s);
    launch_exclusive_scan(b->cnt_all, b->cnt_n, b->scan_tmp, b->scan_total, s);
    launch_batch_counts(b->jobs_d, J, b->counts_d, s);
    HIPCHK(hipMemcpyAsync(b->status_h, b->counts_d, 2 * sizeof(uint32_t) * J, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    for (uint32_t j = 0; j < J; ++j)
      if (b->status_h[J + j]) return fail(SRX_OUT_OF_MEMORY, "clip buffer overflow");
    b->records.assign(b->status_h, b->status_h + J);
    b->counted[b->tile] = b->records;
  }
  uint64_t total = 0;
  for (uint32_t j = 0; j < J; ++j) total += b->records[j];
  if (total >= 0xFFFFFFFFull) return fail(SRX_E_UNSUPPORTED, "batch record count");
  b->capacity = static_cast<uint32_t>(total);
  return SRX_OK;
}

// Per-draw buffers, the job table, the fixed stage maps and the counting pass (vertex shader,
// setup, clipping and the scan of the per-triangle record counts), whose per-job record counts size
// the binning buffers.
int32_t batch_prepare(batch_obj* b) {
  renderer_obj* r = b->r;
  hipStream_t s = r->stream;
  const uint32_t J = static_cast<uint32_t>(b->draws.size());
  size_t n_vbuf = 0, n_rec = 0, n_attr = 0, n_clipv = 0, n_prims = 0, n_tiles = 0;
  uint64_t n_cnt = 0;
  for (auto& d : b->draws) {
    const draw_params& P = d.P;
    const size_t regs = static_cast<size_t>(P.nattr) + 1;
    n_vbuf += (static_cast<size_t>(P.nverts) + 1) * regs;
    n_rec += 15ull * P.prim_count;
    n_attr += 9ull * P.prim_count * regs;
    n_clipv += 5ull * P.prim_count * regs;
    n_prims += P.prim_count;
    n_cnt += 3ull * P.prim_count;
    int32_t gx8, gy8;  // tile arrays sized for the smallest GPU tiles
    gpu_tiles(P, 8, &gx8, &gy8);
    n_tiles += static_cast<size_t>(gx8) * gy8;
  }
  if (n_cnt + 1 >= 0xFFFFFFFFull) return fail(SRX_E_UNSUPPORTED, "batch primitive count");
  b->cnt_n = static_cast<uint32_t>(n_cnt + 1);
  float4 *vbuf, *tri_rec, *tri_attr, *clip_verts;
  uint8_t* clip_flag;
  int32_t* clip_slot;
  uint32_t *clip_hdr, *tiles3, *wide;
  ps_samplers* samplers;
  auto& F = b->fixed;
  RCHK(batch_alloc(b, F, &vbuf, n_vbuf));
  RCHK(batch_alloc(b, F, &tri_rec, n_rec));
  RCHK(batch_alloc(b, F, &tri_attr, n_attr));
  RCHK(batch_alloc(b, F, &clip_verts, n_clipv));
  RCHK(batch_alloc(b, F, &clip_flag, n_prims));
  RCHK(batch_alloc(b, F, &clip_slot, n_prims));
  RCHK(batch_alloc(b, F, &clip_hdr, n_prims));
  RCHK(batch_alloc(b, F, &wide, n_prims));
  RCHK(batch_alloc(b, F, &tiles3, 3 * n_tiles));
  RCHK(batch_alloc(b, F, &samplers, J));
  RCHK(batch_alloc(b, F, &b->cnt_all, b->cnt_n));
  RCHK(batch_alloc(b, F, &b->counters_all, static_cast<size_t>(J) * CT_COUNT));
  RCHK(batch_alloc(b, F, &b->counts_d, 2 * static_cast<size_t>(J)));
  RCHK(batch_alloc(b, F, &b->status_d, 4));
  RCHK(batch_alloc(b, F, &b->jobs_d, J));
  b->scan_total = b->status_d + 1;
  HIPCHK(hipHostMalloc(&b->status_h, std::max<size_t>(2 * static_cast<size_t>(J), 4) * sizeof(uint32_t), hipHostMallocDefault));

  std::vector<ps_samplers> sm(J);
  b->jobs_h.assign(J, batch_job{});
  size_t o_vbuf = 0, o_rec = 0, o_attr = 0, o_clipv = 0, o_prims = 0, o_tiles = 0, o_cnt = 0;
  std::vector<uint32_t> vs_blocks(J), setup_blocks(J), clip_blocks(J);
  for (uint32_t j = 0; j < J; ++j) {
    const draw_params& P = b->draws[j].P;
    batch_job& jb = b->jobs_h[j];
    std::memset(&jb, 0, sizeof jb);
    jb.P = P;
    sm[j] = b->draws[j].SM;
    const size_t regs = static_cast<size_t>(P.nattr) + 1;
    int32_t gx8, gy8;
    gpu_tiles(P, 8, &gx8, &gy8);
    const uint32_t T = static_cast<uint32_t>(gx8 * gy8);
    draw_buffers& B = jb.B;
    B.vbuf = vbuf + o_vbuf;
    B.tri_rec = tri_rec + o_rec;
    B.tri_attr = tri_attr + o_attr;
    B.clip_verts = clip_verts + o_clipv;
    B.clip_flag = clip_flag + o_prims;
    B.clip_slot = clip_slot + o_prims;
    B.clip_hdr = clip_hdr + o_prims;
    B.wide_list = wide + o_prims;
    B.clip_capacity = P.prim_count;
    B.cnt = b->cnt_all + o_cnt;
    B.counters = b->counters_all + static_cast<size_t>(j) * CT_COUNT;
    B.stats = r->stats_d;
    B.samplers = samplers + j;
    B.tstart = tiles3 + o_tiles;
    B.tend = tiles3 + n_tiles + o_tiles;
    B.order = tiles3 + 2 * n_tiles + o_tiles;
    jb.vis_mode = vis_mode(P);
    jb.tiles = static_cast<uint32_t>(P.gtiles_x * P.gtiles_y);
    o_vbuf += (static_cast<size_t>(P.nverts) + 1) * regs;
    o_rec += 15ull * P.prim_count;
    o_attr += 9ull * P.prim_count * regs;
    o_clipv += 5ull * P.prim_count * regs;
    o_prims += P.prim_count;
    o_tiles += T;
    o_cnt += 3ull * P.prim_count;
    vs_blocks[j] = (P.nverts + 1 + 255) / 256;
    setup_blocks[j] = (P.prim_count + 255) / 256;
    clip_blocks[j] = std::min<uint32_t>((P.prim_count + 63) / 64, 128);
  }
  HIPCHK(hipMemcpyAsync(samplers, sm.data(), sizeof(ps_samplers) * J, hipMemcpyHostToDevice, s));

  // fixed stage maps: jobs in recorded order
  std::vector<uint32_t> order(J);
  for (uint32_t j = 0; j < J; ++j) order[j] = j;
  map_image img;
  b->m_vs = img.add(vs_blocks, order);
  b->m_emit = img.add(setup_blocks, order);
  for (int na = 1; na <= 3; ++na) {
    std::vector<uint32_t> sb(J, 0), cb(J, 0);
    bool any = false;
    for (uint32_t j = 0; j < J; ++j)
      if (std::min(std::max(b->draws[j].P.nattr, 1), 3) == na) {
        sb[j] = setup_blocks[j];
        cb[j] = clip_blocks[j];
        any = true;
      }
    if (!any) continue;
    b->m_setup.emplace_back(na, img.add(sb, order));
    b->m_clip.emplace_back(na, img.add(cb, order));
  }
  std::vector<uint32_t> fill_blocks(b->fills.size()), fill_order(b->fills.size());
  std::vector<fill_item> fi(b->fills.size());
  for (size_t i = 0; i < b->fills.size(); ++i) {
    fi[i] = b->fills[i].it;
    fill_blocks[i] = static_cast<uint32_t>((std::max<size_t>(fi[i].words, 1) + 255) / 256);
    fill_order[i] = static_cast<uint32_t>(i);
  }
  b->m_fill = img.add(fill_blocks, fill_order);
  std::vector<uint32_t> res_blocks(b->resolves.size()), res_order(b->resolves.size());
  std::vector<resolve_item> ri(b->resolves.size());
  for (size_t i = 0; i < b->resolves.size(); ++i) {
    ri[i] = b->resolves[i].it;
    res_blocks[i] = static_cast<uint32_t>((static_cast<size_t>(ri[i].src.w) * ri[i].src.h + 255) / 256);
    res_order[i] = static_cast<uint32_t>(i);
  }
  b->m_resolve = img.add(res_blocks, res_order);
  RCHK(batch_alloc(b, F, &b->maps_fixed, img.words.size()));
  HIPCHK(hipMemcpyAsync(b->maps_fixed, img.words.data(), img.words.size() * 4, hipMemcpyHostToDevice, s));
  RCHK(batch_alloc(b, F, &b->fills_d, fi.size()));
  RCHK(batch_alloc(b, F, &b->resolves_d, ri.size()));
  if (!fi.empty()) HIPCHK(hipMemcpyAsync(b->fills_d, fi.data(), fi.size() * sizeof(fill_item), hipMemcpyHostToDevice, s));
  if (!ri.empty()) HIPCHK(hipMemcpyAsync(b->resolves_d, ri.data(), ri.size() * sizeof(resolve_item), hipMemcpyHostToDevice, s));
  b->tile = 16;  // the recorded draws' GPU tiles
  return batch_count(b);
}

// Every stage of the batch on stream s (captured into the batch's graph).
void batch_enqueue(batch_obj* b, hipStream_t s) {
  const uint32_t J = static_cast<uint32_t>(b->draws.size());
  const uint32_t* mf = b->maps_fixed;
  const uint32_t* mt = b->maps_tuned;
  launch_batch_reset(b->counters_all, J * CT_COUNT, b->cnt_all + (b->cnt_n - 1), b->hist_all + (b->hist_n - 1),
                     b->status_d, s);
  launch_fill_batch(b->fills_d, map_image::at(mf, b->m_fill), b->m_fill.blocks, s);
  launch_vs_batch(b->jobs_d, map_image::at(mf, b->m_vs), b->m_vs.blocks, s);
  for (auto& e : b->m_setup) launch_setup_batch(e.first, b->jobs_d, map_image::at(mf, e.second), e.second.blocks, s);
  for (auto& e : b->m_clip) launch_clip_batch(e.first, b->jobs_d, map_image::at(mf, e.second), e.second.blocks, s);
  launch_exclusive_scan(b->cnt_all, b->cnt_n, b->scan_tmp, b->scan_total, s);
  launch_emit_batch(b->jobs_d, map_image::at(mf, b->m_emit), b->m_emit.blocks, b->capacity, s);
  launch_bucket_batch(b->bin_pt, 0, b->jobs_d, map_image::at(mt, b->m_bucket), b->m_bucket.blocks, b->capacity, s);
  launch_exclusive_scan(b->hist_all, b->hist_n, b->scan_tmp, nullptr, s);
  launch_bucket_batch(b->bin_pt, 1, b->jobs_d, map_image::at(mt, b->m_bucket), b->m_bucket.blocks, b->capacity, s);
  launch_range_batch(b->jobs_d, map_image::at(mt, b->m_range), b->m_range.blocks, s);
  launch_tile_order_batch(b->jobs_d, J, s);
  for (auto& c : b->m_vis)
    launch_vis_batch(c.na, c.samples, c.shade, b->vis_threads, b->tile, b->jobs_d, map_image::at(mt, c.m), c.m.blocks, s);
  launch_resolve_batch(b->resolves_d, map_image::at(mf, b->m_resolve), b->m_resolve.blocks, s);
  launch_batch_check(b->jobs_d, J, b->capacity, b->bin_pt, b->status_d, s);
}

// Binning buffers and maps for the tuning (bin_pt, vis_threads), then the graph.
int32_t batch_configure(batch_obj* b, int tile, int bin_pt, int vis_threads) {
  if (!batch_bin_items_valid(bin_pt) || !(vis_threads == 64 || vis_threads == 128 || vis_threads == 256) ||
      !(tile == 8 || tile == 16 || tile == 32))
    return fail(SRX_INVALID_PARAMETER, "batch tuning");
  for (auto& d : b->draws) {
    int32_t gx, gy;
    gpu_tiles(d.P, tile, &gx, &gy);
    if (!vis_batch_supported(d.P.samples, needs_shading(d.P), vis_threads, tile) ||
        static_cast<uint64_t>(gx) * gy >= 0x40000000ull)
      return fail(SRX_E_UNSUPPORTED, "GPU tile size not available for every job of the batch");
  }
  renderer_obj* r = b->r;
  hipStream_t s = r->stream;
  HIPCHK(hipStreamSynchronize(s));
  b->free_tuned();
  b->tile = tile;
  RCHK(batch_count(b));
  b->free_tuned();
  b->bin_pt = bin_pt;
  b->vis_threads = vis_threads;
  const uint32_t J = static_cast<uint32_t>(b->draws.size());
  const uint32_t items = 256u * bin_pt;
  std::vector<uint32_t> bucket_blocks(J), range_blocks(J), order(J);
  uint64_t hist = 0;
  for (uint32_t j = 0; j < J; ++j) {
    batch_job& jb = b->jobs_h[j];
    jb.bucket_blocks = std::max<uint32_t>((b->records[j] + items - 1) / items, 1);
    bucket_blocks[j] = jb.bucket_blocks;
    range_blocks[j] = (jb.tiles + 255) / 256;
    hist += static_cast<uint64_t>(jb.tiles) * jb.bucket_blocks;
    order[j] = j;
  }
  if (hist + 1 >= 0xFFFFFFFFull) return fail(SRX_E_UNSUPPORTED, "batch histogram size");
  b->hist_n = static_cast<uint32_t>(hist + 1);
  uint32_t *keys0, *vals0, *vals1;
  RCHK(batch_alloc(b, b->tuned, &keys0, b->capacity));
  RCHK(batch_alloc(b, b->tuned, &vals0, b->capacity));
  RCHK(batch_alloc(b, b->tuned, &vals1, b->capacity));
  RCHK(batch_alloc(b, b->tuned, &b->hist_all, b->hist_n));
  RCHK(batch_alloc(b, b->tuned, &b->scan_tmp, std::max(scan_tmp_size(b->cnt_n), scan_tmp_size(b->hist_n))));
  uint64_t h = 0;
  for (uint32_t j = 0; j < J; ++j) {
    batch_job& jb = b->jobs_h[j];
    jb.B.keys[0] = keys0;
    jb.B.vals[0] = vals0;
    jb.B.keys[1] = nullptr;
    jb.B.vals[1] = vals1;
    jb.B.hist = b->hist_all + h;
    h += static_cast<uint64_t>(jb.tiles) * jb.bucket_blocks;
  }
  // visibility: one launch per kernel variant, the jobs with the most records per tile first
  std::vector<uint32_t> heavy(J);
  for (uint32_t j = 0; j < J; ++j) heavy[j] = j;
  std::stable_sort(heavy.begin(), heavy.end(), [&](uint32_t x, uint32_t y) {
    return static_cast<uint64_t>(b->records[x]) * b->jobs_h[y].tiles > static_cast<uint64_t>(b->records[y]) * b->jobs_h[x].tiles;
  });
  map_image img;
  b->m_bucket = img.add(bucket_blocks, order);
  b->m_range = img.add(range_blocks, order);
  b->m_vis.clear();
  for (uint32_t j : heavy) {
    const draw_params& P = b->draws[j].P;
    const int na = needs_shading(P) ? std::min(std::max(P.nattr, 1), 3) : 1;
    const bool shade = needs_shading(P);
    bool found = false;
    for (auto& c : b->m_vis) found |= c.na == na && c.samples == P.samples && c.shade == shade;
    if (!found) b->m_vis.push_back(batch_obj::vis_class{na, P.samples, shade, {}});
  }
  for (auto& c : b->m_vis) {
    std::vector<uint32_t> vb(J, 0);
    for (uint32_t j = 0; j < J; ++j) {
      const draw_params& P = b->draws[j].P;
      const bool shade = needs_shading(P);
      const int na = shade ? std::min(std::max(P.nattr, 1), 3) : 1;
      if (na == c.na && P.samples == c.samples && shade == c.shade) vb[j] = b->jobs_h[j].tiles;
    }
    c.m = img.add(vb, heavy);
  }
  RCHK(batch_alloc(b, b->tuned, &b->maps_tuned, img.words.size()));
  HIPCHK(hipMemcpyAsync(b->maps_tuned, img.words.data(), img.words.size() * 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(b->jobs_d, b->jobs_h.data(), sizeof(batch_job) * J, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  batch_enqueue(b, s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  b->graph = g;
  HIPCHK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return SRX_OK;
}

int32_t batch_refuse(renderer_obj* r, int32_t code, const char* msg) { return r->recording->refuse(code, msg); }

}  // namespace srx_rt

using namespace srx_rt;

extern "C" {

int32_t srx_batch_begin(srx_renderer rr) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || r->recording) return SRX_INVALID_PARAMETER;
  RCHK(run_pending_clears(r));  // clears issued before the batch run before it
  r->recording = new batch_obj();
  return SRX_OK;
}
int32_t srx_batch_end(srx_renderer rr, srx_batch* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !out || !r->recording) return SRX_INVALID_PARAMETER;
  batch_obj* b = r->recording;
  r->recording = nullptr;
  if (b->error != SRX_OK) {
    const int32_t e = b->error;
    release(b);
    return e;
  }
  retain(r);
  b->r = r;
  b->owns_renderer = true;
  int32_t rc = batch_prepare(b);
  if (rc == SRX_OK) rc = batch_configure(b, b->tile, b->bin_pt, b->vis_threads);
  if (rc != SRX_OK) {
    release(b);
    return rc;
  }
  *out = reinterpret_cast<srx_batch>(static_cast<object*>(b));
  return SRX_OK;
}
int32_t srx_batch_launch(srx_batch bb) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || !b->exec) return SRX_INVALID_PARAMETER;
  RCHK(run_pending_clears(b->r));  // clears issued before the launch run before it
  HIPCHK(hipGraphLaunch(b->exec, b->r->stream));
  return SRX_OK;
}
int32_t srx_batch_status(srx_batch bb, uint32_t* status) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || !status) return SRX_INVALID_PARAMETER;
  HIPCHK(hipMemcpyAsync(b->status_h, b->status_d, sizeof(uint32_t), hipMemcpyDeviceToHost, b->r->stream));
  HIPCHK(hipStreamSynchronize(b->r->stream));
  *status = b->status_h[0];
  return SRX_OK;
}
int32_t srx_batch_configure(srx_batch bb, int32_t tile, int32_t bin_items, int32_t vis_threads) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || bin_items % 256) return SRX_INVALID_PARAMETER;
  return batch_configure(b, tile, bin_items / 256, vis_threads);
}
int32_t srx_batch_autotune(srx_batch bb, uint32_t reps, srx_batch_tuning* results, size_t capacity, size_t* count) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || reps == 0) return SRX_INVALID_PARAMETER;
  RCHK(run_pending_clears(b->r));
  hipStream_t s = b->r->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  size_t n = 0;
  float best = 0.0f;
  int best_tile = b->tile, best_pt = b->bin_pt, best_th = b->vis_threads;
  int32_t rc = SRX_OK;
  for (int tile : {8, 16, 32}) {
    for (int pt : {8, 16, 32, 64}) {
      for (int th : {64, 128, 256}) {
        rc = batch_configure(b, tile, pt, th);
        if (rc == SRX_E_UNSUPPORTED) {  // a tile size some job cannot take: not a candidate
          rc = SRX_OK;
          continue;
        }
        if (rc != SRX_OK) break;
        hipError_t e = hipGraphLaunch(b->exec, s);  // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (uint32_t i = 0; i < reps && e == hipSuccess; ++i) e = hipGraphLaunch(b->exec, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          rc = hip_fail(e, "srx_batch_autotune");
          break;
        }
        ms /= static_cast<float>(reps);
        if (results && n < capacity) results[n] = srx_batch_tuning{tile, pt * 256, th, ms};
        ++n;
        if (n == 1 || ms < best) {
          best = ms;
          best_tile = tile;
          best_pt = pt;
          best_th = th;
        }
      }
      if (rc != SRX_OK) break;
    }
    if (rc != SRX_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (count) *count = n;
  if (rc != SRX_OK) return rc;
  return batch_configure(b, best_tile, best_pt, best_th);
}
int32_t srx_batch_get_info(srx_batch bb, srx_batch_info* info) {
  auto* b = as<batch_obj>(bb, K_BATCH);
  if (!b || !info) return SRX_INVALID_PARAMETER;
  info->jobs = static_cast<uint32_t>(b->draws.size());
  info->clears = static_cast<uint32_t>(b->fills.size());
  info->resolves = static_cast<uint32_t>(b->resolves.size());
  info->tile = static_cast<uint32_t>(b->tile);
  info->bin_items = static_cast<uint32_t>(b->bin_pt * 256);
  info->vis_threads = static_cast<uint32_t>(b->vis_threads);
  info->vis_launches = static_cast<uint32_t>(b->m_vis.size());
  info->primitives = b->primitives;
  info->records = b->capacity;
  info->device_bytes = b->device_bytes;
  return SRX_OK;
}
}  // extern "C"
