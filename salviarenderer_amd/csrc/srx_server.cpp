// Request server: continuous batching of render requests over the batched kernels.
#include "srx_runtime_internal.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

namespace srx_rt {

// ---------------------------------------------------------------- request server
// Serving loop over the batched kernels: requests (a resident mesh drawn with the pass-through
// vertex shader, Gouraud pixel shader and depth LESS into a fresh target of the requested size and
// sample count, then resolved) are admitted by their device footprint -- geometry scratch, bins
// and targets -- into batches that a dispatcher thread launches on several streams as earlier
// batches are still running (continuous batching).  Meshes and textures stay resident in an LRU
// pool.  Every request's result equals its one-by-one render.

// First-fit allocator over one allocation (offsets 256-byte aligned, free ranges coalesced).
struct range_arena {
  uint8_t* base = nullptr;
  size_t size = 0, used = 0, peak = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
  void init(uint8_t* b, size_t n) {
    base = b;
    size = n;
    free_.clear();
    if (n) free_[0] = n;
  }
  static size_t round(size_t n) { return (std::max<size_t>(n, 1) + 255) & ~static_cast<size_t>(255); }
  size_t largest() const {
    size_t m = 0;
    for (auto& f : free_) m = std::max(m, f.second);
    return m;
  }
  bool alloc(size_t n, size_t* off) {
    n = round(n);
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, at = it->first + n;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used += n;
        peak = std::max(peak, used);
        return true;
      }
    return false;
  }
  void release(size_t off, size_t n) {
    n = round(n);
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct srv_asset {
  uint8_t* d = nullptr;
  size_t bytes = 0;
  uint64_t stamp = 0;  // last use (LRU)
  uint32_t users = 0;  // queued or running requests
  uint32_t n_tris = 0, size = 0;
};

struct srv_job {
  uint64_t ticket = 0;
  srx_request req{};
  int64_t admitted_ns = 0, done_ns = 0;
  uint32_t est_records = 0, batch_jobs = 0;
  bool exact = false;  // est_records is a measured count
  int32_t status = SRX_OK;
  size_t stage_off = 0, stage_bytes = 0;  // host staging of the outputs
  size_t color_bytes = 0, depth_bytes = 0;
};

struct srv_batch {
  std::vector<srv_job> jobs;
  size_t dev_off = 0, dev_bytes = 0, host_off = 0, host_bytes = 0;
  uint32_t* counts_h = nullptr;  // pinned: records[J], clip overflow[J]
  uint32_t capacity = 0;
  std::vector<uint32_t> bucket_blocks;
  int stream = 0;
  hipEvent_t done = nullptr;
};

struct server_obj : object {
  renderer_obj* r = nullptr;  // retained
  srx_server_desc desc{};
  std::vector<hipStream_t> streams;
  std::vector<bool> stream_busy;
  hipStream_t upload = nullptr;
  uint8_t* dev = nullptr;
  uint8_t* host = nullptr;
  range_arena dev_arena, host_arena;
  std::mutex mu;
  std::condition_variable cv;
  std::thread worker;
  bool stop = false;
  int32_t fatal = SRX_OK;
  std::string fatal_msg;
  std::chrono::steady_clock::time_point epoch = std::chrono::steady_clock::now();
  uint64_t next_ticket = 1, clock = 0;
  std::unordered_map<uint64_t, srv_asset> meshes, textures;
  size_t pool_used = 0;
  std::deque<srv_job> queue;
  std::vector<srv_batch> running;
  std::unordered_map<uint64_t, srv_job> finished;  // awaiting srx_server_release
  std::vector<uint64_t> unreported;
  std::unordered_map<uint64_t, uint32_t> records_cache;  // (mesh, tiles) -> exact records
  srx_server_stats stats{};
  std::vector<hipEvent_t> events;
  server_obj() : object(K_SERVER) {}
  ~server_obj() override;
  int64_t now() const {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - epoch).count();
  }
};

// Records depend on the mesh and the target size (the sample count is kept apart too).
uint64_t cache_key(const srx_request& q) {
  return (q.mesh_key * 0x9E3779B97F4A7C15ull) ^ (static_cast<uint64_t>(q.width) << 36 | static_cast<uint64_t>(q.height) << 8 | q.samples);
}

// draw_params of a served request: the state build_draw derives for this pipeline (pass-through
// vertex shader over POSITION/COLOR float4 pairs, Gouraud, cull none, depth LESS with writes,
// stencil off, blend off, full-target viewport).
void serve_params(draw_params& P, const uint8_t* mesh, uint32_t n, uint32_t w, uint32_t h, uint32_t samples,
                  uint8_t* color, uint8_t* ds) {
  std::memset(&P, 0, sizeof P);
  P.vs_id = SRX_VS_PASSTHROUGH_COLOR;
  P.nattr = 1;
  for (int i = 0; i < MAX_ATTRS; ++i) P.attr_mod[i] = SRX_AM_LINEAR;
  const uint64_t bytes = 96ull * n;
  for (int i = 0; i < 2; ++i) {
    input_elem_dev& e = P.in[P.nin++];
    e.base = mesh + 16 * i;
    e.avail = bytes - 16 * i;
    e.stride = 32;
    e.fmt = SRX_FMT_R32G32B32A32_FLOAT;
    e.reg = i;
    e.default_w = i == 0 ? 1.0f : 0.0f;
  }
  P.nverts = 3 * n;
  P.topo = SRX_TOPOLOGY_TRIANGLE_LIST;
  P.prim_count = n;
  P.cull_mode = SRX_CULL_NONE;
  P.vp[2] = static_cast<float>(w);
  P.vp[3] = static_cast<float>(h);
  P.vp[5] = 1.0f;
  P.target_w = static_cast<float>(w);
  P.target_h = static_cast<float>(h);
  P.samples = static_cast<int32_t>(samples);
  static const float p1[1][2] = {{0.5f, 0.5f}};
  static const float p2[2][2] = {{0.25f, 0.25f}, {0.75f, 0.75f}};
  static const float p4[4][2] = {{0.375f, 0.125f}, {0.875f, 0.375f}, {0.125f, 0.625f}, {0.625f, 0.875f}};
  const float(*pat)[2] = samples == 1 ? p1 : samples == 2 ? p2 : p4;
  for (uint32_t s = 0; s < samples; ++s) {
    P.sp[s][0] = pat[s][0];
    P.sp[s][1] = pat[s][1];
  }
  P.tiles_x = static_cast<int32_t>((w + 63) / 64);
  P.tiles_y = static_cast<int32_t>((h + 63) / 64);
  P.gt = 16;
  P.gtiles_x = static_cast<int32_t>((w + 15) / 16);
  P.gtiles_y = static_cast<int32_t>((h + 15) / 16);
  P.color = surface_desc{color, static_cast<int32_t>(w), static_cast<int32_t>(h), static_cast<int32_t>(samples), PF_RGBA8};
  P.ds = surface_desc{ds, static_cast<int32_t>(w), static_cast<int32_t>(h), static_cast<int32_t>(samples), PF_RG32F};
  P.depth_enable = 1;
  P.depth_func = SRX_CMP_LESS;
  P.read_depth = 1;
  P.write_depth = 1;
  P.read_mask = 0xFF;
  P.write_mask = 0xFF;
  P.st_func[0] = P.st_func[1] = SRX_CMP_ALWAYS;
  P.st_pass_op[0] = P.st_pass_op[1] = SRX_SOP_KEEP;
  P.early_z = 1;
  P.blend_id = SRX_BLEND_OFF;
  P.ps_id = SRX_PS_GOURAUD;
}

// Device bytes of one job with `records` bin entries and its share of the batch-wide arrays.
struct job_layout {
  size_t color, ds, resolved, vbuf, tri_rec, tri_attr, clip_verts, clip_flag, clip_slot, clip_hdr, wide, tiles, hist, total;
};
job_layout layout_of(const srx_request& q, uint32_t records, uint32_t bin_items) {
  const size_t px = static_cast<size_t>(q.width) * q.height, n = q.n_tris, T = ((q.width + 15) / 16) * ((q.height + 15) / 16);
  const size_t nb = std::max<size_t>((records + bin_items - 1) / bin_items, 1);
  job_layout L;
  size_t o = 0;
  auto put = [&](size_t bytes) {
    const size_t at = o;
    o += range_arena::round(bytes);
    return at;
  };
  L.color = put(px * q.samples * 4);
  L.ds = put(px * q.samples * 8);
  L.resolved = put(q.samples > 1 ? px * 4 : 0);
  L.vbuf = put((3 * n + 1) * 2 * 16);
  L.tri_rec = put(15 * n * 16);
  L.tri_attr = put(9 * n * 2 * 16);
  L.clip_verts = put(5 * n * 2 * 16);
  L.clip_flag = put(n);
  L.clip_slot = put(4 * n);
  L.clip_hdr = put(4 * n);
  L.wide = put(4 * n);
  L.tiles = put(12 * T);
  L.hist = put(4 * T * nb);
  L.total = o;
  return L;
}

uint32_t estimate_records(server_obj* S, const srx_request& q, bool* exact) {
  auto it = S->records_cache.find(cache_key(q));
  *exact = it != S->records_cache.end();
  if (*exact) return it->second;
  const uint64_t T = ((q.width + 15) / 16) * ((q.height + 15) / 16);
  return static_cast<uint32_t>(std::min<uint64_t>(16ull * q.n_tris + T, 0x7FFFFFFFull));
}

void unuse_assets(server_obj* S, const srx_request& q) {
  auto m = S->meshes.find(q.mesh_key);
  if (m != S->meshes.end() && m->second.users) --m->second.users;
  if (q.texture_key) {
    auto t = S->textures.find(q.texture_key);
    if (t != S->textures.end() && t->second.users) --t->second.users;
  }
}

// Frees least recently used assets nobody waits for until `need` more bytes fit the pool.
int32_t make_room(server_obj* S, size_t need) {
  while (S->pool_used + need > S->desc.pool_bytes) {
    std::unordered_map<uint64_t, srv_asset>* from = nullptr;
    uint64_t key = 0, oldest = UINT64_MAX;
    for (auto* m : {&S->meshes, &S->textures})
      for (auto& a : *m)
        if (!a.second.users && a.second.stamp < oldest) {
          oldest = a.second.stamp;
          key = a.first;
          from = m;
        }
    if (!from) return fail(SRX_OUT_OF_MEMORY, "resident pool full of assets in use");
    srv_asset& a = (*from)[key];
    HIPCHK(hipStreamSynchronize(S->upload));
    HIPCHK(hipFree(a.d));
    S->pool_used -= a.bytes;
    ++S->stats.evictions;
    from->erase(key);
  }
  return SRX_OK;
}

int32_t put_mesh(server_obj* S, uint64_t key, const void* data, uint32_t n) {
  auto it = S->meshes.find(key);
  if (it != S->meshes.end()) {
    if (it->second.n_tris != n) return fail(SRX_INVALID_PARAMETER, "mesh key reused with another triangle count");
    it->second.stamp = ++S->clock;
    return SRX_OK;
  }
  if (!data) return fail(SRX_INVALID_PARAMETER, "mesh not resident and no data given");
  const size_t bytes = 96ull * n;
  RCHK(make_room(S, bytes));
  srv_asset a;
  a.bytes = bytes;
  a.n_tris = n;
  a.stamp = ++S->clock;
  HIPCHK(hipMalloc(&a.d, std::max<size_t>(bytes, 1)));
  HIPCHK(hipMemcpyAsync(a.d, data, bytes, hipMemcpyHostToDevice, S->upload));
  HIPCHK(hipStreamSynchronize(S->upload));
  S->pool_used += bytes;
  ++S->stats.uploads;
  S->meshes[key] = a;
  return SRX_OK;
}

// RGBA8 level 0 of size^2 texels; the box-filtered chain is generated on the device.
int32_t put_texture(server_obj* S, uint64_t key, const void* data, uint32_t size) {
  auto it = S->textures.find(key);
  if (it != S->textures.end()) {
    it->second.stamp = ++S->clock;
    return SRX_OK;
  }
  if (!data || size == 0) return fail(SRX_INVALID_PARAMETER, "texture not resident and no data given");
  size_t bytes = 0;
  for (uint32_t s = size; s; s >>= 1) bytes += static_cast<size_t>(s) * s * 4;
  RCHK(make_room(S, bytes));
  srv_asset a;
  a.bytes = bytes;
  a.size = size;
  a.stamp = ++S->clock;
  HIPCHK(hipMalloc(&a.d, bytes));
  HIPCHK(hipMemcpyAsync(a.d, data, static_cast<size_t>(size) * size * 4, hipMemcpyHostToDevice, S->upload));
  size_t off = 0;
  for (uint32_t s = size; s > 1; s >>= 1) {
    const uint32_t h = s / 2;
    surface_desc src{a.d + off, static_cast<int32_t>(s), static_cast<int32_t>(s), 1, PF_RGBA8};
    off += static_cast<size_t>(s) * s * 4;
    surface_desc dst{a.d + off, static_cast<int32_t>(h), static_cast<int32_t>(h), 1, PF_RGBA8};
    launch_mip(src, dst, SRX_FILTER_LINEAR, S->upload);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(S->upload));
  S->pool_used += bytes;
  ++S->stats.uploads;
  S->textures[key] = a;
  return SRX_OK;
}

// Stage-map words one job adds to a batch: a (first, count) pair per stage and the job id of each
// of its workgroups (vertex, setup/emit, clip, binning, ranges, visibility, two clears, resolve).
uint64_t job_map_words(const srx_request& q, uint32_t est, uint32_t items) {
  const uint64_t n = q.n_tris, T = ((q.width + 15) / 16) * ((q.height + 15) / 16), px = static_cast<uint64_t>(q.width) * q.height;
  const uint64_t nb = std::max<uint64_t>((est + items - 1) / items, 1);
  const uint64_t fill_c = (std::max<uint64_t>(px * q.samples * 4 / 16, 1) + 255) / 256;
  const uint64_t fill_d = (std::max<uint64_t>(px * q.samples * 8 / 16, 1) + 255) / 256;
  return 2 * 12 + (3 * n + 1 + 255) / 256 + (n + 255) / 256 + std::min<uint64_t>((n + 63) / 64, 128) + nb +
         (T + 255) / 256 + T + fill_c + fill_d + (q.samples > 1 ? (px + 255) / 256 : 0);
}

// Device bytes a batch needs besides its jobs' own ranges.
size_t batch_shared_bytes(size_t jobs, uint64_t cnt_n, uint64_t capacity, uint64_t hist_n, uint64_t map_words) {
  const size_t scan = std::max(scan_tmp_size(static_cast<uint32_t>(cnt_n)), scan_tmp_size(static_cast<uint32_t>(hist_n)));
  return range_arena::round(4 * cnt_n) + range_arena::round(4 * CT_COUNT * jobs) + range_arena::round(8 * jobs + 16) +
         3 * range_arena::round(4 * capacity) + range_arena::round(4 * hist_n) + range_arena::round(4 * scan) +
         range_arena::round(sizeof(batch_job) * jobs) + range_arena::round(sizeof(ps_samplers) * jobs) +
         range_arena::round(4 * (map_words + 16)) + range_arena::round(sizeof(fill_item) * 2 * jobs) +
         range_arena::round(sizeof(resolve_item) * jobs);
}
// Pinned bytes of a batch: the tables it uploads, the counts it reads back, and the outputs.
size_t batch_host_bytes(size_t jobs, uint64_t map_words, size_t outputs) {
  return range_arena::round(sizeof(batch_job) * jobs) + range_arena::round(4 * (map_words + 16)) +
         range_arena::round(sizeof(fill_item) * 2 * jobs) + range_arena::round(sizeof(resolve_item) * jobs) +
         range_arena::round(8 * jobs) + outputs;
}
size_t output_bytes(const srx_request& q) {
  const size_t px = static_cast<size_t>(q.width) * q.height;
  return range_arena::round(px * 4) + (q.want_depth ? range_arena::round(px * q.samples * 8) : 0);
}

// Builds and launches one batch from the queue head on stream si (called with the lock held):
// arrived requests are admitted in order while batch slots and arena space last.
int32_t launch_served_batch(server_obj* S, int si) {
  const int64_t now = S->now();
  const uint32_t items = static_cast<uint32_t>(S->desc.bin_items);
  std::vector<srv_job> take;
  std::vector<job_layout> lays;
  size_t dev_jobs = 0, host_out = 0;
  uint64_t cnt_n = 1, cap = 0, hist_n = 1, words = 0;
  const size_t dev_room = S->dev_arena.largest(), host_room = S->host_arena.largest();
  while (!S->queue.empty() && take.size() < S->desc.max_batch_jobs) {
    srv_job& q = S->queue.front();
    if (q.req.arrival_ns > now) break;
    if (!q.exact) q.est_records = estimate_records(S, q.req, &q.exact);  // measured since submission?
    const uint32_t est = q.est_records;
    const job_layout L = layout_of(q.req, est, items);
    const uint64_t n = q.req.n_tris, T = ((q.req.width + 15) / 16) * ((q.req.height + 15) / 16);
    const uint64_t nb = std::max<uint64_t>((est + items - 1) / items, 1);
    const uint64_t w2 = words + job_map_words(q.req, est, items);
    const size_t J2 = take.size() + 1;
    const size_t dev_need = dev_jobs + L.total + batch_shared_bytes(J2, cnt_n + 3 * n, cap + est, hist_n + T * nb, w2);
    const size_t host_need = batch_host_bytes(J2, w2, host_out + output_bytes(q.req));
    if (dev_need > dev_room || host_need > host_room || cap + est > 0xFFFFFFF0ull || cnt_n + 3 * n > 0xFFFFFFF0ull ||
        hist_n + T * nb > 0xFFFFFFF0ull) {
      if (take.empty() && S->running.empty()) {  // can never fit
        q.status = fail(SRX_OUT_OF_MEMORY, "request larger than the server's arenas");
        q.done_ns = now;
        unuse_assets(S, q.req);
        S->unreported.push_back(q.ticket);
        S->finished[q.ticket] = q;
        S->queue.pop_front();
        continue;
      }
      break;
    }
    dev_jobs += L.total;
    host_out += output_bytes(q.req);
    cnt_n += 3 * n;
    cap += est;
    hist_n += T * nb;
    words = w2;
    take.push_back(q);
    lays.push_back(L);
    S->queue.pop_front();
  }
  if (take.empty()) return SRX_OK;
  // Records are laid out job after job, so a job whose estimate falls short pushes the later ones
  // out of the batch's capacity: measured jobs go first, estimated ones last.
  {
    std::vector<size_t> idx(take.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_partition(idx.begin(), idx.end(), [&](size_t i) { return take[i].exact; });
    std::vector<srv_job> t2;
    std::vector<job_layout> l2;
    for (size_t i : idx) {
      t2.push_back(take[i]);
      l2.push_back(lays[i]);
    }
    take.swap(t2);
    lays.swap(l2);
  }
  const uint32_t J = static_cast<uint32_t>(take.size());
  auto give_back = [&] {
    for (auto it = take.rbegin(); it != take.rend(); ++it) S->queue.push_front(*it);
  };
  srv_batch b;
  b.stream = si;
  b.capacity = static_cast<uint32_t>(cap);
  b.dev_bytes = dev_jobs + batch_shared_bytes(J, cnt_n, cap, hist_n, words);
  b.host_bytes = batch_host_bytes(J, words, host_out);
  if (!S->dev_arena.alloc(b.dev_bytes, &b.dev_off)) {
    give_back();
    return SRX_OK;
  }
  if (!S->host_arena.alloc(b.host_bytes, &b.host_off)) {
    S->dev_arena.release(b.dev_off, b.dev_bytes);
    give_back();
    return SRX_OK;
  }
  hipStream_t s = S->streams[si];
  uint8_t* D = S->dev + b.dev_off;
  size_t off = 0;
  auto take_dev = [&](size_t bytes) {
    uint8_t* p = D + off;
    off += range_arena::round(bytes);
    return p;
  };
  // per-job ranges
  std::vector<batch_job> jobs(J);
  std::vector<fill_item> fills;
  std::vector<resolve_item> resolves;
  std::vector<const uint8_t*> out_color(J), out_ds(J);
  for (uint32_t j = 0; j < J; ++j) {
    const srx_request& q = take[j].req;
    const job_layout& L = lays[j];
    uint8_t* base = take_dev(L.total);
    batch_job& jb = jobs[j];
    std::memset(&jb, 0, sizeof jb);
    serve_params(jb.P, S->meshes[q.mesh_key].d, q.n_tris, q.width, q.height, q.samples, base + L.color, base + L.ds);
    jb.tiles = static_cast<uint32_t>(jb.P.gtiles_x * jb.P.gtiles_y);
    jb.vis_mode = vis_mode(jb.P);
    jb.bucket_blocks = std::max<uint32_t>((take[j].est_records + items - 1) / items, 1);
    draw_buffers& B = jb.B;
    B.vbuf = reinterpret_cast<float4*>(base + L.vbuf);
    B.tri_rec = reinterpret_cast<float4*>(base + L.tri_rec);
    B.tri_attr = reinterpret_cast<float4*>(base + L.tri_attr);
    B.clip_verts = reinterpret_cast<float4*>(base + L.clip_verts);
    B.clip_flag = base + L.clip_flag;
    B.clip_slot = reinterpret_cast<int32_t*>(base + L.clip_slot);
    B.clip_hdr = reinterpret_cast<uint32_t*>(base + L.clip_hdr);
    B.wide_list = reinterpret_cast<uint32_t*>(base + L.wide);
    B.clip_capacity = q.n_tris;
    uint32_t* t3 = reinterpret_cast<uint32_t*>(base + L.tiles);
    B.tstart = t3;
    B.tend = t3 + jb.tiles;
    B.order = t3 + 2 * jb.tiles;
    B.stats = S->r->stats_d;
    const size_t px = static_cast<size_t>(q.width) * q.height;
    uint32_t zero = 0;
    float stencil_bits;
    std::memcpy(&stencil_bits, &zero, 4);
    fills.push_back(fill_item{base + L.color, px * q.samples * 4 / 16, static_cast<uint32_t>(px * q.samples * 4 % 16), PF_RGBA8,
                              make_float4(q.clear_color[0], q.clear_color[1], q.clear_color[2], q.clear_color[3])});
    fills.push_back(fill_item{base + L.ds, px * q.samples * 8 / 16, static_cast<uint32_t>(px * q.samples * 8 % 16), PF_RG32F,
                              make_float4(1.0f, stencil_bits, 0.0f, 0.0f)});
    if (q.samples > 1)
      resolves.push_back(resolve_item{jb.P.color, surface_desc{base + L.resolved, static_cast<int32_t>(q.width),
                                                               static_cast<int32_t>(q.height), 1, PF_RGBA8}});
    out_color[j] = base + (q.samples > 1 ? L.resolved : L.color);
    out_ds[j] = base + L.ds;
  }
  // batch-wide arrays
  uint32_t* cnt_all = reinterpret_cast<uint32_t*>(take_dev(4 * cnt_n));
  uint32_t* counters_all = reinterpret_cast<uint32_t*>(take_dev(4 * CT_COUNT * J));
  uint32_t* counts_d = reinterpret_cast<uint32_t*>(take_dev(8 * J + 16));
  uint32_t* keys0 = reinterpret_cast<uint32_t*>(take_dev(4 * cap));
  uint32_t* vals0 = reinterpret_cast<uint32_t*>(take_dev(4 * cap));
  uint32_t* vals1 = reinterpret_cast<uint32_t*>(take_dev(4 * cap));
  uint32_t* hist_all = reinterpret_cast<uint32_t*>(take_dev(4 * hist_n));
  uint32_t* scan_tmp = reinterpret_cast<uint32_t*>(
      take_dev(4 * std::max(scan_tmp_size(static_cast<uint32_t>(cnt_n)), scan_tmp_size(static_cast<uint32_t>(hist_n)))));
  batch_job* jobs_d = reinterpret_cast<batch_job*>(take_dev(sizeof(batch_job) * J));
  ps_samplers* sm_d = reinterpret_cast<ps_samplers*>(take_dev(sizeof(ps_samplers) * J));
  std::vector<uint32_t> vs_b(J), setup_b(J), clip_b(J), bucket_b(J), range_b(J), order(J);
  uint64_t co = 0, ho = 0;
  for (uint32_t j = 0; j < J; ++j) {
    batch_job& jb = jobs[j];
    jb.B.cnt = cnt_all + co;
    jb.B.counters = counters_all + CT_COUNT * j;
    jb.B.samplers = sm_d + j;
    jb.B.keys[0] = keys0;
    jb.B.vals[0] = vals0;
    jb.B.vals[1] = vals1;
    jb.B.hist = hist_all + ho;
    co += 3ull * jb.P.prim_count;
    ho += static_cast<uint64_t>(jb.tiles) * jb.bucket_blocks;
    vs_b[j] = (jb.P.nverts + 1 + 255) / 256;
    setup_b[j] = (jb.P.prim_count + 255) / 256;
    clip_b[j] = std::min<uint32_t>((jb.P.prim_count + 63) / 64, 128);
    bucket_b[j] = jb.bucket_blocks;
    range_b[j] = (jb.tiles + 255) / 256;
    order[j] = j;
  }
  std::vector<uint32_t> heavy(order);
  std::stable_sort(heavy.begin(), heavy.end(), [&](uint32_t x, uint32_t y) {
    return static_cast<uint64_t>(take[x].est_records) * jobs[y].tiles > static_cast<uint64_t>(take[y].est_records) * jobs[x].tiles;
  });
  map_image img;
  const map_image::entry m_vs = img.add(vs_b, order), m_setup = img.add(setup_b, order), m_clip = img.add(clip_b, order);
  const map_image::entry m_bucket = img.add(bucket_b, order), m_range = img.add(range_b, order);
  std::vector<std::pair<int, map_image::entry>> m_vis;
  for (int smp : {1, 2, 4}) {
    std::vector<uint32_t> vb(J, 0);
    bool any = false;
    for (uint32_t j = 0; j < J; ++j)
      if (jobs[j].P.samples == smp) {
        vb[j] = jobs[j].tiles;
        any = true;
      }
    if (any) m_vis.emplace_back(smp, img.add(vb, heavy));
  }
  std::vector<uint32_t> fill_b(fills.size()), fill_o(fills.size()), res_b(resolves.size()), res_o(resolves.size());
  for (size_t i = 0; i < fills.size(); ++i) {
    fill_b[i] = static_cast<uint32_t>((std::max<size_t>(fills[i].words, 1) + 255) / 256);
    fill_o[i] = static_cast<uint32_t>(i);
  }
  for (size_t i = 0; i < resolves.size(); ++i) {
    res_b[i] = static_cast<uint32_t>((static_cast<size_t>(resolves[i].src.w) * resolves[i].src.h + 255) / 256);
    res_o[i] = static_cast<uint32_t>(i);
  }
  const map_image::entry m_fill = img.add(fill_b, fill_o), m_res = img.add(res_b, res_o);
  const size_t nwords = img.words.size();
  uint32_t* maps_d = reinterpret_cast<uint32_t*>(take_dev(4 * (words + 16)));
  fill_item* fills_d = reinterpret_cast<fill_item*>(take_dev(sizeof(fill_item) * 2 * J));
  resolve_item* res_d = reinterpret_cast<resolve_item*>(take_dev(sizeof(resolve_item) * J));
  if (off > b.dev_bytes || nwords > words + 16) {  // the reservation is an upper bound by construction
    S->dev_arena.release(b.dev_off, b.dev_bytes);
    S->host_arena.release(b.host_off, b.host_bytes);
    give_back();
    return fail(SRX_FAILED, "served batch layout exceeds its reservation");
  }
  // pinned staging: tables, read-back counts, outputs
  uint8_t* H = S->host + b.host_off;
  size_t hoff = 0;
  auto take_host = [&](size_t bytes) {
    uint8_t* p = H + hoff;
    hoff += range_arena::round(bytes);
    return p;
  };
  uint8_t* h_jobs = take_host(sizeof(batch_job) * J);
  uint8_t* h_maps = take_host(4 * (words + 16));
  uint8_t* h_fills = take_host(sizeof(fill_item) * 2 * J);
  uint8_t* h_res = take_host(sizeof(resolve_item) * J);
  b.counts_h = reinterpret_cast<uint32_t*>(take_host(8 * J));
  std::memcpy(h_jobs, jobs.data(), sizeof(batch_job) * J);
  std::memcpy(h_maps, img.words.data(), 4 * nwords);
  std::memcpy(h_fills, fills.data(), sizeof(fill_item) * fills.size());
  if (!resolves.empty()) std::memcpy(h_res, resolves.data(), sizeof(resolve_item) * resolves.size());
  HIPCHK(hipMemcpyAsync(jobs_d, h_jobs, sizeof(batch_job) * J, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(maps_d, h_maps, 4 * nwords, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(fills_d, h_fills, sizeof(fill_item) * fills.size(), hipMemcpyHostToDevice, s));
  if (!resolves.empty()) HIPCHK(hipMemcpyAsync(res_d, h_res, sizeof(resolve_item) * resolves.size(), hipMemcpyHostToDevice, s));
  // the pipeline of batch_enqueue, launched directly (every served batch differs, so no graph)
  const uint32_t* mf = maps_d;
  launch_batch_reset(counters_all, CT_COUNT * J, cnt_all + (cnt_n - 1), hist_all + (hist_n - 1), counts_d + 2 * J, s);
  launch_fill_batch(fills_d, map_image::at(mf, m_fill), m_fill.blocks, s);
  launch_vs_batch(jobs_d, map_image::at(mf, m_vs), m_vs.blocks, s);
  launch_setup_batch(1, jobs_d, map_image::at(mf, m_setup), m_setup.blocks, s);
  launch_clip_batch(1, jobs_d, map_image::at(mf, m_clip), m_clip.blocks, s);
  launch_exclusive_scan(cnt_all, static_cast<uint32_t>(cnt_n), scan_tmp, nullptr, s);
  launch_emit_batch(jobs_d, map_image::at(mf, m_setup), m_setup.blocks, b.capacity, s);
  const int pt = static_cast<int>(items / 256);
  launch_bucket_batch(pt, 0, jobs_d, map_image::at(mf, m_bucket), m_bucket.blocks, b.capacity, s);
  launch_exclusive_scan(hist_all, static_cast<uint32_t>(hist_n), scan_tmp, nullptr, s);
  launch_bucket_batch(pt, 1, jobs_d, map_image::at(mf, m_bucket), m_bucket.blocks, b.capacity, s);
  launch_range_batch(jobs_d, map_image::at(mf, m_range), m_range.blocks, s);
  launch_tile_order_batch(jobs_d, J, s);
  for (auto& v : m_vis)
    launch_vis_batch(1, v.first, true, S->desc.vis_threads, 16, jobs_d, map_image::at(mf, v.second), v.second.blocks, s);
  launch_resolve_batch(res_d, map_image::at(mf, m_res), m_res.blocks, s);
  launch_batch_counts(jobs_d, J, counts_d, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(b.counts_h, counts_d, 8 * J, hipMemcpyDeviceToHost, s));
  for (uint32_t j = 0; j < J; ++j) {
    srv_job& t = take[j];
    const srx_request& q = t.req;
    const size_t px = static_cast<size_t>(q.width) * q.height;
    t.color_bytes = px * 4;
    t.depth_bytes = q.want_depth ? px * q.samples * 8 : 0;
    t.stage_off = b.host_off + hoff;
    HIPCHK(hipMemcpyAsync(take_host(t.color_bytes), out_color[j], t.color_bytes, hipMemcpyDeviceToHost, s));
    if (t.depth_bytes) HIPCHK(hipMemcpyAsync(take_host(t.depth_bytes), out_ds[j], t.depth_bytes, hipMemcpyDeviceToHost, s));
    t.stage_bytes = b.host_off + hoff - t.stage_off;
    t.admitted_ns = now;
    t.batch_jobs = J;
  }
  if (S->events.empty()) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    S->events.push_back(e);
  }
  b.done = S->events.back();
  S->events.pop_back();
  HIPCHK(hipEventRecord(b.done, s));
  b.bucket_blocks = bucket_b;
  b.jobs = std::move(take);
  S->stream_busy[si] = true;
  S->running.push_back(std::move(b));
  ++S->stats.batches;
  S->stats.max_batch_jobs = std::max<uint64_t>(S->stats.max_batch_jobs, J);
  return SRX_OK;
}

// A finished batch: results of complete jobs are reported, jobs whose records outgrew their
// estimate go back to the head of the queue with the exact count.
void finish_batch(server_obj* S, srv_batch& b) {
  const uint32_t J = static_cast<uint32_t>(b.jobs.size());
  const uint32_t items = static_cast<uint32_t>(S->desc.bin_items);
  const int64_t now = S->now();
  uint64_t end = 0;
  std::vector<srv_job> retry;
  for (uint32_t j = 0; j < J; ++j) {
    srv_job& t = b.jobs[j];
    const uint32_t rec = b.counts_h[j];
    end += rec;
    const bool ok = end <= b.capacity && rec <= b.bucket_blocks[j] * items && !b.counts_h[J + j];
    S->records_cache[cache_key(t.req)] = rec;
    if (!ok && !b.counts_h[J + j]) {
      S->host_arena.release(t.stage_off, t.stage_bytes);
      t.est_records = rec;
      t.exact = true;
      ++S->stats.retries;
      retry.push_back(t);
      continue;
    }
    t.status = b.counts_h[J + j] ? SRX_OUT_OF_MEMORY : SRX_OK;
    t.done_ns = now;
    unuse_assets(S, t.req);
    ++S->stats.jobs;
    S->unreported.push_back(t.ticket);
    S->finished[t.ticket] = t;
  }
  for (auto it = retry.rbegin(); it != retry.rend(); ++it) S->queue.push_front(*it);
  // the tables and counts precede the outputs in the host region; outputs stay until released
  const size_t outs_begin = J ? b.jobs[0].stage_off - b.host_off : 0;
  S->host_arena.release(b.host_off, outs_begin);
  S->dev_arena.release(b.dev_off, b.dev_bytes);
  S->events.push_back(b.done);
  S->stream_busy[b.stream] = false;
}

void serve_loop(server_obj* S) {
  std::unique_lock<std::mutex> lk(S->mu);
  while (!S->stop) {
    bool progress = false;
    for (size_t i = 0; i < S->running.size();) {
      hipError_t e = hipEventQuery(S->running[i].done);
      if (e == hipSuccess) {
        finish_batch(S, S->running[i]);
        S->running.erase(S->running.begin() + static_cast<long>(i));
        progress = true;
      } else if (e == hipErrorNotReady) {
        ++i;
      } else {
        S->fatal = hip_fail(e, "served batch");
        S->fatal_msg = g_last_error_str;
        S->stop = true;
        break;
      }
    }
    if (progress) S->cv.notify_all();
    for (size_t si = 0; si < S->streams.size() && !S->stop; ++si) {
      if (S->stream_busy[si] || S->queue.empty() || S->queue.front().req.arrival_ns > S->now()) continue;
      const size_t before = S->queue.size();
      int32_t rc = launch_served_batch(S, static_cast<int>(si));
      if (rc != SRX_OK) {
        S->fatal = rc;
        S->fatal_msg = g_last_error_str;
        S->stop = true;
      }
      if (S->queue.size() != before) progress = true;
      if (!S->stream_busy[si]) break;  // nothing admitted: wait for space
    }
    if (!progress) {
      // sleep until a batch may have finished or the next request arrives
      int64_t wait_ns = S->running.empty() ? 2000000 : 20000;
      if (!S->queue.empty() && S->running.size() < S->streams.size())
        wait_ns = std::min<int64_t>(wait_ns, std::max<int64_t>(S->queue.front().req.arrival_ns - S->now(), 0));
      S->cv.wait_for(lk, std::chrono::nanoseconds(std::max<int64_t>(wait_ns, 5000)));
    }
  }
  S->cv.notify_all();
}

server_obj::~server_obj() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stop = true;
  }
  cv.notify_all();
  if (worker.joinable()) worker.join();
  for (auto s : streams) (void)hipStreamSynchronize(s);
  if (upload) (void)hipStreamSynchronize(upload);
  for (auto* m : {&meshes, &textures})
    for (auto& a : *m) (void)hipFree(a.second.d);
  for (auto e : events) (void)hipEventDestroy(e);
  for (auto& b : running)
    if (b.done) (void)hipEventDestroy(b.done);
  if (dev) (void)hipFree(dev);
  if (host) (void)hipHostFree(host);
  for (auto s : streams) (void)hipStreamDestroy(s);
  if (upload) (void)hipStreamDestroy(upload);
  release(r);
}

}  // namespace srx_rt

using namespace srx_rt;

extern "C" {

void srx_default_server_desc(srx_server_desc* d) {
  *d = srx_server_desc{200ull << 30, 64ull << 30, 2ull << 30, 64, 3, 8192, 64};
}
int32_t srx_create_server(srx_renderer rr, const srx_server_desc* d, srx_server* out) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r || !d || !out || d->streams == 0 || d->streams > 8 || d->max_batch_jobs == 0 || !batch_bin_items_valid(d->bin_items / 256) ||
      d->bin_items % 256 || !(d->vis_threads == 64 || d->vis_threads == 128 || d->vis_threads == 256))
    return SRX_INVALID_PARAMETER;
  auto* S = new server_obj();
  S->desc = *d;
  retain(r);
  S->r = r;
  hipError_t e = hipMalloc(&S->dev, std::max<uint64_t>(d->arena_bytes, 256));
  if (e == hipSuccess) e = hipHostMalloc(&S->host, std::max<uint64_t>(d->staging_bytes, 256), hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->upload, hipStreamNonBlocking);
  for (uint32_t i = 0; i < d->streams && e == hipSuccess; ++i) {
    hipStream_t st = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) S->streams.push_back(st);
  }
  if (e != hipSuccess) {
    const int32_t rc = hip_fail(e, "srx_create_server");
    release(S);
    return rc;
  }
  S->stream_busy.assign(S->streams.size(), false);
  S->dev_arena.init(S->dev, d->arena_bytes);
  S->host_arena.init(S->host, d->staging_bytes);
  S->worker = std::thread(serve_loop, S);
  *out = reinterpret_cast<srx_server>(static_cast<object*>(S));
  return SRX_OK;
}
int32_t srx_server_put_mesh(srx_server ss, uint64_t key, const void* vertices, uint32_t n_tris) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || n_tris == 0 || n_tris >= 0x15555555u) return SRX_INVALID_PARAMETER;
  std::lock_guard<std::mutex> lk(S->mu);
  return put_mesh(S, key, vertices, n_tris);
}
int32_t srx_server_put_texture(srx_server ss, uint64_t key, const void* rgba8, uint32_t size) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || key == 0 || size == 0 || size > 16384 || (size & (size - 1))) return SRX_INVALID_PARAMETER;
  std::lock_guard<std::mutex> lk(S->mu);
  return put_texture(S, key, rgba8, size);
}
int32_t srx_server_submit(srx_server ss, const srx_request* q, uint64_t* ticket) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !q || !ticket) return SRX_INVALID_PARAMETER;
  if (!(q->samples == 1 || q->samples == 2 || q->samples == 4) || q->width == 0 || q->height == 0 || q->n_tris == 0 ||
      q->n_tris >= 0x15555555u)
    return SRX_INVALID_PARAMETER;
  if (q->width > 32768 || q->height > 32768) return fail(SRX_E_UNSUPPORTED, "served target wider or taller than 32768");
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->fatal != SRX_OK) return fail(S->fatal, S->fatal_msg.c_str());
  // each asset is held as soon as it is resident, so the next upload cannot evict it
  RCHK(put_mesh(S, q->mesh_key, q->mesh_data, q->n_tris));
  ++S->meshes[q->mesh_key].users;
  if (q->texture_key) {
    const int32_t rc = put_texture(S, q->texture_key, q->texture_data, q->texture_size);
    if (rc != SRX_OK) {
      --S->meshes[q->mesh_key].users;
      return rc;
    }
    ++S->textures[q->texture_key].users;
  }
  srv_job t;
  t.ticket = S->next_ticket++;
  t.req = *q;
  t.req.mesh_data = nullptr;
  t.req.texture_data = nullptr;
  t.est_records = estimate_records(S, *q, &t.exact);
  auto pos = S->queue.end();  // ordered by arrival time, ties in submission order
  while (pos != S->queue.begin() && std::prev(pos)->req.arrival_ns > q->arrival_ns) --pos;
  S->queue.insert(pos, t);
  *ticket = t.ticket;
  S->cv.notify_all();
  return SRX_OK;
}
int32_t srx_server_poll(srx_server ss, srx_completion* out, size_t capacity, size_t* count, int32_t wait_ms) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !count || (capacity && !out)) return SRX_INVALID_PARAMETER;
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->unreported.empty() && wait_ms > 0 && S->fatal == SRX_OK)
    S->cv.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !S->unreported.empty() || S->stop; });
  size_t n = 0;
  while (n < capacity && !S->unreported.empty()) {
    const uint64_t tk = S->unreported.front();
    S->unreported.erase(S->unreported.begin());
    auto it = S->finished.find(tk);
    if (it == S->finished.end()) continue;  // released before it was reported
    const srv_job& t = it->second;
    srx_completion& c = out[n++];
    c.ticket = tk;
    c.status = t.status;
    c.batch_jobs = t.batch_jobs;
    c.arrival_ns = t.req.arrival_ns;
    c.admitted_ns = t.admitted_ns;
    c.done_ns = t.done_ns;
    c.color = t.stage_bytes ? S->host + t.stage_off : nullptr;
    c.color_bytes = t.stage_bytes ? t.color_bytes : 0;
    c.depth = t.depth_bytes && t.stage_bytes ? S->host + t.stage_off + range_arena::round(t.color_bytes) : nullptr;
    c.depth_bytes = c.depth ? t.depth_bytes : 0;
  }
  *count = n;
  if (S->fatal != SRX_OK) return fail(S->fatal, S->fatal_msg.c_str());
  return SRX_OK;
}
int32_t srx_server_release(srx_server ss, uint64_t ticket) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S) return SRX_INVALID_PARAMETER;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->finished.find(ticket);
  if (it == S->finished.end()) return SRX_INVALID_PARAMETER;
  if (it->second.stage_bytes) S->host_arena.release(it->second.stage_off, it->second.stage_bytes);
  S->finished.erase(it);
  S->unreported.erase(std::remove(S->unreported.begin(), S->unreported.end(), ticket), S->unreported.end());
  S->cv.notify_all();
  return SRX_OK;
}
int64_t srx_server_now(srx_server ss) {
  auto* S = as<server_obj>(ss, K_SERVER);
  return S ? S->now() : -1;
}
int32_t srx_server_get_stats(srx_server ss, srx_server_stats* st) {
  auto* S = as<server_obj>(ss, K_SERVER);
  if (!S || !st) return SRX_INVALID_PARAMETER;
  std::lock_guard<std::mutex> lk(S->mu);
  *st = S->stats;
  st->pool_bytes = S->pool_used;
  st->resident_meshes = S->meshes.size();
  st->resident_textures = S->textures.size();
  st->arena_peak_bytes = S->dev_arena.peak;
  st->queued = S->queue.size();
  st->running = S->running.size();
  return SRX_OK;
}
}  // extern "C"
