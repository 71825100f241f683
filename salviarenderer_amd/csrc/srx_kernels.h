// Host-side launchers of the pipeline kernels (implemented in srx_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srx_device.h"

namespace srx {


// Device statistics slots (pipeline_statistics / internal_statistics, async_object.h:90-175)
enum stat_slot {
  ST_IA_VERTICES = 0, ST_IA_PRIMITIVES, ST_VS_INVOCATIONS, ST_CINVOCATIONS, ST_CPRIMITIVES,
  ST_PS_INVOCATIONS, ST_BACKEND_PIXELS, ST_COUNT
};
// Setup counters
// CT_WIDE: triangles whose records the clip launch counts by wave (their ids in wide_list)
// CT_CLASS: GPU tiles per record-count class (16 slots), CT_CURSOR: their placement cursors
enum counter_slot { CT_CLIP_PRIMS = 0, CT_CLIP_SLOTS, CT_CLIP_OVERFLOW, CT_TOTAL, CT_WIDE, CT_CLASS, CT_CURSOR = CT_CLASS + 16,
                    CT_CLIP_DONE = CT_CURSOR + 16, CT_COUNT };

// Per-draw device buffers (owned by the renderer, grown on demand).
struct draw_buffers {
  float4* vbuf;        // (nverts+1) * (nattr+1) clip-space vertex-shader outputs
  float4* tri_rec;     // 3*prims * 5: edges, bbox, depth plane of each clipped triangle
  float4* tri_attr;    // 3*prims * 3*(nattr+1): v0, ddx, ddy of every register
  uint32_t* vis;       // target w*h*samples: winning triangle per sample (opaque early-Z draws)
  const ps_samplers* samplers;  // device copy of the bound samplers
  uint32_t* cnt;       // 3*prims: GPU-tile records per clipped triangle (batches: then its exclusive scan)
  uint32_t* bsum;      // single draws: records of each 256-primitive block's triangles (null in batches)
  uint4* pairs;        // single draws, per primitive: fan triangle 0's first four (tile << 1 | full)
                       // records as setup found them, x = 0xFFFFFFFF when not kept (null in batches)
  uint8_t* clip_flag;  // prims
  int32_t* clip_slot;  // prims
  float4* clip_verts;  // clip_capacity * 5 * (nattr+1), screen space
  uint32_t* clip_hdr;  // clip_capacity
  uint32_t clip_capacity;
  uint32_t* wide_list; // prims: triangles over many 64x64 tiles, counted by the clip launch
  uint32_t* keys[2];   // records: GPU tile id
  uint32_t* vals[2];   // records: triangle id << 1 | tile-level accept
  const uint32_t* order;  // GPU tiles, most records first (workgroup launch order)
  uint4* vis_entries;     // visibility work table (vis_table_bounds)
  uint4* vis_merges;
  uint32_t* vis_counts;   // entries, merges, partial slots
  unsigned long long* vis_partial;  // partial per-sample keys of split tiles
  uint32_t* tstart;    // GPU tiles
  uint32_t* tile_count;  // 2 * GPU tiles: records per tile, then the scatter's per-tile cursors
  uint32_t* tend;
  uint32_t* hist;      // radix histogram 256 * blocks
  uint32_t* scan_tmp;  // scan partial sums
  uint32_t* counters;  // CT_COUNT
  unsigned long long* stats;    // ST_COUNT
  unsigned long long* partial;  // per-workgroup statistics partials (>= max(setup blocks, GPU tiles))
};

// ---------------------------------------------------------------- batched jobs
// Workgroup -> job map of one batched launch: job[b] for every workgroup b; range[j] = (first
// workgroup, workgroups) of job j (no workgroups: the job has no work in this launch).
struct stage_map {
  const uint32_t* job;
  const uint2* range;
};
// One draw of a batch: its parameters and its slices of the batch's buffers.  B.keys/B.vals are the
// batch-wide record arrays; B.cnt is the job's slice of the batch-wide per-triangle counts (followed
// by the next job's slice or the trailing total slot); B.hist likewise.
struct batch_job {
  draw_params P;
  draw_buffers B;
  int32_t vis_mode;
  uint32_t tiles;          // GPU tiles
  uint32_t bucket_blocks;  // binning workgroups; the histogram is tiles x bucket_blocks
  uint32_t pad;
};
// A clear as a 16-byte pattern fill of `value` converted to format pf (texel sizes 4, 8, 16):
// `words` 16-byte stores, then `tail` bytes (whole texels).  A depth/stencil clear is the RG32F
// value (depth, stencil bits).
struct fill_item {
  uint8_t* data;
  size_t words;
  uint32_t tail;
  int32_t pf;
  float4 value;
};
struct resolve_item {
  surface_desc src, dst;
};

void launch_batch_reset(uint32_t* counters, uint32_t n, uint32_t* cnt_tail, uint32_t* hist_tail, uint32_t* status,
                        hipStream_t s);
void launch_fill_batch(const fill_item* items, stage_map m, uint32_t blocks, hipStream_t s);
void launch_resolve_batch(const resolve_item* items, stage_map m, uint32_t blocks, hipStream_t s);
void launch_vs_batch(const batch_job* jobs, stage_map m, uint32_t blocks, hipStream_t s);
void launch_setup_batch(int na, const batch_job* jobs, stage_map m, uint32_t blocks, hipStream_t s);
void launch_clip_batch(int na, const batch_job* jobs, stage_map m, uint32_t blocks, hipStream_t s);
void launch_emit_batch(const batch_job* jobs, stage_map m, uint32_t blocks, uint32_t capacity, hipStream_t s);
// binning with pt records per thread (256 * pt per workgroup): phase 0 histograms, 1 scatter
bool batch_bin_items_valid(int pt);
void launch_bucket_batch(int pt, int phase, const batch_job* jobs, stage_map m, uint32_t blocks, uint32_t capacity,
                         hipStream_t s);
void launch_range_batch(const batch_job* jobs, stage_map m, uint32_t blocks, hipStream_t s);
void launch_tile_order_batch(const batch_job* jobs, uint32_t njobs, hipStream_t s);
// threads: 64, 128 or 256 lanes per GPU tile; gt: GPU tile side 8, 16 or 32 (8 and 32 for single-sample
// draws).  False for a combination without a kernel.
bool vis_batch_supported(int samples, bool shade, int threads, int gt);
bool launch_vis_batch(int na, int samples, bool shade, int threads, int gt, const batch_job* jobs, stage_map m,
                      uint32_t blocks, hipStream_t s);
void launch_batch_check(const batch_job* jobs, uint32_t njobs, uint32_t capacity, int pt, uint32_t* status, hipStream_t s);
// out[j]: records of job j; out[njobs + j]: its clip-overflow counter
void launch_batch_counts(const batch_job* jobs, uint32_t njobs, uint32_t* out, hipStream_t s);
int vis_mode(const draw_params& P);

hipError_t upload_ewa_table(const float* table256, hipStream_t s);

void launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
// The vertex launch of a single draw; it also zeroes the draw's counters and ntiles words of tiles.
void launch_vs(const draw_params& P, float4* vbuf, uint32_t* counters, uint32_t* tiles, size_t ntiles, hipStream_t s);
void launch_setup(const draw_params& P, const draw_buffers& B, hipStream_t s);
void launch_clip(const draw_params& P, const draw_buffers& B, hipStream_t s);
// In-place exclusive scan of n u32; writes the total to *total_dev.  tmp holds >= scan_tmp_size(n).
size_t scan_tmp_size(uint32_t n);
void launch_exclusive_scan(uint32_t* data, uint32_t n, uint32_t* tmp, uint32_t* total_dev, hipStream_t s);
void launch_emit(const draw_params& P, const draw_buffers& B, uint32_t capacity, hipStream_t s);
// Stable LSD radix sort of (keys, vals) by the low `bits` bits of keys; returns the buffer index
// (0 or 1) that holds the result.
int launch_sort(const draw_buffers& B, uint32_t n, int bits, hipStream_t s);
// Sorts the GPU tiles by record count, heaviest first, into O.vals[result]; O supplies keys/vals
// scratch (>= tiles each) and shares hist/scan_tmp with B.
// GPU tile scheduling order, heaviest record-count class first.  `counted`: the bucket sort's range
// pass has already counted the tiles per class.
void launch_tile_order(const draw_params& P, const draw_buffers& B, uint32_t* order, bool counted, hipStream_t s);
// One-pass (unstable within a tile) bucket sort of the n emitted records by GPU tile into vals[1],
// with the tile ranges; false when the tile count exceeds its LDS histogram (then radix-sort).
// Counting sort of the records by GPU tile (order-independent draws) into vals[1] with the tile ranges
// in tstart/tend, sized for `cap` records; the kernels read the record count from
// B.counters[CT_TOTAL] (at most cap are sorted).  B.tile_count must be zero.  With `order`, the
// tiles' scheduling order (as launch_tile_order writes it) is written too.
bool bucket_sort_applicable(const draw_params& P);
bool launch_bucket_sort(const draw_params& P, const draw_buffers& B, uint32_t cap, uint32_t* order, hipStream_t s);
uint32_t bucket_hist_size(uint32_t T, uint32_t n);
// Whether launch_vis will take the draw (its visibility mode exists and the tiles are 16x16).
bool vis_applicable(const draw_params& P);
void launch_tile_ranges(const draw_params& P, const draw_buffers& B, int which, uint32_t n, hipStream_t s);
void launch_raster(const draw_params& P, const draw_buffers& B, int which, hipStream_t s);
// Order-independent visibility for opaque early-Z draws without statistics; false if not applicable.
// Order-independent visibility (and, with shade, the deferred pixel shader) of an opaque early-Z
// draw; false when the draw needs the in-order kernel.
bool launch_vis(const draw_params& P, const draw_buffers& B, int which, bool shade, uint32_t records, hipStream_t s);
// Upper bounds of the visibility work table for T GPU tiles and `records` binned records: entries
// (first launch), merges (second launch) and partial-key slots (256 * samples keys each).
void vis_table_bounds(uint32_t T, uint32_t records, uint32_t* entries, uint32_t* merges, uint32_t* slots);
bool needs_shading(const draw_params& P);
// A SASL pixel shader whose colour is written: such draws run the in-order raster kernel.
bool sasl_shades(const draw_params& P);
void launch_sasl_quads(const sasl_program* prog, const ps_samplers* SM, uint32_t has_mask, const float4* in, float4* out,
                       uint32_t nquads, hipStream_t s);
void launch_shade(const draw_params& P, const draw_buffers& B, hipStream_t s);

void launch_clear(const surface_desc& s, float4 value, hipStream_t st);
// flags: 1 depth, 2 stencil
void launch_clear_ds(const surface_desc& s, uint32_t flags, float depth, uint32_t stencil, hipStream_t st);
void launch_resolve(const surface_desc& src, const surface_desc& dst, hipStream_t st);
void launch_mip(const surface_desc& src, const surface_desc& dst, int filter, hipStream_t st);

}  // namespace srx
