// Launch parameters shared by the host orchestration and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gr_device.h"

namespace gr {

// Device memory of one batch. The host uploads frames, objects and the zeroed
// counters (stats, frame_overflow, batch counters) with one copy.
struct LaunchParams {
    const DevObject* objs;
    int nobj;
    const DevFrame* frames;
    int nframes;
    const DevTexture* textures;
    int face_blocks;         // k_face / k_bin_fill blocks
    int total_tiles;         // raster workgroups
    int raster_persistent;   // 1: k_raster as a fixed grid striding over the tiles
    int tw_log2, th_log2;    // raster tile size
    uint32_t* tile_counts;   // [total_tiles]; zero between draws (k_raster re-zeroes)
    uint32_t* tile_offsets;  // [total_tiles]
    uint32_t* tile_cursor;   // [total_tiles]
    uint32_t* tile_slab;     // [total_tiles]: slab of a split tile or kNoSlab
    RasterItem* items;       // [max_items]: later parts of split tiles
    uint32_t max_items;
    unsigned long long* slabs;  // [max_slabs][tile pixels] keys; zero between draws
    uint32_t max_slabs;
    uint32_t part_refs;      // tile-list entries per raster part
    uint32_t* frame_overflow;// [nframes]
    uint32_t* wg_count;      // [face_blocks]
    uint32_t* rec_base;      // [face_blocks]: first overflow record of each face block
    uint32_t* rec_pool;      // [kRecPools * kPoolStride]: pool counters; zero between draws
    uint32_t rec_pools;      // overflow pools in use (1..kRecPools)
    uint32_t pool_recs;      // records per pool: pool j is
                             // [ovf_begin + j * pool_recs, ovf_begin + (j + 1) * pool_recs)
    uint32_t ovf_begin;      // face_blocks * kRecsPerBlock
    uint32_t* refs;          // [refs_capacity]
    unsigned long long refs_capacity;
    unsigned long long* stats;  // [nframes * kNumStats]
    DevBatchCounters* counters;
    DevControl* ctrl;
    // XCD-aware face-block order: xcd_face[x * (nframes + 1) + f] = face blocks
    // of frames < f that XCD x runs (blocks x, x + 8, ... of every frame); the
    // grid is 8 * the longest XCD list.
    const int* xcd_face;
    int face_grid;
};

// Host words carried to device memory by one k_put launch (its argument).
constexpr int kPutWords = 496;
struct PutChunk {
    uint32_t n, pad;
    uint32_t w[kPutWords];
};

// Prefix tables of the XCD-aware order for per-frame unit counts `n`.
void xcd_tables(const int* n, int nframes, int* table, int* grid);

void launch_face_normals(const float* pos, const int32_t* vi, int F, float* fn, hipStream_t s);
void launch_face(const LaunchParams& p, hipStream_t s);
void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
void launch_check_div(uint64_t n, uint64_t seed, unsigned long long* mismatches, float* first_bad,
                      hipStream_t s);
void launch_bin_scan(const LaunchParams& p, hipStream_t s);
void launch_overlay(const LaunchParams& p, hipStream_t s);
void launch_decorate(const LaunchParams& p, hipStream_t s);
void launch_fog(void* rgba, const void* depth, uint64_t n, float fog_start, float fog_end,
                uint32_t color, hipStream_t s);
void launch_bin_fill(const LaunchParams& p, hipStream_t s);
bool launch_raster_parts(const LaunchParams& p, hipStream_t s);
bool launch_raster(const LaunchParams& p, hipStream_t s);
void launch_fill32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
// Adds an order-independent digest of `words` 32-bit words at p to *out.
void launch_checksum(const void* p, size_t words, unsigned long long* out, hipStream_t s);
// Copies `bytes` host bytes to dst (4-byte aligned) through kernel arguments.
void launch_put(void* dst, const void* src, size_t bytes, hipStream_t s);

}  // namespace gr
