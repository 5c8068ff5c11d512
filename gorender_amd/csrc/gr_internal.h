// Library-internal access to a context for the serving loop (gr_serve.cpp).
// Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "gorender_hip.h"

namespace gr_internal {

// gr_draw_batch without the argument checks: queue one batch on the context
// stream. The stats of its frames then live at last_stats_device() until the
// next batch is queued.
gr_status enqueue_batch(gr_context* c, const gr_batch_frame* frames, int32_t nf,
                        const gr_object* objs, int32_t no);
hipStream_t context_stream(gr_context* c);
int context_device(gr_context* c);
const void* last_stats_device(gr_context* c);
size_t stats_bytes_per_frame();
void decode_stats(const void* frame_stats, gr_stats* out);
// Whether the frame had an overlay line too long to draw (GR_ERR_UNSUPPORTED).
bool frame_long_lines(const void* frame_stats);
// Sticky control words (device) and their size; decode a host copy of them.
void* control_device(gr_context* c);
size_t control_bytes();
void decode_control(const void* host_copy, bool* overflow, unsigned long long* refs_needed,
                    bool* unsupported);
// Tile-list capacity for the next batch, after an overflow that needed `needed`.
void grow_refs_to(gr_context* c, unsigned long long needed);
// The checks a frame must pass to be drawn (frame fields, mesh and texture
// handles, normal indices under Gouraud shading, slot bounds); the status
// and message a batch holding it would fail with.
gr_status check_request(gr_context* c, const gr_frame* frame, const gr_object* objects,
                        int32_t num_objects);
// Keep every batch on pipe 0 (the serving loop overlaps its own readback).
void set_one_pipe(gr_context* c, bool on);

// The no-allocation draw's fixed device array (gr_static.hip): resolve its
// address on `device` (current), and look it up.
gr_status static_scratch_init(int device);
void* static_scratch(int device, size_t* bytes);
// One frame drawn with every per-draw buffer carved from [base, base + bytes):
// no allocation of any kind on the way.
gr_status enqueue_noalloc(gr_context* c, const gr_frame* frame, const gr_object* objs, int32_t no,
                          void* d_rgba, void* d_depth, void* base, size_t bytes);

}  // namespace gr_internal
