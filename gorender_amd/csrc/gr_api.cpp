// C ABI of the render path: context, resident meshes/textures in HBM, and the
// per-draw orchestration of the five kernels on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "gorender_hip.h"
#include "gr_error.h"
#include "gr_internal.h"
#include "gr_device.h"
#include "gr_launch.h"

namespace {

thread_local std::string g_last_error;

gr_status fail(gr_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

#define GR_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(e_ == hipErrorOutOfMemory ? GR_ERR_OUT_OF_MEMORY : GR_ERR_DEVICE, \
                        std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct MeshRes {
    bool live = false;
    gr::DevMesh dm{};
    std::vector<void*> allocs;
    int64_t bytes = 0;
    bool ni_valid = true;
    std::vector<int32_t> textures;  // distinct texture handles referenced by faces
};

struct TexRes {
    bool live = false;
    gr::DevTexture dt{};
    void* texels = nullptr;
    int64_t bytes = 0;
};

struct Staging {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};

struct TimedLaunch {
    const char* name;
    hipEvent_t a, b;
};

struct Pipe {
    hipStream_t stream = nullptr;
    // batch: frames | objects | stats | counters | frame_overflow | XCD tables,
    // uploaded with one copy per draw (the counters as zeros).
    DevBuf batch, rrec, srec, wrec, tile_counts, tiles, wgcount, refs, ctrl;
    DevBuf items;  // raster work items (later parts of split tiles)
    DevBuf slabs;  // depth keys of split tiles; zero between draws
    DevBuf okey;   // overlay keys of the frames that draw overlays; zero between draws
    DevBuf orec;   // overlay records per triangle record (batches with overlay frames)
    DevBuf rbase;  // record pool counters (zero between draws), then each face block's first
                   // overflow record
    hipEvent_t raster_done = nullptr;  // after this pipe's last k_raster
    bool raster_pending = false;
};

}  // namespace



struct gr_context {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::vector<MeshRes> meshes;
    std::vector<TexRes> textures;
    DevBuf d_textures;
    bool tex_dirty = true;

    // Per-batch device state, twice: consecutive batches alternate between
    // the two pipes (two streams), so one batch's face and binning kernels
    // overlap the previous batch's raster. Each k_raster waits for the
    // previous batch's, so frames are written in submission order.
    Pipe pipes[2];
    int last_pipe = -1;  // pipe of the last batch
    bool one_pipe = false;  // synchronous draws: pipe 0 only
    hipStream_t own1 = nullptr;  // pipe 1's stream
    DevBuf out_rgba, out_depth;
    Staging ring[4];
    int ring_i = 0;
    void* h_ctrl = nullptr;   // pinned: sticky control words
    void* h_frame = nullptr;  // pinned: gr_draw's colour + depth on their way to the caller
    size_t h_frame_cap = 0;
    int32_t last_frames = 0;
    const unsigned long long* last_stats = nullptr;  // device stats of the last draw
    // Control words of the no-allocation draw (in its fixed device array),
    // once this context has used it.
    gr::DevControl* noalloc_ctrl = nullptr;
    // Live batch graphs (their control words are checked by gr_synchronize).
    std::vector<gr_batch_graph*> graphs;
    const uint32_t* last_overflow = nullptr;  // device overflow flags of the last draw's frames
    const gr::DevBatchCounters* last_counters = nullptr;  // device counters of the last enqueue

    // The present path (gr_present_draw / gr_present_wait): two frames in
    // flight, each drawn into device buffers and copied to pinned host memory.
    struct PresentSlot {
        DevBuf rgba, depth;
        void* host = nullptr;     // pinned: colour, then depth
        size_t host_cap = 0;
        void* h_meta = nullptr;   // pinned: stats of the frame, then its overflow flag
        hipEvent_t copied = nullptr;
        bool pending = false;
        bool with_depth = false;
        gr_frame frame{};
        std::vector<gr_object> objs;
    } present[2];
    int present_next = 0;  // slot of the next gr_present_draw
    int present_head = 0;  // oldest queued slot

    gr_config cfg{32, 32, 1 << 22, 4, 512, 2, 1};

    bool timing = false;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int64_t>> kernel_ms;
};

namespace {

int ilog2(int v) {
    int r = 0;
    while ((1 << r) < v) r++;
    return r;
}

gr_status sync_all(gr_context* c) {
    GR_HIP(hipStreamSynchronize(c->stream));
    for (Pipe& pp : c->pipes)
        if (pp.stream && pp.stream != c->stream) GR_HIP(hipStreamSynchronize(pp.stream));
    return GR_OK;
}

// Order work the caller queues on the context stream (fills, copies of its
// buffers) with the draws: it starts after the last batch's raster, and the
// next batch's raster starts after it.
gr_status after_draws(gr_context* c) {
    if (c->last_pipe >= 0 && c->pipes[c->last_pipe].raster_pending &&
        c->pipes[c->last_pipe].stream != c->stream)
        GR_HIP(hipStreamWaitEvent(c->stream, c->pipes[c->last_pipe].raster_done, 0));
    return GR_OK;
}
gr_status before_draws(gr_context* c) {
    Pipe& p0 = c->pipes[0];
    p0.stream = c->stream;
    if (!p0.raster_done) GR_HIP(hipEventCreateWithFlags(&p0.raster_done, hipEventDisableTiming));
    GR_HIP(hipEventRecord(p0.raster_done, c->stream));
    p0.raster_pending = true;
    c->last_pipe = 0;
    return GR_OK;
}

gr_status ensure(gr_context* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (b.cap >= bytes) return GR_OK;
    if (b.p) {
        gr_status st = sync_all(c);
        if (st) return st;
        GR_HIP(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    size_t cap = std::max(bytes, b.cap + b.cap / 2);
    cap = (cap + 4095) & ~size_t(4095);
    GR_HIP(hipMalloc(&b.p, cap));
    b.cap = cap;
    return GR_OK;
}

gr_status staging_get(gr_context* c, size_t bytes, void** out, int* slot) {
    int i = c->ring_i;
    c->ring_i = (c->ring_i + 1) % 4;
    Staging& s = c->ring[i];
    if (s.pending) {
        GR_HIP(hipEventSynchronize(s.ev));
        s.pending = false;
    }
    if (s.cap < bytes) {
        if (s.host) GR_HIP(hipHostFree(s.host));
        s.host = nullptr;
        size_t cap = std::max(bytes, size_t(1) << 16);
        GR_HIP(hipHostMalloc(&s.host, cap, hipHostMallocDefault));
        s.cap = cap;
    }
    if (!s.ev) GR_HIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    *out = s.host;
    *slot = i;
    return GR_OK;
}

gr_status staging_mark(gr_context* c, int slot, hipStream_t s) {
    GR_HIP(hipEventRecord(c->ring[slot].ev, s));
    c->ring[slot].pending = true;
    return GR_OK;
}

gr_status upload(gr_context* c, void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return GR_OK;
    void* h;
    int slot;
    gr_status st = staging_get(c, bytes, &h, &slot);
    if (st) return st;
    std::memcpy(h, src, bytes);
    GR_HIP(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s));
    return staging_mark(c, slot, s);
}

hipEvent_t pool_event(gr_context* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

template <typename F>
void timed_launch(gr_context* c, const char* name, hipStream_t s, F f) {
    if (!c->timing) {
        f();
        return;
    }
    hipEvent_t a = pool_event(c), b = pool_event(c);
    if (!a || !b) {
        f();
        return;
    }
    (void)hipEventRecord(a, s);
    f();
    (void)hipEventRecord(b, s);
    c->timed.push_back({name, a, b});
}

void collect_times(gr_context* c) {
    for (auto& t : c->timed) {
        float ms = 0.f;
        if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            auto& e = c->kernel_ms[t.name];
            e.first += ms;
            e.second += 1;
        }
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->timed.clear();
}

uint32_t pack_rgba(const uint8_t c[4]) {
    return (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16) | ((uint32_t)c[3] << 24);
}

// renderer.go:50-76
void ref_tile_bounds(int tile, int num_tiles, int w, int h, float out[4]) {
    if (num_tiles == 1) {
        out[0] = 0.f;
        out[1] = 0.f;
        out[2] = (float)w;
        out[3] = (float)h;
        return;
    }
    uint64_t ntx = (uint64_t)std::sqrt((double)num_tiles);
    uint64_t nty = ((uint64_t)num_tiles + ntx - 1) / ntx;
    uint64_t tw = ((uint64_t)w + ntx - 1) / ntx;
    uint64_t th = ((uint64_t)h + nty - 1) / nty;
    float sx = (float)(((uint64_t)tile % ntx) * tw);
    float sy = (float)(((uint64_t)tile / ntx) * th);
    float ex = sx + (float)tw, ey = sy + (float)th;
    if (ex > (float)w) ex = (float)w;
    if (ey > (float)h) ey = (float)h;
    out[0] = sx;
    out[1] = sy;
    out[2] = ex;
    out[3] = ey;
}

bool valid_tile(int v) { return v == 8 || v == 16 || v == 32 || v == 64; }

gr_status validate_frame(const gr_frame& f) {
    if (f.width < 1 || f.height < 1 || f.width > gr::kMaxDim || f.height > gr::kMaxDim)
        return fail(GR_ERR_INVALID_ARGUMENT, "frame size must be within 1..16384");

    if (f.ref_tiles != 1 && f.ref_tiles != 16)
        return fail(GR_ERR_INVALID_ARGUMENT, "ref_tiles must be 1 or 16");
    return GR_OK;
}

// Go builtin min/max semantics on floats (mesh.go:33-38).
float go_min(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a < b) return a;
    if (b < a) return b;
    return std::signbit(a) ? a : b;
}
float go_max(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a > b) return a;
    if (b > a) return b;
    return std::signbit(a) ? b : a;
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Whether a mesh has at least one face per 4 pixels of its bounding box
// projected on screen (false when the box crosses the plane W = 0 or a
// projection is not finite). k_face counts such meshes' small tile rectangles
// through LDS; the estimate only steers that choice, never the result.
bool dense_on_screen(const float* mvp, const float bbox[8][4], int F, int W, int H) {
    float x0 = 1.f, x1 = -1.f, y0 = 1.f, y1 = -1.f;
    int neg = 0;
    for (int k = 0; k < 8; k++) {
        const float* c = bbox[k];
        const float X = mvp[0] * c[0] + mvp[1] * c[1] + mvp[2] * c[2] + mvp[3] * c[3];
        const float Y = mvp[4] * c[0] + mvp[5] * c[1] + mvp[6] * c[2] + mvp[7] * c[3];
        const float Wc = mvp[12] * c[0] + mvp[13] * c[1] + mvp[14] * c[2] + mvp[15] * c[3];
        neg += Wc < 0.f;
        if (!std::isfinite(X / Wc) || !std::isfinite(Y / Wc)) return false;
        x0 = std::min(x0, X / Wc);
        x1 = std::max(x1, X / Wc);
        y0 = std::min(y0, Y / Wc);
        y1 = std::max(y1, Y / Wc);
    }
    if (neg != 0 && neg != 8) return false;
    x0 = std::max(x0, -1.f), x1 = std::min(x1, 1.f);
    y0 = std::max(y0, -1.f), y1 = std::min(y1, 1.f);
    const double area = std::max(0.0, (double)(x1 - x0) * W / 2) * std::max(0.0, (double)(y1 - y0) * H / 2);
    return (double)F * 4.0 >= std::max(area, 1.0);
}

// Device memory for the depth keys of split tiles.
constexpr unsigned long long kSlabBudget = 64ull << 20;

// Tile-list entries per batch: offsets and cursors are 32-bit.
constexpr int64_t kMaxRefs = 0xFFFFFFFFll;

// Objects of one no-allocation draw (their descriptors are built on the stack).
constexpr int kNoAllocMaxObjects = 16;

// Everything that makes one frame of a batch invalid, checked before any
// launch. The serving loop runs it when a request is submitted, so one bad
// request cannot fail a batch it would share.
gr_status check_frame(gr_context* c, const gr_frame& fr, const gr_object* objs, int n) {
    gr_status st = validate_frame(fr);
    if (st) return st;
    size_t blocks = 0;
    for (int k = 0; k < n; k++) {
        const gr_object& ob = objs[k];
        if (ob.mesh < 0 || ob.mesh >= (int)c->meshes.size() || !c->meshes[ob.mesh].live)
            return fail(GR_ERR_INVALID_HANDLE, "invalid mesh handle");
        const MeshRes& mr = c->meshes[ob.mesh];
        for (int32_t th : mr.textures)
            if (th < 0 || th >= (int)c->textures.size() || !c->textures[th].live)
                return fail(GR_ERR_INVALID_HANDLE, "mesh references a destroyed texture");
        const bool gouraud =
            (fr.flags & GR_LIGHTING) && !(fr.flags & GR_FLAT_SHADING) && mr.dm.N > 0;
        if (gouraud && !mr.ni_valid)
            return fail(GR_ERR_INDEX_OUT_OF_RANGE,
                        "normal index out of range under Gouraud shading");
        blocks += (size_t)((mr.dm.F + gr::kFacesPerWg - 1) / gr::kFacesPerWg);
    }
    if (blocks * gr::kSlotsPerWg >= 0xFFFFFFFFull)
        return fail(GR_ERR_INVALID_ARGUMENT, "too many faces in one frame");
    if ((fr.flags & (GR_SHOW_EDGES | GR_SHOW_VERTICES)) &&
        blocks * gr::kSlotsPerWg >= 0x7FFFFFFFull)
        return fail(GR_ERR_UNSUPPORTED, "overlays on a frame with more than 2^31 slots");
    return GR_OK;
}

// Host-side description of one batch: the device structs (buffer offsets in
// place of buffer pointers), the header layout and the sizes of the per-batch
// buffers. Objects must be laid out frame by frame.
struct Plan {
    int nf = 0, nobj = 0, fblocks = 0, tiles = 0;
    int twl = 0, thl = 0;
    size_t okey_words = 0;
    bool decorate = false;  // a frame asks for tile borders or the crosshair
    bool raster_persistent = false;  // k_raster as a persistent grid (gr_config.raster_grid 2)
    gr::DevFrame* frames = nullptr;  // [nf], caller storage
    gr::DevObject* dobj = nullptr;   // [nobj], caller storage
    int* xcd = nullptr;              // [8 * (nf + 1)], caller storage
    int face_grid = 0;
    size_t off_objs = 0, off_stats = 0, off_counters = 0, off_over = 0, off_xcd = 0;
    size_t header_bytes = 0;
};

size_t xcd_ints(int nf) { return 8 * ((size_t)nf + 1); }

// Validates the batch and fills the plan. `frames`, `dobj` and `xcd` hold at
// least nf, no and xcd_ints(nf) entries.
gr_status make_plan(gr_context* c, const gr_batch_frame* bf, int nf, const gr_object* objs,
                    int no, int TW, int TH, gr::DevFrame* frames, gr::DevObject* dobj, int* xcd,
                    Plan& P) {
    if (nf <= 0) return fail(GR_ERR_INVALID_ARGUMENT, "no frames");
    P = Plan{};
    P.nf = nf;
    P.twl = ilog2(TW);
    P.raster_persistent = c->cfg.raster_grid == 2;
    P.thl = ilog2(TH);
    P.frames = frames;
    P.dobj = dobj;
    P.xcd = xcd;
    int nobj = 0;
    for (int f = 0; f < nf; f++) {
        const gr_frame& fr = bf[f].frame;
        if (bf[f].num_objects < 0 || bf[f].first_object < 0 ||
            bf[f].first_object + bf[f].num_objects > no)
            return fail(GR_ERR_INVALID_ARGUMENT, "frame object range out of bounds");
        gr_status st = check_frame(c, fr, objs + bf[f].first_object, bf[f].num_objects);
        if (st) return st;
        if (!bf[f].d_rgba || !bf[f].d_depth)
            return fail(GR_ERR_INVALID_ARGUMENT, "null device output");
        gr::DevFrame& d = frames[f];
        std::memset(&d, 0, sizeof d);
        d.W = fr.width;
        d.H = fr.height;
        d.flags = fr.flags;
        d.ref_tiles = fr.ref_tiles;
        std::memcpy(d.screen, fr.screen, sizeof d.screen);
        std::memcpy(d.light, fr.light, sizeof d.light);
        d.ambient = fr.ambient;
        d.diffuse = fr.diffuse;
        d.znear = fr.znear;
        d.zfar = fr.zfar;
        d.face_color = pack_rgba(fr.face_color);
        d.clear_color = pack_rgba(fr.clear_color);
        d.dot_color = pack_rgba(fr.dot_color);
        d.dot_step = fr.dot_step;
        for (int t = 0; t < fr.ref_tiles; t++)
            ref_tile_bounds(t, fr.ref_tiles, fr.width, fr.height, d.ref_tile_rect[t]);
        d.tiles_x = (fr.width + TW - 1) / TW;
        d.tiles_y = (fr.height + TH - 1) / TH;
        d.tile_begin = P.tiles;
        P.tiles += d.tiles_x * d.tiles_y;
        d.face_block_begin = P.fblocks;
        d.rgba = (uint8_t*)bf[f].d_rgba;
        d.depth = (float*)bf[f].d_depth;
        // Overlay key buffer offset (bytes + 1; 0 = no overlays).
        if (fr.flags & (GR_SHOW_TILE_BOUNDS | GR_SHOW_CROSSHAIR)) P.decorate = true;
        d.okey = nullptr;
        if (fr.flags & (GR_SHOW_EDGES | GR_SHOW_VERTICES)) {
            d.okey = (uint32_t*)(uintptr_t)(P.okey_words * 4 + 1);
            P.okey_words += (size_t)fr.width * fr.height;
        }
        int seq = 0;
        d.obj_begin = nobj;
        d.obj_count = bf[f].num_objects;
        for (int k = 0; k < bf[f].num_objects; k++) {
            const gr_object& ob = objs[bf[f].first_object + k];
            const MeshRes& mr = c->meshes[ob.mesh];
            gr::DevObject& o = dobj[nobj++];
            std::memset(&o, 0, sizeof o);
            o.m = mr.dm;
            std::memcpy(o.mvp, ob.mvp, sizeof o.mvp);
            std::memcpy(o.world, ob.world, sizeof o.world);
            o.frame = f;
            o.gouraud = (fr.flags & GR_LIGHTING) && !(fr.flags & GR_FLAT_SHADING) && mr.dm.N > 0;
            o.dense = dense_on_screen(ob.mvp, mr.dm.bbox, mr.dm.F, fr.width, fr.height) ? 1 : 0;
            o.vis = gr_box_visibility(ob.mvp, &mr.dm.bbox[0][0], fr.znear, fr.zfar);
            o.face_blocks = (mr.dm.F + gr::kFacesPerWg - 1) / gr::kFacesPerWg;
            o.face_block = P.fblocks;
            o.seq_block = seq;
            P.fblocks += o.face_blocks;
            seq += o.face_blocks;
        }
        d.face_blocks = P.fblocks - d.face_block_begin;
        // Set to the batch's record arrays by write_header (orec: 1 = overlays).
        d.orec = d.okey ? (gr::OverlayRec*)(uintptr_t)1 : nullptr;
        d.rrec = nullptr;
        d.srec = nullptr;
        d.wrec = nullptr;
        d.rbase = nullptr;
    }
    P.nobj = nobj;
    P.off_objs = align256((size_t)nf * sizeof(gr::DevFrame));
    P.off_stats = P.off_objs + align256((size_t)nobj * sizeof(gr::DevObject));
    P.off_counters = P.off_stats + align256((size_t)nf * gr::kNumStats * 8);
    P.off_over = P.off_counters + align256(sizeof(gr::DevBatchCounters));
    P.off_xcd = P.off_over + align256((size_t)nf * 4);
    P.header_bytes = P.off_xcd + align256(xcd_ints(nf) * sizeof(int));
    int nface[64];
    std::vector<int> nface_big;
    int* nfp = nface;
    if (nf > 64) {
        nface_big.resize(nf);
        nfp = nface_big.data();
    }
    for (int f = 0; f < nf; f++) nfp[f] = frames[f].face_blocks;
    gr::xcd_tables(nfp, nf, xcd, &P.face_grid);
    return GR_OK;
}

// Words of the record-pool counters.
constexpr size_t kPoolWords = (size_t)gr::kRecPools * gr::kPoolStride;

// Device buffers of one batch: a pipe's, a batch graph's, or carved from the
// fixed scratch array of the no-allocation draw.
struct Bufs {
    char* header = nullptr;  // frames | objects | stats | counters | frame_overflow | XCD tables
    // Triangle records [face blocks * kRecsPerBlock + rec_pools * pool_recs]:
    // each face block's own, then the overflow pools, claimed by the blocks
    // that need more as they run; rbase [face blocks] holds each block's first
    // overflow record, rec_pool the pool counters (zero between draws).
    gr::RasterRec* rrec = nullptr;
    gr::ShadeRec* srec = nullptr;
    gr::WideRec* wrec = nullptr;
    uint32_t rec_pools = 1, pool_recs = 0;
    uint32_t* rbase = nullptr;
    uint32_t* rec_pool = nullptr;
    uint32_t* tile_counts = nullptr;  // zero between draws (k_raster re-zeroes)
    uint32_t* tiles = nullptr;        // offsets | cursors | slabs
    gr::RasterItem* items = nullptr;
    uint32_t max_items = 0;
    unsigned long long* slabs = nullptr;  // zero between draws
    uint32_t max_slabs = 0;
    uint32_t part_refs = 0;
    uint32_t* wgcount = nullptr;
    uint32_t* refs = nullptr;
    unsigned long long refs_cap = 0;
    gr::DevControl* ctrl = nullptr;  // sticky until read
    uint32_t* okey = nullptr;        // zero between draws
    gr::OverlayRec* orec = nullptr;
    const gr::DevTexture* textures = nullptr;
};

// Split tiles: a tile is split only when its list is longer than part_refs,
// so the items of a batch number at most refs_cap / part_refs and the split
// tiles at most refs_cap / (part_refs + 1); slabs are further capped by a
// memory budget (a tile without a slab is rasterised whole).
void split_limits(unsigned long long refs_cap, uint32_t part_refs, size_t slab_bytes,
                  uint32_t* max_items, uint32_t* max_slabs) {
    *max_items = (uint32_t)std::min<unsigned long long>(refs_cap / part_refs + 1, 0xFFFFFFF0ull);
    *max_slabs = (uint32_t)std::max<unsigned long long>(
        1, std::min<unsigned long long>(refs_cap / (part_refs + 1ull) + 1, kSlabBudget / slab_bytes));
}

// The header as uploaded: device pointers patched in, counters zero.
void write_header(const Plan& P, const Bufs& B, void* h) {
    std::memset(h, 0, P.header_bytes);
    gr::DevFrame* hf = (gr::DevFrame*)h;
    std::memcpy(hf, P.frames, (size_t)P.nf * sizeof(gr::DevFrame));
    for (int f = 0; f < P.nf; f++) {
        gr::DevFrame& d = hf[f];
        if (d.okey) d.okey = (uint32_t*)((char*)B.okey + ((uintptr_t)d.okey - 1));
        if (d.orec) d.orec = B.orec;
        d.ctrl = B.ctrl;
        d.rrec = B.rrec;
        d.srec = B.srec;
        d.wrec = B.wrec;
        d.rbase = B.rbase + d.face_block_begin;
    }
    if (P.nobj) std::memcpy((char*)h + P.off_objs, P.dobj, (size_t)P.nobj * sizeof(gr::DevObject));
    std::memcpy((char*)h + P.off_xcd, P.xcd, xcd_ints(P.nf) * sizeof(int));
}

gr::LaunchParams launch_params(const Plan& P, const Bufs& B) {
    char* bb = B.header;
    gr::LaunchParams p{};
    p.objs = (const gr::DevObject*)(bb + P.off_objs);
    p.nobj = P.nobj;
    p.frames = (const gr::DevFrame*)bb;
    p.nframes = P.nf;
    p.textures = B.textures;
    p.face_blocks = P.fblocks;
    p.total_tiles = P.tiles;
    p.tw_log2 = P.twl;
    p.th_log2 = P.thl;
    p.tile_counts = B.tile_counts;
    p.tile_offsets = B.tiles;
    p.tile_cursor = B.tiles + P.tiles;
    p.tile_slab = B.tiles + 2 * (size_t)P.tiles;
    p.items = B.items;
    p.max_items = B.max_items;
    p.slabs = B.slabs;
    p.max_slabs = B.max_slabs;
    p.part_refs = B.part_refs;
    p.frame_overflow = (uint32_t*)(bb + P.off_over);
    p.wg_count = B.wgcount;
    p.rec_base = B.rbase;
    p.rec_pool = B.rec_pool;
    p.rec_pools = B.rec_pools;
    p.pool_recs = B.pool_recs;
    p.ovf_begin = (uint32_t)((size_t)P.fblocks * gr::kRecsPerBlock);
    p.refs = B.refs;
    p.refs_capacity = B.refs_cap;
    p.stats = (unsigned long long*)(bb + P.off_stats);
    p.counters = (gr::DevBatchCounters*)(bb + P.off_counters);
    p.ctrl = B.ctrl;
    p.xcd_face = (const int*)(bb + P.off_xcd);
    p.face_grid = P.face_grid;
    p.raster_persistent = P.raster_persistent ? 1 : 0;
    return p;
}

// The kernels of one batch on stream s. `raster_after` (may be null): an event
// the raster waits for, so that frames are written in submission order.
gr_status launch_plan(gr_context* c, const Plan& P, const gr::LaunchParams& p, hipStream_t s,
                      hipEvent_t raster_after) {
    timed_launch(c, "k_face", s, [&] { gr::launch_face(p, s); });
    timed_launch(c, "k_bin_scan", s, [&] { gr::launch_bin_scan(p, s); });
    timed_launch(c, "k_bin_fill", s, [&] { gr::launch_bin_fill(p, s); });
    if (P.okey_words) timed_launch(c, "k_overlay", s, [&] { gr::launch_overlay(p, s); });
    bool ok = true;
    timed_launch(c, "k_raster_parts", s, [&] { ok = gr::launch_raster_parts(p, s); });
    if (raster_after) GR_HIP(hipStreamWaitEvent(s, raster_after, 0));
    if (ok) timed_launch(c, "k_raster", s, [&] { ok = gr::launch_raster(p, s); });
    // Drawn over the finished frame, before the event the next batch's raster
    // waits for (it may write the same buffers).
    if (ok && P.decorate) timed_launch(c, "k_decorate", s, [&] { gr::launch_decorate(p, s); });
    hipError_t le = hipGetLastError();
    if (!ok || le != hipSuccess) {
        // Leave the tile and record-pool counters zeroed for the next draw.
        gr::launch_fill32(p.tile_counts, 0u, (size_t)P.tiles, s);
        gr::launch_fill32(p.rec_pool, 0u, kPoolWords, s);
        if (!ok) return fail(GR_ERR_INVALID_ARGUMENT, "unsupported tile size");
        return fail(GR_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(le));
    }
    return GR_OK;
}

// Allocates (or grows) buffer b and zero-fills a new allocation on stream s.
gr_status ensure_zeroed(gr_context* c, DevBuf& b, size_t bytes, hipStream_t s) {
    if (b.cap >= bytes && b.p) return GR_OK;
    gr_status st = ensure(c, b, bytes);
    if (st) return st;
    gr::launch_fill32((uint32_t*)b.p, 0u, b.cap / 4, s);
    GR_HIP(hipGetLastError());
    return GR_OK;
}

// Overflow pools of a batch: one per face block up to kRecPools, together one
// record per kOvfRefs tile-list entries of the capacity (grown on demand for
// both) and each at most what its blocks can claim beyond their own ranges.
void record_pools(const Plan& P, unsigned long long refs_cap, uint32_t* pools, uint32_t* per) {
    const unsigned long long np = (unsigned long long)std::min(gr::kRecPools, std::max(P.fblocks, 1));
    const unsigned long long most = ((unsigned long long)std::max(P.fblocks, 1) + np - 1) / np *
                                    (unsigned long long)(gr::kSlotsPerWg - gr::kRecsPerBlock);
    *pools = (uint32_t)np;
    *per = (uint32_t)std::max<unsigned long long>(1, std::min(refs_cap / gr::kOvfRefs / np, most));
}

// Records of a batch: each face block's own, then the overflow pools.
size_t records_of(const Plan& P, uint32_t pools, uint32_t per) {
    return (size_t)P.fblocks * gr::kRecsPerBlock + (size_t)pools * per;
}

// Device buffers of a plan in a set of growable buffers (a pipe's or a batch
// graph's); new allocations that must start at zero are zero-filled on s.
gr_status size_bufs(gr_context* c, const Plan& P, Pipe& Q, unsigned long long refs_cap,
                    uint32_t part_refs, int TW, int TH, hipStream_t s, Bufs& B) {
    gr_status st;
    uint32_t pools, per;
    record_pools(P, refs_cap, &pools, &per);
    const size_t recs = records_of(P, pools, per);
    if ((st = ensure(c, Q.batch, P.header_bytes))) return st;
    if ((st = ensure(c, Q.rrec, recs * sizeof(gr::RasterRec)))) return st;
    if ((st = ensure(c, Q.srec, recs * sizeof(gr::ShadeRec)))) return st;
    if ((st = ensure(c, Q.wrec, recs * sizeof(gr::WideRec)))) return st;
    if ((st = ensure_zeroed(c, Q.rbase, (kPoolWords + std::max(P.fblocks, 1)) * sizeof(uint32_t), s)))
        return st;
    if ((st = ensure_zeroed(c, Q.tile_counts, (size_t)P.tiles * 4, s))) return st;
    if ((st = ensure(c, Q.tiles, 3 * (size_t)P.tiles * sizeof(uint32_t)))) return st;
    const size_t slab_bytes = (size_t)TW * TH * 8;
    uint32_t max_items, max_slabs;
    split_limits(refs_cap, part_refs, slab_bytes, &max_items, &max_slabs);
    if ((st = ensure(c, Q.items, (size_t)max_items * sizeof(gr::RasterItem)))) return st;
    if ((st = ensure_zeroed(c, Q.slabs, (size_t)max_slabs * slab_bytes, s))) return st;
    if ((st = ensure(c, Q.wgcount, std::max(P.fblocks, 1) * sizeof(uint32_t)))) return st;
    if ((st = ensure(c, Q.refs, refs_cap * sizeof(uint32_t)))) return st;
    if ((st = ensure_zeroed(c, Q.ctrl, sizeof(gr::DevControl), s))) return st;
    if (P.okey_words) {
        if ((st = ensure(c, Q.orec, recs * sizeof(gr::OverlayRec)))) return st;
        if ((st = ensure_zeroed(c, Q.okey, P.okey_words * 4, s))) return st;
    }
    B.header = (char*)Q.batch.p;
    B.rrec = (gr::RasterRec*)Q.rrec.p;
    B.srec = (gr::ShadeRec*)Q.srec.p;
    B.wrec = (gr::WideRec*)Q.wrec.p;
    B.rec_pools = pools;
    B.pool_recs = per;
    B.rec_pool = (uint32_t*)Q.rbase.p;
    B.rbase = B.rec_pool + kPoolWords;
    B.tile_counts = (uint32_t*)Q.tile_counts.p;
    B.tiles = (uint32_t*)Q.tiles.p;
    B.items = (gr::RasterItem*)Q.items.p;
    B.max_items = max_items;
    B.slabs = (unsigned long long*)Q.slabs.p;
    B.max_slabs = max_slabs;
    B.part_refs = part_refs;
    B.wgcount = (uint32_t*)Q.wgcount.p;
    B.refs = (uint32_t*)Q.refs.p;
    B.refs_cap = refs_cap;
    B.ctrl = (gr::DevControl*)Q.ctrl.p;
    B.okey = (uint32_t*)Q.okey.p;
    B.orec = (gr::OverlayRec*)Q.orec.p;
    B.textures = (const gr::DevTexture*)c->d_textures.p;
    return GR_OK;
}

// Texture descriptor table, rebuilt after textures changed (read by the raster
// of a batch that may still be running on the other pipe: wait for it first).
gr_status refresh_textures(gr_context* c, hipStream_t s) {
    if (!c->tex_dirty) return GR_OK;
    std::vector<gr::DevTexture> tt(std::max<size_t>(1, c->textures.size()));
    for (size_t i = 0; i < c->textures.size(); i++) tt[i] = c->textures[i].dt;
    gr_status st;
    if ((st = sync_all(c))) return st;
    if ((st = ensure(c, c->d_textures, tt.size() * sizeof(gr::DevTexture)))) return st;
    if ((st = upload(c, c->d_textures.p, tt.data(), tt.size() * sizeof(gr::DevTexture), s)))
        return st;
    c->tex_dirty = false;
    return GR_OK;
}

unsigned long long refs_capacity_of(const gr_context* c) {
    return (unsigned long long)std::min<int64_t>(std::max<int64_t>(c->cfg.refs_capacity, 1024),
                                                 kMaxRefs);
}

// The whole pipeline for a batch of frames on the next pipe; outputs are
// device pointers.
gr_status enqueue(gr_context* c, const gr_batch_frame* bf, int nf, const gr_object* objs,
                  int no) {
    const int TW = c->cfg.tile_w, TH = c->cfg.tile_h;
    std::vector<gr::DevFrame> frames(std::max(nf, 1));
    std::vector<gr::DevObject> dobj(std::max(no, 1));
    std::vector<int> xcd(xcd_ints(std::max(nf, 1)));
    Plan P;
    gr_status st = make_plan(c, bf, nf, objs, no, TW, TH, frames.data(), dobj.data(), xcd.data(), P);
    if (st) return st;

    // The pipe of this batch: the other one than the last batch's when the
    // context runs on its own streams with two pipes, else pipe 0 on c->stream.
    const bool two = c->stream == c->own && c->cfg.pipelines >= 2 && !c->one_pipe;
    const int pi = two ? (c->last_pipe + 1) & 1 : 0;
    Pipe& Q = c->pipes[pi];
    Q.stream = pi == 0 ? c->stream : c->own1;
    const hipStream_t s = Q.stream;

    if ((st = refresh_textures(c, s))) return st;
    Bufs B;
    if ((st = size_bufs(c, P, Q, refs_capacity_of(c), (uint32_t)c->cfg.tile_split, TW, TH, s, B)))
        return st;
    {
        void* h;
        int slot;
        if ((st = staging_get(c, P.header_bytes, &h, &slot))) return st;
        write_header(P, B, h);
        GR_HIP(hipMemcpyAsync(B.header, h, P.header_bytes, hipMemcpyHostToDevice, s));
        if ((st = staging_mark(c, slot, s))) return st;
    }
    const gr::LaunchParams p = launch_params(P, B);
    // Frames are written in submission order: this raster starts after the
    // previous batch's (on the other pipe).
    hipEvent_t after = nullptr;
    if (c->last_pipe >= 0 && c->last_pipe != pi && c->pipes[c->last_pipe].raster_pending)
        after = c->pipes[c->last_pipe].raster_done;
    if ((st = launch_plan(c, P, p, s, after))) return st;
    if (!Q.raster_done) GR_HIP(hipEventCreateWithFlags(&Q.raster_done, hipEventDisableTiming));
    GR_HIP(hipEventRecord(Q.raster_done, s));
    Q.raster_pending = true;
    c->last_pipe = pi;
    c->last_frames = nf;
    c->last_stats = p.stats;
    c->last_overflow = p.frame_overflow;
    c->last_counters = p.counters;
    return GR_OK;
}

// Copy the stats of the last draw (device) to `out` after waiting for it.
gr_status read_stats(gr_context* c, std::vector<unsigned long long>& out) {
    out.assign((size_t)c->last_frames * gr::kNumStats, 0ull);
    if (!c->last_stats || out.empty()) return GR_OK;
    gr_status st = sync_all(c);
    if (st) return st;
    GR_HIP(hipMemcpy(out.data(), c->last_stats, out.size() * 8, hipMemcpyDeviceToHost));
    return GR_OK;
}

void to_gr_stats(const unsigned long long* q, gr_stats* s) {
    s->tpf = (int64_t)q[gr::kStatTpf];
    s->triangles = (int64_t)q[gr::kStatTriangles];
    s->fragments_covered = (int64_t)q[gr::kStatCovered];
    s->pixels_shaded = (int64_t)q[gr::kStatShaded];
    s->faces_front = (int64_t)q[gr::kStatFront];
    s->clip_overflow = (int64_t)q[gr::kStatClipOverflow];
    s->refs = (int64_t)q[gr::kStatRefs];
    s->tiles_split = (int64_t)q[gr::kStatSplit];
}

// Wait for every pipe, then read and clear the sticky control words of both.
// Reads the sticky control words at d (device) after the work on stream s, and
// clears them when set.
gr_status read_ctrl(gr_context* c, void* d, hipStream_t s, gr::DevControl* ctl) {
    GR_HIP(hipMemcpyAsync(c->h_ctrl, d, sizeof *ctl, hipMemcpyDeviceToHost, s));
    GR_HIP(hipStreamSynchronize(s));
    std::memcpy(ctl, c->h_ctrl, sizeof *ctl);
    if (ctl->unsupported || ctl->overflow) {
        gr::launch_fill32((uint32_t*)d, 0u, sizeof *ctl / 4, s);
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(s));
    }
    return GR_OK;
}

gr_status graphs_check(gr_context* c, bool* unsupported, bool* overflow,
                       unsigned long long* needed);

// Wait for every pipe, then read and clear the sticky control words of both,
// of the batch graphs and of the no-allocation draw.
gr_status sync_and_check(gr_context* c, unsigned long long* needed) {
    gr_status st = sync_all(c);
    if (st) return st;
    if (c->timing) collect_times(c);
    *needed = 0;
    bool unsupported = false, overflow = false;
    for (Pipe& pp : c->pipes) {
        if (!pp.ctrl.p) continue;
        gr::DevControl ctl;
        if ((st = read_ctrl(c, pp.ctrl.p, pp.stream, &ctl))) return st;
        unsupported = unsupported || ctl.unsupported;
        overflow = overflow || ctl.overflow;
        if (ctl.overflow) *needed = std::max(*needed, ctl.refs_needed);
    }
    if ((st = graphs_check(c, &unsupported, &overflow, needed))) return st;
    bool noalloc_overflow = false;
    if (c->noalloc_ctrl) {
        gr::DevControl ctl;
        if ((st = read_ctrl(c, c->noalloc_ctrl, c->stream, &ctl))) return st;
        noalloc_overflow = ctl.overflow != 0;
    }
    if (unsupported) return fail(GR_ERR_UNSUPPORTED, "overlay line longer than 65536 pixels");
    if (noalloc_overflow && !overflow)
        return fail(GR_ERR_CAPACITY,
                    "tile lists or triangle records of a no-allocation draw exceeded its "
                    "fixed scratch array");
    if (overflow) return GR_ERR_CAPACITY;
    return GR_OK;
}

void grow_refs(gr_context* c, unsigned long long needed) {
    int64_t want = std::min<int64_t>((int64_t)needed + (int64_t)needed / 4 + 1024, kMaxRefs);
    if (want > c->cfg.refs_capacity) c->cfg.refs_capacity = want;
}

gr_status upload_array(gr_context* c, MeshRes& mr, const void* src, size_t bytes, void** out) {
    void* d = nullptr;
    GR_HIP(hipMalloc(&d, std::max<size_t>(bytes, 4)));
    mr.allocs.push_back(d);
    mr.bytes += (int64_t)bytes;
    GR_HIP(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice));
    *out = d;
    return GR_OK;
}

void free_mesh(MeshRes& mr) {
    for (void* p : mr.allocs) (void)hipFree(p);
    mr.allocs.clear();
    mr.live = false;
}

bool is_pow2(int32_t x) { return x != 0 && (x & (x - 1)) == 0; }

// Queues the draw of a present slot's frame and the copies of its colour (and
// depth), stats and overflow flag to pinned host memory, on the stream of the
// pipe that draws it (the next frame runs on the other pipe meanwhile).
gr_status present_enqueue(gr_context* c, gr_context::PresentSlot& ps) {
    const size_t npx = (size_t)ps.frame.width * ps.frame.height;
    gr_batch_frame bf;
    bf.frame = ps.frame;
    bf.first_object = 0;
    bf.num_objects = (int32_t)ps.objs.size();
    bf.d_rgba = ps.rgba.p;
    bf.d_depth = ps.depth.p;
    gr_status st = enqueue(c, &bf, 1, ps.objs.data(), (int)ps.objs.size());
    if (st) return st;
    const hipStream_t s = c->pipes[c->last_pipe].stream;
    // The pixels by a copy kernel writing the pinned pages (the other pipe's
    // draw runs meanwhile, and DMA copies then reach about half the link's
    // rate); the few words of stats and flags by DMA.
    gr::launch_copy(ps.host, ps.rgba.p, npx * 4, s);
    if (ps.with_depth) gr::launch_copy((char*)ps.host + npx * 4, ps.depth.p, npx * 4, s);
    GR_HIP(hipGetLastError());
    GR_HIP(hipMemcpyAsync(ps.h_meta, c->last_stats, gr::kNumStats * 8, hipMemcpyDeviceToHost, s));
    GR_HIP(hipMemcpyAsync((char*)ps.h_meta + gr::kNumStats * 8, c->last_overflow, 4,
                          hipMemcpyDeviceToHost, s));
    GR_HIP(hipEventRecord(ps.copied, s));
    ps.pending = true;
    return GR_OK;
}

void free_pipe_bufs(Pipe& pp) {
    DevBuf* pb[] = {&pp.batch, &pp.rrec,    &pp.srec, &pp.wrec,  &pp.tile_counts,
                    &pp.tiles, &pp.wgcount, &pp.refs, &pp.ctrl,  &pp.items,
                    &pp.slabs, &pp.okey,    &pp.orec,  &pp.rbase};
    for (DevBuf* b : pb) {
        if (b->p) (void)hipFree(b->p);
        *b = DevBuf{};
    }
}

}  // namespace

// One batch's launch sequence captured in two HIP graphs, with its own device
// buffers: counter zeroing, k_face, k_bin_scan, k_bin_fill, k_overlay and
// k_raster_parts in the geometry graph; k_raster and k_decorate in the raster
// graph, so that a replay's raster can wait for the previous batch's between
// the two graph launches while its geometry overlaps that raster.
struct gr_batch_graph {
    gr_context* c = nullptr;
    int TW = 32, TH = 32;
    uint32_t part_refs = 512;
    unsigned long long refs_cap = 0;
    std::vector<gr_batch_frame> bf;  // as created (outputs, object ranges)
    std::vector<gr_object> objs;
    std::vector<gr::DevFrame> frames;
    std::vector<gr::DevObject> dobj;
    std::vector<int> xcd;
    Plan P;
    Pipe buf;          // the graph's buffers (its stream field is unused)
    DevBuf textures;   // texture table as at creation
    Bufs B;
    gr::LaunchParams p{};
    hipStream_t cap = nullptr;  // capture (and first-run) stream
    hipGraph_t geo = nullptr, ras = nullptr;
    hipGraphExec_t geo_x = nullptr, ras_x = nullptr;
    hipEvent_t done = nullptr;   // after the last replay's raster
    hipStream_t last = nullptr;  // stream of the last replay
    bool regrow = false;         // a replay overflowed its tile lists
    unsigned long long needed = 0;
};

namespace {

void graph_drop_exec(gr_batch_graph* g) {
    if (g->geo_x) (void)hipGraphExecDestroy(g->geo_x);
    if (g->ras_x) (void)hipGraphExecDestroy(g->ras_x);
    if (g->geo) (void)hipGraphDestroy(g->geo);
    if (g->ras) (void)hipGraphDestroy(g->ras);
    g->geo_x = g->ras_x = nullptr;
    g->geo = g->ras = nullptr;
}

// Capture of one stage on the graph's capture stream; the capture is always
// ended, so a failed launch leaves the stream usable.
template <typename F>
gr_status capture(gr_batch_graph* g, hipGraph_t* out, hipGraphExec_t* exec, F launches) {
    GR_HIP(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
    const bool ok = launches();
    const hipError_t le = hipGetLastError();
    hipGraph_t gr = nullptr;
    const hipError_t ee = hipStreamEndCapture(g->cap, &gr);
    if (!ok || le != hipSuccess || ee != hipSuccess || !gr) {
        if (gr) (void)hipGraphDestroy(gr);
        if (!ok) return fail(GR_ERR_INVALID_ARGUMENT, "unsupported tile size");
        return fail(GR_ERR_DEVICE, std::string("graph capture: ") +
                                       hipGetErrorString(le != hipSuccess ? le : ee));
    }
    *out = gr;
    GR_HIP(hipGraphInstantiate(exec, gr, nullptr, nullptr, 0));
    return GR_OK;
}

// (Re)builds the graph's buffers and header at its capacity and captures its
// two stages.
gr_status graph_build(gr_batch_graph* g) {
    gr_context* c = g->c;
    gr_status st;
    graph_drop_exec(g);
    if ((st = size_bufs(c, g->P, g->buf, g->refs_cap, g->part_refs, g->TW, g->TH, g->cap, g->B)))
        return st;
    g->B.textures = (const gr::DevTexture*)g->textures.p;
    std::vector<unsigned char> h(g->P.header_bytes);
    write_header(g->P, g->B, h.data());
    GR_HIP(hipMemcpyAsync(g->B.header, h.data(), h.size(), hipMemcpyHostToDevice, g->cap));
    GR_HIP(hipStreamSynchronize(g->cap));
    g->p = launch_params(g->P, g->B);
    const gr::LaunchParams& p = g->p;
    const Plan& P = g->P;
    const hipStream_t s = g->cap;
    // The per-replay counters (stats, batch counters, frame overflow flags)
    // start at zero; the frames, objects and XCD tables stay as uploaded.
    char* zero0 = g->B.header + P.off_stats;
    const size_t zero_words = (P.off_xcd - P.off_stats) / 4;
    if ((st = capture(g, &g->geo, &g->geo_x, [&] {
             gr::launch_fill32((uint32_t*)zero0, 0u, zero_words, s);
             gr::launch_face(p, s);
             gr::launch_bin_scan(p, s);
             gr::launch_bin_fill(p, s);
             if (P.okey_words) gr::launch_overlay(p, s);
             return gr::launch_raster_parts(p, s);
         })))
        return st;
    if ((st = capture(g, &g->ras, &g->ras_x, [&] {
             if (!gr::launch_raster(p, s)) return false;
             if (P.decorate) gr::launch_decorate(p, s);
             return true;
         })))
        return st;
    return GR_OK;
}

// Reads and clears the control words of every batch graph (after sync_all); a
// graph whose replay overflowed regrows at its next launch.
gr_status graphs_check(gr_context* c, bool* unsupported, bool* overflow,
                       unsigned long long* needed) {
    for (gr_batch_graph* g : c->graphs) {
        if (!g->B.ctrl) continue;
        gr::DevControl ctl;
        gr_status st = read_ctrl(c, g->B.ctrl, g->cap, &ctl);
        if (st) return st;
        *unsupported = *unsupported || ctl.unsupported;
        if (ctl.overflow) {
            *overflow = true;
            *needed = std::max(*needed, ctl.refs_needed);
            g->regrow = true;
            g->needed = std::max(g->needed, ctl.refs_needed);
        }
    }
    return GR_OK;
}

unsigned long long grown(unsigned long long needed) {
    return (unsigned long long)std::min<int64_t>((int64_t)needed + (int64_t)needed / 4 + 1024,
                                                 kMaxRefs);
}

gr_status graph_launch(gr_batch_graph* g) {
    gr_context* c = g->c;
    gr_status st;
    if (g->regrow) {
        if ((st = sync_all(c))) return st;
        g->refs_cap = std::max(g->refs_cap, grown(g->needed));
        if ((st = graph_build(g))) return st;
        g->regrow = false;
        g->needed = 0;
    }
    const bool two = c->stream == c->own && c->cfg.pipelines >= 2 && !c->one_pipe;
    const int pi = two ? (c->last_pipe + 1) & 1 : 0;
    Pipe& Q = c->pipes[pi];
    Q.stream = pi == 0 ? c->stream : c->own1;
    const hipStream_t s = Q.stream;
    // The previous replay of this graph (its buffers) and the previous batch's
    // raster (frame order) come first.
    if (g->last && g->last != s) GR_HIP(hipStreamWaitEvent(s, g->done, 0));
    GR_HIP(hipGraphLaunch(g->geo_x, s));
    if (c->last_pipe >= 0 && c->last_pipe != pi && c->pipes[c->last_pipe].raster_pending)
        GR_HIP(hipStreamWaitEvent(s, c->pipes[c->last_pipe].raster_done, 0));
    GR_HIP(hipGraphLaunch(g->ras_x, s));
    if (!Q.raster_done) GR_HIP(hipEventCreateWithFlags(&Q.raster_done, hipEventDisableTiming));
    GR_HIP(hipEventRecord(Q.raster_done, s));
    GR_HIP(hipEventRecord(g->done, s));
    Q.raster_pending = true;
    g->last = s;
    c->last_pipe = pi;
    c->last_frames = g->P.nf;
    c->last_stats = g->p.stats;
    return GR_OK;
}

void graph_free(gr_batch_graph* g) {
    graph_drop_exec(g);
    free_pipe_bufs(g->buf);
    if (g->textures.p) (void)hipFree(g->textures.p);
    if (g->done) (void)hipEventDestroy(g->done);
    if (g->cap) (void)hipStreamDestroy(g->cap);
    delete g;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

namespace gr_internal {

gr_status fail(gr_status s, const std::string& msg) { return ::fail(s, msg); }

gr_status enqueue_batch(gr_context* c, const gr_batch_frame* frames, int32_t nf,
                        const gr_object* objs, int32_t no) {
    try {
        GR_HIP(hipSetDevice(c->device));
        return enqueue(c, frames, nf, objs, no);
    } catch (...) {
        return ::fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}
hipStream_t context_stream(gr_context* c) {
    return c->last_pipe >= 0 && c->pipes[c->last_pipe].stream ? c->pipes[c->last_pipe].stream
                                                              : c->stream;
}
int context_device(gr_context* c) { return c->device; }
const void* last_stats_device(gr_context* c) { return c->last_stats; }
size_t stats_bytes_per_frame() { return gr::kNumStats * sizeof(unsigned long long); }
void decode_stats(const void* frame_stats, gr_stats* out) {
    to_gr_stats((const unsigned long long*)frame_stats, out);
}
bool frame_long_lines(const void* frame_stats) {
    return ((const unsigned long long*)frame_stats)[gr::kStatLongLines] != 0;
}
void* control_device(gr_context* c) { return c->last_pipe >= 0 ? c->pipes[c->last_pipe].ctrl.p : nullptr; }
size_t control_bytes() { return sizeof(gr::DevControl); }
void decode_control(const void* host_copy, bool* overflow, unsigned long long* refs_needed,
                    bool* unsupported) {
    gr::DevControl ctl;
    std::memcpy(&ctl, host_copy, sizeof ctl);
    *overflow = ctl.overflow != 0;
    *refs_needed = ctl.refs_needed;
    *unsupported = ctl.unsupported != 0;
}
void grow_refs_to(gr_context* c, unsigned long long needed) { grow_refs(c, needed); }
void set_one_pipe(gr_context* c, bool on) { c->one_pipe = on; }
gr_status check_request(gr_context* c, const gr_frame* frame, const gr_object* objects,
                        int32_t num_objects) {
    return check_frame(c, *frame, objects, num_objects);
}

// The no-allocation draw. Layout of [base, base + bytes): control words (sticky
// until gr_synchronize reads them), batch header, texture table, tile counters,
// tile offsets/cursors/slabs, face-block counts and first records, then what
// remains shared between triangle records and tile-list entries, as many of
// each (fewer records when the frame has fewer slots). Tiles are never split
// (no slabs or work items).
// Host data reaches the array through kernel arguments (k_put) and zeroing is
// a fill kernel, so nothing here allocates memory of any kind, and the launch
// sequence holds no memset or host-buffer copy.
gr_status enqueue_noalloc(gr_context* c, const gr_frame* frame, const gr_object* objs, int32_t no,
                          void* d_rgba, void* d_depth, void* base, size_t bytes) {
    try {
        GR_HIP(hipSetDevice(c->device));
        if (no > kNoAllocMaxObjects)
            return ::fail(GR_ERR_INVALID_ARGUMENT, "the no-allocation draw takes at most 16 objects");
        if (frame->flags & (GR_SHOW_EDGES | GR_SHOW_VERTICES))
            return ::fail(GR_ERR_UNSUPPORTED,
                          "edge and vertex overlays need per-pixel keys: use gr_draw_device");
        const int TW = c->cfg.tile_w, TH = c->cfg.tile_h;
        gr_batch_frame bf;
        bf.frame = *frame;
        bf.first_object = 0;
        bf.num_objects = no;
        bf.d_rgba = d_rgba;
        bf.d_depth = d_depth;
        gr::DevFrame frames[1];
        gr::DevObject dobj[kNoAllocMaxObjects];
        int xcd[16];
        Plan P;
        gr_status st = make_plan(c, &bf, 1, objs, no, TW, TH, frames, dobj, xcd, P);
        if (st) return st;

        const size_t ntex = std::max<size_t>(1, c->textures.size());
        size_t off = 0;
        auto take = [&](size_t n) {
            const size_t o = off;
            off += align256(std::max<size_t>(n, 4));
            return (char*)base + o;
        };
        Bufs B;
        B.ctrl = (gr::DevControl*)take(sizeof(gr::DevControl));
        B.header = take(P.header_bytes);
        B.textures = (const gr::DevTexture*)take(ntex * sizeof(gr::DevTexture));
        B.tile_counts = (uint32_t*)take((size_t)P.tiles * 4);
        B.tiles = (uint32_t*)take(3 * (size_t)P.tiles * 4);
        B.wgcount = (uint32_t*)take((size_t)std::max(P.fblocks, 1) * 4);
        B.rec_pool = (uint32_t*)take(kPoolWords * 4);
        B.rbase = (uint32_t*)take((size_t)std::max(P.fblocks, 1) * 4);
        constexpr size_t kRecBytes =
            sizeof(gr::RasterRec) + sizeof(gr::ShadeRec) + sizeof(gr::WideRec);
        const size_t min_refs = 1u << 16;
        // Every block's own records, at least one overflow record per pool and
        // 2^16 tile-list entries; three record arrays and the entries each
        // start on a 256-byte boundary.
        const size_t own = (size_t)P.fblocks * gr::kRecsPerBlock;
        const size_t need = off + 4 * 256 + min_refs * 4 + (own + gr::kRecPools) * kRecBytes;
        if (need > bytes)
            return ::fail(GR_ERR_INVALID_ARGUMENT,
                          "frame too large for the no-allocation draw: needs " +
                              std::to_string(need >> 20) +
                              " MiB of its 32 MiB device array (use gr_draw_device)");
        // The rest: kOvfRefs tile-list entries per overflow record.
        const size_t rest = bytes - off - 4 * 256 - own * kRecBytes;
        record_pools(P, rest / (kRecBytes + 4 * gr::kOvfRefs) * gr::kOvfRefs, &B.rec_pools,
                     &B.pool_recs);
        const size_t recs = records_of(P, B.rec_pools, B.pool_recs);
        B.rrec = (gr::RasterRec*)take(recs * sizeof(gr::RasterRec));
        B.srec = (gr::ShadeRec*)take(recs * sizeof(gr::ShadeRec));
        B.wrec = (gr::WideRec*)take(recs * sizeof(gr::WideRec));
        B.refs = (uint32_t*)((char*)base + off);
        B.refs_cap = std::min<unsigned long long>((bytes - off) / 4, (unsigned long long)kMaxRefs);
        B.part_refs = 0xFFFFFFFFu;  // no split tiles
        B.items = nullptr;
        B.max_items = 0;
        B.slabs = nullptr;
        B.max_slabs = 0;

        const hipStream_t s = c->stream;
        // The previous draw of the context may run on the other pipe: this
        // one's buffers are its own, but its frame is written after that one's.
        hipEvent_t after = nullptr;
        if (c->last_pipe == 1 && c->pipes[1].raster_pending && c->pipes[1].stream != s)
            after = c->pipes[1].raster_done;
        // Header and texture table through kernel arguments, from the stack.
        unsigned char hbuf[8192];
        if (P.header_bytes > sizeof hbuf)
            return ::fail(GR_ERR_INVALID_ARGUMENT, "batch header too large");
        write_header(P, B, hbuf);
        gr::launch_put(B.header, hbuf, P.header_bytes, s);
        {
            constexpr size_t kChunk = (gr::kPutWords * 4) / sizeof(gr::DevTexture);
            gr::DevTexture tt[kChunk];
            for (size_t t0 = 0; t0 < ntex; t0 += kChunk) {
                const size_t n = std::min(kChunk, ntex - t0);
                for (size_t i = 0; i < n; i++)
                    tt[i] = t0 + i < c->textures.size() ? c->textures[t0 + i].dt : gr::DevTexture{};
                gr::launch_put((char*)B.textures + t0 * sizeof(gr::DevTexture), tt,
                               n * sizeof(gr::DevTexture), s);
            }
        }
        // Tile and record-pool counters start at zero (the array is shared).
        gr::launch_fill32(B.tile_counts, 0u, (size_t)P.tiles, s);
        gr::launch_fill32(B.rec_pool, 0u, kPoolWords, s);
        GR_HIP(hipGetLastError());
        const gr::LaunchParams p = launch_params(P, B);
        if ((st = launch_plan(c, P, p, s, after))) return st;
        Pipe& Q = c->pipes[0];
        Q.stream = s;
        GR_HIP(hipEventRecord(Q.raster_done, s));
        Q.raster_pending = true;
        c->last_pipe = 0;
        c->last_frames = 1;
        c->last_stats = p.stats;
        c->noalloc_ctrl = B.ctrl;
        return GR_OK;
    } catch (...) {
        return ::fail(GR_ERR_DEVICE, "unexpected exception");
    }
}

}  // namespace gr_internal

extern "C" {

const char* gr_last_error(void) { return g_last_error.c_str(); }
int32_t gr_abi_version(void) { return GR_ABI_VERSION; }

gr_status gr_create(int32_t device, gr_context** out) {
    try {
        if (!out) return fail(GR_ERR_INVALID_ARGUMENT, "null out");
        *out = nullptr;
        int n = 0;
        GR_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) return fail(GR_ERR_INVALID_ARGUMENT, "no such device");
        GR_HIP(hipSetDevice(device));
        gr_context* c = new gr_context();
        c->device = device;
        if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return fail(GR_ERR_DEVICE, "hipStreamCreate failed");
        }
        c->stream = c->own;
        if (hipStreamCreateWithFlags(&c->own1, hipStreamNonBlocking) != hipSuccess) {
            (void)hipStreamDestroy(c->own);
            delete c;
            return fail(GR_ERR_DEVICE, "hipStreamCreate failed");
        }
        if (hipHostMalloc(&c->h_ctrl, sizeof(gr::DevControl), hipHostMallocDefault) != hipSuccess) {
            (void)hipStreamDestroy(c->own1);
            (void)hipStreamDestroy(c->own);
            delete c;
            return fail(GR_ERR_OUT_OF_MEMORY, "pinned allocation failed");
        }
        // Set-up of the no-allocation draw: its array's address on this device,
        // and the event its raster records (created here, not in the draw).
        gr_status sst = gr_internal::static_scratch_init(device);
        if (!sst && hipEventCreateWithFlags(&c->pipes[0].raster_done, hipEventDisableTiming) !=
                        hipSuccess)
            sst = fail(GR_ERR_DEVICE, "hipEventCreate failed");
        if (sst) {
            gr_destroy(c);
            return sst;
        }
        *out = c;
        return GR_OK;
    } catch (const std::bad_alloc&) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(GR_ERR_DEVICE, "unexpected exception");
    }
}

void gr_destroy(gr_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)sync_all(c);
    for (gr_batch_graph* g : c->graphs) graph_free(g);
    c->graphs.clear();
    for (auto& m : c->meshes) free_mesh(m);
    for (auto& t : c->textures)
        if (t.texels) (void)hipFree(t.texels);
    (void)sync_all(c);
    DevBuf* bufs[] = {&c->d_textures, &c->out_rgba, &c->out_depth};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (Pipe& pp : c->pipes) {
        free_pipe_bufs(pp);
        if (pp.raster_done) (void)hipEventDestroy(pp.raster_done);
    }
    if (c->h_frame) (void)hipHostFree(c->h_frame);
    for (auto& ps : c->present) {
        if (ps.rgba.p) (void)hipFree(ps.rgba.p);
        if (ps.depth.p) (void)hipFree(ps.depth.p);
        if (ps.host) (void)hipHostFree(ps.host);
        if (ps.h_meta) (void)hipHostFree(ps.h_meta);
        if (ps.copied) (void)hipEventDestroy(ps.copied);
    }
    if (c->own1) (void)hipStreamDestroy(c->own1);
    for (auto& s : c->ring) {
        if (s.host) (void)hipHostFree(s.host);
        if (s.ev) (void)hipEventDestroy(s.ev);
    }
    for (auto& t : c->timed) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    if (c->h_ctrl) (void)hipHostFree(c->h_ctrl);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
}

gr_status gr_set_config(gr_context* c, const gr_config* cfg) {
    if (!c || !cfg) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    gr_config n = c->cfg;
    if (cfg->tile_w) n.tile_w = cfg->tile_w;
    if (cfg->tile_h) n.tile_h = cfg->tile_h;
    if (cfg->refs_capacity) n.refs_capacity = cfg->refs_capacity;
    if (cfg->raster_waves) n.raster_waves = cfg->raster_waves;
    if (cfg->tile_split) n.tile_split = cfg->tile_split;
    if (cfg->pipelines) n.pipelines = cfg->pipelines;
    if (cfg->raster_grid) n.raster_grid = cfg->raster_grid;
    if (n.raster_grid != 1 && n.raster_grid != 2)
        return fail(GR_ERR_INVALID_ARGUMENT, "raster_grid must be 1 or 2");
    if (!valid_tile(n.tile_w) || !valid_tile(n.tile_h))
        return fail(GR_ERR_INVALID_ARGUMENT, "tile sides must be 8, 16, 32 or 64");
    if (n.raster_waves != 4) return fail(GR_ERR_INVALID_ARGUMENT, "raster_waves must be 4");
    if (n.pipelines != 1 && n.pipelines != 2)
        return fail(GR_ERR_INVALID_ARGUMENT, "pipelines must be 1 or 2");
    if (n.refs_capacity < 1024 || n.refs_capacity > kMaxRefs)
        return fail(GR_ERR_INVALID_ARGUMENT, "refs_capacity must be in [1024, 2^32 - 1]");
    if (n.tile_split < 64 || n.tile_split > (1 << 24))
        return fail(GR_ERR_INVALID_ARGUMENT, "tile_split must be in [64, 2^24]");
    c->cfg = n;
    return GR_OK;
}

gr_status gr_get_config(gr_context* c, gr_config* cfg) {
    if (!c || !cfg) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    *cfg = c->cfg;
    return GR_OK;
}

gr_status gr_set_stream(gr_context* c, void* s) {
    if (!c) return fail(GR_ERR_INVALID_ARGUMENT, "null context");
    gr_status st = sync_all(c);
    if (st) return st;
    c->stream = s ? (hipStream_t)s : c->own;
    return GR_OK;
}

void* gr_get_stream(gr_context* c) { return c ? (void*)c->stream : nullptr; }

gr_status gr_texture_create(gr_context* c, const gr_texture_desc* d, const uint8_t* rgba,
                            int32_t* out) {
    try {
        if (!c || !d || !out) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
        if (d->type != GR_TEXTURE_SOLID && d->type != GR_TEXTURE_IMAGE &&
            d->type != GR_TEXTURE_IMAGE_FAST)
            return fail(GR_ERR_INVALID_ARGUMENT, "unknown texture type");
        GR_HIP(hipSetDevice(c->device));
        TexRes t;
        t.live = true;
        t.dt.type = d->type;
        t.dt.w = d->width;
        t.dt.h = d->height;
        t.dt.wf = (float)d->width;
        t.dt.hf = (float)d->height;
        t.dt.scale = d->scale;
        t.dt.color = pack_rgba(d->color);
        t.dt.texels = nullptr;
        if (d->type == GR_TEXTURE_IMAGE || d->type == GR_TEXTURE_IMAGE_FAST) {
            if (d->width < 1 || d->height < 1 || !rgba)
                return fail(GR_ERR_INVALID_ARGUMENT, "image texture needs size and texels");
            if ((int64_t)d->width * d->height > ((int64_t)1 << 28))
                return fail(GR_ERR_INVALID_ARGUMENT, "texture larger than 2^28 texels");
            if (d->type == GR_TEXTURE_IMAGE_FAST && !(is_pow2(d->width) && is_pow2(d->height)))
                return fail(GR_ERR_INVALID_ARGUMENT, "IMAGE_FAST needs power-of-two sizes");
            size_t bytes = (size_t)d->width * d->height * 4;
            GR_HIP(hipMalloc(&t.texels, bytes));
            hipError_t e = hipMemcpy(t.texels, rgba, bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                (void)hipFree(t.texels);
                return fail(GR_ERR_DEVICE, hipGetErrorString(e));
            }
            t.dt.texels = (const uint32_t*)t.texels;
            t.bytes = (int64_t)bytes;
        }
        c->textures.push_back(t);
        c->tex_dirty = true;
        *out = (int32_t)c->textures.size() - 1;
        return GR_OK;
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

gr_status gr_texture_destroy(gr_context* c, int32_t h) {
    if (!c || h < 0 || h >= (int32_t)c->textures.size() || !c->textures[h].live)
        return fail(GR_ERR_INVALID_HANDLE, "invalid texture handle");
    gr_status st = sync_all(c);
    if (st) return st;
    if (c->textures[h].texels) GR_HIP(hipFree(c->textures[h].texels));
    c->textures[h] = TexRes{};
    c->tex_dirty = true;
    return GR_OK;
}

gr_status gr_mesh_create(gr_context* c, const gr_mesh_desc* d, int32_t* out) {
    try {
        if (!c || !d || !out) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
        GR_HIP(hipSetDevice(c->device));
        const int32_t V = d->num_vertices, N = d->num_vertex_normals, F = d->num_faces;
        if (V < 1 || !d->vertices)
            return fail(GR_ERR_INVALID_ARGUMENT, "mesh needs at least one vertex");
        if (N < 0 || (N > 0 && !d->vertex_normals) || F < 0 || (F > 0 && !d->vertex_indices))
            return fail(GR_ERR_INVALID_ARGUMENT, "inconsistent mesh arrays");
        for (int64_t k = 0; k < 3 * (int64_t)F; k++)
            if (d->vertex_indices[k] < 0 || d->vertex_indices[k] >= V)
                return fail(GR_ERR_INDEX_OUT_OF_RANGE, "vertex index out of range");

        MeshRes mr;
        mr.live = true;
        gr::DevMesh& m = mr.dm;
        std::memset(&m, 0, sizeof m);
        m.V = V;
        m.N = N;
        m.F = F;
        gr_status st;
        // Packed X,Y,Z positions (one 12-byte gather per face corner); W kept
        // only when some W != 1.
        {
            std::vector<float> xyz(3 * (size_t)V), w(V);
            bool w1 = true;
            for (int32_t i = 0; i < V; i++) {
                for (int k = 0; k < 3; k++) xyz[3 * (size_t)i + k] = d->vertices[4 * (size_t)i + k];
                w[i] = d->vertices[4 * (size_t)i + 3];
                uint32_t wb;
                std::memcpy(&wb, &w[i], 4);
                w1 = w1 && wb == 0x3F800000u;
            }
            void* p;
            if ((st = upload_array(c, mr, xyz.data(), 12 * (size_t)V, &p))) goto failed;
            m.pos = (const float*)p;
            if (!w1) {
                if ((st = upload_array(c, mr, w.data(), 4 * (size_t)V, &p))) goto failed;
                m.pw = (const float*)p;
            }
            // boundingBox (mesh.go:28-51)
            float mn[3] = {xyz[0], xyz[1], xyz[2]}, mx[3] = {xyz[0], xyz[1], xyz[2]};
            for (int32_t i = 0; i < V; i++) {
                for (int k = 0; k < 3; k++) {
                    mn[k] = go_min(mn[k], xyz[3 * (size_t)i + k]);
                    mx[k] = go_max(mx[k], xyz[3 * (size_t)i + k]);
                }
            }
            int k = 0;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                    for (int cc = 0; cc < 2; cc++) {
                        m.bbox[k][0] = a ? mx[0] : mn[0];
                        m.bbox[k][1] = b ? mx[1] : mn[1];
                        m.bbox[k][2] = cc ? mx[2] : mn[2];
                        m.bbox[k][3] = 1.0f;
                        k++;
                    }
        }
        if (N > 0) {
            std::vector<float> xyz(3 * (size_t)N), w(N);
            bool w1 = true;
            for (int32_t i = 0; i < N; i++) {
                for (int k = 0; k < 3; k++)
                    xyz[3 * (size_t)i + k] = d->vertex_normals[4 * (size_t)i + k];
                w[i] = d->vertex_normals[4 * (size_t)i + 3];
                uint32_t wb;
                std::memcpy(&wb, &w[i], 4);
                w1 = w1 && wb == 0x3F800000u;
            }
            void* p;
            if ((st = upload_array(c, mr, xyz.data(), 12 * (size_t)N, &p))) goto failed;
            m.nrm = (const float*)p;
            if (!w1) {
                if ((st = upload_array(c, mr, w.data(), 4 * (size_t)N, &p))) goto failed;
                m.nw = (const float*)p;
            }
        }
        if (F > 0) {
            void* p;
            if ((st = upload_array(c, mr, d->vertex_indices, 12 * (size_t)F, &p))) goto failed;
            m.vi = (const int32_t*)p;
            // Normal indices: a Go zero Face has all zeros when none are given.
            std::vector<int32_t> ni(3 * (size_t)F, 0);
            if (d->normal_indices) std::memcpy(ni.data(), d->normal_indices, ni.size() * 4);
            mr.ni_valid = true;
            if (N > 0)
                for (int32_t v : ni) mr.ni_valid = mr.ni_valid && v >= 0 && v < N;
            bool same = std::memcmp(ni.data(), d->vertex_indices, ni.size() * 4) == 0;
            if (!same) {
                if ((st = upload_array(c, mr, ni.data(), ni.size() * 4, &p))) goto failed;
                m.ni = (const int32_t*)p;
            }
            // UVs: a per-vertex table when every corner agrees with one.
            if (d->uvs) {
                std::vector<float> tab(2 * (size_t)V, 0.f);
                std::vector<uint8_t> seen(V, 0);
                bool per_vertex = true;
                for (int64_t k = 0; k < 3 * (int64_t)F && per_vertex; k++) {
                    int32_t v = d->vertex_indices[k];
                    uint32_t nu, nv, ou, ov;
                    std::memcpy(&nu, &d->uvs[2 * k], 4);
                    std::memcpy(&nv, &d->uvs[2 * k + 1], 4);
                    if (!seen[v]) {
                        seen[v] = 1;
                        tab[2 * v] = d->uvs[2 * k];
                        tab[2 * v + 1] = d->uvs[2 * k + 1];
                    } else {
                        std::memcpy(&ou, &tab[2 * v], 4);
                        std::memcpy(&ov, &tab[2 * v + 1], 4);
                        per_vertex = nu == ou && nv == ov;
                    }
                }
                if (per_vertex) {
                    if ((st = upload_array(c, mr, tab.data(), tab.size() * 4, &p))) goto failed;
                    m.uv_vert = (const float*)p;
                } else {
                    if ((st = upload_array(c, mr, d->uvs, 24 * (size_t)F, &p))) goto failed;
                    m.uv_face = (const float*)p;
                }
            }
            // Texture per face, or one texture for the whole mesh.
            m.mesh_tex = -1;
            if (d->face_textures) {
                bool uniform = true;
                for (int32_t f = 0; f < F; f++) {
                    int32_t t = d->face_textures[f];
                    if (t < -1 || t >= (int32_t)c->textures.size() || (t >= 0 && !c->textures[t].live)) {
                        st = fail(GR_ERR_INVALID_HANDLE, "invalid texture handle in face_textures");
                        goto failed;
                    }
                    uniform = uniform && t == d->face_textures[0];
                    if (t >= 0 && std::find(mr.textures.begin(), mr.textures.end(), t) == mr.textures.end())
                        mr.textures.push_back(t);
                }
                if (uniform) {
                    m.mesh_tex = d->face_textures[0];
                } else {
                    if ((st = upload_array(c, mr, d->face_textures, 4 * (size_t)F, &p))) goto failed;
                    m.face_tex = (const int32_t*)p;
                }
            }
            // Face normals (mesh.go:53-60) as given, or computed on the GPU.
            if (d->face_normals) {
                if ((st = upload_array(c, mr, d->face_normals, 16 * (size_t)F, &p))) goto failed;
                m.fn = (const float*)p;
            } else {
                void* fnp = nullptr;
                if (hipMalloc(&fnp, 16 * (size_t)F) != hipSuccess) {
                    st = fail(GR_ERR_OUT_OF_MEMORY, "hipMalloc face normals");
                    goto failed;
                }
                mr.allocs.push_back(fnp);
                mr.bytes += 16 * (int64_t)F;
                gr::launch_face_normals(m.pos, m.vi, F, (float*)fnp, c->stream);
                if (hipStreamSynchronize(c->stream) != hipSuccess || hipGetLastError() != hipSuccess) {
                    st = fail(GR_ERR_DEVICE, "face normal kernel failed");
                    goto failed;
                }
                m.fn = (const float*)fnp;
            }
        }
        c->meshes.push_back(mr);
        *out = (int32_t)c->meshes.size() - 1;
        return GR_OK;
    failed:
        free_mesh(mr);
        return st;
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

gr_status gr_mesh_destroy(gr_context* c, int32_t h) {
    if (!c || h < 0 || h >= (int32_t)c->meshes.size() || !c->meshes[h].live)
        return fail(GR_ERR_INVALID_HANDLE, "invalid mesh handle");
    gr_status st = sync_all(c);
    if (st) return st;
    free_mesh(c->meshes[h]);
    c->meshes[h] = MeshRes{};
    return GR_OK;
}

int64_t gr_mesh_device_bytes(gr_context* c, int32_t h) {
    if (!c || h < 0 || h >= (int32_t)c->meshes.size() || !c->meshes[h].live) return -1;
    return c->meshes[h].bytes;
}

void gr_frame_defaults(gr_frame* f, int32_t w, int32_t h) {
    if (!f) return;
    std::memset(f, 0, sizeof *f);
    f->width = w;
    f->height = h;
    f->flags = GR_DEFAULT_FLAGS;
    f->ref_tiles = 16;
    // NewScreenMatrix (matrix.go:108-118)
    float hw = (float)w / 2.0f, hh = (float)h / 2.0f;
    f->screen[0] = hw;
    f->screen[3] = hw;
    f->screen[5] = hh;
    f->screen[7] = hh;
    f->screen[10] = 0.5f;
    f->screen[11] = 0.5f;
    f->screen[15] = 1.0f;
    // Vec3{-1,1,1}.Normalize() (renderer.go:265, vector.go:92-94)
    float len = (float)std::sqrt((double)(1.0f + 1.0f + 1.0f));
    f->light[0] = -1.0f / len;
    f->light[1] = 1.0f / len;
    f->light[2] = 1.0f / len;
    f->ambient = 0.5f;
    f->diffuse = 0.5f;
    f->znear = 0.0f;
    f->zfar = 50.0f;
    const uint8_t face[4] = {200, 200, 200, 255}, clr[4] = {50, 50, 50, 255},
                  dot[4] = {100, 100, 100, 255};
    std::memcpy(f->face_color, face, 4);
    std::memcpy(f->clear_color, clr, 4);
    std::memcpy(f->dot_color, dot, 4);
    f->dot_step = 10;
}

gr_status gr_draw_batch(gr_context* c, const gr_batch_frame* frames, int32_t nf,
                        const gr_object* objs, int32_t no) {
    try {
        if (!c || !frames || (no > 0 && !objs)) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
        GR_HIP(hipSetDevice(c->device));
        return enqueue(c, frames, nf, objs, no);
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

gr_status gr_draw_device(gr_context* c, const gr_frame* frame, const gr_object* objs,
                         int32_t no, void* d_rgba, void* d_depth) {
    if (!frame) return fail(GR_ERR_INVALID_ARGUMENT, "null frame");
    gr_batch_frame bf;
    bf.frame = *frame;
    bf.first_object = 0;
    bf.num_objects = no;
    bf.d_rgba = d_rgba;
    bf.d_depth = d_depth;
    return gr_draw_batch(c, &bf, 1, objs, no);
}

gr_status gr_synchronize(gr_context* c) {
    if (!c) return fail(GR_ERR_INVALID_ARGUMENT, "null context");
    if (!c->h_ctrl) return sync_all(c);
    unsigned long long needed = 0;
    gr_status st = sync_and_check(c, &needed);
    if (st == GR_ERR_CAPACITY) {
        if (!needed) return st;  // the no-allocation draw's fixed array: message set
        grow_refs(c, needed);
        return fail(GR_ERR_CAPACITY, "tile-list capacity exceeded; redraw the batch");
    }
    return st;
}

gr_status gr_draw(gr_context* c, const gr_frame* frame, const gr_object* objs, int32_t no,
                  uint8_t* rgba, float* depth, gr_stats* stats) {
    try {
        if (!c || !frame || (no > 0 && !objs)) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
        GR_HIP(hipSetDevice(c->device));
        gr_status st = validate_frame(*frame);
        if (st) return st;
        size_t npx = (size_t)frame->width * frame->height;
        if ((st = ensure(c, c->out_rgba, npx * 4))) return st;
        if ((st = ensure(c, c->out_depth, npx * 4))) return st;
        // The caller's buffers are pageable: the frame goes to pinned memory
        // by a copy kernel (near the PCIe link's rate), then to them by a host
        // copy, instead of a synchronous device-to-pageable copy.
        if (c->h_frame_cap < npx * 8) {
            if (c->h_frame) GR_HIP(hipHostFree(c->h_frame));
            c->h_frame = nullptr;
            c->h_frame_cap = 0;
            GR_HIP(hipHostMalloc(&c->h_frame, npx * 8, hipHostMallocDefault));
            c->h_frame_cap = npx * 8;
        }
        char* const staged = (char*)c->h_frame;
        // Synchronous: each frame is waited for, so there is nothing to overlap.
        struct OnePipe {
            gr_context* c;
            ~OnePipe() { c->one_pipe = false; }
        } one{c};
        c->one_pipe = true;
        for (int attempt = 0; attempt < 4; attempt++) {
            st = gr_draw_device(c, frame, objs, no, c->out_rgba.p, c->out_depth.p);
            if (st) return st;
            // The copies follow the draw on its stream; a draw that overflowed
            // is drawn again and its copies are discarded.
            if (rgba) gr::launch_copy(staged, c->out_rgba.p, npx * 4, c->stream);
            if (depth) gr::launch_copy(staged + npx * 4, c->out_depth.p, npx * 4, c->stream);
            GR_HIP(hipGetLastError());
            unsigned long long needed = 0;
            st = sync_and_check(c, &needed);
            if (st == GR_ERR_CAPACITY) {
                grow_refs(c, needed);
                continue;
            }
            if (st) return st;
            if (rgba) std::memcpy(rgba, staged, npx * 4);
            if (depth) std::memcpy(depth, staged + npx * 4, npx * 4);
            if (stats) {
                std::vector<unsigned long long> sv;
                if ((st = read_stats(c, sv))) return st;
                to_gr_stats(sv.data(), stats);
            }
            return GR_OK;
        }
        return fail(GR_ERR_CAPACITY, "tile-list capacity could not be satisfied");
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

gr_status gr_present_draw(gr_context* c, const gr_frame* frame, const gr_object* objs,
                          int32_t no, int32_t with_depth) {
    try {
        if (!c || !frame || no < 0 || (no > 0 && !objs))
            return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
        if (with_depth != 0 && with_depth != 1)
            return fail(GR_ERR_INVALID_ARGUMENT, "with_depth must be 0 or 1");
        GR_HIP(hipSetDevice(c->device));
        gr_context::PresentSlot& ps = c->present[c->present_next];
        if (ps.pending)
            return fail(GR_ERR_INVALID_ARGUMENT,
                        "two frames in flight: gr_present_wait for the oldest first");
        gr_status st = validate_frame(*frame);
        if (st) return st;
        const size_t npx = (size_t)frame->width * frame->height;
        if ((st = ensure(c, ps.rgba, npx * 4)) || (st = ensure(c, ps.depth, npx * 4))) return st;
        const size_t hb = npx * (with_depth ? 8 : 4);
        if (ps.host_cap < hb) {
            if (ps.host) GR_HIP(hipHostFree(ps.host));
            ps.host = nullptr;
            ps.host_cap = 0;
            GR_HIP(hipHostMalloc(&ps.host, hb, hipHostMallocDefault));
            ps.host_cap = hb;
        }
        if (!ps.h_meta) GR_HIP(hipHostMalloc(&ps.h_meta, gr::kNumStats * 8 + 16, hipHostMallocDefault));
        if (!ps.copied) GR_HIP(hipEventCreateWithFlags(&ps.copied, hipEventDisableTiming));
        ps.frame = *frame;
        ps.objs.assign(objs, objs + no);
        ps.with_depth = with_depth != 0;
        if ((st = present_enqueue(c, ps))) return st;
        c->present_next ^= 1;
        return GR_OK;
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

gr_status gr_present_wait(gr_context* c, const uint8_t** rgba, const float** depth,
                          gr_stats* stats) {
    try {
        if (!c) return fail(GR_ERR_INVALID_ARGUMENT, "null context");
        GR_HIP(hipSetDevice(c->device));
        gr_context::PresentSlot& ps = c->present[c->present_head];
        if (!ps.pending) return fail(GR_ERR_INVALID_ARGUMENT, "no frame queued");
        for (int attempt = 0;; attempt++) {
            GR_HIP(hipEventSynchronize(ps.copied));
            uint32_t over = 0;
            std::memcpy(&over, (char*)ps.h_meta + gr::kNumStats * 8, 4);
            if (!over) break;
            // The frame's tile lists overflowed: grow them and draw it again
            // (after every queued draw, so the next frame keeps its order).
            unsigned long long needed = 0;
            gr_status st = sync_and_check(c, &needed);
            if (st && st != GR_ERR_CAPACITY) return st;
            if (attempt == 3)
                return fail(GR_ERR_CAPACITY, "tile-list capacity could not be satisfied");
            grow_refs(c, needed);
            if ((st = present_enqueue(c, ps))) return st;
        }
        ps.pending = false;
        c->present_head ^= 1;
        const unsigned long long* q = (const unsigned long long*)ps.h_meta;
        if (stats) to_gr_stats(q, stats);
        if (rgba) *rgba = (const uint8_t*)ps.host;
        if (depth)
            *depth = ps.with_depth
                         ? (const float*)((const char*)ps.host +
                                          (size_t)ps.frame.width * ps.frame.height * 4)
                         : nullptr;
        if (q[gr::kStatLongLines])
            return fail(GR_ERR_UNSUPPORTED, "overlay line longer than 65536 pixels");
        return GR_OK;
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

gr_status gr_last_stats(gr_context* c, gr_stats* stats, int32_t max_frames, int32_t* nf) {
    if (!c || !stats) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<unsigned long long> sv;
    gr_status st = read_stats(c, sv);
    if (st) return st;
    int32_t n = std::min(max_frames, c->last_frames);
    for (int32_t f = 0; f < n; f++) to_gr_stats(sv.data() + (size_t)f * gr::kNumStats, &stats[f]);
    if (nf) *nf = n;
    return GR_OK;
}

gr_status gr_device_alloc(gr_context* c, size_t bytes, void** out) {
    if (!c || !out) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    GR_HIP(hipSetDevice(c->device));
    GR_HIP(hipMalloc(out, std::max<size_t>(bytes, 4)));
    return GR_OK;
}

gr_status gr_device_free(gr_context* c, void* p) {
    if (!c) return fail(GR_ERR_INVALID_ARGUMENT, "null context");
    if (!p) return GR_OK;
    gr_status st = sync_all(c);
    if (st) return st;
    GR_HIP(hipFree(p));
    return GR_OK;
}

gr_status gr_device_fill(gr_context* c, void* p, int32_t value, size_t bytes) {
    if (!c || (!p && bytes)) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    if (!bytes) return GR_OK;
    gr_status st = after_draws(c);
    if (st) return st;
    GR_HIP(hipMemsetAsync(p, value, bytes, c->stream));
    return before_draws(c);
}

gr_status gr_device_copy(gr_context* c, void* dst, const void* src, size_t bytes) {
    if (!c || ((!dst || !src) && bytes)) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    GR_HIP(hipSetDevice(c->device));
    gr_status st = after_draws(c);
    if (st) return st;
    timed_launch(c, "k_copy", c->stream, [&] { gr::launch_copy(dst, src, bytes, c->stream); });
    GR_HIP(hipGetLastError());
    return before_draws(c);
}

gr_status gr_fog(gr_context* c, void* rgba, const void* depth, int32_t width, int32_t height,
                 float fog_start, float fog_end, const uint8_t color[4]) {
    if (!c || !rgba || !depth || !color) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    if (width < 1 || height < 1 || width > gr::kMaxDim || height > gr::kMaxDim)
        return fail(GR_ERR_INVALID_ARGUMENT, "frame size must be within 1..16384");
    GR_HIP(hipSetDevice(c->device));
    gr_status st = after_draws(c);
    if (st) return st;
    const uint32_t packed = pack_rgba(color);
    timed_launch(c, "k_fog", c->stream, [&] {
        gr::launch_fog(rgba, depth, (uint64_t)width * (uint64_t)height, fog_start, fog_end, packed,
                       c->stream);
    });
    GR_HIP(hipGetLastError());
    return before_draws(c);
}

gr_status gr_check_exact_division(gr_context* c, uint64_t count, uint64_t seed,
                                  uint64_t* mismatches, float* first_mismatch) {
    if (!c || !mismatches) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    GR_HIP(hipSetDevice(c->device));
    GR_HIP(hipStreamSynchronize(c->stream));
    void* buf = nullptr;
    GR_HIP(hipMalloc(&buf, 32));
    struct Free {
        void* p;
        ~Free() { (void)hipFree(p); }
    } guard{buf};
    GR_HIP(hipMemsetAsync(buf, 0, 32, c->stream));
    gr::launch_check_div(count, seed, (unsigned long long*)buf, (float*)((char*)buf + 16),
                         c->stream);
    GR_HIP(hipGetLastError());
    unsigned char host[32];
    GR_HIP(hipMemcpyAsync(host, buf, 32, hipMemcpyDeviceToHost, c->stream));
    GR_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(mismatches, host, 8);
    if (first_mismatch) std::memcpy(first_mismatch, host + 16, 16);
    return GR_OK;
}

gr_status gr_device_checksum(gr_context* c, const void* p, size_t bytes, void* digest) {
    if (!c || !digest || (!p && bytes)) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    if (bytes % 4) return fail(GR_ERR_INVALID_ARGUMENT, "checksum over whole 32-bit words only");
    GR_HIP(hipSetDevice(c->device));
    gr_status st = after_draws(c);
    if (st) return st;
    gr::launch_checksum(p, bytes / 4, (unsigned long long*)digest, c->stream);
    GR_HIP(hipGetLastError());
    return before_draws(c);
}

gr_status gr_copy_to_host(gr_context* c, void* dst, const void* src, size_t bytes) {
    if (!c || ((!dst || !src) && bytes)) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    gr_status st = sync_all(c);
    if (st) return st;
    if (bytes) GR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GR_OK;
}

gr_status gr_set_kernel_timing(gr_context* c, int32_t enabled) {
    if (!c) return fail(GR_ERR_INVALID_ARGUMENT, "null context");
    gr_status st = sync_all(c);
    if (st) return st;
    collect_times(c);
    c->timing = enabled != 0;
    return GR_OK;
}

int32_t gr_kernel_times(gr_context* c, const char** names, double* total_ms, int64_t* launches,
                        int32_t max) {
    if (!c) return 0;
    (void)sync_all(c);
    collect_times(c);
    int32_t i = 0;
    static std::mutex mu;
    static std::map<std::string, std::string> interned;
    for (auto& kv : c->kernel_ms) {
        if (i >= max) break;
        std::lock_guard<std::mutex> g(mu);
        auto it = interned.emplace(kv.first, kv.first).first;
        if (names) names[i] = it->second.c_str();
        if (total_ms) total_ms[i] = kv.second.first;
        if (launches) launches[i] = kv.second.second;
        i++;
    }
    return i;
}

void gr_reset_kernel_times(gr_context* c) {
    if (!c) return;
    (void)sync_all(c);
    collect_times(c);
    c->kernel_ms.clear();
}

gr_status gr_batch_graph_create(gr_context* c, const gr_batch_frame* frames, int32_t nf,
                                const gr_object* objs, int32_t no, gr_batch_graph** out) {
    if (!c || !frames || !out || (no > 0 && !objs))
        return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    gr_batch_graph* g = nullptr;
    try {
        GR_HIP(hipSetDevice(c->device));
        gr_status st = sync_all(c);
        if (st) return st;
        g = new gr_batch_graph();
        g->c = c;
        g->TW = c->cfg.tile_w;
        g->TH = c->cfg.tile_h;
        g->part_refs = (uint32_t)c->cfg.tile_split;
        g->refs_cap = refs_capacity_of(c);
        g->bf.assign(frames, frames + std::max(nf, 0));
        g->objs.assign(objs, objs + std::max(no, 0));
        g->frames.resize(std::max(nf, 1));
        g->dobj.resize(std::max(no, 1));
        g->xcd.resize(xcd_ints(std::max(nf, 1)));
        st = make_plan(c, frames, nf, objs, no, g->TW, g->TH, g->frames.data(), g->dobj.data(),
                       g->xcd.data(), g->P);
        if (!st && (hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking) != hipSuccess ||
                    hipEventCreateWithFlags(&g->done, hipEventDisableTiming) != hipSuccess))
            st = fail(GR_ERR_DEVICE, "graph stream or event");
        if (!st) {
            std::vector<gr::DevTexture> tt(std::max<size_t>(1, c->textures.size()));
            for (size_t i = 0; i < c->textures.size(); i++) tt[i] = c->textures[i].dt;
            st = ensure(c, g->textures, tt.size() * sizeof(gr::DevTexture));
            if (!st && hipMemcpy(g->textures.p, tt.data(), tt.size() * sizeof(gr::DevTexture),
                                 hipMemcpyHostToDevice) != hipSuccess)
                st = fail(GR_ERR_DEVICE, "graph texture table");
        }
        // Capture, then a first replay on the capture stream that sizes the
        // tile lists (grown and captured again while they overflow).
        for (int attempt = 0; !st; attempt++) {
            if ((st = graph_build(g))) break;
            if (hipGraphLaunch(g->geo_x, g->cap) != hipSuccess ||
                hipGraphLaunch(g->ras_x, g->cap) != hipSuccess) {
                st = fail(GR_ERR_DEVICE, "graph launch");
                break;
            }
            gr::DevControl ctl;
            if ((st = read_ctrl(c, g->B.ctrl, g->cap, &ctl))) break;
            if (ctl.unsupported) {
                st = fail(GR_ERR_UNSUPPORTED, "overlay line longer than 65536 pixels");
                break;
            }
            if (!ctl.overflow) break;
            if (attempt == 3) {
                st = fail(GR_ERR_CAPACITY, "tile-list capacity could not be satisfied");
                break;
            }
            g->refs_cap = std::max(g->refs_cap, grown(ctl.refs_needed));
        }
        if (st) {
            graph_free(g);
            return st;
        }
        c->graphs.push_back(g);
        *out = g;
        return GR_OK;
    } catch (...) {
        if (g) graph_free(g);
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

gr_status gr_batch_graph_update(gr_batch_graph* g, const gr_object* objs, int32_t no) {
    if (!g || (no > 0 && !objs)) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    if (no != (int32_t)g->objs.size())
        return fail(GR_ERR_INVALID_ARGUMENT, "update must give the graph's number of objects");
    for (int32_t k = 0; k < no; k++)
        if (objs[k].mesh != g->objs[k].mesh)
            return fail(GR_ERR_INVALID_ARGUMENT, "update must keep each object's mesh");
    try {
        gr_context* c = g->c;
        GR_HIP(hipSetDevice(c->device));
        g->objs.assign(objs, objs + no);
        Plan P;
        gr_status st = make_plan(c, g->bf.data(), (int)g->bf.size(), g->objs.data(), no, g->TW,
                                 g->TH, g->frames.data(), g->dobj.data(), g->xcd.data(), P);
        if (st) return st;
        g->P = P;
        // The header is read by the graph's last replay: wait for it first.
        if (g->last) GR_HIP(hipEventSynchronize(g->done));
        std::vector<unsigned char> h(g->P.header_bytes);
        write_header(g->P, g->B, h.data());
        GR_HIP(hipMemcpyAsync(g->B.header, h.data(), h.size(), hipMemcpyHostToDevice, g->cap));
        GR_HIP(hipStreamSynchronize(g->cap));
        return GR_OK;
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

gr_status gr_batch_graph_launch(gr_batch_graph* g) {
    if (!g) return fail(GR_ERR_INVALID_ARGUMENT, "null graph");
    try {
        GR_HIP(hipSetDevice(g->c->device));
        return graph_launch(g);
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

void gr_batch_graph_destroy(gr_batch_graph* g) {
    if (!g) return;
    gr_context* c = g->c;
    (void)hipSetDevice(c->device);
    (void)sync_all(c);
    (void)hipStreamSynchronize(g->cap);
    c->graphs.erase(std::remove(c->graphs.begin(), c->graphs.end(), g), c->graphs.end());
    graph_free(g);
}

gr_status gr_autotune(gr_context* c, const gr_batch_frame* frames, int32_t nf,
                      const gr_object* objs, int32_t no, int32_t reps, gr_tune_result* results,
                      int32_t max_results, int32_t* num_results, gr_config* best) {
    if (!c || !frames || nf <= 0 || (no > 0 && !objs) || reps < 1)
        return fail(GR_ERR_INVALID_ARGUMENT, "null argument, no frames or reps < 1");
    if (num_results) *num_results = 0;
    const gr_config saved = c->cfg;
    const bool saved_one = c->one_pipe;
    hipEvent_t ea = nullptr, eb = nullptr;
    unsigned long long* d_sum = nullptr;
    struct Restore {
        gr_context* c;
        bool one;
        hipEvent_t* a;
        hipEvent_t* b;
        unsigned long long** sum;
        ~Restore() {
            c->one_pipe = one;
            if (*a) (void)hipEventDestroy(*a);
            if (*b) (void)hipEventDestroy(*b);
            if (*sum) (void)hipFree(*sum);
        }
    } restore{c, saved_one, &ea, &eb, &d_sum};
    try {
        GR_HIP(hipSetDevice(c->device));
        gr_status st = sync_all(c);
        if (st) return st;
        GR_HIP(hipEventCreate(&ea));
        GR_HIP(hipEventCreate(&eb));
        GR_HIP(hipMalloc((void**)&d_sum, 8));
        c->one_pipe = true;
        static const int kTiles[4] = {8, 16, 32, 64};
        static const int kSplits[4] = {256, 512, 1024, 4096};
        bool have_best = false, have_sum = false;
        double best_ms = 0.0;
        unsigned long long first_sum = 0;
        gr_config best_cfg = saved;
        int32_t nres = 0;
        for (int tw : kTiles)
            for (int th : kTiles)
                for (int split : kSplits) {
                    c->cfg = saved;
                    c->cfg.tile_w = tw;
                    c->cfg.tile_h = th;
                    c->cfg.tile_split = split;
                    // Untimed launch: grows the tile lists while they overflow,
                    // then sizes them to the batch's references plus a quarter.
                    unsigned long long needed = 0;
                    for (int attempt = 0;; attempt++) {
                        if ((st = enqueue(c, frames, nf, objs, no))) return st;
                        st = sync_and_check(c, &needed);
                        if (st != GR_ERR_CAPACITY) break;
                        if (attempt == 3)
                            return fail(GR_ERR_CAPACITY, "tile-list capacity could not be satisfied");
                        grow_refs(c, needed);
                    }
                    if (st) return st;
                    std::vector<unsigned long long> sv;
                    if ((st = read_stats(c, sv))) return st;
                    unsigned long long refs = 0;
                    for (int f = 0; f < nf; f++) refs += sv[(size_t)f * gr::kNumStats + gr::kStatRefs];
                    // The capacity also holds the triangle records the face blocks claimed.
                    gr::DevBatchCounters bc;
                    GR_HIP(hipMemcpy(&bc, c->last_counters, sizeof bc, hipMemcpyDeviceToHost));
                    refs = std::max(refs, bc.recs_used);
                    c->cfg.refs_capacity = std::max<int64_t>(1 << 16, (int64_t)grown(refs));
                    GR_HIP(hipEventRecord(ea, c->stream));
                    for (int r = 0; r < reps; r++)
                        if ((st = enqueue(c, frames, nf, objs, no))) return st;
                    GR_HIP(hipEventRecord(eb, c->stream));
                    if ((st = sync_and_check(c, &needed))) return st;
                    float ms = 0.f;
                    GR_HIP(hipEventElapsedTime(&ms, ea, eb));
                    // Digest of every frame's colour and depth.
                    gr::launch_fill32((uint32_t*)d_sum, 0u, 2, c->stream);
                    for (int f = 0; f < nf; f++) {
                        const size_t px = (size_t)frames[f].frame.width * frames[f].frame.height;
                        gr::launch_checksum(frames[f].d_rgba, px, d_sum, c->stream);
                        gr::launch_checksum(frames[f].d_depth, px, d_sum, c->stream);
                    }
                    GR_HIP(hipGetLastError());
                    unsigned long long sum = 0;
                    GR_HIP(hipMemcpyAsync(c->h_ctrl, d_sum, 8, hipMemcpyDeviceToHost, c->stream));
                    GR_HIP(hipStreamSynchronize(c->stream));
                    std::memcpy(&sum, c->h_ctrl, 8);
                    const double per = (double)ms / reps;
                    if (results && nres < max_results) {
                        gr_tune_result& tr = results[nres];
                        tr.tile_w = tw;
                        tr.tile_h = th;
                        tr.tile_split = split;
                        tr.raster_waves = c->cfg.raster_waves;
                        tr.ms_per_launch = per;
                        tr.checksum = sum;
                        nres++;
                    }
                    if (!have_sum) {
                        first_sum = sum;
                        have_sum = true;
                    } else if (sum != first_sum) {
                        c->cfg = saved;
                        if (num_results) *num_results = nres;
                        return fail(GR_ERR_DEVICE, "autotune: the image depends on the "
                                                   "configuration (tile " +
                                                       std::to_string(tw) + "x" +
                                                       std::to_string(th) + ", split " +
                                                       std::to_string(split) + ")");
                    }
                    if (!have_best || per < best_ms) {
                        have_best = true;
                        best_ms = per;
                        best_cfg = c->cfg;
                    }
                }
        best_cfg.pipelines = saved.pipelines;
        c->cfg = best_cfg;
        if (best) *best = best_cfg;
        if (num_results) *num_results = nres;
        return GR_OK;
    } catch (...) {
        c->cfg = saved;
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

}  // extern "C"
