// Device-side data layout of the render path (shared by host launch code and
// kernels). All structs are plain data copied to HBM per draw.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gr {

// Faces handled by one geometry workgroup, and the output-triangle slots it
// reserves: a face yields at most 7 fan triangles after clipping against 6
// planes (clipping.go:8-10), so slot = wg * kSlotsPerWg + local prefix keeps the
// reference's (object, face, fan) draw order as a plain integer order.
constexpr int kFacesPerWg = 256;
constexpr int kMaxFan = 7;
constexpr int kSlotsPerWg = kFacesPerWg * kMaxFan;  // 1792
constexpr int kMaxClip = 9;                         // maxClipPoints (clipping.go:10)

// Triangle records: face block b (of the batch) owns records
// [b * kRecsPerBlock, (b + 1) * kRecsPerBlock), enough for every face of the
// block unclipped plus 64 more records. A block that may need more (it
// reserves kMaxFan records for each face that reaches the full clipper)
// claims the rest from the overflow records, split into kRecPools pools
// (block b from pool b % pools), each with its own counter kPoolStride words
// from the next so that the claims of a launch do not serialise on one
// address. The overflow records number one per kOvfRefs tile-list entries of
// the batch's capacity, which grows on demand.
constexpr int kRecsPerBlock = kFacesPerWg + 64;  // 320
constexpr int kRecPools = 64;
constexpr int kPoolStride = 64;
constexpr int kOvfRefs = 16;
constexpr int kMaxDim = 16384;                      // framebuffer side limit

// Coordinates within +-kFastCoord keep every edge value of a pixel inside a
// 16384^2 framebuffer below 2^31, so int32 arithmetic equals the reference's
// int64 (rasterizer.go:107-117) exactly.
constexpr int kFastCoord = 16384;

// Box visibility (clipping.go:22-26).
enum : int32_t { kBoxOutside = 0, kBoxIntersect = 1, kBoxInside = 2 };

// Flags mirror include/gorender_hip.h.
enum : uint32_t {
    kShowFaces = 1u << 0,
    kBackfaceCulling = 1u << 1,
    kLighting = 1u << 2,
    kFlatShading = 1u << 3,
    kFrustumClipping = 1u << 4,
    kShowTextures = 1u << 5,
    kShowEdges = 1u << 6,
    kShowVertices = 1u << 7,
    kShowTileBounds = 1u << 8,
    kShowCrosshair = 1u << 9,
};

struct DevMesh {
    int32_t V, N, F;
    int32_t mesh_tex;        // texture used by every face when face_tex == nullptr
    const float* pos;        // [V*3] packed X,Y,Z (one 12-byte gather per corner)
    const float* pw;         // [V] W, nullptr when every W == 1
    const float* nrm;        // [N*3] packed vertex normals
    const float* nw;         // [N] W, nullptr when every W == 1
    const int32_t* vi;       // [F*3]
    const int32_t* ni;       // [F*3]; nullptr when identical to vi
    const float* uv_vert;    // [V*2] per-vertex UVs, or nullptr
    const float* uv_face;    // [F*6] per-corner UVs, or nullptr
    const int32_t* face_tex; // [F] texture slot per face or nullptr
    const float* fn;         // [F*4] face normals
    float bbox[8][4];        // boundingBox corners (mesh.go:28-51)
};

struct DevObject {
    DevMesh m;
    float mvp[16];
    float world[16];
    int32_t frame;          // index in the batch
    int32_t gouraud;        // vertex-normal lighting in use
    int32_t face_block;     // first k_face block (batch-global)
    int32_t face_blocks;    // number of k_face blocks
    int32_t seq_block;      // first face block index within the frame (draw order)
    int32_t dense;          // >= 1 face per 4 pixels of the projected bounding box
    int32_t vis;            // BoxVisibility of the object (kBoxOutside/Intersect/Inside)
};

struct DevTexture {
    int32_t type, w, h, pad;
    float wf, hf, scale;
    uint32_t color;
    const uint32_t* texels;
};

struct DevControl;

struct DevFrame {
    int32_t W, H;
    uint32_t flags;
    int32_t ref_tiles;
    float screen[16];
    float light[3];
    float ambient, diffuse, znear, zfar;
    uint32_t face_color, clear_color, dot_color;
    int32_t dot_step;
    float ref_tile_rect[16][4];  // reference TPF grid (renderer.go:50-76)
    int32_t tiles_x, tiles_y;    // raster tiles of this frame
    int32_t tile_begin;          // first tile in the batch tile arrays
    int32_t face_block_begin;    // first k_face block of this frame
    int32_t face_blocks;         // k_face blocks of this frame
    int32_t obj_begin;           // first DevObject of this frame (objects are by frame)
    int32_t obj_count;
    uint8_t* rgba;               // device outputs
    float* depth;
    uint32_t* okey;              // overlay keys [H*W] or null: (slot+1) << 1 | vertex
    struct OverlayRec* orec;     // per triangle record when okey is set
    DevControl* ctrl;
    // Triangle records of the batch. A slot (the draw-order number of a
    // triangle: frame-local face block * kSlotsPerWg + its offset in the
    // block) has its record in the block's own range (offsets below
    // kRecsPerBlock) or at rbase[block] + offset - kRecsPerBlock.
    struct RasterRec* rrec;
    struct ShadeRec* srec;
    struct WideRec* wrec;
    const uint32_t* rbase;       // first overflow record of each face block of this frame
};

// Per-triangle raster set-up, 80 bytes: what the visibility pass reads. The
// four divisors and their reciprocals come first (two aligned 16-byte reads).
struct alignas(16) RasterRec {
    float sf;                    // float32(fx12+fx20+fx01), constant per triangle
    float z0, z1, z2;            // clip-space W of the vertices (renderer.go:369)
    float rsf, rz0, rz1, rz2;    // RN(1/divisor) when every divisor allows the
                                 // reciprocal-based exact quotient, else all 0
    int16_t bx0, bx1, by0, by1;  // clamped pixel bounding box; bx0 > bx1 = empty
    uint32_t wide;               // 1: coordinates beyond kFastCoord, use WideRec
    int32_t e[9];                // A01,B01,C01', A12,B12,C12', A20,B20,C20'
};
static_assert(sizeof(RasterRec) == 80, "RasterRec is five 16-byte pieces");

// Per-triangle shading inputs, 48 bytes: with pieces 3 and 4 of the RasterRec
// (the edge coefficients of edges 12 and 20), all the deferred shading of a
// pixel reads, unless its zRec is zero.
struct alignas(16) ShadeRec {
    float uz[3];  // u_i / z_i (rasterizer.go:132-137)
    float vz[3];
    float in[3];  // intensities
    int32_t tex;  // texture slot or -1
    float sf;     // the RasterRec's sf
    uint32_t rsf; // its rsf (0: IEEE division), or kShadeWide for a WideRec triangle
};
constexpr uint32_t kShadeWide = 0x7FC0DEADu;  // a NaN: never a reciprocal

// Screen positions of a post-clip triangle for the debug overlays, 32 bytes.
struct alignas(16) OverlayRec {
    float sx[3], sy[3];
    uint32_t binned;  // drawn by drawProjection (renderer.go:166-217)
    uint32_t pad;
};

// Exact int64 coordinates for triangles outside the int32 fast range.
struct alignas(16) WideRec {
    int64_t x[3], y[3];
    uint32_t adj;  // top-left adjustment bits of edges 01, 12, 20
    uint32_t pad[3];
};

// Per-frame counters (device), in order of gr_stats.
enum : int32_t {
    kStatTpf = 0,
    kStatTriangles,
    kStatCovered,
    kStatShaded,
    kStatFront,
    kStatClipOverflow,
    kStatRefs,
    kStatSplit,  // tiles whose list was split across raster workgroups
    kStatLongLines,  // overlay lines too long to draw (not in gr_stats)
    kNumStats
};

// Per-batch counters, uploaded as zeros with the descriptors of every draw.
struct DevBatchCounters {
    unsigned long long refs_used;  // tile-list entries allocated so far
    uint32_t n_items;              // raster work items (later parts of split tiles)
    uint32_t next_item;            // next work item to take
    uint32_t n_slabs;              // slabs handed to split tiles
    uint32_t pad;
    unsigned long long recs_used;  // triangle records the face blocks claimed (k_bin_scan)
};

// Tile not split (no slab); also the end of the raster work queue.
constexpr uint32_t kNoSlab = 0xFFFFFFFFu;

// Raster work item: a later part (>= 1) of a split tile.
struct RasterItem {
    uint32_t tile, part;
};

// Sticky control words (cleared by the host when it reads them): set when a
// frame of any batch needed more tile-list entries or triangle records than
// were allocated (one capacity sizes both).
struct DevControl {
    unsigned long long refs_needed;  // the capacity that would have sufficed
    uint32_t overflow;
    uint32_t unsupported;  // an overlay line longer than kMaxOverlayLine
};

// Longest overlay line drawn (pixels); longer ones make the draw unsupported.
constexpr int64_t kMaxOverlayLine = 1 << 16;

}  // namespace gr
