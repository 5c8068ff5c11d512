// gfx950 kernels of the render path.
//
//   k_face      per face: MVP transform of the three corners, backface test,
//               lighting, frustum clipping (wave-cooperative Sutherland-Hodgman
//               with ballot/popcount compaction), perspective divide + screen
//               transform, full-screen bin test, raster set-up. One workgroup
//               = 256 faces = 1792 slots in draw order. The workgroup then
//               counts its tile references in LDS and adds them to the global
//               tile counts.
//   k_bin_scan  one workgroup per frame: tile counts -> tile-list offsets.
//   k_bin_fill  per face workgroup: scatter slots into tile lists, with one
//               global atomic per distinct tile per workgroup.
//   k_raster    one workgroup per screen tile: triangles staged in LDS,
//               visibility with a 64-bit (depth, draw-sequence) max key in LDS,
//               one deferred shading pass and a single write of colour + depth.
//
// Every float32 expression keeps the reference's evaluation order; the build
// uses -ffp-contract=off and correctly rounded division/sqrt.
#pragma clang fp contract(off)


#include <algorithm>
#include <cstddef>
#include <cstring>

#include "gr_device.h"
#include "gr_launch.h"

namespace gr {

// ---------------------------------------------------------------------------
// Go conversion semantics
// ---------------------------------------------------------------------------

// int(float32) on amd64 (CVTTSS2SQ): truncation, NaN/out of range -> INT64_MIN.
__device__ __forceinline__ int64_t go_int(float x) {
    if (!(x >= -9223372036854775808.0f && x < 9223372036854775808.0f)) return INT64_MIN;
    if (x > -2147483648.0f && x < 2147483648.0f) return (int64_t)(int32_t)x;
    return (int64_t)x;
}

// uint8(float32): CVTTSS2SL then the low byte; NaN/out of int32 range -> 0.
// One compare: -2^31 itself converts to INT32_MIN, whose low byte is 0 too.
__device__ __forceinline__ uint32_t go_uint8(float x) {
    return __builtin_fabsf(x) < 2147483648.0f ? (uint32_t)(int32_t)x & 0xFFu : 0u;
}

__device__ __forceinline__ bool f_signbit(float a) { return (__float_as_uint(a) >> 31) != 0; }

// Go builtin min/max on floats: NaN wins, min(-0,+0) = -0.
__device__ __forceinline__ float go_min(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a < b) return a;
    if (b < a) return b;
    return f_signbit(a) ? a : b;
}
__device__ __forceinline__ float go_max(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a > b) return a;
    if (b > a) return b;
    return f_signbit(a) ? b : a;
}

__device__ __forceinline__ int64_t wadd(int64_t a, int64_t b) {
    return (int64_t)((uint64_t)a + (uint64_t)b);
}
__device__ __forceinline__ int64_t wsub(int64_t a, int64_t b) {
    return (int64_t)((uint64_t)a - (uint64_t)b);
}
__device__ __forceinline__ int64_t wmul(int64_t a, int64_t b) {
    return (int64_t)((uint64_t)a * (uint64_t)b);
}

// Load and store of one value through a global-address-space pointer: the
// pointers of DevFrame and DevMesh are generic, and a generic (flat) access
// also counts against the LDS wait counter, so a later LDS wait would wait for
// it too.
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
    *(__attribute__((address_space(1))) T*)p = v;
}
// Load of a 16-byte-aligned record through a global-address-space pointer
// (global_load rather than flat, so LDS traffic is not waited on with it).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ T gload(const T* p) {
    static_assert(sizeof(T) % 16 == 0 && alignof(T) >= 16, "16-byte records only");
    typedef const __attribute__((address_space(1))) u32x4* gptr;
    const gptr src = (gptr)p;
    u32x4 tmp[sizeof(T) / 16];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 16); i++) tmp[i] = src[i];
    T out;
    __builtin_memcpy(&out, tmp, sizeof(T));
    return out;
}

// Store of a 16-byte-aligned record through a global-address-space pointer.
template <typename T>
__device__ __forceinline__ void gstore(T* p, const T& v) {
    static_assert(sizeof(T) % 16 == 0 && alignof(T) >= 16, "16-byte records only");
    typedef __attribute__((ext_vector_type(4))) uint32_t v4;
    typedef __attribute__((address_space(1))) v4* gptr;
    v4 tmp[sizeof(T) / 16];
    __builtin_memcpy(tmp, &v, sizeof(T));
    const gptr dst = (gptr)p;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 16); i++) dst[i] = tmp[i];
}

// A*x + B*y + C in wrapping 32-bit arithmetic, for A, B, x, y that fit 24
// bits signed, as two multiply-adds (v_mad_i32_i24). The empty asm keeps the
// partial sum whole, so that the compiler does not turn the two products and
// two additions into two multiplies and a three-operand add.
__device__ __forceinline__ uint32_t edge24(int A, int x, int B, int y, uint32_t C) {
    uint32_t u = (uint32_t)__mul24(B, y) + C;
    asm("" : "+v"(u));
    return (uint32_t)__mul24(A, x) + u;
}

// ---------------------------------------------------------------------------
// Exact float32 division from a correctly rounded reciprocal
// ---------------------------------------------------------------------------

// |d| in [2^-40, 2^41): quotients a/d of numerators in [2^-70, 2^70] then keep
// every intermediate of div_rcp in the normal range.
__device__ __forceinline__ bool rcp_range(float d) {
    const uint32_t e = (__float_as_uint(d) >> 23) & 0xFFu;
    return e - (127u - 40u) <= 80u;
}

// RN(a / d) given y = RN(1 / d). q0 = RN(a*y) is within 1.5 ulp of a/d; one
// residual correction brings it within 1 ulp; with y the correctly rounded
// reciprocal and q within 1 ulp, the residual a - d*q is exact and
// RN(q + r*y) = RN(a/d) (Markstein, IBM J. Res. Dev. 34(1), 1990; the
// Handbook of Floating-Point Arithmetic states it for any precision p).
// Domain: d with rcp_range(d), a == +0 or 2^-70 <= |a| <= 2^70 (a == -0
// would come back as +0 for d > 0).
__device__ __forceinline__ float div_rcp(float a, float d, float y) {
    float q = a * y;
    float r = __builtin_fmaf(-d, q, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-d, q, a);
    return __builtin_fmaf(r, y, q);
}

// ---------------------------------------------------------------------------
// Vector helpers in reference order
// ---------------------------------------------------------------------------

// asm_purego.go:9-19 / matrix.go:167-173: ((m0*x + m1*y) + m2*z) + m3*w per row.
__device__ __forceinline__ float4 xform(const float* m, float x, float y, float z, float w) {
    float4 r;
    r.x = m[0] * x + m[1] * y + m[2] * z + m[3] * w;
    r.y = m[4] * x + m[5] * y + m[6] * z + m[7] * w;
    r.z = m[8] * x + m[9] * y + m[10] * z + m[11] * w;
    r.w = m[12] * x + m[13] * y + m[14] * z + m[15] * w;
    return r;
}

// vector.go:115-125 (Vec4.Length + Divide, W dropped by ToVec3), then the
// Lambert term of renderer.go:331 / 336-337.
__device__ __forceinline__ float lambert(float4 n, const float* light, float ambient,
                                         float diffuse) {
    float len = __builtin_sqrtf(n.x * n.x + n.y * n.y + n.z * n.z + n.w * n.w);
    float x = n.x / len, y = n.y / len, z = n.z / len;
    float d = x * light[0] + y * light[1] + z * light[2];
    return ambient + d * diffuse;
}

// Clip-space vertex with its attributes.
struct CV {
    float x, y, z, w, u, v, i;
};

// clipping.go:93-129 plane (point, normal) for index p in draw order
// Left, Right, Top, Bottom, Near, Far.
__device__ __forceinline__ void plane_of(int p, float zn, float zf, float4& P, float4& N) {
    switch (p) {
    case 0: P = make_float4(-1.f, 0.f, 0.f, 1.f); N = make_float4(1.f, 0.f, 0.f, 1.f); break;
    case 1: P = make_float4(1.f, 0.f, 0.f, 1.f); N = make_float4(-1.f, 0.f, 0.f, 1.f); break;
    case 2: P = make_float4(0.f, -1.f, 0.f, 1.f); N = make_float4(0.f, 1.f, 0.f, 1.f); break;
    case 3: P = make_float4(0.f, 1.f, 0.f, 1.f); N = make_float4(0.f, -1.f, 0.f, 1.f); break;
    case 4: P = make_float4(0.f, 0.f, zn, 1.f); N = make_float4(0.f, 0.f, -1.f, 0.f); break;
    default: P = make_float4(0.f, 0.f, zf, 1.f); N = make_float4(0.f, 0.f, 1.f, 0.f); break;
    }
}

// clipping.go:74-76: (q - P) . N <= 0
__device__ __forceinline__ bool inside(const float4& P, const float4& N, const CV& q) {
    float d = (q.x - P.x) * N.x + (q.y - P.y) * N.y + (q.z - P.z) * N.z + (q.w - P.w) * N.w;
    return d <= 0.f;
}

// clipping.go:79-86 with q0 = A, q1 = B, then lerp32 / lerpUV (clipping.go:156-165).
__device__ __forceinline__ CV intersect(const float4& P, const float4& N, const CV& A,
                                        const CV& B) {
    float ux = B.x - A.x, uy = B.y - A.y, uz = B.z - A.z, uw = B.w - A.w;
    float wx = A.x - P.x, wy = A.y - P.y, wz = A.z - P.z, ww = A.w - P.w;
    float d = N.x * ux + N.y * uy + N.z * uz + N.w * uw;
    float n = -(N.x * wx + N.y * wy + N.z * wz + N.w * ww);
    float f = n / d;
    CV r;
    r.x = A.x + ux * f;
    r.y = A.y + uy * f;
    r.z = A.z + uz * f;
    r.w = A.w + uw * f;
    r.i = A.i + (B.i - A.i) * f;
    r.u = A.u + (B.u - A.u) * f;
    r.v = A.v + (B.v - A.v) * f;
    return r;
}

// Inside bits of a vertex for the 6 planes (bit p = inside plane p).
__device__ __forceinline__ uint32_t inside_mask(const CV& q, float zn, float zf) {
    uint32_t m = 0;
#pragma unroll
    for (int p = 0; p < 6; p++) {
        float4 P, N;
        plane_of(p, zn, zf, P, N);
        m |= inside(P, N, q) ? (1u << p) : 0u;
    }
    return m;
}

// renderer.go:246-250 (clip space, before the divide).
__device__ __forceinline__ bool facing_camera(const float4& v0, const float4& v1,
                                              const float4& v2) {
    float ax = v1.x - v0.x, ay = v1.y - v0.y, az = v1.z - v0.z;
    float bx = v2.x - v0.x, by = v2.y - v0.y, bz = v2.z - v0.z;
    float nx = ay * bz - az * by;
    float ny = az * bx - ax * bz;
    float nz = ax * by - ay * bx;
    float d = nx * (0.f - v0.x) + ny * (0.f - v0.y) + nz * (0.f - v0.z);
    return d > 0.f;
}

// Units x, x + 8, x + 16, ... of a frame with n units run on XCD x.
__host__ __device__ __forceinline__ int xcd_count(int x, int n) {
    return n > x ? (n - x + 7) >> 3 : 0;
}

template <typename F>
__device__ __forceinline__ int last_le(int n, int b, F key);

// Workgroup h -> (frame, unit within the frame) in the XCD-aware order: the
// hardware places workgroup h on XCD h % 8, and XCD x takes units x, x + 8, ...
// of every frame, frame after frame, so a unit's inputs stay in one XCD's L2
// across the frames of a batch while neighbouring units (whose cost varies
// together) spread over all XCDs. Returns false for the padding workgroups.
__device__ __forceinline__ bool xcd_unit(const int* table, int nframes, int h, int& f,
                                         int& local) {
    const int x = h & 7, k = h >> 3;
    const int* P = table + x * (nframes + 1);
    if (k >= P[nframes]) return false;
    f = last_le(nframes, k, [&](int i) { return P[i]; });
    local = x + 8 * (k - P[f]);
    return true;
}

// Binary search: last index i in [0, n) with key(i) <= b.
template <typename F>
__device__ __forceinline__ int last_le(int n, int b, F key) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (key(mid) <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// The wave's total of a per-lane count that only lane 0 holds in the common
// case (the raster's covered-fragment count is wave-uniform except on the
// int64 path): the full reduction runs only when another lane holds a count.
__device__ __forceinline__ unsigned long long wave_total(unsigned long long v, int lane) {
    if (__ballot(v != 0ull && lane != 0) == 0ull)
        return __builtin_amdgcn_readfirstlane((unsigned)v) |
               ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(v >> 32)) << 32);
    return wave_sum_u64(v);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return (1ull << lane) - 1ull; }

__device__ __forceinline__ CV shfl_cv(const CV& a, int src) {
    CV r;
    r.x = __shfl(a.x, src, 64);
    r.y = __shfl(a.y, src, 64);
    r.z = __shfl(a.z, src, 64);
    r.w = __shfl(a.w, src, 64);
    r.u = __shfl(a.u, src, 64);
    r.v = __shfl(a.v, src, 64);
    r.i = __shfl(a.i, src, 64);
    return r;
}

// ---------------------------------------------------------------------------
// Wave-cooperative ClipTriangle (clipping.go:167-236)
// ---------------------------------------------------------------------------

// Every lane passes the same triangle (a, b, c). Per plane pass, lane k owns
// edge (k, (k+1) % n) of the polygon and contributes [X, A], [A], [X] or []
// exactly as the reference's loop body; ballots give each lane its output
// position. The polygon ping-pongs between buf0 and buf1 in LDS (one pair per
// wave) and ends in buf0 after the 6 passes. Returns the final vertex count
// (uniform), 0 when clipped away, -1 when a pass would exceed 9 vertices.
__device__ int wave_clip(CV* buf0, CV* buf1, const CV& a, const CV& b, const CV& c, float zn,
                         float zf, int lane) {
    if (lane == 0) buf0[0] = a;
    if (lane == 1) buf0[1] = b;
    if (lane == 2) buf0[2] = c;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int n = 3;
    CV* in = buf0;
    CV* out = buf1;
    for (int p = 0; p < 6; p++) {
        float4 P, N;
        plane_of(p, zn, zf, P, N);
        const bool act = lane < n;
        const int ai = act ? (lane + 1 == n ? 0 : lane + 1) : 0;
        const CV A = in[ai];
        const CV B = in[act ? lane : 0];
        const bool ia = act && inside(P, N, A);
        const bool ib = act && inside(P, N, B);
        const bool one = ia || ib;     // at least one output vertex
        const bool two = ia && !ib;    // [X, A]
        const uint64_t b1 = __ballot(one), b2 = __ballot(two);
        const uint64_t lt = lanemask_lt(lane);
        const int pos = __popcll(b1 & lt) + __popcll(b2 & lt);
        const int m = __popcll(b1) + __popcll(b2);
        if (m > kMaxClip) return -1;
        if (m == 0) return 0;
        if (ia != ib) {
            CV X = intersect(P, N, A, B);
            out[pos] = X;
            if (two) out[pos + 1] = A;
        } else if (one) {
            out[pos] = A;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        n = m;
        CV* t = in;
        in = out;
        out = t;
    }
    return n;
}

// ---------------------------------------------------------------------------
// Per-lane clipping of a triangle that straddles one plane
// ---------------------------------------------------------------------------

__device__ __forceinline__ CV sel(bool c, const CV& a, const CV& b) {
    CV r;
    r.x = c ? a.x : b.x;
    r.y = c ? a.y : b.y;
    r.z = c ? a.z : b.z;
    r.w = c ? a.w : b.w;
    r.u = c ? a.u : b.u;
    r.v = c ? a.v : b.v;
    r.i = c ? a.i : b.i;
    return r;
}

// Append v at position m of q[0..3] with register selects (no dynamic index).
__device__ __forceinline__ void append4(CV q[4], int& m, const CV& v) {
#pragma unroll
    for (int k = 0; k < 4; k++) q[k] = sel(m == k, v, q[k]);
    m++;
}

// ClipTriangle for a triangle whose three vertices are inside the planes before
// p and straddle plane p (at least one vertex inside). Passes before p only
// rotate the triangle left by one (clipping.go:204-224 with every vertex
// inside), so plane p sees it rotated by p mod 3; the pass over p runs the
// reference's edge loop. Each later plane is then either passed by every
// vertex (a pure rotation), failed by every vertex (the reference returns 0 at
// that pass, clipping.go:227-229), or straddled: only the last case needs the
// full clipper. Returns -1 for it, 0 when clipped away, otherwise the vertex
// count (3 or 4) of the final polygon in q.
__device__ int clip_one_plane(int p, const CV& c0, const CV& c1, const CV& c2, float zn,
                              float zf, CV q[4]) {
    const int r = p % 3;
    const CV R0 = sel(r == 0, c0, sel(r == 1, c1, c2));
    const CV R1 = sel(r == 0, c1, sel(r == 1, c2, c0));
    const CV R2 = sel(r == 0, c2, sel(r == 1, c0, c1));
    float4 P, N;
    plane_of(p, zn, zf, P, N);
    const bool i0 = inside(P, N, R0), i1 = inside(P, N, R1), i2 = inside(P, N, R2);
    q[0] = q[1] = q[2] = q[3] = R0;
    int m = 0;
    // edges b = 0, 1, 2 with A = R[(b+1) % 3], B = R[b]
    const CV As[3] = {R1, R2, R0}, Bs[3] = {R0, R1, R2};
    const bool iA[3] = {i1, i2, i0}, iB[3] = {i0, i1, i2};
#pragma unroll
    for (int b = 0; b < 3; b++) {
        if (iA[b]) {
            if (!iB[b]) append4(q, m, intersect(P, N, As[b], Bs[b]));
            append4(q, m, As[b]);
        } else if (iB[b]) {
            append4(q, m, intersect(P, N, As[b], Bs[b]));
        }
    }
    int rot = 0;
#pragma unroll
    for (int pl = 0; pl < 6; pl++) {
        if (pl <= p) continue;
        float4 P2, N2;
        plane_of(pl, zn, zf, P2, N2);
        int nin = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) nin += (k < m && inside(P2, N2, q[k])) ? 1 : 0;
        if (nin == 0) return 0;
        if (nin != m) return -1;
        rot++;
    }
    // `rot` more left rotations of an m-gon.
    rot %= m;
    CV t[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        // t[k] = q[(k + rot) % m]
        const int s3 = (k + rot) % 3, s4 = (k + rot) % 4;
        const int s = m == 3 ? s3 : s4;
        t[k] = sel(s == 0, q[0], sel(s == 1, q[1], sel(s == 2, q[2], q[3])));
    }
#pragma unroll
    for (int k = 0; k < 4; k++) q[k] = t[k];
    return m;
}

// ---------------------------------------------------------------------------
// Triangle set-up
// ---------------------------------------------------------------------------

// Per-lane counters of k_face packed in one word (one register for the whole
// kernel): TPF references in bits 0-15, binned triangles 16-23, front faces
// 24-27, clip overflows 28-31. A lane owns kFacesPerLane faces in pass 1 (at
// most 4 front faces) and, over at most 4 rounds of pass 2, at most 4 clip
// overflows and ~40 triangles of at most 16 references each: below these
// widths, and a wave's sums of the 16-bit fields stay below 2^16.
struct FaceStats {
    uint32_t packed;
};

// ---------------------------------------------------------------------------
// Debug overlays (renderer.go:191-216): lines and 3x3 rects write colour only,
// after their triangle's face, in draw order. Each overlay pixel keeps the max
// of (slot + 1) << 1 | vertex; the raster's shading pass shows the overlay
// where it comes at or after the depth winner (within a triangle the vertex
// rects are drawn last).
// ---------------------------------------------------------------------------

constexpr uint32_t kEdgeColor = 0xFF000000u;    // {0,0,0,255} (renderer.go:19)
constexpr uint32_t kVertexColor = 0xFF00A1FFu;  // {255,161,0,255} (renderer.go:18)
constexpr uint32_t kWhite = 0xFFFFFFFFu;        // edges without faces (renderer.go:194)

// utils.go:11-21: minInt maps to maxInt.
__device__ __forceinline__ int64_t go_abs(int64_t v) {
    if (v < 0) v = wsub(0, v);
    if (v < 0) v = wsub(v, 1);
    return v;
}

// FrameBuffer.Pixel (rasterizer.go:25-30): only the flat index is checked.
__device__ __forceinline__ void ov_pixel(const DevFrame& fr, int64_t x, int64_t y, uint32_t code) {
    const int64_t idx = wadd(wmul(y, (int64_t)fr.W), x);
    if (idx > 0 && idx < (int64_t)fr.W * fr.H) atomicMax(&fr.okey[idx], code);
}

// FrameBuffer.Rect (rasterizer.go:54-63) with width = height = 3.
__device__ void ov_rect(const DevFrame& fr, int64_t x, int64_t y, uint32_t code) {
    if (x >= fr.W || y >= fr.H) return;
    for (int64_t py = y; py < wadd(y, 3); py = wadd(py, 1))
        for (int64_t px = x; px < wadd(x, 3); px = wadd(px, 1)) ov_pixel(fr, px, py, code);
}

// FrameBuffer.Line (rasterizer.go:65-79): a float32 DDA accumulated step by step.
// Returns false, drawing nothing, for a line longer than kMaxOverlayLine.
__device__ bool ov_line(const DevFrame& fr, int64_t x0, int64_t y0, int64_t x1, int64_t y1,
                        uint32_t code) {
    const int64_t dx = wsub(x1, x0), dy = wsub(y1, y0);
    const int64_t ax = go_abs(dx), ay = go_abs(dy);
    const int64_t side = ax > ay ? ax : ay;
    if (side > kMaxOverlayLine) {
        atomicOr(&fr.ctrl->unsupported, 1u);
        return false;
    }
    const float xs = (float)dx / (float)side, ys = (float)dy / (float)side;
    float cx = (float)x0, cy = (float)y0;
    for (int64_t i = 0; i <= side; i++) {
        ov_pixel(fr, go_int(cx), go_int(cy), code);
        cx += xs;
        cy += ys;
    }
    return true;
}

// drawProjection's overlays for one binned triangle (renderer.go:191-216).
// Returns false if a line of the triangle was too long to draw.
__device__ bool draw_overlays(const DevFrame& fr, const OverlayRec& o, uint32_t slot) {
    const float* sx = o.sx;
    const float* sy = o.sy;
    int64_t ix[3], iy[3];
    for (int j = 0; j < 3; j++) {
        ix[j] = go_int(sx[j]);
        iy[j] = go_int(sy[j]);
    }
    const uint32_t code = (slot + 1u) << 1;
    bool ok = true;
    if (fr.flags & kShowEdges) {
        ok = ov_line(fr, ix[0], iy[0], ix[1], iy[1], code) & ok;
        ok = ov_line(fr, ix[1], iy[1], ix[2], iy[2], code) & ok;
        ok = ov_line(fr, ix[2], iy[2], ix[0], iy[0], code) & ok;
        if (fr.flags & kShowFaces) {
            const float mx = (sx[0] + sx[1] + sx[2]) / 3.0f, my = (sy[0] + sy[1] + sy[2]) / 3.0f;
            ov_rect(fr, wsub(go_int(mx), 1), wsub(go_int(my), 1), code);
        }
    }
    if (fr.flags & kShowVertices)
        for (int j = 0; j < 3; j++) ov_rect(fr, wsub(ix[j], 1), wsub(iy[j], 1), code | 1u);
    return ok;
}

// Tile references of a triangle's tile rectangle, counted for the list scan.
// Dense meshes put thousands of small triangles into the same tile; counted one
// global atomic each, they serialise on that counter. One-tile rectangles, and
// for dense meshes (at least one face per 4 pixels of the projected bounding
// box) rectangles of up to 2x2 tiles, are therefore counted in a per-workgroup
// LDS table of (tile, count) entries, flushed with one global atomic per entry
// when the workgroup ends (a tile that finds its entry held by another tile is
// counted in global memory). Other rectangles, whose tiles see few references
// each, add one global atomic per tile without the table's round trips.
constexpr int kTileTable = 256;
constexpr uint32_t kNoTile = 0xFFFFFFFFu;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
struct TileTally {
    lds_u32* tag;
    lds_u32* cnt;
    uint32_t* global;
    bool dense;
};

__device__ __forceinline__ void tally_tile(const TileTally& tally, uint32_t t) {
    const uint32_t h = t & (uint32_t)(kTileTable - 1);
    uint32_t old = kNoTile;  // the entry's previous tag when the exchange fails
    __hip_atomic_compare_exchange_strong(tally.tag + h, &old, t, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == kNoTile || old == t)
        __hip_atomic_fetch_add(tally.cnt + h, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else
        atomicAdd(&tally.global[t], 1u);
}

__device__ __forceinline__ void count_tiles(const DevFrame& fr, const TileTally& tally, int tx0,
                                            int tx1, int ty0, int ty1) {
    if (tx0 == tx1 && ty0 == ty1) {
        tally_tile(tally, (uint32_t)(fr.tile_begin + ty0 * fr.tiles_x + tx0));
        return;
    }
    const bool small = tally.dense && tx1 - tx0 <= 1 && ty1 - ty0 <= 1;
    for (int ty = ty0; ty <= ty1; ty++)
        for (int tx = tx0; tx <= tx1; tx++) {
            const uint32_t t = (uint32_t)(fr.tile_begin + ty * fr.tiles_x + tx);
            if (small)
                tally_tile(tally, t);
            else
                atomicAdd(&tally.global[t], 1u);
        }
}

// Writes the set-up of one post-clip triangle into record ri (renderer.go:361-394
// then the per-triangle part of rasterizer.go:98-137). Returns the TPF
// references in the low 16 bits and 1 << 16 when the triangle was binned.
__device__ uint32_t emit_triangle(const DevFrame& fr, const CV& a, const CV& b, const CV& c,
                                  int tex, uint32_t ri, const TileTally& tally, int twl,
                                  int thl) {
    uint32_t tpf = 0, tris = 0;
    const CV vv[3] = {a, b, c};
    float sx[3], sy[3], zw[3];
    // The screen x and y rows use z/w only through screen[2] and screen[6]: when
    // both are zero, z/w matters only by being infinite or NaN (0 * z/w), and the
    // sign of a zero term never changes what the rasteriser does with sx, sy
    // (int(), comparisons). w/w is 1 for finite nonzero w, else NaN.
    const bool zcol = fr.screen[2] != 0.f || fr.screen[6] != 0.f;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const float w = vv[j].w, z = vv[j].z;
        const float px = vv[j].x / w, py = vv[j].y / w;
        const bool wok = __builtin_isfinite(w) && w != 0.f;
        const float pw = wok ? 1.0f : __builtin_nanf("");
        float pz = 0.0f;
        if (zcol || !(__builtin_isfinite(z) && __builtin_fabsf(z) < __builtin_fabsf(w) * 0x1p100f))
            pz = z / w;
        sx[j] = fr.screen[0] * px + fr.screen[1] * py + fr.screen[2] * pz + fr.screen[3] * pw;
        sy[j] = fr.screen[4] * px + fr.screen[5] * py + fr.screen[6] * pz + fr.screen[7] * pw;
        zw[j] = w;
    }
    float fminx = go_min(go_min(sx[0], sx[1]), sx[2]);
    float fmaxx = go_max(go_max(sx[0], sx[1]), sx[2]);
    float fminy = go_min(go_min(sy[0], sy[1]), sy[2]);
    float fmaxy = go_max(go_max(sy[0], sy[1]), sy[2]);
    // renderer.go:226-244 against the reference TPF grid. Its tiles are a
    // product grid (x bounds from tile % n, y bounds from tile / n,
    // renderer.go:61-67), so the overlapping tiles are the overlapping columns
    // times the overlapping rows.
    {
        const int ng = fr.ref_tiles == 16 ? 4 : 1;
        uint32_t cols = 0, rows = 0;
        for (int c = 0; c < ng; c++) {
            const float* rx = fr.ref_tile_rect[c];
            const float* ry = fr.ref_tile_rect[c * ng];
            cols += (uint32_t)(fmaxx >= rx[0] && fminx <= rx[2]);
            rows += (uint32_t)(fmaxy >= ry[1] && fminy <= ry[3]);
        }
        tpf = cols * rows;
    }
    RasterRec rr;
    rr.bx0 = 1;
    rr.bx1 = 0;
    rr.by0 = 1;
    rr.by1 = 0;
    rr.wide = 0;
    rr.sf = 0.f;
    rr.z0 = rr.z1 = rr.z2 = 0.f;
    rr.rsf = rr.rz0 = rr.rz1 = rr.rz2 = 0.f;
#pragma unroll
    for (int e = 0; e < 9; e++) rr.e[e] = 0;
    float W = (float)fr.W, H = (float)fr.H;
    bool binned = fmaxx >= 0.f && fminx <= W && fmaxy >= 0.f && fminy <= H;
    if (fr.orec) {
        // Screen positions for the overlay pass (binned triangles only draw them).
        OverlayRec o;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            o.sx[j] = sx[j];
            o.sy[j] = sy[j];
        }
        o.binned = binned ? 1u : 0u;
        o.pad = 0;
        gstore(&fr.orec[ri], o);
    }
    if (binned) {
        tris = 1;
        int64_t ix[3], iy[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            ix[j] = go_int(sx[j]);
            iy[j] = go_int(sy[j]);
        }
        // rasterizer.go:99-104 with tile (0,0)-(W,H) of the sequential renderer.
        int64_t mnx = max(min(min(ix[0], ix[1]), ix[2]), (int64_t)0);
        int64_t mxx = min(max(max(ix[0], ix[1]), ix[2]), (int64_t)fr.W - 1);
        int64_t mny = max(min(min(iy[0], iy[1]), iy[2]), (int64_t)0);
        int64_t mxy = min(max(max(iy[0], iy[1]), iy[2]), (int64_t)fr.H - 1);
        if ((fr.flags & kShowFaces) && mnx <= mxx && mny <= mxy) {
            rr.bx0 = (int16_t)mnx;
            rr.bx1 = (int16_t)mxx;
            rr.by0 = (int16_t)mny;
            rr.by1 = (int16_t)mxy;
            rr.z0 = zw[0];
            rr.z1 = zw[1];
            rr.z2 = zw[2];
            bool fast = true;
#pragma unroll
            for (int j = 0; j < 3; j++)
                fast = fast && ix[j] >= -kFastCoord && ix[j] <= kFastCoord &&
                       iy[j] >= -kFastCoord && iy[j] <= kFastCoord;
            // Edge functions f_e(x,y) = A*x + B*y + C (rasterizer.go:107-117) with the
            // top-left adjustment folded into C (rasterizer.go:120-129). The pixel
            // sum fx12 + fx20 + fx01 is the sum of the C's (the A and B terms
            // cancel).
            const int i0[3] = {0, 1, 2}, i1[3] = {1, 2, 0};
            if (fast) {
                // |x|, |y| <= 2^14: 24-bit products, and every value (the sum of
                // the three C's included) fits int32.
                int32_t X[3], Y[3], csum = 0;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    X[j] = (int32_t)ix[j];
                    Y[j] = (int32_t)iy[j];
                }
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    const int p = i0[e], q = i1[e];
                    const int32_t A = Y[p] - Y[q], B = X[q] - X[p];
                    int32_t C = __mul24(X[p], Y[q]) - __mul24(X[q], Y[p]);
                    if (!(B > 0 || (B == 0 && A > 0))) C -= 1;
                    rr.e[3 * e + 0] = A;
                    rr.e[3 * e + 1] = B;
                    rr.e[3 * e + 2] = C;
                    csum += C;
                }
                rr.sf = (float)csum;
                // A box of at most 2x2 pixels none of which passes the edge
                // test (the raster's, rasterizer.go:147) yields no fragment: it
                // stays out of the tile lists. Dense meshes far from the camera
                // emit mostly such sub-pixel triangles.
                if (mxx - mnx <= 1 && mxy - mny <= 1) {
                    int32_t any = 0;
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int32_t x = (int32_t)min(mnx + (k & 1), mxx);
                        const int32_t y = (int32_t)min(mny + (k >> 1), mxy);
                        const int32_t f0 = (int32_t)edge24(rr.e[0], x, rr.e[1], y, rr.e[2]);
                        const int32_t f1 = (int32_t)edge24(rr.e[3], x, rr.e[4], y, rr.e[5]);
                        const int32_t f2 = (int32_t)edge24(rr.e[6], x, rr.e[7], y, rr.e[8]);
                        any |= f0 & f1 & f2;
                    }
                    if (any >= 0) {
                        rr.bx0 = 1;
                        rr.bx1 = 0;
                        rr.by0 = 1;
                        rr.by1 = 0;
                    }
                }
            } else {
                // Wrapping int64 arithmetic as the reference's int.
                int64_t csum = 0;
                uint32_t adj = 0;
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    const int p = i0[e], q = i1[e];
                    const int64_t A = wsub(iy[p], iy[q]);
                    const int64_t B = wsub(ix[q], ix[p]);
                    int64_t C = wsub(wmul(ix[p], iy[q]), wmul(ix[q], iy[p]));
                    if (!(B > 0 || (B == 0 && A > 0))) {
                        C = wsub(C, 1);
                        adj |= 1u << e;
                    }
                    csum = wadd(csum, C);
                }
                rr.sf = (float)csum;
                rr.wide = 1;
                WideRec wr;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    wr.x[j] = ix[j];
                    wr.y[j] = iy[j];
                }
                wr.adj = adj;
                wr.pad[0] = wr.pad[1] = wr.pad[2] = 0;
                gstore(&fr.wrec[ri], wr);
            }
            ShadeRec sr;
            // u, v as clipped; finalize_record divides them by z (rasterizer.go:132-137).
            sr.uz[0] = a.u;
            sr.vz[0] = a.v;
            sr.uz[1] = b.u;
            sr.vz[1] = b.v;
            sr.uz[2] = c.u;
            sr.vz[2] = c.v;
            sr.in[0] = a.i;
            sr.in[1] = b.i;
            sr.in[2] = c.i;
            sr.tex = tex;
            sr.sf = rr.sf;
            sr.rsf = rr.wide ? kShadeWide : 0u;  // finalize_record sets a reciprocal
            gstore(&fr.srec[ri], sr);
        }
    }
    gstore(&fr.rrec[ri], rr);
    if (rr.bx0 <= rr.bx1)
        count_tiles(fr, tally, (int)rr.bx0 >> twl, (int)rr.bx1 >> twl, (int)rr.by0 >> thl,
                    (int)rr.by1 >> thl);
    return tpf | (tris << 16);
}

// A slot that holds no triangle (record ri): an empty box, never binned or
// finalized.
__device__ __forceinline__ void empty_record(const DevFrame& fr, uint32_t ri) {
    RasterRec rr;
    rr.sf = rr.z0 = rr.z1 = rr.z2 = 0.f;
    rr.rsf = rr.rz0 = rr.rz1 = rr.rz2 = 0.f;
    rr.bx0 = 1;
    rr.bx1 = 0;
    rr.by0 = 1;
    rr.by1 = 0;
    rr.wide = 0;
#pragma unroll
    for (int e = 0; e < 9; e++) rr.e[e] = 0;
    gstore(&fr.rrec[ri], rr);
    if (fr.orec) {
        OverlayRec o = {};
        gstore(&fr.orec[ri], o);
    }
}

// Record of offset `off` of face block gb (of the batch) whose overflow
// records start at ovf.
__device__ __forceinline__ uint32_t block_rec(uint32_t gb, uint32_t ovf, uint32_t off) {
    return off < (uint32_t)kRecsPerBlock ? gb * (uint32_t)kRecsPerBlock + off
                                         : ovf + (off - (uint32_t)kRecsPerBlock);
}

// Record of slot `slot` of a frame (the block's overflow start is read only
// past its own range).
__device__ __forceinline__ uint32_t rec_index(const DevFrame& fr, uint32_t slot) {
    const uint32_t blk = slot / (uint32_t)kSlotsPerWg, off = slot % (uint32_t)kSlotsPerWg;
    return off < (uint32_t)kRecsPerBlock
               ? ((uint32_t)fr.face_block_begin + blk) * (uint32_t)kRecsPerBlock + off
               : gld(&fr.rbase[blk]) + (off - (uint32_t)kRecsPerBlock);
}

// General faces clipped per round by one wave; each fills kMaxFan lanes.
constexpr int kStage = 64 / kMaxFan;

__device__ __forceinline__ void add_emit(FaceStats& st, uint32_t e) {
    st.packed += (e & 0xFFFFu) + ((e >> 16) << 16);
}

// ---------------------------------------------------------------------------
// Workgroup tile binning helpers
// ---------------------------------------------------------------------------

constexpr int kLdsTiles = 1024;  // LDS tile counters per workgroup
// One wave per k_bin_fill workgroup: a face block emits a few dozen triangles
// in typical frames, and the workgroup's time is its chain of dependent memory
// round trips, so more resident workgroups per CU is what makes it faster.
constexpr int kFillThreads = 64;

struct TileRect {
    int tx0, tx1, ty0, ty1;
};

// Tile range of a raster record's box (empty: tx0 > tx1), from the record's
// piece 2 (bx0 | bx1 << 16, by0 | by1 << 16, wide, A01).
__device__ __forceinline__ TileRect box_tiles(const uint4& p2, int twl, int thl) {
    const int bx0 = (int)(int16_t)(p2.x & 0xFFFFu), bx1 = (int)(int16_t)(p2.x >> 16);
    const int by0 = (int)(int16_t)(p2.y & 0xFFFFu), by1 = (int)(int16_t)(p2.y >> 16);
    TileRect t;
    if (bx0 > bx1) {
        t.tx0 = 1;
        t.tx1 = 0;
        t.ty0 = 1;
        t.ty1 = 0;
    } else {
        t.tx0 = bx0 >> twl;
        t.tx1 = bx1 >> twl;
        t.ty0 = by0 >> thl;
        t.ty1 = by1 >> thl;
    }
    return t;
}

__device__ __forceinline__ TileRect rec_tiles(const RasterRec* r, int twl, int thl) {
    return box_tiles(gload(reinterpret_cast<const uint4*>(r) + 2), twl, thl);
}

// ---------------------------------------------------------------------------
// k_face
// ---------------------------------------------------------------------------

// Per-triangle divisions only the raster stage needs, done once per record
// after the face stage: u/z and v/z (rasterizer.go:132-137; the face stage
// stores u and v) and the reciprocals of the divisors for the exact quotient
// path, whose numerators are nonzero int32 edge values and barycentrics with
// |alpha|, |beta| >= 2^-41 and gamma = +0 or |gamma| >= 2^-64 (left 0, i.e.
// IEEE division, for int64 triangles and divisors outside rcp_range).
__device__ __forceinline__ TileRect finalize_record(RasterRec* r, ShadeRec* sr, int twl, int thl) {
    const float4 d = gload(reinterpret_cast<const float4*>(r));
    const uint4 p2 = gload(reinterpret_cast<const uint4*>(r) + 2);
    const TileRect t = box_tiles(p2, twl, thl);
    if (t.tx0 > t.tx1) return t;  // not rasterised; no shading inputs were written
    // Pieces 0 and 1 of the shading record: uz[0..2], vz[0] | vz[1], vz[2], in[0], in[1].
    static_assert(offsetof(ShadeRec, vz) == 12 && offsetof(ShadeRec, in) == 24,
                  "shading record pieces");
    float4 s0 = gload(reinterpret_cast<const float4*>(sr));
    float4 s1 = gload(reinterpret_cast<const float4*>(sr) + 1);
    s0.x = s0.x / d.y;  // uz[0]
    s0.w = s0.w / d.y;  // vz[0]
    s0.y = s0.y / d.z;  // uz[1]
    s1.x = s1.x / d.z;  // vz[1]
    s0.z = s0.z / d.w;  // uz[2]
    s1.y = s1.y / d.w;  // vz[2]
    gstore(reinterpret_cast<float4*>(sr), s0);
    gstore(reinterpret_cast<float4*>(sr) + 1, s1);
    if (!p2.z && rcp_range(d.x) && rcp_range(d.y) && rcp_range(d.z) && rcp_range(d.w)) {
        const float4 rc = make_float4(1.0f / d.x, 1.0f / d.y, 1.0f / d.z, 1.0f / d.w);
        gstore(reinterpret_cast<float4*>(r) + 1, rc);
        gst(&sr->rsf, __float_as_uint(rc.x));  // the shading pass reads it from here
    }
    return t;
}

// Exclusive block scan of small counts (256 threads = 4 waves).
// Exclusive scan over the workgroup of per-face fan sizes, which are 0, 1, 2
// or kMaxFan: the in-wave prefix is counted from three ballots (no cross-lane
// shuffles, whose lane addresses would need registers at the point where
// k_face holds the most), the wave totals go through LDS.
__device__ __forceinline__ int fan_exclusive_scan(int v, int* lds4, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t b1 = __ballot(v == 1), b2 = __ballot(v == 2), b7 = __ballot(v == kMaxFan);
    const uint64_t lt = lanemask_lt(lane);
    const int x = __popcll(b1 & lt) + 2 * __popcll(b2 & lt) + kMaxFan * __popcll(b7 & lt);
    if (lane == 0) lds4[wave] = __popcll(b1) + 2 * __popcll(b2) + kMaxFan * __popcll(b7);
    __syncthreads();
    int base = 0;
    total = 0;
    for (int w = 0; w < 4; w++) {
        const int t = lds4[w];
        if (w < wave) base += t;
        total += t;
    }
    return base + x;
}

__device__ __forceinline__ int block_exclusive_scan(int v, int* lds4, int& total) {
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds4[wave] = x;
    __syncthreads();
    int base = 0;
    total = 0;
    for (int w = 0; w < 4; w++) {
        int t = lds4[w];
        if (w < wave) base += t;
        total += t;
    }
    return base + x - v;
}

// Clip-space corner k of a face: MVP * (x, y, z, w) (renderer.go:303-304,318-320).
__device__ __forceinline__ float4 corner(const DevObject& o, int v) {
    const float* p = o.m.pos + 3 * (size_t)v;
    float w = o.m.pw ? gld(&o.m.pw[v]) : 1.0f;
    return xform(o.mvp, gld(p), gld(p + 1), gld(p + 2), w);
}

// Gouraud intensity of normal n: World * normal, 4-D normalise (renderer.go:308-310,328-333).
__device__ __forceinline__ float normal_intensity(const DevObject& o, const DevFrame& fr, int n) {
    const float* q = o.m.nrm + 3 * (size_t)n;
    float w = o.m.nw ? gld(&o.m.nw[n]) : 1.0f;
    return lambert(xform(o.world, gld(q), gld(q + 1), gld(q + 2), w), fr.light, fr.ambient,
                   fr.diffuse);
}

// LDS writes of one wave made visible to the wave's other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One face block (kFacesPerWg faces of one object) per single-wave workgroup,
// kFacesPerLane faces per lane. A block's few survivors keep one wave busy
// through the long per-survivor chain of pass 2 (dependent loads, lighting,
// clipping, set-up), so the workgroup is that one wave: no sibling waves wait
// at barriers holding registers, and every resident wave works.
constexpr int kFacesPerLane = kFacesPerWg / 64;

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_face(const DevObject* __restrict__ objs, int nobj,
                                              const DevFrame* __restrict__ frames,
                                              uint32_t* __restrict__ tile_counts,
                                              uint32_t* __restrict__ wg_count,
                                              unsigned long long* __restrict__ stats, int twl,
                                              int thl, const int* __restrict__ xcd_face,
                                              int nframes, uint32_t* __restrict__ rec_base,
                                              uint32_t* __restrict__ rec_pool, uint32_t npools,
                                              uint32_t pool_recs, uint32_t ovf_begin,
                                              DevControl* __restrict__ ctrl,
                                              uint32_t* __restrict__ frame_overflow) {
    static_assert(kFacesPerWg == 64 * kFacesPerLane, "whole faces per lane");
    // Survivors in face order: face offset | first clip plane << 8.
    __shared__ uint32_t lds_sface[kFacesPerWg];
    __shared__ CV wpoly[2][kMaxClip];
    __shared__ CV lds_stage[kStage][kMaxClip];  // clipped polygons of general faces
    __shared__ int lds_stage_k[kStage], lds_stage_tex[kStage];
    __shared__ int lds_stage_off[kStage];  // the face's first offset in the block
    __shared__ uint32_t lds_tile_tag[kTileTable], lds_tile_cnt[kTileTable];

    int bf, blocal;
    if (!xcd_unit(xcd_face, nframes, (int)blockIdx.x, bf, blocal)) return;
    const int lane = (int)threadIdx.x;
    for (int i = lane; i < kTileTable; i += 64) {
        lds_tile_tag[i] = kNoTile;
        lds_tile_cnt[i] = 0u;
    }
    // Search only this frame's objects (one step for one object per frame).
    const int b = frames[bf].face_block_begin + blocal;
    const int ob0 = frames[bf].obj_begin;
    const int oi =
        ob0 + last_le(frames[bf].obj_count, b, [&](int i) { return objs[ob0 + i].face_block; });
    const DevObject& o = objs[oi];
    const TileTally tally{(lds_u32*)lds_tile_tag, (lds_u32*)lds_tile_cnt, tile_counts,
                          o.dense != 0};
    const DevMesh& m = o.m;
    const DevFrame& fr = frames[o.frame];
    const int wg_in_obj = b - o.face_block;
    const uint32_t flags = fr.flags;
    const int vis = o.vis;  // BoxVisibility, per object on the host

    FaceStats st = {0u};

    // Pass 1, faces lane, lane + 64, ... of the block: clip-space corners,
    // backface test and plane classification. The survivors (front faces the
    // clipper does not reject outright) are compacted in face order into LDS,
    // so the lighting, clipping and set-up below run on dense lanes.
    int nsurv = 0;      // wave-uniform
    uint32_t nrec = 0;  // records the survivors can need: 1, or kMaxFan when clipped
    if (vis != kBoxOutside) {
        // All the lane's vertex indices, then all its corners, are loaded
        // before any is used: lanes past the mesh's last face read its last
        // face (in bounds) and are not counted.
        const int face0 = wg_in_obj * kFacesPerWg + lane;
        int iv[kFacesPerLane][3];
#pragma unroll
        for (int k = 0; k < kFacesPerLane; k++) {
            const int face = min(face0 + 64 * k, m.F - 1);
            iv[k][0] = gld(&m.vi[3 * face + 0]);
            iv[k][1] = gld(&m.vi[3 * face + 1]);
            iv[k][2] = gld(&m.vi[3 * face + 2]);
        }
        float4 cv[kFacesPerLane][3];
#pragma unroll
        for (int k = 0; k < kFacesPerLane; k++)
#pragma unroll
            for (int c = 0; c < 3; c++) cv[k][c] = corner(o, iv[k][c]);
#pragma unroll
        for (int k = 0; k < kFacesPerLane; k++) {
            const int fo = k * 64 + lane;  // face offset in the block
            const float4 v0 = cv[k][0], v1 = cv[k][1], v2 = cv[k][2];
            bool survivor = false;
            int p = 6;
            if (face0 + 64 * k < m.F &&
                (!(flags & kBackfaceCulling) || facing_camera(v0, v1, v2))) {
                st.packed += 1u << 24;
                uint32_t m0 = 0x3Fu, m1 = 0x3Fu, m2 = 0x3Fu;
                if ((flags & kFrustumClipping) && vis != kBoxInside) {
                    m0 = inside_mask(CV{v0.x, v0.y, v0.z, v0.w, 0.f, 0.f, 0.f}, fr.znear, fr.zfar);
                    m1 = inside_mask(CV{v1.x, v1.y, v1.z, v1.w, 0.f, 0.f, 0.f}, fr.znear, fr.zfar);
                    m2 = inside_mask(CV{v2.x, v2.y, v2.z, v2.w, 0.f, 0.f, 0.f}, fr.znear, fr.zfar);
                }
                const uint32_t all_in = m0 & m1 & m2;
                // First plane not passed by every vertex (6 if none). Earlier
                // planes leave the triangle untouched; when this one rejects
                // every vertex the pass yields an empty polygon.
                p = __builtin_ctz(~all_in | 0x40u);
                survivor = !(p < 6 && !(((m0 | m1 | m2) >> p) & 1u));
            }
            const uint64_t sm = __ballot(survivor);
            if (survivor)
                lds_sface[nsurv + __popcll(sm & lanemask_lt(lane))] = (uint32_t)fo | ((uint32_t)p << 8);
            nsurv += __popcll(sm);
            nrec += (uint32_t)__popcll(sm) + (kMaxFan - 1) * (uint32_t)__popcll(__ballot(survivor && p < 6));
        }
    }
    wave_sync();

    // Records beyond the block's own range (sized for the largest fan of every
    // clipped survivor): a contiguous range of its overflow pool, claimed now.
    // Past the pool's end the block draws nothing and the batch reports an
    // overflow with a capacity that gives every pool this much (the caller
    // grows the capacity and draws it again); its frame is marked so that its
    // raster draws the background only.
    uint32_t ovf = 0u;  // wave-uniform
    if (nsurv > 0 && nrec > (uint32_t)kRecsPerBlock) {
        const uint32_t extra = nrec - (uint32_t)kRecsPerBlock;
        const uint32_t pool = (uint32_t)b % npools;
        uint32_t r = 0u;
        if (lane == 0) r = atomicAdd(&rec_pool[pool * kPoolStride], extra);
        r = (uint32_t)__builtin_amdgcn_readfirstlane((int)r);
        const unsigned long long end = (unsigned long long)r + extra;
        if (end > pool_recs) {
            if (lane == 0) {
                atomicMax(&ctrl->refs_needed, end * npools * kOvfRefs);
                atomicOr(&ctrl->overflow, 1u);
                atomicOr(&frame_overflow[o.frame], 1u);
            }
            nsurv = 0;
        } else {
            ovf = ovf_begin + pool * pool_recs + r;
        }
    }
    if (lane == 0) rec_base[b] = ovf;

    // Pass 2, survivor r0 + lane, 64 a round: lighting, texture coordinates,
    // clipping. 0: nothing, 1: emit the triangle as is, 2: full clipper in the
    // wave, 3: one-plane polygon. q[0..nq-1] holds the polygon to emit (the
    // triangle itself for modes 1 and 2), so only one set of corners stays
    // live across the emits.
    uint32_t run = 0;  // slots of the earlier rounds (wave-uniform)
    for (int r0 = 0; r0 < nsurv; r0 += 64) {
        int mode = 0;
        int n = 0;
        CV q[4];
        int nq = 0;
        int tex = -1;
        if (r0 + lane < nsurv) {
            CV c0, c1, c2;
            const uint32_t e = lds_sface[r0 + lane];
            const int sf = wg_in_obj * kFacesPerWg + (int)(e & 0xFFu);
            const int sp = (int)(e >> 8);
            const int i0 = gld(&m.vi[3 * sf + 0]), i1 = gld(&m.vi[3 * sf + 1]),
                      i2 = gld(&m.vi[3 * sf + 2]);
            // The corners again (the same arithmetic as pass 1).
            const float4 w0 = corner(o, i0), w1 = corner(o, i1), w2 = corner(o, i2);
            c0 = CV{w0.x, w0.y, w0.z, w0.w, 0.f, 0.f, 0.f};
            c1 = CV{w1.x, w1.y, w1.z, w1.w, 0.f, 0.f, 0.f};
            c2 = CV{w2.x, w2.y, w2.z, w2.w, 0.f, 0.f, 0.f};
            float in0, in1, in2;
            if (flags & kLighting) {
                if (o.gouraud) {
                    int n0 = i0, n1 = i1, n2 = i2;
                    if (m.ni) {
                        n0 = gld(&m.ni[3 * sf + 0]);
                        n1 = gld(&m.ni[3 * sf + 1]);
                        n2 = gld(&m.ni[3 * sf + 2]);
                    }
                    in0 = normal_intensity(o, fr, n0);
                    in1 = normal_intensity(o, fr, n1);
                    in2 = normal_intensity(o, fr, n2);
                } else {
                    const float* f = m.fn + 4 * (size_t)sf;
                    float4 t = xform(o.world, gld(f), gld(f + 1), gld(f + 2), gld(f + 3));
                    in0 = in1 = in2 = lambert(t, fr.light, fr.ambient, fr.diffuse);
                }
            } else {
                in0 = in1 = in2 = fr.ambient;
            }
            c0.i = in0;
            c1.i = in1;
            c2.i = in2;
            if (m.uv_vert) {
                c0.u = gld(&m.uv_vert[2 * i0]);
                c0.v = gld(&m.uv_vert[2 * i0 + 1]);
                c1.u = gld(&m.uv_vert[2 * i1]);
                c1.v = gld(&m.uv_vert[2 * i1 + 1]);
                c2.u = gld(&m.uv_vert[2 * i2]);
                c2.v = gld(&m.uv_vert[2 * i2 + 1]);
            } else if (m.uv_face) {
                const float* uv = m.uv_face + 6 * (size_t)sf;
                c0.u = gld(uv);
                c0.v = gld(uv + 1);
                c1.u = gld(uv + 2);
                c1.v = gld(uv + 3);
                c2.u = gld(uv + 4);
                c2.v = gld(uv + 5);
            }
            if (flags & kShowTextures) tex = m.face_tex ? gld(&m.face_tex[sf]) : m.mesh_tex;
            if (sp == 6) {
                // Every pass only rotates the triangle; after 6 passes it is
                // unchanged (clipping.go:199-232); no clipping: as is.
                mode = 1;
                n = 1;
                nq = 3;
                q[0] = c0;
                q[1] = c1;
                q[2] = c2;
            } else {
                nq = clip_one_plane(sp, c0, c1, c2, fr.znear, fr.zfar, q);
                if (nq > 0) {
                    mode = 3;
                    n = nq - 2;
                } else if (nq < 0) {
                    nq = 0;
                    mode = 2;  // counted below, one face at a time
                    q[0] = c0;
                    q[1] = c1;
                    q[2] = c2;
                }
            }
        }

        // Faces for the full clipper reserve the largest fan (kMaxFan slots):
        // the slots their polygon leaves unused get empty records, which are
        // never binned, so slots stay in draw order and each face is clipped
        // once. The prefix of the fan sizes (0, 1, 2 or kMaxFan) in face order
        // is counted from three ballots.
        if (mode == 2) n = kMaxFan;
        const uint64_t b1 = __ballot(n == 1), b2 = __ballot(n == 2), b7 = __ballot(n == kMaxFan);
        const uint64_t lt = lanemask_lt(lane);
        // The lane's slots start at offset off of the block's.
        const uint32_t off = run + (uint32_t)(__popcll(b1 & lt) + 2 * __popcll(b2 & lt) +
                                              kMaxFan * __popcll(b7 & lt));
        run += (uint32_t)(__popcll(b1) + 2 * __popcll(b2) + kMaxFan * __popcll(b7));
        // Fan of q (Polygon.Triangulate, clipping.go:44-62): (P0, P1, P2),
        // then (P0, P2, P3) for a quad.
        if (mode == 1 || mode == 3)
            add_emit(st, emit_triangle(fr, q[0], q[1], q[2], tex, block_rec(b, ovf, off), tally,
                                       twl, thl));
        if (mode == 3 && nq == 4)
            add_emit(st, emit_triangle(fr, q[0], q[2], q[3], tex, block_rec(b, ovf, off + 1),
                                       tally, twl, thl));

        // General faces, kStage at a time: the wave clips each one into its
        // stage (clipping.go:167-236), then one lane per fan slot emits the
        // triangle (P0, P_t+1, P_t+2) of Polygon.Triangulate
        // (clipping.go:44-62) or an empty record.
        CV* buf0 = wpoly[0];
        CV* buf1 = wpoly[1];
        for (uint64_t gen = __ballot(mode == 2); gen;) {
            int nst = 0;
            for (; gen && nst < kStage; gen &= gen - 1, nst++) {
                const int L = __builtin_ctzll(gen);
                const CV a = shfl_cv(q[0], L), bb = shfl_cv(q[1], L), c = shfl_cv(q[2], L);
                const int texL = __shfl(tex, L, 64);
                const int sL = __shfl((int)off, L, 64);
                const int k = wave_clip(buf0, buf1, a, bb, c, fr.znear, fr.zfar, lane);
                if (k < 0 && lane == L) st.packed += 1u << 28;
                if (lane < k) lds_stage[nst][lane] = buf0[lane];
                if (lane == 0) {
                    lds_stage_k[nst] = k;
                    lds_stage_tex[nst] = texL;
                    lds_stage_off[nst] = sL;
                }
                wave_sync();
            }
            const int r = lane / kMaxFan, t = lane - r * kMaxFan;
            if (r < nst) {
                const int k = lds_stage_k[r];
                const uint32_t ri = block_rec(b, ovf, (uint32_t)lds_stage_off[r] + (uint32_t)t);
                const CV* P = lds_stage[r];
                if (t < k - 2)
                    add_emit(st, emit_triangle(fr, P[0], P[t + 1], P[t + 2], lds_stage_tex[r],
                                               ri, tally, twl, thl));
                else
                    empty_record(fr, ri);
            }
            wave_sync();
        }
    }
    if (lane == 0) wg_count[b] = run;

    // Counters, two wave sums of 16-bit fields (TPF references | binned
    // triangles; front faces | clip overflows), then one global atomic per
    // counter per block.
    {
        const uint32_t c0 = wave_sum_u32((st.packed & 0xFFFFu) | (((st.packed >> 16) & 0xFFu) << 16));
        const uint32_t c1 = wave_sum_u32(((st.packed >> 24) & 0xFu) | ((st.packed >> 28) << 16));
        if (lane < 4) {
            const uint32_t w = lane < 2 ? c0 : c1;
            const unsigned long long t = (w >> ((lane & 1) * 16)) & 0xFFFFu;
            const int idx[4] = {kStatTpf, kStatTriangles, kStatFront, kStatClipOverflow};
            if (t) atomicAdd(&stats[o.frame * kNumStats + idx[lane]], t);
        }
    }
    // Every count is in the table: flush it.
    wave_sync();
    for (int i = lane; i < kTileTable; i += 64) {
        const uint32_t v = lds_tile_cnt[i];
        if (v) atomicAdd(&tile_counts[lds_tile_tag[i]], v);
    }
}

// ---------------------------------------------------------------------------
// k_overlay: debug overlays of every binned triangle (frames that ask for them)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_overlay(const DevObject* __restrict__ objs, int nobj,
                                                 const DevFrame* __restrict__ frames,
                                                 const uint32_t* __restrict__ wg_count,
                                                 unsigned long long* __restrict__ stats) {
    const int b = blockIdx.x;
    const int oi = last_le(nobj, b, [&](int i) { return objs[i].face_block; });
    const DevObject& o = objs[oi];
    const DevFrame& fr = frames[o.frame];
    if (!fr.orec) return;
    const uint32_t cnt = wg_count[b];
    const uint32_t blk = (uint32_t)(o.seq_block + (b - o.face_block));  // in the frame
    const uint32_t slot0 = blk * kSlotsPerWg;
    const uint32_t ovf = cnt > (uint32_t)kRecsPerBlock ? gld(&fr.rbase[blk]) : 0u;
    for (uint32_t i = threadIdx.x; i < cnt; i += 256) {
        const OverlayRec rec = fr.orec[block_rec(b, ovf, i)];
        // A frame with a line too long is counted on its own, besides the
        // batch's control word, so a caller can tell which frames failed.
        if (rec.binned && !draw_overlays(fr, rec, slot0 + i))
            atomicAdd(&stats[o.frame * kNumStats + kStatLongLines], 1ull);
    }
}

// ---------------------------------------------------------------------------
// k_decorate: the end of Draw outside demo mode (renderer.go:476-480)
// ---------------------------------------------------------------------------

// FrameBuffer.Pixel (rasterizer.go:25-30) straight into the colour buffer:
// only the flat index is checked, so an x beyond a row wraps into the next.
__device__ __forceinline__ void fb_pixel(const DevFrame& fr, int64_t x, int64_t y, uint32_t c) {
    const int64_t idx = wadd(wmul(y, fr.W), x);
    if (idx > 0 && idx < (int64_t)fr.W * fr.H) ((uint32_t*)fr.rgba)[idx] = c;
}

// FrameBuffer.Line (rasterizer.go:65-79): the float32 DDA, one lane per line.
// The lines here stay within the frame's bounds (at most 16385 steps).
__device__ void fb_line(const DevFrame& fr, int64_t x0, int64_t y0, int64_t x1, int64_t y1,
                        uint32_t c) {
    const int64_t dx = wsub(x1, x0), dy = wsub(y1, y0);
    const int64_t ax = go_abs(dx), ay = go_abs(dy);
    const int64_t side = ax > ay ? ax : ay;
    const float xs = (float)dx / (float)side, ys = (float)dy / (float)side;
    float cx = (float)x0, cy = (float)y0;
    for (int64_t i = 0; i <= side; i++) {
        fb_pixel(fr, go_int(cx), go_int(cy), c);
        cx += xs;
        cy += ys;
    }
}

// One workgroup per frame: the reference tiles' borders (drawTilesBoundaries,
// renderer.go:422-434), four lanes per tile, in red; then the crosshair
// (CrossHair, rasterizer.go:209-217) in yellow. Colour only, after the frame.
__global__ __launch_bounds__(64) void k_decorate(const DevFrame* __restrict__ frames) {
    const DevFrame& fr = frames[blockIdx.x];
    if (!(fr.flags & (kShowTileBounds | kShowCrosshair))) return;  // uniform per block
    const int t = threadIdx.x;
    if ((fr.flags & kShowTileBounds) && t < 4 * fr.ref_tiles) {
        constexpr uint32_t kRed = 0xFF0000FFu;  // {255, 0, 0, 255}
        const float* r = fr.ref_tile_rect[t >> 2];
        const int64_t sx = go_int(r[0]), sy = go_int(r[1]), ex = go_int(r[2]), ey = go_int(r[3]);
        switch (t & 3) {
        case 0: fb_line(fr, sx, sy, ex, sy, kRed); break;
        case 1: fb_line(fr, ex, sy, ex, ey, kRed); break;
        case 2: fb_line(fr, ex, ey, sx, ey, kRed); break;
        default: fb_line(fr, sx, ey, sx, sy, kRed); break;
        }
    }
    __syncthreads();  // the crosshair is drawn over the borders
    if ((fr.flags & kShowCrosshair) && t < 4) {
        constexpr uint32_t kYellow = 0xFF00FFFFu;  // {255, 255, 0, 255}
        constexpr int64_t size = 5, offset = 3;
        const int64_t x = fr.W / 2, y = fr.H / 2;
        switch (t) {
        case 0: fb_line(fr, x - size, y, x - offset, y, kYellow); break;
        case 1: fb_line(fr, x + offset, y, x + size, y, kYellow); break;
        case 2: fb_line(fr, x, y - size, x, y - offset, kYellow); break;
        default: fb_line(fr, x, y + offset, x, y + size, kYellow); break;
        }
    }
}

// FrameBuffer.Fog (rasterizer.go:185-207) with blendRGBA's
// uint8(float32(a)*(1-f) + float32(b)*f) per channel.
__global__ __launch_bounds__(256) void k_fog(uint32_t* __restrict__ rgba,
                                             const float* __restrict__ depth, uint64_t n,
                                             float fog_start, float fog_end, uint32_t color) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * 256u) {
        const float d = depth[i];
        if (d >= fog_start) continue;
        if (d <= fog_end) {
            rgba[i] = color;
            continue;
        }
        const float f = 1 - ((fog_end - d) / (fog_end - fog_start));
        const uint32_t p = rgba[i];
        uint32_t out = 0;
        for (int k = 0; k < 4; k++) {
            const float a = (float)((p >> (8 * k)) & 0xFFu), b = (float)((color >> (8 * k)) & 0xFFu);
            const float u = a * (1 - f);
            const float v = b * f;
            out |= (go_uint8(u + v) & 0xFFu) << (8 * k);
        }
        rgba[i] = out;
    }
}

// ---------------------------------------------------------------------------
// k_bin_scan: tile counts -> list offsets, one workgroup per frame
// ---------------------------------------------------------------------------

// A kernel of its own: a last-block scan inside k_face would need a device-scope
// release in every face workgroup, which on a multi-XCD part writes back and
// invalidates L2 each time.
__global__ __launch_bounds__(1024) void k_bin_scan(const DevFrame* __restrict__ frames,
                                                   uint32_t* __restrict__ tile_counts,
                                                   uint32_t* __restrict__ tile_offsets,
                                                   uint32_t* __restrict__ tile_cursor,
                                                   uint32_t* __restrict__ frame_overflow,
                                                   unsigned long long* __restrict__ stats,
                                                   DevBatchCounters* __restrict__ counters,
                                                   DevControl* __restrict__ ctrl,
                                                   unsigned long long refs_capacity,
                                                   uint32_t* __restrict__ tile_slab,
                                                   RasterItem* __restrict__ items,
                                                   uint32_t max_items, uint32_t max_slabs,
                                                   uint32_t part_refs,
                                                   uint32_t* __restrict__ rec_pool) {
    __shared__ unsigned long long wsum[16];
    __shared__ unsigned long long base_s;
    const int f = blockIdx.x;
    const DevFrame& fr = frames[f];
    const int ntiles = fr.tiles_x * fr.tiles_y;
    uint32_t* cnt = tile_counts + fr.tile_begin;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    // k_face is done with the record pools: their total to the batch counters,
    // and the counters back to zero for the next draw.
    static_assert(kRecPools % 64 == 0 && kRecPools <= 1024, "whole waves read the pool counters");
    if (f == 0 && threadIdx.x < kRecPools) {
        uint32_t* pc = &rec_pool[threadIdx.x * kPoolStride];
        const uint32_t v = *pc;
        *pc = 0u;
        const unsigned long long t = wave_sum_u64(v);
        if (lane == 0) atomicAdd(&counters->recs_used, t);
    }

    unsigned long long local = 0;
    for (int i = threadIdx.x; i < ntiles; i += 1024) local += cnt[i];
    local = wave_sum_u64(local);
    if (lane == 0) wsum[wave] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int w = 0; w < 16; w++) tot += wsum[w];
        const unsigned long long base = atomicAdd(&counters->refs_used, tot);
        stats[f * kNumStats + kStatRefs] = tot;
        if (base + tot > refs_capacity) {
            frame_overflow[f] = 1u;
            atomicOr(&ctrl->overflow, 1u);
            atomicMax(&ctrl->refs_needed, base + tot);
        }
        base_s = base;
    }
    __syncthreads();
    const bool over = frame_overflow[f] != 0u;
    unsigned long long run = base_s;
    for (int c = 0; c < ntiles; c += 1024) {
        const int i = c + threadIdx.x;
        const uint32_t v = i < ntiles ? cnt[i] : 0u;
        unsigned long long x = v;
        for (int s = 1; s < 64; s <<= 1) {
            unsigned long long y = __shfl_up(x, s, 64);
            if (lane >= s) x += y;
        }
        __syncthreads();
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        unsigned long long wbase = 0, ctot = 0;
        for (int w = 0; w < 16; w++) {
            if (w < wave) wbase += wsum[w];
            ctot += wsum[w];
        }
        if (i < ntiles) {
            const unsigned long long offv = run + wbase + x - v;
            tile_offsets[fr.tile_begin + i] = (uint32_t)offv;
            tile_cursor[fr.tile_begin + i] = (uint32_t)offv;
            if (over) cnt[i] = 0u;  // overflowed frame: raster draws background only
            if (!over && v > part_refs) {
                // Split the list into parts of part_refs entries: part 0 stays
                // with the tile's workgroup, the others become work items. With
                // no slab left the tile stays whole (slower, same result).
                const uint32_t np = (v - 1u) / part_refs + 1u;
                uint32_t slab = atomicAdd(&counters->n_slabs, 1u);
                if (slab < max_slabs) {
                    atomicAdd(&stats[f * kNumStats + kStatSplit], 1ull);
                    // Items of all split tiles number at most refs / part_refs,
                    // and max_items is sized for the whole tile-list capacity.
                    const uint32_t ib = atomicAdd(&counters->n_items, np - 1u);
                    for (uint32_t k = 1; k < np; k++)
                        if (ib + k - 1u < max_items)
                            items[ib + k - 1u] = RasterItem{(uint32_t)(fr.tile_begin + i), k};
                } else {
                    slab = kNoSlab;
                }
                tile_slab[fr.tile_begin + i] = slab;
            }
        }
        run += ctot;
    }
}

// ---------------------------------------------------------------------------
// k_bin_fill
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kFillThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_bin_fill(const DevObject* __restrict__ objs, int nobj,
                                                  const DevFrame* __restrict__ frames,
                                                  const uint32_t* __restrict__ wg_count,
                                                  const uint32_t* __restrict__ frame_overflow,
                                                  uint32_t* __restrict__ tile_cursor,
                                                  uint32_t* __restrict__ refs, int twl,
                                                  int thl) {
    __shared__ uint32_t lcnt[kLdsTiles];
    __shared__ int lrect[4];
    const int b = blockIdx.x;
    const int oi = last_le(nobj, b, [&](int i) { return objs[i].face_block; });
    const DevObject& o = objs[oi];
    if (frame_overflow[o.frame]) return;
    const DevFrame& fr = frames[o.frame];
    const uint32_t cnt = wg_count[b];
    if (cnt == 0) return;
    // The block's slots start at slot0; its records are block_rec(b, ovf, i).
    const uint32_t blk = (uint32_t)(o.seq_block + (b - o.face_block));  // in the frame
    const uint32_t slot0 = blk * kSlotsPerWg;
    const uint32_t ovf = cnt > (uint32_t)kRecsPerBlock ? gld(&fr.rbase[blk]) : 0u;
    // Tile ranges of the block's first kFillKeep * kFillThreads records stay in
    // registers from the finalising pass to the placing pass (the records are
    // read once); later records, if any, are read again in each pass.
    constexpr int kFillKeep = 4;
    TileRect tr[kFillKeep];
    if (threadIdx.x == 0) {
        lrect[0] = 0x7FFFFFFF;
        lrect[1] = -1;
        lrect[2] = 0x7FFFFFFF;
        lrect[3] = -1;
    }
    int ax = 0x7FFFFFFF, bx = -1, ay = 0x7FFFFFFF, by = -1;
    auto fold = [&](const TileRect& t) {
        if (t.tx0 > t.tx1) return;
        ax = min(ax, t.tx0);
        bx = max(bx, t.tx1);
        ay = min(ay, t.ty0);
        by = max(by, t.ty1);
    };
#pragma unroll
    for (int k = 0; k < kFillKeep; k++) {
        const uint32_t i = threadIdx.x + (uint32_t)(k * kFillThreads);
        tr[k] = TileRect{1, 0, 1, 0};
        if (i < cnt) {
            const uint32_t ri = block_rec(b, ovf, i);
            tr[k] = finalize_record(&fr.rrec[ri], &fr.srec[ri], twl, thl);
            fold(tr[k]);
        }
    }
    for (uint32_t i = threadIdx.x + kFillKeep * kFillThreads; i < cnt; i += kFillThreads) {
        const uint32_t ri = block_rec(b, ovf, i);
        fold(finalize_record(&fr.rrec[ri], &fr.srec[ri], twl, thl));
    }
    // Every tile range of record i, from registers or read again.
    auto each_range = [&](auto&& fn) {
#pragma unroll
        for (int k = 0; k < kFillKeep; k++) {
            const uint32_t i = threadIdx.x + (uint32_t)(k * kFillThreads);
            if (i < cnt) fn(tr[k], i);
        }
        for (uint32_t i = threadIdx.x + kFillKeep * kFillThreads; i < cnt; i += kFillThreads)
            fn(rec_tiles(&fr.rrec[block_rec(b, ovf, i)], twl, thl), i);
    };
    // The union of the ranges, by LDS atomics.
    __syncthreads();
    if (bx >= 0) {
        atomicMin(&lrect[0], ax);
        atomicMax(&lrect[1], bx);
        atomicMin(&lrect[2], ay);
        atomicMax(&lrect[3], by);
    }
    __syncthreads();
    const int rx0 = lrect[0], rx1 = lrect[1], ry0 = lrect[2], ry1 = lrect[3];
    if (rx1 < 0) return;
    const int rw = rx1 - rx0 + 1, rh = ry1 - ry0 + 1;
    if (rw * rh > kLdsTiles) {
        // Very large rectangle: one global atomic per reference.
        each_range([&](const TileRect& t, uint32_t i) {
            for (int ty = t.ty0; ty <= t.ty1; ty++)
                for (int tx = t.tx0; tx <= t.tx1; tx++) {
                    const uint32_t pos =
                        atomicAdd(&tile_cursor[fr.tile_begin + ty * fr.tiles_x + tx], 1u);
                    refs[pos] = slot0 + i;
                }
        });
        return;
    }
    for (int i = threadIdx.x; i < rw * rh; i += kFillThreads) lcnt[i] = 0u;
    __syncthreads();
    each_range([&](const TileRect& t, uint32_t) {
        for (int ty = t.ty0; ty <= t.ty1; ty++)
            for (int tx = t.tx0; tx <= t.tx1; tx++)
                atomicAdd(&lcnt[(ty - ry0) * rw + (tx - rx0)], 1u);
    });
    __syncthreads();
    // One global reservation per touched tile; lcnt becomes the write cursor.
    for (int i = threadIdx.x; i < rw * rh; i += kFillThreads) {
        const uint32_t v = lcnt[i];
        if (v)
            lcnt[i] = atomicAdd(&tile_cursor[fr.tile_begin + (ry0 + i / rw) * fr.tiles_x + rx0 +
                                            i % rw],
                                v);
    }
    __syncthreads();
    each_range([&](const TileRect& t, uint32_t i) {
        for (int ty = t.ty0; ty <= t.ty1; ty++)
            for (int tx = t.tx0; tx <= t.tx1; tx++) {
                const uint32_t pos = atomicAdd(&lcnt[(ty - ry0) * rw + (tx - rx0)], 1u);
                refs[pos] = slot0 + i;
            }
    });
}

// ---------------------------------------------------------------------------
// k_raster
// ---------------------------------------------------------------------------

// Key ordering: reference pixels keep the LAST fragment whose depth equals the
// maximum depth of all fragments >= -1 (rasterizer.go:156 with the >= test
// and z cleared to -1, rasterizer.go:37). Max over (depth, draw sequence) with
// -0 folded into +0 (they compare equal in the reference) selects it.
__device__ __forceinline__ unsigned long long depth_key(float z, uint32_t seq) {
    uint32_t b = __float_as_uint(z);
    if (b == 0x80000000u) b = 0u;
    uint32_t ord = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ((unsigned long long)ord << 32) | seq;
}

constexpr unsigned long long kBackgroundKey = 0x407FFFFFull << 32;  // depth -1, seq 0

// Barycentrics and the reciprocal depth of one inside pixel (rasterizer.go:149-153)
// from its edge values f12, f20: d = (sf, z0, z1, z2), rc = their reciprocals
// or zeros (IEEE division path).
struct Bary {
    float alpha, beta, gamma, zrec;
};

// The reciprocal path of bary(), for a triangle whose reciprocals are set.
__device__ __forceinline__ Bary bary_rcp(float f12f, float f20f, const float4& d,
                                         const float4& rc) {
    Bary b;
    b.alpha = div_rcp(f12f, d.x, rc.x);
    b.beta = div_rcp(f20f, d.x, rc.x);
    b.gamma = 1.0f - b.alpha - b.beta;
    b.zrec = -(div_rcp(b.alpha, d.y, rc.y) + div_rcp(b.beta, d.z, rc.z) +
               div_rcp(b.gamma, d.w, rc.w));
    return b;
}

__device__ __forceinline__ Bary bary(float f12f, float f20f, const float4& d, const float4& rc) {
    Bary b;
    if (rc.x != 0.f) {
        b = bary_rcp(f12f, f20f, d, rc);
    } else {
        b.alpha = f12f / d.x;
        b.beta = f20f / d.x;
        b.gamma = 1.0f - b.alpha - b.beta;
        b.zrec = -(b.alpha / d.y + b.beta / d.z + b.gamma / d.w);
    }
    return b;
}

// Edge values f01, f12, f20 of pixel (x, y) of a triangle beyond the int32
// range, in wrapping int64 arithmetic (rasterizer.go:107-129).
__device__ __forceinline__ void wide_edges(const WideRec* wr, int x, int y, int64_t f[3]) {
    const int i0[3] = {0, 1, 2}, i1[3] = {1, 2, 0};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int p = i0[k], q = i1[k];
        int64_t A = wsub(wr->y[p], wr->y[q]);
        int64_t B = wsub(wr->x[q], wr->x[p]);
        int64_t C = wsub(wmul(wr->x[p], wr->y[q]), wmul(wr->x[q], wr->y[p]));
        if (wr->adj >> k & 1u) C = wsub(C, 1);
        f[k] = wadd(wadd(wmul(A, x), wmul(B, y)), C);
    }
}

// Low 32 bits of int(v) (CVTTSS2SQ).
__device__ __forceinline__ int32_t go_int_lo(float v) {
    if (__builtin_fabsf(v) < 2147483648.0f) return (int32_t)v;
    return (int32_t)go_int(v);
}

// int(v) % m, Go's truncated remainder, for 0 < m < 2^31.
__device__ __forceinline__ int32_t go_int_rem(float v, int32_t m) {
    if (__builtin_fabsf(v) < 2147483648.0f) return (int32_t)v % m;
    return (int32_t)(go_int(v) % (int64_t)m);
}

// texture.go:69-88. Textures hold at most 2^28 texels, so indexes fit int32.
__device__ __forceinline__ uint32_t sample(const DevTexture& t, float u, float v) {
    switch (t.type) {
    case 0:
        return t.color;
    case 2: {
        const int32_t x = go_int_lo((1 - u) * t.scale * t.wf) & (t.w - 1);
        const int32_t y = go_int_lo(v * t.scale * t.hf) & (t.h - 1);
        return gld(&t.texels[y * t.w + x]);
    }
    case 1: {
        const int32_t x = go_int_rem((1 - u) * t.scale * t.wf, t.w);
        const int32_t y = go_int_rem(v * t.scale * t.hf, t.h);
        int32_t idx = y * t.w + x;
        if (idx < 0) idx = 0;
        return gld(&t.texels[idx]);
    }
    default:
        return 0xFFFF00FFu;  // magenta {255,0,255,255}
    }
}

// rasterizer.go:81-88
__device__ __forceinline__ uint32_t color_intensity(uint32_t c, float k) {
    uint32_t r = go_uint8((float)(c & 0xFFu) * k);
    uint32_t g = go_uint8((float)((c >> 8) & 0xFFu) * k);
    uint32_t b = go_uint8((float)((c >> 16) & 0xFFu) * k);
    return r | (g << 8) | (b << 16) | (c & 0xFF000000u);
}

__device__ __forceinline__ int sgpr(int v) { return __builtin_amdgcn_readfirstlane(v); }


// The raster stage reads a staged RasterRec as five 16-byte pieces: divisors,
// reciprocals, then (box x, box y, wide, A01), (B01, C01, A12, B12),
// (C12, A20, B20, C20).
static_assert(offsetof(RasterRec, bx0) == 32 && offsetof(RasterRec, by0) == 36 &&
                  offsetof(RasterRec, wide) == 40 && offsetof(RasterRec, e) == 44 &&
                  sizeof(RasterRec) == 80,
              "raster records are staged as five uint4 pieces");


// Exact depth of `count` queued fragments of one wave (rasterizer.go:149-156),
// one per lane, and the (depth, sequence) max into the tile key.
template <int TW>
__device__ __forceinline__ void frag_depth(const uint32_t (*q)[128], uint32_t head,
                                           uint32_t count, const uint4* lrec,
                                           const uint32_t* lslot, unsigned long long* key,
                                           int lane) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if ((uint32_t)lane < count) {
        const uint32_t k = (head + (uint32_t)lane) & 127u;
        const uint32_t pk = q[0][k];
        const int32_t f12 = (int32_t)q[1][k], f20 = (int32_t)q[2][k];
        const uint32_t i = pk >> 12, px_idx = pk & 0xFFFu;
        const uint4 d = lrec[i * 5], rc = lrec[i * 5 + 1];
        const float4 df = make_float4(__uint_as_float(d.x), __uint_as_float(d.y),
                                      __uint_as_float(d.z), __uint_as_float(d.w));
        const float4 rf = make_float4(__uint_as_float(rc.x), __uint_as_float(rc.y),
                                      __uint_as_float(rc.z), __uint_as_float(rc.w));
        // Almost every triangle has reciprocals: when all of the wave's do, the
        // branch is taken once for the wave instead of per lane.
        const Bary b = __ballot(rf.x == 0.f) == 0ull ? bary_rcp((float)f12, (float)f20, df, rf)
                                                      : bary((float)f12, (float)f20, df, rf);
        if (b.zrec >= -1.0f) atomicMax(&key[px_idx], depth_key(b.zrec, lslot[i] + 1u));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of k_raster beyond the key array: 64 staged records (80 B) and slots,
// the per-wave fragment rings and the stats reduction.
constexpr int kRasterChunk = 64;
constexpr int kRasterLdsFixed = kRasterChunk * (80 + 4) + 4 * 3 * 128 * 4 + 4 * 2 * 8;

// Waves per SIMD the tile's LDS allows (one wave per SIMD per workgroup), at
// most 8: the register budget follows from it (512 / waves VGPRs).
constexpr int raster_waves(int npx) {
    return (160 * 1024) / (npx * 8 + kRasterLdsFixed) >= 8
               ? 8
               : (160 * 1024) / (npx * 8 + kRasterLdsFixed);
}

// Walk form of a record for one tile (piece 2 of the staged record): the box
// clamped to the tile as one word (x0 - X0 | y0 - Y0 << 6 | x1 - x0 << 12 |
// height << 18 | log2 lanes per row << 25 | wide << 28; height 0 when the box
// misses the tile) and, on the int32 path, the three edge values at the
// clamped box's corner (x0, y0), A*x0 + B*y0 + C (rasterizer.go:107-129 at the
// first pixel of the walk). Inputs are the record's pieces 2..4: box x, box y,
// wide, A01 | B01, C01, A12, B12 | C12, A20, B20, C20.
constexpr uint32_t kWalkWide = 1u << 28;
template <int TW, int TH>
__device__ __forceinline__ uint4 walk_form(const uint4& w2, const uint4& w3, const uint4& w4,
                                           int X0, int Y0) {
    const int x0 = max((int)(int16_t)(w2.x & 0xFFFFu), X0);
    const int x1 = min((int)(int16_t)(w2.x >> 16), X0 + TW - 1);
    const int y0 = max((int)(int16_t)(w2.y & 0xFFFFu), Y0);
    const int y1 = min((int)(int16_t)(w2.y >> 16), Y0 + TH - 1);
    if (x1 < x0 || y1 < y0) return make_uint4(0u, 0u, 0u, 0u);
    const int bw = x1 - x0 + 1;
    int l2 = bw <= 1 ? 0 : 32 - __builtin_clz((uint32_t)(bw - 1));
    if (l2 > 6) l2 = 6;
    uint32_t info = (uint32_t)(x0 - X0) | (uint32_t)(y0 - Y0) << 6 | (uint32_t)(x1 - x0) << 12 |
                    (uint32_t)(y1 - y0 + 1) << 18 | (uint32_t)l2 << 25;
    if (w2.z) return make_uint4(info | kWalkWide, 0u, 0u, 0u);
    // |A|, |B| <= 2^15 and |x|, |y| < 2^14 on the int32 path: exact 24-bit
    // products, and the sums fit int32.
    const int f0 = __mul24((int)w2.w, x0) + __mul24((int)w3.x, y0) + (int)w3.y;
    const int f1 = __mul24((int)w3.z, x0) + __mul24((int)w3.w, y0) + (int)w4.x;
    const int f2 = __mul24((int)w4.y, x0) + __mul24((int)w4.z, y0) + (int)w4.w;
    return make_uint4(info, (uint32_t)f0, (uint32_t)f1, (uint32_t)f2);
}

// Resolves n tile-list entries (slots) into the tile's depth keys: the
// visibility part of rasterizer.go:139-156 for every listed triangle. Returns
// this lane's share of the covered-fragment count (summed over the wave).
template <int TW, int TH>
__device__ __forceinline__ unsigned long long raster_list(
    const DevFrame& fr, int X0, int Y0, const uint32_t* __restrict__ list, uint32_t cnt,
    unsigned long long* key, uint4* lrec, uint32_t* lslot, uint32_t (*q)[128], int lane,
    int wave) {
    constexpr int CHUNK = kRasterChunk;  // triangles staged in LDS at a time
    static_assert(TW * TH <= 4096 && CHUNK <= (1 << 20),
                  "fragment codes hold the pixel index in 12 bits");
    static_assert(TW <= 64 && TH <= 64, "a clamped box fits one column block of lanes");
    unsigned long long covered_lane = 0;  // wide path, per lane
    uint32_t covered_wave = 0;            // int32 path, wave-uniform
    for (uint32_t base = 0; base < cnt; base += CHUNK) {
        const uint32_t n = min(cnt - base, (uint32_t)CHUNK);
        // Stage the chunk: slots (one coalesced read), then the raster records,
        // all loads of the chunk in flight together. Waves 1 and 2 copy the
        // divisors and reciprocals (pieces 0 and 1) of record t; wave 0 reads
        // the box and edge coefficients of record t and writes their walk
        // form in place of pieces 2..4 (once per record, a lane each, rather
        // than once per record in every wave's scalar unit).
        __syncthreads();
        if (threadIdx.x < n) lslot[threadIdx.x] = list[base + threadIdx.x];
        __syncthreads();
        {
            const uint32_t t = threadIdx.x & 63u;
            if (wave < 3 && t < n) {
                const uint4* src = reinterpret_cast<const uint4*>(&fr.rrec[rec_index(fr, lslot[t])]);
                if (wave > 0) {
                    lrec[t * 5 + wave - 1] = gload(src + wave - 1);
                } else {
                    const uint4 w2 = gload(src + 2), w3 = gload(src + 3), w4 = gload(src + 4);
                    lrec[t * 5 + 2] = walk_form<TW, TH>(w2, w3, w4, X0, Y0);
                    lrec[t * 5 + 3] = make_uint4(w2.w, w3.z, w4.y, w3.x);  // A01 A12 A20 B01
                    lrec[t * 5 + 4] = make_uint4(w3.w, w4.z, 0u, 0u);      // B12 B20
                }
            }
        }
        __syncthreads();
        // Per wave, one triangle at a time: lanes cover an lpr x (64/lpr) pixel
        // block of its clamped box, stepping the int32 edge values by additions
        // (rasterizer.go:139-183 walks them the same way); inside fragments go
        // to an LDS ring by ballot and their exact depth runs on full waves.
        uint32_t qhead = 0, qtail = 0;  // wave-uniform ring positions
        for (uint32_t i = wave; i < n; i += 4) {
            // The walk form of the record (see walk_form): the box word moves
            // to a scalar register, the coefficients stay in vector registers
            // (the same value in every lane).
            const uint4 w2 = lrec[i * 5 + 2], w3 = lrec[i * 5 + 3], w4 = lrec[i * 5 + 4];
            const uint32_t info = (uint32_t)sgpr((int)w2.x);
            const int bh = (int)((info >> 18) & 127u);
            if (bh == 0) continue;  // the box misses the tile
            const int ox = (int)(info & 63u), oy = (int)((info >> 6) & 63u);
            const int dxm = (int)((info >> 12) & 63u), l2 = (int)((info >> 25) & 7u);
            const int x0 = X0 + ox, x1 = x0 + dxm, y0 = Y0 + oy, y1 = y0 + bh - 1;
            const int lpr = 1 << l2, rows = 64 >> l2;
            const int lx = lane & (lpr - 1), ly = lane >> l2;
            if (info & kWalkWide) {
                // Coordinates beyond the int32 range: direct int64 evaluation.
                const uint32_t slot = (uint32_t)sgpr((int)lslot[i]);
                const WideRec* wr = &fr.wrec[rec_index(fr, slot)];
                const uint4 d4 = lrec[i * 5], r4 = lrec[i * 5 + 1];
                const float4 rp[2] = {
                    make_float4(__uint_as_float(d4.x), __uint_as_float(d4.y),
                                __uint_as_float(d4.z), __uint_as_float(d4.w)),
                    make_float4(__uint_as_float(r4.x), __uint_as_float(r4.y),
                                __uint_as_float(r4.z), __uint_as_float(r4.w))};
                for (int yb = y0; yb <= y1; yb += rows) {
                    const int y = yb + ly;
                    for (int xb = x0; xb <= x1; xb += lpr) {
                        const int x = xb + lx;
                        if (y > y1 || x > x1) continue;
                        int64_t f[3];
                        wide_edges(wr, x, y, f);
                        if (!(f[0] < 0 && f[1] < 0 && f[2] < 0)) continue;
                        covered_lane++;
                        const Bary b = bary((float)f[1], (float)f[2], rp[0], rp[1]);
                        if (b.zrec >= -1.0f)
                            atomicMax(&key[(y - Y0) * TW + (x - X0)],
                                      depth_key(b.zrec, slot + 1u));
                    }
                }
                continue;
            }
            // Edge values of the lane's pixel: the value at the clamped box's
            // corner (x0, y0) plus A*lx + B*ly; |A|, |B| <= 2^15 and lx, ly
            // < 64 on this path, so the 24-bit products are exact and every
            // edge value fits int32 (|x|, |y| < 2^14).
            const int A0 = (int)w3.x, A1 = (int)w3.y, A2 = (int)w3.z;
            const int B0 = (int)w3.w, B1 = (int)w4.x, B2 = (int)w4.y;
            int f0 = (int)edge24(A0, lx, B0, ly, w2.y);
            int f1 = (int)edge24(A1, lx, B1, ly, w2.z);
            int f2 = (int)edge24(A2, lx, B2, ly, w2.w);
            const int sy0 = (int)((uint32_t)B0 << (6 - l2));  // B * rows
            const int sy1 = (int)((uint32_t)B1 << (6 - l2));
            const int sy2 = (int)((uint32_t)B2 << (6 - l2));
            // Fragment code: list index << 12 | the pixel's index in the tile
            // (y * TW + x, below 4096 for every tile size), exact in every lane
            // that enqueues (its pixel lies inside the tile).
            const uint32_t prow = (i << 12) + (uint32_t)(oy * TW + ox) + (uint32_t)(ly * TW + lx);
            // One walk step: lanes with all < 0 hold an inside fragment (every
            // edge value < 0, rasterizer.go:147, and the pixel inside the box);
            // they go to the wave's ring by ballot, and every 64 queued
            // fragments get their exact depth.
            auto step = [&](int all, int g1, int g2, uint32_t pk) {
                const uint64_t mask = __ballot(all < 0);
                if (mask) {
                    if (all < 0) {
                        const uint32_t k =
                            __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(
                                                          (uint32_t)mask, qtail)) &
                            127u;
                        q[0][k] = pk;
                        q[1][k] = (uint32_t)g1;
                        q[2][k] = (uint32_t)g2;
                    }
                    const uint32_t m = (uint32_t)__popcll(mask);
                    qtail += m;
                    covered_wave += m;
                    if (qtail - qhead >= 64u) {
                        frag_depth<TW>(q, qhead, 64u, lrec, lslot, key, lane);
                        qhead += 64u;
                    }
                }
            };
            // The box is at most TW <= 64 pixels wide, so one column block of
            // lanes covers it: a flat loop over row blocks. The lane's column
            // is inside the box iff gx < 0; only the last block can hold rows
            // beyond it.
            const int gx = lx - dxm - 1;
            const int nfull = bh >> (6 - l2);  // blocks of `rows` whole rows
            uint32_t pk = prow;
            for (int sb = 0; sb < nfull; sb++) {
                step(f0 & f1 & f2 & gx, f1, f2, pk);
                f0 += sy0;
                f1 += sy1;
                f2 += sy2;
                pk += (uint32_t)rows * (uint32_t)TW;
            }
            const int rest = bh - (nfull << (6 - l2));
            if (rest > 0) {
                const int gy = ly - rest;  // negative iff the lane's row is inside
                step(f0 & f1 & f2 & gx & gy, f1, f2, pk);
            }
        }
        while (qtail != qhead) {
            const uint32_t cnt_q = min(qtail - qhead, 64u);
            frag_depth<TW>(q, qhead, cnt_q, lrec, lslot, key, lane);
            qhead += cnt_q;
        }
    }
    return covered_lane + (lane == 0 ? (unsigned long long)covered_wave : 0ull);
}

// Deferred shading of the winning fragment of each pixel of the tile, one
// write per pixel. Returns the wave's count of shaded pixels.
//
// Background colours are read as values and the dot test is evaluated
// without short-circuit branches: hipcc 7.2 (clang 22) turned the
// short-circuit form into a select of field addresses that dropped the
// y % step condition when the loop runs once per thread.
template <int TW, int TH>
__device__ __forceinline__ unsigned long long shade_tile(const DevFrame& fr, int X0, int Y0,
                                                         const unsigned long long* key,
                                                         const DevTexture* __restrict__ textures) {
    constexpr int NPX = TW * TH;
    const uint32_t dot_color = fr.dot_color, clear_color = fr.clear_color;
    const int step = fr.dot_step;
    const int step1 = step >= 1 ? step : 1;
    unsigned long long shaded = 0;  // wave-uniform
    for (int i = threadIdx.x; i < NPX; i += 256) {
        const int x = X0 + (i % TW), y = Y0 + (i / TW);
        const bool inframe = x < fr.W && y < fr.H;
        const unsigned long long k = key[i];
        const uint32_t seq = (uint32_t)k;
        shaded += (unsigned long long)__popcll(__ballot(inframe && seq != 0u));
        if (!inframe) continue;
        uint32_t color;
        float depth;
        if (seq == 0u) {
            // DotGrid (rasterizer.go:46-52): x, y in {step, 2*step, ...}
            const uint32_t dot = (uint32_t)(step >= 1) & (uint32_t)(x >= step) &
                                 (uint32_t)(y >= step) & (uint32_t)(x % step1 == 0) &
                                 (uint32_t)(y % step1 == 0);
            color = dot ? dot_color : clear_color;
            depth = -1.0f;
        } else {
            const uint32_t ri = rec_index(fr, seq - 1u);
            // The shading record and pieces 3 and 4 of the raster record (B01,
            // C01, A12, B12 | C12, A20, B20, C20), in one round trip.
            const uint4* rq = reinterpret_cast<const uint4*>(&fr.rrec[ri]);
            const ShadeRec s = gload(&fr.srec[ri]);
            const uint4 e3 = gload(rq + 3), e4 = gload(rq + 4);
            // The int32 edge values are formed before the wide test; a wide
            // record replaces them (its int32 values are never used, and the
            // sums wrap as unsigned).
            float f12f = (float)(int32_t)edge24((int)e3.z, x, (int)e3.w, y, e4.x);
            float f20f = (float)(int32_t)edge24((int)e4.y, x, (int)e4.z, y, e4.w);
            // The key holds zRec's bits (with -0 folded into +0): only a zero
            // zRec is recomputed; alpha and beta are as in bary().
            const uint32_t ord = (uint32_t)(k >> 32);
            const float zkey = __uint_as_float((ord & 0x80000000u) ? (ord & 0x7FFFFFFFu) : ~ord);
            // rsf is 0 or kShadeWide (IEEE division) unless the raster record
            // holds reciprocals.
            const bool common = zkey != 0.0f && s.rsf != 0u && s.rsf != kShadeWide;
            Bary b;
            if (__ballot(!common) == 0ull) {
                // Every pixel of the wave: int32 edge values, reciprocals, and
                // zRec from the key (one branch for the wave, not per lane).
                const float rsf = __uint_as_float(s.rsf);
                b.alpha = div_rcp(f12f, s.sf, rsf);
                b.beta = div_rcp(f20f, s.sf, rsf);
                b.gamma = 1.0f - b.alpha - b.beta;
                b.zrec = zkey;
            } else {
                if (s.rsf == kShadeWide) {
                    int64_t f[3];
                    wide_edges(&fr.wrec[ri], x, y, f);
                    f12f = (float)f[1];
                    f20f = (float)f[2];
                }
                if (zkey != 0.0f) {
                    const float rsf = __uint_as_float(s.rsf);
                    if (s.rsf != 0u && s.rsf != kShadeWide) {
                        b.alpha = div_rcp(f12f, s.sf, rsf);
                        b.beta = div_rcp(f20f, s.sf, rsf);
                    } else {
                        b.alpha = f12f / s.sf;
                        b.beta = f20f / s.sf;
                    }
                    b.gamma = 1.0f - b.alpha - b.beta;
                    b.zrec = zkey;
                } else {
                    const float4 d = gload(reinterpret_cast<const float4*>(rq)),
                                 rc = gload(reinterpret_cast<const float4*>(rq) + 1);
                    b = bary(f12f, f20f, d, rc);
                }
            }
            // rasterizer.go:158-170
            float u = (b.alpha * s.uz[0] + b.beta * s.uz[1] + b.gamma * s.uz[2]) / b.zrec;
            float v = (b.alpha * s.vz[0] + b.beta * s.vz[1] + b.gamma * s.vz[2]) / b.zrec;
            float inten = b.alpha * s.in[0] + b.beta * s.in[1] + b.gamma * s.in[2];
            uint32_t c;
            const int tex0 = sgpr(s.tex);
            if (__ballot(s.tex != tex0) == 0ull) {
                // One texture (or none) for the wave's pixels: its descriptor
                // is read once, with scalar loads.
                if (tex0 < 0) {
                    c = fr.face_color;
                } else {
                    static_assert(sizeof(DevTexture) == 40, "ten words");
                    typedef const __attribute__((address_space(4))) uint32_t* cword;
                    const cword src = (cword)(textures + tex0);
                    uint32_t w[10];
#pragma unroll
                    for (int k = 0; k < 10; k++) w[k] = src[k];
                    DevTexture t;
                    __builtin_memcpy(&t, w, sizeof(DevTexture));
                    c = sample(t, u, v);
                }
            } else {
                c = s.tex < 0 ? fr.face_color : sample(textures[s.tex], u, v);
            }
            color = color_intensity(c, inten);
            depth = b.zrec;
        }
        const size_t idx = (size_t)y * fr.W + x;
        if (fr.okey) {
            // Overlay keys are zero between draws: each pixel's is read once.
            const uint32_t ok = gld(&fr.okey[idx]);
            if (ok) {
                gst(&fr.okey[idx], 0u);
                if ((ok >> 1) >= seq)
                    color = (ok & 1u) ? kVertexColor : ((fr.flags & kShowFaces) ? kEdgeColor : kWhite);
            }
        }
        gst(&((uint32_t*)fr.rgba)[idx], color);
        gst(&fr.depth[idx], depth);
    }
    return shaded;
}

// A tile whose list holds more than part_refs entries, and that k_bin_scan gave
// a slab, is split into parts of part_refs entries. Parts 1.. are work items
// that k_raster_parts resolves first, any workgroup taking the next item from
// the batch queue: each resolves its entries into LDS keys and merges them into
// the tile's slab in HBM with a 64-bit max (the (depth, sequence) resolve does
// not depend on order). k_raster then resolves part 0 of every tile, folds in
// the slab of a split tile (and clears it) and shades.
template <int TW, int TH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(raster_waves(TW * TH))))
void k_raster_parts(const DevFrame* __restrict__ frames, int nframes,
                    const uint32_t* __restrict__ tile_counts,
                    const uint32_t* __restrict__ tile_offsets,
                    const uint32_t* __restrict__ tile_slab, const uint32_t* __restrict__ refs,
                    unsigned long long* __restrict__ stats,
                    DevBatchCounters* __restrict__ counters, const RasterItem* __restrict__ items,
                    uint32_t max_items, unsigned long long* __restrict__ slabs,
                    uint32_t part_refs) {
    constexpr int NPX = TW * TH;
    __shared__ unsigned long long key[NPX];
    __shared__ uint4 lrec[kRasterChunk * 5];
    __shared__ uint32_t lslot[kRasterChunk];
    __shared__ uint32_t qbuf[4][3][128];
    __shared__ unsigned long long red;  // covered fragments of the item
    __shared__ RasterItem litem;
    const int lane = threadIdx.x & 63, wave = sgpr(threadIdx.x >> 6);
    // Written by k_bin_scan (an earlier kernel).
    const uint32_t n_items = min(counters->n_items, max_items);
    if (n_items == 0u) return;
    // Every branch around the barriers of this loop is wave-uniform (scalar):
    // thread-level branches here (a thread-0 fetch, an exit on a value read
    // from LDS) let the compiler restructure the loop so that thread 0 leaves
    // it while the other threads go round the barriers again, which hangs.
    for (;;) {
        if (wave == 0) {
            // Lane 0 fetches the next item; the wave publishes it.
            uint32_t it = kNoSlab;
            if (lane == 0) {
                it = __hip_atomic_load(&counters->next_item, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                if (it < n_items)
                    it = __hip_atomic_fetch_add(&counters->next_item, 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            }
            it = (uint32_t)sgpr((int)it);
            const RasterItem w = it < n_items ? items[it] : RasterItem{kNoSlab, 0u};
            if (lane == 0) {
                litem = w;
                red = 0ull;
            }
        }
        __syncthreads();
        const uint32_t wtile = (uint32_t)sgpr((int)litem.tile);
        const uint32_t wpart = (uint32_t)sgpr((int)litem.part);
        if (wtile == kNoSlab) return;
        const int tile = (int)wtile;
        const int fi = last_le(nframes, tile, [&](int i) { return frames[i].tile_begin; });
        const DevFrame& fr = frames[fi];
        const int t = tile - fr.tile_begin;
        const int X0 = (t % fr.tiles_x) * TW, Y0 = (t / fr.tiles_x) * TH;
        const uint32_t cnt = tile_counts[tile], beg = wpart * part_refs;
        for (int i = threadIdx.x; i < NPX; i += 256) key[i] = kBackgroundKey;
        // raster_list's first barrier orders the initialisation (and the read of
        // litem) before what follows.
        unsigned long long covered =
            raster_list<TW, TH>(fr, X0, Y0, refs + tile_offsets[tile] + beg,
                                min(cnt - beg, part_refs), key, lrec, lslot, qbuf[wave], lane,
                                wave);
        __syncthreads();
        unsigned long long* slab = slabs + (size_t)tile_slab[tile] * NPX;
        for (int i = threadIdx.x; i < NPX; i += 256) {
            const unsigned long long k = key[i];
            if ((uint32_t)k != 0u) atomicMax(&slab[i], k);
        }
        const unsigned long long cov = wave_total(covered, lane);
        if (lane == 0 && cov) atomicAdd(&red, cov);
        __syncthreads();
        if (wave == 0 && lane == 0 && red) atomicAdd(&stats[fi * kNumStats + kStatCovered], red);
        // litem and red are read before the next item overwrites them.
        __syncthreads();
    }
}

// Screen tile g: part 0 of its list (all of it when the tile is not split),
// then the deferred shading. Every thread of the workgroup calls it.
template <int TW, int TH>
__device__ __forceinline__ void raster_tile(
    int g, const DevFrame* __restrict__ frames, int nframes, uint32_t* __restrict__ tile_counts,
    const uint32_t* __restrict__ tile_offsets, const uint32_t* __restrict__ tile_slab,
    const uint32_t* __restrict__ refs, const DevTexture* __restrict__ textures,
    unsigned long long* __restrict__ stats, unsigned long long* __restrict__ slabs,
    uint32_t part_refs, unsigned long long* key, uint4* lrec, uint32_t* lslot,
    uint32_t (*qbuf)[3][128], unsigned long long (*red)[2], int lane, int wave) {
    constexpr int NPX = TW * TH;
    const int fi = last_le(nframes, g, [&](int i) { return frames[i].tile_begin; });
    const DevFrame& fr = frames[fi];
    const int t = g - fr.tile_begin;
    const int X0 = (t % fr.tiles_x) * TW, Y0 = (t / fr.tiles_x) * TH;

    for (int i = threadIdx.x; i < NPX; i += 256) key[i] = kBackgroundKey;

    const uint32_t cnt = tile_counts[g], off = tile_offsets[g];
    const uint32_t slab_i = cnt > part_refs ? tile_slab[g] : kNoSlab;
    // raster_list's first barrier orders the initialisation before its use.
    const unsigned long long covered = raster_list<TW, TH>(
        fr, X0, Y0, refs + off, slab_i != kNoSlab ? part_refs : cnt, key, lrec, lslot, qbuf[wave],
        lane, wave);
    __syncthreads();
    if (slab_i != kNoSlab) {
        // Fold in the later parts (merged by k_raster_parts; 0 = no fragment)
        // and leave the slab cleared for the next draw.
        unsigned long long* slab = slabs + (size_t)slab_i * NPX;
        for (int i = threadIdx.x; i < NPX; i += 256) {
            const unsigned long long k = slab[i];
            if (k > key[i]) key[i] = k;
            slab[i] = 0ull;
        }
        __syncthreads();
    }
    // Tile counts are zero between draws: the face stage of the next draw
    // accumulates into them again.
    if (threadIdx.x == 0 && cnt) tile_counts[g] = 0u;

    const unsigned long long shaded = shade_tile<TW, TH>(fr, X0, Y0, key, textures);
    const unsigned long long cov = wave_total(covered, lane);
    if (lane == 0) {
        red[wave][0] = cov;
        red[wave][1] = shaded;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const unsigned long long s = red[0][threadIdx.x] + red[1][threadIdx.x] +
                                     red[2][threadIdx.x] + red[3][threadIdx.x];
        if (s) atomicAdd(&stats[fi * kNumStats + (threadIdx.x ? kStatShaded : kStatCovered)], s);
    }
}

// XCD-aware tile order. Workgroup h runs on XCD h mod 8; each XCD takes runs
// of kTileRun consecutive tiles (runs x, x + 8, x + 16, ... for XCD x), so
// neighbouring tiles, which share triangle records, read them through one L2.
constexpr int kTileRun = 128;
__device__ __forceinline__ int xcd_tile(int h) {
    const int x = h & 7, j = h >> 3;
    return ((j / kTileRun) * 8 + x) * kTileRun + j % kTileRun;
}

// One workgroup per screen tile (grid: total_tiles rounded up to whole runs
// of all XCDs; workgroups past the last tile return at once).
template <int TW, int TH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(raster_waves(TW * TH))))
void k_raster(const DevFrame* __restrict__ frames, int nframes,
              uint32_t* __restrict__ tile_counts, const uint32_t* __restrict__ tile_offsets,
              const uint32_t* __restrict__ tile_slab, const uint32_t* __restrict__ refs,
              const DevTexture* __restrict__ textures, unsigned long long* __restrict__ stats,
              unsigned long long* __restrict__ slabs, uint32_t part_refs, int total_tiles) {
    const int g = xcd_tile((int)blockIdx.x);
    if (g >= total_tiles) return;
    __shared__ unsigned long long key[TW * TH];
    __shared__ uint4 lrec[kRasterChunk * 5];  // the staged raster records
    __shared__ uint32_t lslot[kRasterChunk];
    __shared__ uint32_t qbuf[4][3][128];  // per-wave ring: pixel, f12, f20
    __shared__ unsigned long long red[4][2];  // per wave: covered fragments, shaded pixels
    const int lane = threadIdx.x & 63, wave = sgpr(threadIdx.x >> 6);
    raster_tile<TW, TH>(g, frames, nframes, tile_counts, tile_offsets, tile_slab, refs, textures,
                        stats, slabs, part_refs, key, lrec, lslot, qbuf, red, lane, wave);
}

// Persistent form: a fixed grid (every CU at the raster's occupancy) in which
// workgroup w draws tiles w, w + grid, w + 2 * grid, ...
template <int TW, int TH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(raster_waves(TW * TH))))
void k_raster_persistent(const DevFrame* __restrict__ frames, int nframes,
                         uint32_t* __restrict__ tile_counts,
                         const uint32_t* __restrict__ tile_offsets,
                         const uint32_t* __restrict__ tile_slab, const uint32_t* __restrict__ refs,
                         const DevTexture* __restrict__ textures,
                         unsigned long long* __restrict__ stats,
                         unsigned long long* __restrict__ slabs, uint32_t part_refs,
                         int total_tiles) {
    __shared__ unsigned long long key[TW * TH];
    __shared__ uint4 lrec[kRasterChunk * 5];
    __shared__ uint32_t lslot[kRasterChunk];
    __shared__ uint32_t qbuf[4][3][128];
    __shared__ unsigned long long red[4][2];
    const int lane = threadIdx.x & 63, wave = sgpr(threadIdx.x >> 6);
    for (int g = (int)blockIdx.x; g < total_tiles; g += (int)gridDim.x) {
        raster_tile<TW, TH>(g, frames, nframes, tile_counts, tile_offsets, tile_slab, refs,
                            textures, stats, slabs, part_refs, key, lrec, lslot, qbuf, red, lane,
                            wave);
        // red and the LDS tile state are rewritten by the next tile.
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Mesh set-up: face normals as NewMesh computes them (mesh.go:53-60)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_face_normals(const float* __restrict__ pos,
                                                      const int32_t* __restrict__ vi, int F,
                                                      float* __restrict__ fn) {
    int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= F) return;
    const float* a = pos + 3 * (size_t)vi[3 * f];
    const float* b = pos + 3 * (size_t)vi[3 * f + 1];
    const float* c = pos + 3 * (size_t)vi[3 * f + 2];
    float ax = b[0] - a[0], ay = b[1] - a[1], az = b[2] - a[2];
    float bx = c[0] - a[0], by = c[1] - a[1], bz = c[2] - a[2];
    float nx = ay * bz - az * by;
    float ny = az * bx - ax * bz;
    float nz = ax * by - ay * bx;
    float len = __builtin_sqrtf(nx * nx + ny * ny + nz * nz);
    fn[4 * (size_t)f + 0] = nx / len;
    fn[4 * (size_t)f + 1] = ny / len;
    fn[4 * (size_t)f + 2] = nz / len;
    fn[4 * (size_t)f + 3] = 1.0f;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

void launch_face_normals(const float* pos, const int32_t* vi, int F, float* fn, hipStream_t s) {
    if (F <= 0) return;
    hipLaunchKernelGGL(k_face_normals, dim3((F + 255) / 256), dim3(256), 0, s, pos, vi, F, fn);
}

void launch_face(const LaunchParams& p, hipStream_t s) {
    if (p.face_blocks == 0) return;
    hipLaunchKernelGGL(k_face, dim3(p.face_grid), dim3(64), 0, s, p.objs, p.nobj, p.frames,
                       p.tile_counts, p.wg_count, p.stats, p.tw_log2, p.th_log2, p.xcd_face,
                       p.nframes, p.rec_base, p.rec_pool, p.rec_pools, p.pool_recs,
                       p.ovf_begin, p.ctrl, p.frame_overflow);
}

void xcd_tables(const int* n, int nframes, int* table, int* grid) {
    int longest = 0;
    for (int x = 0; x < 8; x++) {
        int* P = table + x * (nframes + 1);
        P[0] = 0;
        for (int f = 0; f < nframes; f++)
            P[f + 1] = P[f] + xcd_count(x, n[f]);
        longest = std::max(longest, P[nframes]);
    }
    *grid = 8 * longest;
}

void launch_decorate(const LaunchParams& p, hipStream_t s) {
    if (p.nframes == 0) return;
    hipLaunchKernelGGL(k_decorate, dim3(p.nframes), dim3(64), 0, s, p.frames);
}

void launch_fog(void* rgba, const void* depth, uint64_t n, float fog_start, float fog_end,
                uint32_t color, hipStream_t s) {
    if (n == 0) return;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fog, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)rgba,
                       (const float*)depth, n, fog_start, fog_end, color);
}

void launch_overlay(const LaunchParams& p, hipStream_t s) {
    if (p.face_blocks == 0) return;
    hipLaunchKernelGGL(k_overlay, dim3(p.face_blocks), dim3(256), 0, s, p.objs, p.nobj, p.frames,
                       p.wg_count, p.stats);
}

void launch_bin_scan(const LaunchParams& p, hipStream_t s) {
    hipLaunchKernelGGL(k_bin_scan, dim3(p.nframes), dim3(1024), 0, s, p.frames, p.tile_counts,
                       p.tile_offsets, p.tile_cursor, p.frame_overflow, p.stats, p.counters,
                       p.ctrl, p.refs_capacity, p.tile_slab, p.items, p.max_items, p.max_slabs,
                       p.part_refs, p.rec_pool);
}

void launch_bin_fill(const LaunchParams& p, hipStream_t s) {
    if (p.face_blocks == 0) return;
    hipLaunchKernelGGL(k_bin_fill, dim3(p.face_blocks), dim3(kFillThreads), 0, s, p.objs, p.nobj,
                       p.frames, p.wg_count, p.frame_overflow, p.tile_cursor, p.refs,
                       p.tw_log2, p.th_log2);
}

// Self-check of div_rcp against IEEE division: pair i is drawn from SplitMix64(seed
// + i) in four families over the documented domain (|d| in [2^-40, 2^41),
// |a| in [2^-70, 2^70] or +0): random floats; quotients next to a rounding
// midpoint; divisors with extreme mantissas (all ones, powers of two); integer
// numerators over integer divisors as the raster stage divides them.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ float make_float(uint32_t sign, int exp, uint32_t mant) {
    return __uint_as_float((sign << 31) | ((uint32_t)(exp + 127) << 23) | (mant & 0x7FFFFFu));
}

__global__ __launch_bounds__(256) void k_check_div(uint64_t n, uint64_t seed,
                                                   unsigned long long* mismatches,
                                                   float* first_bad) {
    unsigned long long bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * 256) {
        const uint64_t r0 = splitmix64(seed + 2 * i), r1 = splitmix64(seed + 2 * i + 1);
        const int fam = (int)(i & 3);
        float a, d;
        const int ed = (int)(r0 % 81) - 40;
        if (fam == 0) {
            d = make_float((uint32_t)(r0 >> 63), ed, (uint32_t)(r0 >> 8));
            const int ea = (int)(r1 % 141) - 70;
            a = ((r1 >> 40) & 255) == 0 ? 0.0f
                                        : make_float((uint32_t)(r1 >> 63), ea, (uint32_t)(r1 >> 8));
        } else if (fam == 1) {
            // a = RN(d * m) with m = a float plus half an ulp: a/d lies next to m.
            d = make_float((uint32_t)(r0 >> 63), ed, (uint32_t)(r0 >> 8));
            const int em = (int)(r1 % 61) - 30;
            const double m = (double)make_float(0, em, (uint32_t)(r1 >> 8)) +
                             ldexp(1.0, em - 24) * ((r1 >> 62) & 1 ? 1.0 : -1.0);
            a = (float)((double)d * m);
            if (__builtin_fabsf(a) < 0x1p-70f || __builtin_fabsf(a) > 0x1p70f) a = 1.0f;
        } else if (fam == 2) {
            const uint32_t mant = (r0 >> 40) & 1 ? 0x7FFFFFu : ((r0 >> 41) & 1 ? 0u : 1u);
            d = make_float((uint32_t)(r0 >> 63), ed, mant);
            const int ea = (int)(r1 % 141) - 70;
            a = make_float((uint32_t)(r1 >> 63), ea, (uint32_t)(r1 >> 8));
        } else {
            d = (float)(int32_t)(((r0 >> 16) & 0x7FFFFFFFu) | 1u) * ((r0 >> 63) ? -1.0f : 1.0f);
            a = (float)(int32_t)(((r1 >> 16) & 0x7FFFFFFFu) | 1u) * ((r1 >> 63) ? -1.0f : 1.0f);
            if (((r1 >> 8) & 3) == 0) a = (float)(int32_t)((r1 >> 20) & 0xFFF) + 1.0f;
        }
        const float y = 1.0f / d;
        const float want = a / d;
        const float got = div_rcp(a, d, y);
        if (__float_as_uint(want) != __float_as_uint(got)) {
            if (atomicAdd(mismatches, 1ull) == 0ull) {
                first_bad[0] = a;
                first_bad[1] = d;
                first_bad[2] = want;
                first_bad[3] = got;
            }
            bad++;
        }
    }
    (void)bad;
}

void launch_check_div(uint64_t n, uint64_t seed, unsigned long long* mismatches, float* first_bad,
                      hipStream_t s) {
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 16384);
    if (blocks)
        hipLaunchKernelGGL(k_check_div, dim3((unsigned)blocks), dim3(256), 0, s, n, seed,
                           mismatches, first_bad);
}

// Plain device-to-device copy: 16-byte vector copies when both pointers and the
// size allow it, else bytes. Also the known-size kernel that HBM counters are
// calibrated against.
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src,
                                                uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy8(const uint8_t* __restrict__ src,
                                               uint8_t* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    const bool vec = ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0) && bytes % 16 == 0;
    const size_t n = vec ? bytes / 16 : bytes;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    if (vec)
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, s,
                           (const uint4*)src, (uint4*)dst, n);
    else
        hipLaunchKernelGGL(k_copy8, dim3((unsigned)blocks), dim3(256), 0, s,
                           (const uint8_t*)src, (uint8_t*)dst, n);
}

// Fills n 32-bit words with v: the zeroing of tile counters, slabs and control
// words (a kernel rather than hipMemsetAsync, so that the launch sequence of a
// batch can be captured in a graph).
__global__ __launch_bounds__(256) void k_fill32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = v;
}

void launch_fill32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
}

// Order-independent 64-bit digest of n words: the sum over i of a mix of
// (word i, i), accumulated into *out (so one digest covers several buffers).
__global__ __launch_bounds__(256) void k_checksum(const uint32_t* __restrict__ p, size_t n,
                                                  unsigned long long* __restrict__ out) {
    unsigned long long acc = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        unsigned long long x = ((unsigned long long)p[i] << 32) ^ (unsigned long long)i;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        acc += x ^ (x >> 31);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, acc);
}

void launch_checksum(const void* p, size_t words, unsigned long long* out, hipStream_t s) {
    if (words == 0) return;
    const size_t blocks = std::min<size_t>((words + 255) / 256, 2048);
    hipLaunchKernelGGL(k_checksum, dim3((unsigned)blocks), dim3(256), 0, s, (const uint32_t*)p,
                       words, out);
}

// Writes c.n words (at most kPutWords) carried by the kernel's arguments to
// dst: host data reaches device memory without a host buffer that must outlive
// the call, and without an allocation.
__global__ __launch_bounds__(256) void k_put(uint32_t* __restrict__ dst, PutChunk c) {
    for (uint32_t i = threadIdx.x; i < c.n; i += 256) dst[i] = c.w[i];
}

void launch_put(void* dst, const void* src, size_t bytes, hipStream_t s) {
    const size_t words = (bytes + 3) / 4;
    for (size_t w0 = 0; w0 < words; w0 += kPutWords) {
        PutChunk c;
        c.n = (uint32_t)std::min<size_t>(kPutWords, words - w0);
        c.pad = 0;
        const size_t b0 = w0 * 4, nb = std::min<size_t>((size_t)c.n * 4, bytes - b0);
        std::memset(c.w, 0, sizeof c.w);
        std::memcpy(c.w, (const char*)src + b0, nb);
        hipLaunchKernelGGL(k_put, dim3(1), dim3(256), 0, s, (uint32_t*)dst + w0, c);
    }
}

template <int TW, int TH>
static void raster_t(const LaunchParams& p, hipStream_t s, bool parts) {
    if (parts) {
        // Workgroups of the parts kernel: enough to fill every CU at the
        // raster's occupancy, but no more than the batch has tiles (a single
        // frame rarely splits many); with no work items they return at once,
        // and any number of them drains the queue.
        const uint32_t grid = std::min<uint32_t>(
            std::min<uint32_t>(p.max_items, 256u * (uint32_t)raster_waves(TW * TH)),
            std::max<uint32_t>(256u, (uint32_t)p.total_tiles));
        if (grid)
            hipLaunchKernelGGL((k_raster_parts<TW, TH>), dim3(grid), dim3(256), 0, s, p.frames,
                               p.nframes, p.tile_counts, p.tile_offsets, p.tile_slab, p.refs,
                               p.stats, p.counters, p.items, p.max_items, p.slabs, p.part_refs);
        return;
    }
    if (p.raster_persistent) {
        const uint32_t grid =
            std::min<uint32_t>((uint32_t)p.total_tiles, 256u * (uint32_t)raster_waves(TW * TH));
        hipLaunchKernelGGL((k_raster_persistent<TW, TH>), dim3(grid), dim3(256), 0, s, p.frames,
                           p.nframes, p.tile_counts, p.tile_offsets, p.tile_slab, p.refs,
                           p.textures, p.stats, p.slabs, p.part_refs, p.total_tiles);
        return;
    }
    constexpr int kRound = 8 * kTileRun;
    const uint32_t grid = (uint32_t)((p.total_tiles + kRound - 1) / kRound * kRound);
    hipLaunchKernelGGL((k_raster<TW, TH>), dim3(grid), dim3(256), 0, s, p.frames, p.nframes,
                       p.tile_counts, p.tile_offsets, p.tile_slab, p.refs, p.textures, p.stats,
                       p.slabs, p.part_refs, p.total_tiles);
}

static bool raster_dispatch(const LaunchParams& p, hipStream_t s, bool parts) {
    if (p.total_tiles == 0) return true;
#define GR_RASTER_CASE(W, H)                          \
    if (p.tw_log2 == W && p.th_log2 == H) {           \
        raster_t<(1 << W), (1 << H)>(p, s, parts);    \
        return true;                                  \
    }
    GR_RASTER_CASE(3, 3) GR_RASTER_CASE(3, 4) GR_RASTER_CASE(3, 5) GR_RASTER_CASE(3, 6)
    GR_RASTER_CASE(4, 3) GR_RASTER_CASE(4, 4) GR_RASTER_CASE(4, 5) GR_RASTER_CASE(4, 6)
    GR_RASTER_CASE(5, 3) GR_RASTER_CASE(5, 4) GR_RASTER_CASE(5, 5) GR_RASTER_CASE(5, 6)
    GR_RASTER_CASE(6, 3) GR_RASTER_CASE(6, 4) GR_RASTER_CASE(6, 5) GR_RASTER_CASE(6, 6)
#undef GR_RASTER_CASE
    return false;
}

bool launch_raster_parts(const LaunchParams& p, hipStream_t s) { return raster_dispatch(p, s, true); }
bool launch_raster(const LaunchParams& p, hipStream_t s) { return raster_dispatch(p, s, false); }

}  // namespace gr
