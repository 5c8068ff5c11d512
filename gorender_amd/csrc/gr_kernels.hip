// gfx950 kernels of the render path.
//
//   k_vertex    MVP transform of positions (SoA, coalesced), World transform +
//               4-D normalise + Lambert term of vertex normals, box visibility.
//   k_face      per face: backface test, lighting, frustum clipping, perspective
//               divide + screen transform, full-screen bin test, raster set-up,
//               tile counting. One workgroup = 256 faces = 1792 ordered slots.
//   k_bin_scan  per frame: exclusive scan of tile counts -> tile-list offsets.
//   k_bin_fill  per face workgroup: scatter slots into tile lists.
//   k_raster    one workgroup per screen tile: visibility in LDS with a 64-bit
//               (depth, draw-sequence) max key, then one deferred shading pass
//               and a single coalesced write of colour + depth.
//
// Every float32 expression keeps the reference's evaluation order; the build
// uses -ffp-contract=off and correctly rounded division/sqrt.
#pragma clang fp contract(off)

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
__device__ __forceinline__ uint32_t go_uint8(float x) {
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return 0u;
    return (uint32_t)(int32_t)x & 0xFFu;
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

// vector.go:115-125 (Vec4.Length + Divide), then renderer.go:331/336 Lambert term.
__device__ __forceinline__ float lambert(float4 n, const float* light, float ambient,
                                         float diffuse) {
    float len = __builtin_sqrtf(n.x * n.x + n.y * n.y + n.z * n.z + n.w * n.w);
    float x = n.x / len, y = n.y / len, z = n.z / len;
    float d = x * light[0] + y * light[1] + z * light[2];
    return ambient + d * diffuse;
}

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

// One Sutherland-Hodgman pass (clipping.go:199-232) from `in` (n vertices) into
// `out`; returns the new count, or -1 when a 10th vertex would be added.
__device__ __noinline__ int clip_pass(int p, float zn, float zf, const CV* in, int n, CV* out) {
    float4 P, N;
    plane_of(p, zn, zf, P, N);
    int m = 0;
    for (int b = 0; b < n; b++) {
        int a = (b + 1 == n) ? 0 : b + 1;
        CV A = in[a], B = in[b];
        bool ia = inside(P, N, A);
        if (ia) {
            if (!inside(P, N, B)) {
                if (m >= kMaxClip) return -1;
                out[m++] = intersect(P, N, A, B);
            }
            if (m >= kMaxClip) return -1;
            out[m++] = A;
        } else if (inside(P, N, B)) {
            if (m >= kMaxClip) return -1;
            out[m++] = intersect(P, N, A, B);
        }
    }
    return m;
}

// ClipTriangle (clipping.go:167-236) for a triangle already in poly[0..2];
// the final polygon is left in poly. Returns its vertex count, 0 when clipped
// away, -1 on overflow.
__device__ __noinline__ int clip_general(CV* poly, float zn, float zf) {
    CV tmp[kMaxClip];
    int n = 3;
    for (int p = 0; p < 6; p += 2) {
        n = clip_pass(p, zn, zf, poly, n, tmp);
        if (n <= 0) return n;
        n = clip_pass(p + 1, zn, zf, tmp, n, poly);
        if (n <= 0) return n;
    }
    return n;
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

// ---------------------------------------------------------------------------
// k_vertex
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_vertex(const DevObject* __restrict__ objs, int nobj,
                                                const DevFrame* __restrict__ frames,
                                                int vertex_blocks, uint32_t* __restrict__ zero_u32,
                                                int zero_u32_n,
                                                unsigned long long* __restrict__ zero_u64,
                                                int zero_u64_n) {
    int b = blockIdx.x;
    if (b >= vertex_blocks) {
        // Clear the batch counters (tile counts, per-frame flags, stats).
        int zb = b - vertex_blocks;
        int nz = gridDim.x - vertex_blocks;
        for (int i = zb * 256 + threadIdx.x; i < zero_u32_n; i += nz * 256) zero_u32[i] = 0u;
        for (int i = zb * 256 + threadIdx.x; i < zero_u64_n; i += nz * 256) zero_u64[i] = 0ull;
        return;
    }
    int oi = last_le(nobj, b, [&](int i) { return objs[i].vtx_block; });
    const DevObject& o = objs[oi];
    const DevMesh& m = o.m;
    if (b < o.nrm_block) {
        int v = (b - o.vtx_block) * 256 + threadIdx.x;
        if (v < m.V) {
            float w = m.pw ? m.pw[v] : 1.0f;
            o.clip[v] = xform(o.mvp, m.px[v], m.py[v], m.pz[v], w);
        }
        if (b == o.vtx_block && threadIdx.x == 0) {
            // renderer.go:268-275 + clipping.go:131-154 on the transformed box.
            const DevFrame& fr = frames[o.frame];
            int vis = kBoxInside;
            for (int p = 0; p < 6 && vis == kBoxInside; p++) {
                float4 P, N;
                plane_of(p, fr.znear, fr.zfar, P, N);
                float np = N.x * P.x + N.y * P.y + N.z * P.z + N.w * P.w;
                int outside = 0;
                for (int k = 0; k < 8; k++) {
                    float4 c = xform(o.mvp, m.bbox[k][0], m.bbox[k][1], m.bbox[k][2],
                                     m.bbox[k][3]);
                    float d = N.x * c.x + N.y * c.y + N.z * c.z + N.w * c.w;
                    if (d - np > 0.f) outside++;
                }
                if (outside == 8) vis = kBoxOutside;
                else if (outside > 0) vis = kBoxIntersect;
            }
            *o.vis = vis;
        }
    } else {
        int n = (b - o.nrm_block) * 256 + threadIdx.x;
        if (n < m.N) {
            const DevFrame& fr = frames[o.frame];
            float w = m.nw ? m.nw[n] : 1.0f;
            float4 t = xform(o.world, m.nx[n], m.ny[n], m.nz[n], w);
            o.nint[n] = lambert(t, fr.light, fr.ambient, fr.diffuse);
        }
    }
}

// ---------------------------------------------------------------------------
// k_face
// ---------------------------------------------------------------------------

// Exclusive block scan of small counts (256 threads = 4 waves).
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

struct FaceStats {
    unsigned long long tpf, tris, front, overflow;
};

// Writes the set-up of one post-clip triangle into its slot (renderer.go:361-394
// then the per-triangle part of rasterizer.go:98-137).
__device__ void emit_triangle(const DevFrame& fr, const CV& a, const CV& b, const CV& c,
                              int tex, uint32_t slot, int tw_log2, int th_log2,
                              uint32_t* __restrict__ tile_counts, FaceStats& st) {
    const CV vv[3] = {a, b, c};
    float sx[3], sy[3], zw[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        float w = vv[j].w;
        float px = vv[j].x / w, py = vv[j].y / w, pz = vv[j].z / w, pw = w / w;
        sx[j] = fr.screen[0] * px + fr.screen[1] * py + fr.screen[2] * pz + fr.screen[3] * pw;
        sy[j] = fr.screen[4] * px + fr.screen[5] * py + fr.screen[6] * pz + fr.screen[7] * pw;
        zw[j] = w;
    }
    float fminx = go_min(go_min(sx[0], sx[1]), sx[2]);
    float fmaxx = go_max(go_max(sx[0], sx[1]), sx[2]);
    float fminy = go_min(go_min(sy[0], sy[1]), sy[2]);
    float fmaxy = go_max(go_max(sy[0], sy[1]), sy[2]);
    // renderer.go:226-244 against the reference TPF grid.
    for (int t = 0; t < fr.ref_tiles; t++) {
        const float* r = fr.ref_tile_rect[t];
        if (fmaxx >= r[0] && fminx <= r[2] && fmaxy >= r[1] && fminy <= r[3]) st.tpf++;
    }
    RasterRec rr;
    rr.bx0 = 1;
    rr.bx1 = 0;
    rr.by0 = 1;
    rr.by1 = 0;
    rr.wide = 0;
    rr.sf = 0.f;
    rr.z0 = rr.z1 = rr.z2 = 0.f;
#pragma unroll
    for (int e = 0; e < 9; e++) rr.e[e] = 0;
    float W = (float)fr.W, H = (float)fr.H;
    bool binned = fmaxx >= 0.f && fminx <= W && fmaxy >= 0.f && fminy <= H;
    if (binned) {
        st.tris++;
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
            // top-left adjustment folded into C (rasterizer.go:120-129).
            const int i0[3] = {0, 1, 2}, i1[3] = {1, 2, 0};
            int64_t csum = 0;
            uint32_t adj = 0;
            int64_t A64[3], B64[3], C64[3];
#pragma unroll
            for (int e = 0; e < 3; e++) {
                int p = i0[e], q = i1[e];
                int64_t A = wsub(iy[p], iy[q]);
                int64_t B = wsub(ix[q], ix[p]);
                int64_t C = wsub(wmul(ix[p], iy[q]), wmul(ix[q], iy[p]));
                bool keep = B > 0 || (B == 0 && A > 0);
                if (!keep) {
                    C = wsub(C, 1);
                    adj |= 1u << e;
                }
                A64[e] = A;
                B64[e] = B;
                C64[e] = C;
                csum = wadd(csum, C);
            }
            // Order of the reference sum (fx12 + fx20) + fx01 is irrelevant for
            // wrapping integers.
            rr.sf = (float)csum;
            if (fast) {
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    rr.e[3 * e + 0] = (int32_t)A64[e];
                    rr.e[3 * e + 1] = (int32_t)B64[e];
                    rr.e[3 * e + 2] = (int32_t)C64[e];
                }
            } else {
                rr.wide = 1;
#pragma unroll
                for (int e = 0; e < 9; e++) rr.e[e] = 0;
                WideRec wr;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    wr.x[j] = ix[j];
                    wr.y[j] = iy[j];
                }
                wr.adj = adj;
                wr.pad[0] = wr.pad[1] = wr.pad[2] = 0;
                fr.wrec[slot] = wr;
            }
            ShadeRec sr;
            // rasterizer.go:132-137
            sr.uz[0] = a.u / zw[0];
            sr.vz[0] = a.v / zw[0];
            sr.uz[1] = b.u / zw[1];
            sr.vz[1] = b.v / zw[1];
            sr.uz[2] = c.u / zw[2];
            sr.vz[2] = c.v / zw[2];
            sr.in[0] = a.i;
            sr.in[1] = b.i;
            sr.in[2] = c.i;
            sr.tex = tex;
            sr.pad[0] = sr.pad[1] = 0;
            fr.srec[slot] = sr;
            int tx0 = (int)mnx >> tw_log2, tx1 = (int)mxx >> tw_log2;
            int ty0 = (int)mny >> th_log2, ty1 = (int)mxy >> th_log2;
            for (int ty = ty0; ty <= ty1; ty++)
                for (int tx = tx0; tx <= tx1; tx++)
                    atomicAdd(&tile_counts[fr.tile_begin + ty * fr.tiles_x + tx], 1u);
        }
    }
    fr.rrec[slot] = rr;
}

__global__ __launch_bounds__(256) void k_face(const DevObject* __restrict__ objs, int nobj,
                                              const DevFrame* __restrict__ frames,
                                              uint32_t* __restrict__ tile_counts,
                                              uint32_t* __restrict__ wg_count,
                                              unsigned long long* __restrict__ stats,
                                              int tw_log2, int th_log2) {
    __shared__ int lds_scan[4];
    __shared__ unsigned long long lds_st[4][4];
    int b = blockIdx.x;
    int oi = last_le(nobj, b, [&](int i) { return objs[i].face_block; });
    const DevObject& o = objs[oi];
    const DevMesh& m = o.m;
    const DevFrame& fr = frames[o.frame];
    int wg_in_obj = b - o.face_block;
    int face = wg_in_obj * kFacesPerWg + threadIdx.x;
    uint32_t flags = fr.flags;
    int vis = *o.vis;

    FaceStats st = {0, 0, 0, 0};
    int mode = 0;  // 0: nothing, 1: triangle as is, 2: clipped polygon
    int n = 0;
    CV poly[kMaxClip];
    int tex = -1;

    if (face < m.F && vis != kBoxOutside) {
        int i0 = m.vi[3 * face + 0], i1 = m.vi[3 * face + 1], i2 = m.vi[3 * face + 2];
        float4 v0 = o.clip[i0], v1 = o.clip[i1], v2 = o.clip[i2];
        if (!(flags & kBackfaceCulling) || facing_camera(v0, v1, v2)) {
            st.front++;
            float in0, in1, in2;
            if (flags & kLighting) {
                if (o.gouraud) {
                    int n0 = i0, n1 = i1, n2 = i2;
                    if (m.ni) {
                        n0 = m.ni[3 * face + 0];
                        n1 = m.ni[3 * face + 1];
                        n2 = m.ni[3 * face + 2];
                    }
                    in0 = o.nint[n0];
                    in1 = o.nint[n1];
                    in2 = o.nint[n2];
                } else {
                    const float* f = m.fn + 4 * face;
                    float4 t = xform(o.world, f[0], f[1], f[2], f[3]);
                    in0 = in1 = in2 = lambert(t, fr.light, fr.ambient, fr.diffuse);
                }
            } else {
                in0 = in1 = in2 = fr.ambient;
            }
            float u0 = 0.f, w0 = 0.f, u1 = 0.f, w1 = 0.f, u2 = 0.f, w2 = 0.f;
            if (m.uv_vert) {
                u0 = m.uv_vert[2 * i0];
                w0 = m.uv_vert[2 * i0 + 1];
                u1 = m.uv_vert[2 * i1];
                w1 = m.uv_vert[2 * i1 + 1];
                u2 = m.uv_vert[2 * i2];
                w2 = m.uv_vert[2 * i2 + 1];
            } else if (m.uv_face) {
                const float* uv = m.uv_face + 6 * face;
                u0 = uv[0];
                w0 = uv[1];
                u1 = uv[2];
                w1 = uv[3];
                u2 = uv[4];
                w2 = uv[5];
            }
            if (flags & kShowTextures) tex = m.face_tex ? m.face_tex[face] : m.mesh_tex;
            poly[0] = CV{v0.x, v0.y, v0.z, v0.w, u0, w0, in0};
            poly[1] = CV{v1.x, v1.y, v1.z, v1.w, u1, w1, in1};
            poly[2] = CV{v2.x, v2.y, v2.z, v2.w, u2, w2, in2};
            if ((flags & kFrustumClipping) && vis != kBoxInside) {
                uint32_t m0 = inside_mask(poly[0], fr.znear, fr.zfar);
                uint32_t m1 = inside_mask(poly[1], fr.znear, fr.zfar);
                uint32_t m2 = inside_mask(poly[2], fr.znear, fr.zfar);
                uint32_t all_in = m0 & m1 & m2;
                if (all_in == 0x3Fu) {
                    // Every pass only rotates the triangle; after 6 passes it is
                    // unchanged (clipping.go:199-232).
                    mode = 1;
                    n = 1;
                } else {
                    int p = __builtin_ctz(~all_in & 0x3Fu);
                    uint32_t any_in = (m0 | m1 | m2) >> p & 1u;
                    if (!any_in) {
                        // Earlier planes left the triangle untouched; this one
                        // rejects every vertex (polygon2.Count == 0).
                        n = 0;
                    } else {
                        int k = clip_general(poly, fr.znear, fr.zfar);
                        if (k < 0) {
                            st.overflow++;
                            n = 0;
                        } else {
                            mode = 2;
                            n = k >= 3 ? k - 2 : 0;
                        }
                    }
                }
            } else {
                mode = 1;
                n = 1;
            }
        }
    }

    int total;
    int off = block_exclusive_scan(n, lds_scan, total);
    uint32_t slot0 = (uint32_t)(o.seq_block + wg_in_obj) * kSlotsPerWg + (uint32_t)off;
    if (mode == 1) {
        emit_triangle(fr, poly[0], poly[1], poly[2], tex, slot0, tw_log2, th_log2, tile_counts,
                      st);
    } else if (mode == 2) {
        for (int i = 0; i < n; i++)
            emit_triangle(fr, poly[0], poly[i + 1], poly[i + 2], tex, slot0 + i, tw_log2,
                          th_log2, tile_counts, st);
    }
    if (threadIdx.x == 0) wg_count[b] = (uint32_t)total;

    // Block-reduce the counters, one atomic per counter per block.
    unsigned long long s0 = wave_sum_u64(st.tpf), s1 = wave_sum_u64(st.tris);
    unsigned long long s2 = wave_sum_u64(st.front), s3 = wave_sum_u64(st.overflow);
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        lds_st[wave][0] = s0;
        lds_st[wave][1] = s1;
        lds_st[wave][2] = s2;
        lds_st[wave][3] = s3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long t = lds_st[0][threadIdx.x] + lds_st[1][threadIdx.x] +
                               lds_st[2][threadIdx.x] + lds_st[3][threadIdx.x];
        const int idx[4] = {kStatTpf, kStatTriangles, kStatFront, kStatClipOverflow};
        if (t) atomicAdd(&stats[o.frame * kNumStats + idx[threadIdx.x]], t);
    }
}

// ---------------------------------------------------------------------------
// k_bin_scan: tile counts -> list offsets, per frame
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(1024) void k_bin_scan(const DevFrame* __restrict__ frames,
                                                   uint32_t* __restrict__ tile_counts,
                                                   uint32_t* __restrict__ tile_offsets,
                                                   uint32_t* __restrict__ tile_cursor,
                                                   uint32_t* __restrict__ frame_overflow,
                                                   unsigned long long* __restrict__ stats,
                                                   unsigned long long* __restrict__ refs_used,
                                                   DevControl* __restrict__ ctrl,
                                                   unsigned long long refs_capacity) {
    __shared__ unsigned long long wsum[16];
    __shared__ unsigned long long base_s;
    const DevFrame& fr = frames[blockIdx.x];
    int ntiles = fr.tiles_x * fr.tiles_y;
    uint32_t* cnt = tile_counts + fr.tile_begin;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    // Pass 1: total refs of the frame.
    unsigned long long local = 0;
    for (int i = threadIdx.x; i < ntiles; i += 1024) local += cnt[i];
    local = wave_sum_u64(local);
    if (lane == 0) wsum[wave] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int w = 0; w < 16; w++) tot += wsum[w];
        unsigned long long base = atomicAdd(refs_used, tot);
        stats[blockIdx.x * kNumStats + kStatRefs] = tot;
        if (base + tot > refs_capacity) {
            frame_overflow[blockIdx.x] = 1u;
            atomicOr(&ctrl->overflow, 1u);
            atomicMax(&ctrl->refs_needed, base + tot);
        }
        base_s = base;
    }
    __syncthreads();
    bool over = frame_overflow[blockIdx.x] != 0u;
    unsigned long long run = base_s;
    // Pass 2: exclusive scan in chunks of 1024.
    for (int c = 0; c < ntiles; c += 1024) {
        int i = c + threadIdx.x;
        uint32_t v = i < ntiles ? cnt[i] : 0u;
        unsigned long long x = v;
        for (int o = 1; o < 64; o <<= 1) {
            unsigned long long y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
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
            unsigned long long off = run + wbase + x - v;
            tile_offsets[fr.tile_begin + i] = (uint32_t)off;
            tile_cursor[fr.tile_begin + i] = (uint32_t)off;
            if (over) cnt[i] = 0u;  // overflowed frame: raster draws background only
        }
        run += ctot;
    }
}

// ---------------------------------------------------------------------------
// k_bin_fill
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_bin_fill(const DevObject* __restrict__ objs, int nobj,
                                                  const DevFrame* __restrict__ frames,
                                                  const uint32_t* __restrict__ wg_count,
                                                  const uint32_t* __restrict__ frame_overflow,
                                                  uint32_t* __restrict__ tile_cursor,
                                                  uint32_t* __restrict__ refs, int tw_log2,
                                                  int th_log2) {
    int b = blockIdx.x;
    int oi = last_le(nobj, b, [&](int i) { return objs[i].face_block; });
    const DevObject& o = objs[oi];
    if (frame_overflow[o.frame]) return;
    const DevFrame& fr = frames[o.frame];
    uint32_t cnt = wg_count[b];
    uint32_t slot0 = (uint32_t)(o.seq_block + (b - o.face_block)) * kSlotsPerWg;
    for (uint32_t i = threadIdx.x; i < cnt; i += 256) {
        uint32_t slot = slot0 + i;
        const RasterRec& r = fr.rrec[slot];
        int bx0 = r.bx0, bx1 = r.bx1, by0 = r.by0, by1 = r.by1;
        if (bx0 > bx1) continue;
        int tx0 = bx0 >> tw_log2, tx1 = bx1 >> tw_log2;
        int ty0 = by0 >> th_log2, ty1 = by1 >> th_log2;
        for (int ty = ty0; ty <= ty1; ty++)
            for (int tx = tx0; tx <= tx1; tx++) {
                uint32_t pos = atomicAdd(&tile_cursor[fr.tile_begin + ty * fr.tiles_x + tx], 1u);
                refs[pos] = slot;
            }
    }
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

struct EdgeEval {
    float alpha, beta, gamma, zrec;
    bool in;
};

// Pixel (x, y) of triangle r (rasterizer.go:147-153), evaluating the edge
// functions directly, which equals the reference's incremental int64 walk.
__device__ __forceinline__ EdgeEval eval_pixel(const RasterRec& r, const WideRec* wr, int x,
                                               int y) {
    EdgeEval e;
    float f12f, f20f;
    if (!r.wide) {
        int32_t f01 = r.e[0] * x + r.e[1] * y + r.e[2];
        int32_t f12 = r.e[3] * x + r.e[4] * y + r.e[5];
        int32_t f20 = r.e[6] * x + r.e[7] * y + r.e[8];
        e.in = f01 < 0 && f12 < 0 && f20 < 0;
        f12f = (float)f12;
        f20f = (float)f20;
    } else {
        int64_t f[3];
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
        e.in = f[0] < 0 && f[1] < 0 && f[2] < 0;
        f12f = (float)f[1];
        f20f = (float)f[2];
    }
    e.alpha = f12f / r.sf;
    e.beta = f20f / r.sf;
    e.gamma = 1.0f - e.alpha - e.beta;
    e.zrec = -(e.alpha / r.z0 + e.beta / r.z1 + e.gamma / r.z2);
    return e;
}

// texture.go:69-88
__device__ __forceinline__ uint32_t sample(const DevTexture& t, float u, float v) {
    switch (t.type) {
    case 0:
        return t.color;
    case 2: {
        int64_t x = go_int((1 - u) * t.scale * t.wf) & (int64_t)(t.w - 1);
        int64_t y = go_int(v * t.scale * t.hf) & (int64_t)(t.h - 1);
        return t.texels[y * t.w + x];
    }
    case 1: {
        int64_t x = go_int((1 - u) * t.scale * t.wf) % (int64_t)t.w;
        int64_t y = go_int(v * t.scale * t.hf) % (int64_t)t.h;
        int64_t idx = y * t.w + x;
        if (idx < 0) idx = 0;
        return t.texels[idx];
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

template <int TW, int TH>
__global__ __launch_bounds__(256) void k_raster(const DevFrame* __restrict__ frames, int nframes,
                                                const uint32_t* __restrict__ tile_counts,
                                                const uint32_t* __restrict__ tile_offsets,
                                                const uint32_t* __restrict__ refs,
                                                const DevTexture* __restrict__ textures,
                                                unsigned long long* __restrict__ stats) {
    constexpr int NPX = TW * TH;
    __shared__ unsigned long long key[NPX];
    __shared__ unsigned long long red[4][2];
    int g = blockIdx.x;
    int fi = last_le(nframes, g, [&](int i) { return frames[i].tile_begin; });
    const DevFrame& fr = frames[fi];
    int t = g - fr.tile_begin;
    int X0 = (t % fr.tiles_x) * TW, Y0 = (t / fr.tiles_x) * TH;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    for (int i = threadIdx.x; i < NPX; i += 256) key[i] = kBackgroundKey;
    __syncthreads();

    uint32_t cnt = tile_counts[g], off = tile_offsets[g];
    unsigned long long covered = 0;
    for (uint32_t i = wave; i < cnt; i += 4) {
        uint32_t slot = refs[off + i];
        const RasterRec r = fr.rrec[slot];
        const WideRec* wr = r.wide ? &fr.wrec[slot] : nullptr;
        int x0 = max((int)r.bx0, X0), x1 = min((int)r.bx1, X0 + TW - 1);
        int y0 = max((int)r.by0, Y0), y1 = min((int)r.by1, Y0 + TH - 1);
        int bw = x1 - x0 + 1;
        int lpr_log2 = bw <= 1 ? 0 : 32 - __builtin_clz((uint32_t)(bw - 1));
        if (lpr_log2 > 6) lpr_log2 = 6;
        int lpr = 1 << lpr_log2, rows = 64 >> lpr_log2;
        int lx = lane & (lpr - 1), ly = lane >> lpr_log2;
        for (int yb = y0; yb <= y1; yb += rows) {
            int y = yb + ly;
            for (int xb = x0; xb <= x1; xb += lpr) {
                int x = xb + lx;
                if (y > y1 || x > x1) continue;
                EdgeEval e = eval_pixel(r, wr, x, y);
                if (!e.in) continue;
                covered++;
                if (e.zrec >= -1.0f)
                    atomicMax(&key[(y - Y0) * TW + (x - X0)], depth_key(e.zrec, slot + 1u));
            }
        }
    }
    __syncthreads();

    // Deferred shading of the winning fragment of each pixel, one write per pixel.
    // Background colours are read as values and the dot test is evaluated
    // without short-circuit branches: hipcc 7.2 (clang 22) turned the
    // short-circuit form into a select of field addresses that dropped the
    // y % step condition when the loop runs once per thread.
    const uint32_t dot_color = fr.dot_color, clear_color = fr.clear_color;
    const int step = fr.dot_step;
    const int step1 = step >= 1 ? step : 1;
    unsigned long long shaded = 0;
    for (int i = threadIdx.x; i < NPX; i += 256) {
        int x = X0 + (i % TW), y = Y0 + (i / TW);
        if (x >= fr.W || y >= fr.H) continue;
        unsigned long long k = key[i];
        uint32_t seq = (uint32_t)k;
        uint32_t color;
        float depth;
        if (seq == 0u) {
            // DotGrid (rasterizer.go:46-52): x, y in {step, 2*step, ...}
            uint32_t dot = (uint32_t)(step >= 1) & (uint32_t)(x >= step) &
                           (uint32_t)(y >= step) & (uint32_t)(x % step1 == 0) &
                           (uint32_t)(y % step1 == 0);
            color = dot ? dot_color : clear_color;
            depth = -1.0f;
        } else {
            shaded++;
            uint32_t slot = seq - 1u;
            const RasterRec r = fr.rrec[slot];
            const ShadeRec s = fr.srec[slot];
            const WideRec* wr = r.wide ? &fr.wrec[slot] : nullptr;
            EdgeEval e = eval_pixel(r, wr, x, y);
            // rasterizer.go:158-170
            float u = (e.alpha * s.uz[0] + e.beta * s.uz[1] + e.gamma * s.uz[2]) / e.zrec;
            float v = (e.alpha * s.vz[0] + e.beta * s.vz[1] + e.gamma * s.vz[2]) / e.zrec;
            float inten = e.alpha * s.in[0] + e.beta * s.in[1] + e.gamma * s.in[2];
            uint32_t c = s.tex < 0 ? fr.face_color : sample(textures[s.tex], u, v);
            color = color_intensity(c, inten);
            depth = e.zrec;
        }
        size_t idx = (size_t)y * fr.W + x;
        ((uint32_t*)fr.rgba)[idx] = color;
        fr.depth[idx] = depth;
    }

    covered = wave_sum_u64(covered);
    shaded = wave_sum_u64(shaded);
    if (lane == 0) {
        red[wave][0] = covered;
        red[wave][1] = shaded;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned long long s = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                               red[3][threadIdx.x];
        if (s) atomicAdd(&stats[fi * kNumStats + (threadIdx.x ? kStatShaded : kStatCovered)], s);
    }
}

// ---------------------------------------------------------------------------
// Mesh set-up: face normals as NewMesh computes them (mesh.go:53-60)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_face_normals(const float* __restrict__ px,
                                                      const float* __restrict__ py,
                                                      const float* __restrict__ pz,
                                                      const int32_t* __restrict__ vi, int F,
                                                      float* __restrict__ fn) {
    int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= F) return;
    int a = vi[3 * f], b = vi[3 * f + 1], c = vi[3 * f + 2];
    float ax = px[b] - px[a], ay = py[b] - py[a], az = pz[b] - pz[a];
    float bx = px[c] - px[a], by = py[c] - py[a], bz = pz[c] - pz[a];
    float nx = ay * bz - az * by;
    float ny = az * bx - ax * bz;
    float nz = ax * by - ay * bx;
    float len = __builtin_sqrtf(nx * nx + ny * ny + nz * nz);
    fn[4 * f + 0] = nx / len;
    fn[4 * f + 1] = ny / len;
    fn[4 * f + 2] = nz / len;
    fn[4 * f + 3] = 1.0f;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

void launch_face_normals(const float* px, const float* py, const float* pz, const int32_t* vi,
                         int F, float* fn, hipStream_t s) {
    if (F <= 0) return;
    hipLaunchKernelGGL(k_face_normals, dim3((F + 255) / 256), dim3(256), 0, s, px, py, pz, vi, F,
                       fn);
}

void launch_vertex(const LaunchParams& p, hipStream_t s) {
    int zero_blocks = 64;
    hipLaunchKernelGGL(k_vertex, dim3(p.vertex_blocks + zero_blocks), dim3(256), 0, s, p.objs,
                       p.nobj, p.frames, p.vertex_blocks, p.zero_u32, p.zero_u32_n, p.zero_u64,
                       p.zero_u64_n);
}

void launch_face(const LaunchParams& p, hipStream_t s) {
    if (p.face_blocks == 0) return;
    hipLaunchKernelGGL(k_face, dim3(p.face_blocks), dim3(256), 0, s, p.objs, p.nobj, p.frames,
                       p.tile_counts, p.wg_count, p.stats, p.tw_log2, p.th_log2);
}

void launch_bin_scan(const LaunchParams& p, hipStream_t s) {
    hipLaunchKernelGGL(k_bin_scan, dim3(p.nframes), dim3(1024), 0, s, p.frames, p.tile_counts,
                       p.tile_offsets, p.tile_cursor, p.frame_overflow, p.stats, p.refs_used,
                       p.ctrl, p.refs_capacity);
}

void launch_bin_fill(const LaunchParams& p, hipStream_t s) {
    if (p.face_blocks == 0) return;
    hipLaunchKernelGGL(k_bin_fill, dim3(p.face_blocks), dim3(256), 0, s, p.objs, p.nobj,
                       p.frames, p.wg_count, p.frame_overflow, p.tile_cursor, p.refs,
                       p.tw_log2, p.th_log2);
}

template <int TW, int TH>
static void raster_t(const LaunchParams& p, hipStream_t s) {
    hipLaunchKernelGGL((k_raster<TW, TH>), dim3(p.total_tiles), dim3(256), 0, s, p.frames,
                       p.nframes, p.tile_counts, p.tile_offsets, p.refs, p.textures, p.stats);
}

bool launch_raster(const LaunchParams& p, hipStream_t s) {
    if (p.total_tiles == 0) return true;
#define GR_RASTER_CASE(W, H)                          \
    if (p.tw_log2 == W && p.th_log2 == H) {           \
        raster_t<(1 << W), (1 << H)>(p, s);           \
        return true;                                  \
    }
    GR_RASTER_CASE(3, 3) GR_RASTER_CASE(3, 4) GR_RASTER_CASE(3, 5) GR_RASTER_CASE(3, 6)
    GR_RASTER_CASE(4, 3) GR_RASTER_CASE(4, 4) GR_RASTER_CASE(4, 5) GR_RASTER_CASE(4, 6)
    GR_RASTER_CASE(5, 3) GR_RASTER_CASE(5, 4) GR_RASTER_CASE(5, 5) GR_RASTER_CASE(5, 6)
    GR_RASTER_CASE(6, 3) GR_RASTER_CASE(6, 4) GR_RASTER_CASE(6, 5) GR_RASTER_CASE(6, 6)
#undef GR_RASTER_CASE
    return false;
}

}  // namespace gr
