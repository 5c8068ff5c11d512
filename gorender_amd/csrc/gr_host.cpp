// Host-side float32 math of the reference's per-object set-up (matrix.go,
// vector.go, renderer.go:114-121,255-265). These run once per object per frame
// on the CPU in the reference too; they build the float[16] matrices the draw
// call consumes. Compiled with -ffp-contract=off on x86-64 SSE, the same
// float32 rounding as Go gc amd64.
#include <cmath>
#include <cstring>

#include "gorender_hip.h"

namespace {

struct M4 {
    float m[4][4];
};

M4 identity() {
    M4 r;
    std::memset(&r, 0, sizeof r);
    r.m[0][0] = r.m[1][1] = r.m[2][2] = r.m[3][3] = 1.0f;
    return r;
}

// matrix.go:155-165
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 4; k++) acc += a.m[i][k] * b.m[k][j];
            r.m[i][j] = acc;
        }
    return r;
}

// math32.go:15-25
float sin32(float x) { return (float)std::sin((double)x); }
float cos32(float x) { return (float)std::cos((double)x); }
float tan32(float x) { return (float)std::tan((double)x); }
float sqrt32(float x) { return (float)std::sqrt((double)x); }

M4 rot_x(float a) {  // matrix.go:32-44
    M4 r = identity();
    if (a == 0) return r;
    float s = sin32(a), c = cos32(a);
    r.m[1][1] = c;
    r.m[1][2] = -s;
    r.m[2][1] = s;
    r.m[2][2] = c;
    return r;
}

M4 rot_y(float a) {  // matrix.go:46-58
    M4 r = identity();
    if (a == 0) return r;
    float s = sin32(a), c = cos32(a);
    r.m[0][0] = c;
    r.m[0][2] = s;
    r.m[2][0] = -s;
    r.m[2][2] = c;
    return r;
}

M4 rot_z(float a) {  // matrix.go:60-72
    M4 r = identity();
    if (a == 0) return r;
    float s = sin32(a), c = cos32(a);
    r.m[0][0] = c;
    r.m[0][1] = -s;
    r.m[1][0] = s;
    r.m[1][1] = c;
    return r;
}

struct V3 {
    float x, y, z;
};

V3 cross(V3 a, V3 b) {  // vector.go:80-86
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 normalize(V3 v) {  // vector.go:76-78,92-94
    float l = sqrt32(v.x * v.x + v.y * v.y + v.z * v.z);
    return V3{v.x / l, v.y / l, v.z / l};
}

void store(const M4& a, float* out) { std::memcpy(out, a.m, 64); }
M4 load(const float* in) {
    M4 r;
    std::memcpy(r.m, in, 64);
    return r;
}

// BoxVisibility of an object (renderer.go:268-275, clipping.go:131-154): its
// bounding-box corners through MVP in the transform's order
// (asm_purego.go:9-19), then per plane in Left, Right, Top, Bottom, Near, Far
// order the corners with N.c - N.P > 0 (clipping.go:69-71) are outside; the
// first plane with every corner outside gives Outside, the first with some
// outside gives Intersect, none gives Inside (0, 1, 2). float32 throughout,
// no contraction.
int box_visibility(const float* m, const float bbox[8][4], float zn, float zf) {
    float c[8][4];
    for (int k = 0; k < 8; k++) {
        const float x = bbox[k][0], y = bbox[k][1], z = bbox[k][2], w = bbox[k][3];
        for (int r = 0; r < 4; r++)
            c[k][r] = m[4 * r] * x + m[4 * r + 1] * y + m[4 * r + 2] * z + m[4 * r + 3] * w;
    }
    // clipping.go:93-129: (point, normal) per plane.
    const float P[6][4] = {{-1, 0, 0, 1}, {1, 0, 0, 1}, {0, -1, 0, 1},
                           {0, 1, 0, 1},  {0, 0, zn, 1}, {0, 0, zf, 1}};
    const float N[6][4] = {{1, 0, 0, 1}, {-1, 0, 0, 1}, {0, 1, 0, 1},
                           {0, -1, 0, 1}, {0, 0, -1, 0}, {0, 0, 1, 0}};
    for (int p = 0; p < 6; p++) {
        const float np = N[p][0] * P[p][0] + N[p][1] * P[p][1] + N[p][2] * P[p][2] +
                         N[p][3] * P[p][3];
        int outside = 0;
        for (int k = 0; k < 8; k++) {
            const float d = N[p][0] * c[k][0] + N[p][1] * c[k][1] + N[p][2] * c[k][2] +
                            N[p][3] * c[k][3];
            outside += d - np > 0.f;
        }
        if (outside == 8) return 0;
        if (outside > 0) return 1;
    }
    return 2;
}

}  // namespace

extern "C" {

void gr_matrix_multiply(const float* a, const float* b, float* out) {
    store(mul(load(a), load(b)), out);
}

// matrix.go:74-88 (NewRotationMatrix inside NewWorldMatrix)
void gr_matrix_world(const float scale[3], const float rotation[3], const float translation[3],
                     float* out) {
    M4 s = identity();
    s.m[0][0] = scale[0];
    s.m[1][1] = scale[1];
    s.m[2][2] = scale[2];
    M4 r = identity();
    r = mul(r, rot_x(rotation[0]));
    r = mul(r, rot_y(rotation[1]));
    r = mul(r, rot_z(rotation[2]));
    M4 t = identity();
    t.m[0][3] = translation[0];
    t.m[1][3] = translation[1];
    t.m[2][3] = translation[2];
    M4 m = identity();
    m = mul(s, m);
    m = mul(r, m);
    m = mul(t, m);
    store(m, out);
}

// matrix.go:133-144
void gr_matrix_view(const float eye[3], const float direction[3], const float up[3], float* out) {
    V3 e{eye[0], eye[1], eye[2]};
    V3 z = normalize(V3{direction[0], direction[1], direction[2]});
    V3 x = normalize(cross(V3{up[0], up[1], up[2]}, z));
    V3 y = normalize(cross(z, x));
    M4 m;
    const V3 rows[3] = {x, y, z};
    for (int i = 0; i < 3; i++) {
        m.m[i][0] = rows[i].x;
        m.m[i][1] = rows[i].y;
        m.m[i][2] = rows[i].z;
        m.m[i][3] = -dot(rows[i], e);
    }
    m.m[3][0] = m.m[3][1] = m.m[3][2] = 0.0f;
    m.m[3][3] = 1.0f;
    store(m, out);
}

// matrix.go:92-106
void gr_matrix_perspective(float fov, float aspect, float znear, float zfar, float* out) {
    float th = tan32(fov / 2.0f);
    M4 m;
    std::memset(&m, 0, sizeof m);
    m.m[0][0] = 1.0f / (aspect * th);
    m.m[1][1] = 1.0f / th;
    float m22 = (zfar + znear) / (znear - zfar);
    float m23 = (2.0f * zfar * znear) / (znear - zfar);
    m.m[2][2] = -m22;
    m.m[2][3] = -m23;
    m.m[3][2] = -1.0f;
    store(m, out);
}

// matrix.go:108-118
void gr_matrix_screen(int32_t width, int32_t height, float* out) {
    M4 m;
    std::memset(&m, 0, sizeof m);
    float hw = (float)width / 2.0f, hh = (float)height / 2.0f;
    m.m[0][0] = hw;
    m.m[0][3] = hw;
    m.m[1][1] = hh;
    m.m[1][3] = hh;
    m.m[2][2] = 0.5f;
    m.m[2][3] = 0.5f;
    m.m[3][3] = 1.0f;
    store(m, out);
}

// renderer.go:259-262: ((I * P) * V) * W
void gr_matrix_mvp(const float* perspective, const float* view, const float* world, float* out) {
    M4 m = identity();
    m = mul(m, load(perspective));
    m = mul(m, load(view));
    m = mul(m, load(world));
    store(m, out);
}

// Vec3.Normalize (vector.go:92-94), e.g. the light direction of renderer.go:265.
void gr_vec3_normalize(const float in[3], float out[3]) {
    V3 v = normalize(V3{in[0], in[1], in[2]});
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}

}  // extern "C"

extern "C" int32_t gr_box_visibility(const float* mvp, const float* bbox, float znear, float zfar) {
    if (!mvp || !bbox) return -1;
    return box_visibility(mvp, reinterpret_cast<const float(*)[4]>(bbox), znear, zfar);
}

// NewImageTexture picks the fast (power-of-two) type when both sizes are
// powers of two (texture.go:40-42, isPowerOfTwo utils.go:7-9).
extern "C" int32_t gr_texture_kind_for_size(int32_t w, int32_t h) {
    auto pow2 = [](int32_t x) { return x != 0 && (x & (x - 1)) == 0; };
    return (pow2(w) && pow2(h)) ? GR_TEXTURE_IMAGE_FAST : GR_TEXTURE_IMAGE;
}
