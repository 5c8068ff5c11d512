/*
 * refcpu.c — CPU restatement of gorender's render path. TEST INFRASTRUCTURE ONLY:
 * the product library never links, loads or calls it. Tests, the smoke check and
 * bench.py's cpu_baseline leg use it as the checker.
 *
 * Written from the reference's behaviour (sequential draw order,
 * renderer.go:466-474), one function per reference function, in the reference's
 * float32 evaluation order. Build: -O2 -ffp-contract=off -fno-fast-math on
 * x86-64 (SSE scalar float: RNE, no FMA, denormals kept).
 *
 * Pinning (tests/test_oracle_*.py):
 *   rc_go_int / rc_go_uint8 / rc_go_min / rc_go_max  closed-form special values
 *   rc_transform_batch          exact-integer closed forms, association-order KAT
 *   rc_transform_batch_pairwise oracle/_ref (reference c/matrix_amd64.c) + c/main.c KAT
 *   matrix builders             closed forms (identity short-cut, exact integer products);
 *                               trig values parity unpinned (libm vs Go math)
 *   rc_face_normals / rc_bounding_box  axis-aligned closed forms
 *   rc_box_visibility           hand-classified boxes
 *   rc_clip_triangle            exact-rational Sutherland-Hodgman (Python fractions),
 *                               identity/all-outside closed forms, invariants
 *   rc_facing_camera            exact small-integer triple products
 *   rc_triangle                 brute-force direct int edge evaluation, shared-edge
 *                               coverage invariants, exact right-triangle barycentrics
 *   rc_sample / rc_color_intensity  exhaustive grids against closed forms
 *   rc_draw                     order-independence of per-pixel (max depth, last draw)
 *                               resolution, brute-force pixel recomputation
 */
#include "refcpu.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Go conversion semantics                                                    */
/* ------------------------------------------------------------------------- */

/* int(float32) on amd64 is CVTTSS2SQ: truncation toward zero; NaN and values
 * outside [-2^63, 2^63) give INT64_MIN (renderer.go:182-184, texture.go:75-80). */
int64_t rc_go_int(float x) {
    if (!(x >= -9223372036854775808.0f && x < 9223372036854775808.0f)) return INT64_MIN;
    return (int64_t)x;
}

/* uint8(float32) is CVTTSS2SL then the low 8 bits; NaN or out of int32 range
 * gives 0x80000000 whose low byte is 0 (rasterizer.go:83-85). */
uint8_t rc_go_uint8(float x) {
    int32_t i;
    if (!(x >= -2147483648.0f && x < 2147483648.0f))
        i = INT32_MIN;
    else
        i = (int32_t)x;
    return (uint8_t)((uint32_t)i & 0xFFu);
}

/* Go builtin min/max for floats: any NaN gives NaN; min(-0,+0) = -0. */
float rc_go_min(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a < b) return a;
    if (b < a) return b;
    return signbit(a) ? a : b;
}

float rc_go_max(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a > b) return a;
    if (b > a) return b;
    return signbit(a) ? b : a;
}

static int64_t i64_min(int64_t a, int64_t b) { return a < b ? a : b; }
static int64_t i64_max(int64_t a, int64_t b) { return a > b ? a : b; }

/* Go int arithmetic wraps; do it in uint64 to avoid C signed-overflow UB. */
static int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
static int64_t wsub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
static int64_t wmul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }

/* ------------------------------------------------------------------------- */
/* vector.go / math32.go                                                      */
/* ------------------------------------------------------------------------- */

/* math32.go:11-13 */
float rc_sqrt32(float x) { return (float)sqrt((double)x); }

static rc_vec3 v3_sub(rc_vec3 a, rc_vec3 b) {
    rc_vec3 r = {a.x - b.x, a.y - b.y, a.z - b.z};
    return r;
}

/* vector.go:80-86 */
static rc_vec3 v3_cross(rc_vec3 a, rc_vec3 b) {
    rc_vec3 r;
    r.x = a.y * b.z - a.z * b.y;
    r.y = a.z * b.x - a.x * b.z;
    r.z = a.x * b.y - a.y * b.x;
    return r;
}

/* vector.go:88-90 */
static float v3_dot(rc_vec3 a, rc_vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* vector.go:76-78, 92-94 */
rc_vec3 rc_vec3_normalize(rc_vec3 v) {
    float len = rc_sqrt32(v.x * v.x + v.y * v.y + v.z * v.z);
    rc_vec3 r = {v.x / len, v.y / len, v.z / len};
    return r;
}

static rc_vec4 v4_sub(rc_vec4 a, rc_vec4 b) {
    rc_vec4 r = {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w};
    return r;
}

static rc_vec4 v4_add(rc_vec4 a, rc_vec4 b) {
    rc_vec4 r = {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
    return r;
}

/* vector.go:107-109 */
static rc_vec4 v4_mul(rc_vec4 a, float s) {
    rc_vec4 r = {a.x * s, a.y * s, a.z * s, a.w * s};
    return r;
}

/* vector.go:103-105 */
static rc_vec4 v4_div(rc_vec4 a, float s) {
    rc_vec4 r = {a.x / s, a.y / s, a.z / s, a.w / s};
    return r;
}

/* vector.go:111-113 */
static float v4_dot(rc_vec4 a, rc_vec4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

/* vector.go:115-125 */
rc_vec4 rc_vec4_normalize(rc_vec4 v) {
    float len = rc_sqrt32(v.x * v.x + v.y * v.y + v.z * v.z + v.w * v.w);
    return v4_div(v, len);
}

static rc_vec3 v4_xyz(rc_vec4 v) {
    rc_vec3 r = {v.x, v.y, v.z};
    return r;
}

/* ------------------------------------------------------------------------- */
/* matrix.go / asm_purego.go                                                  */
/* ------------------------------------------------------------------------- */

rc_matrix rc_identity_matrix(void) {
    rc_matrix m;
    memset(&m, 0, sizeof m);
    m.m[0][0] = m.m[1][1] = m.m[2][2] = m.m[3][3] = 1.0f;
    return m;
}

/* matrix.go:155-165: res starts at +0 and accumulates k = 0..3 in order. */
rc_matrix rc_matrix_multiply(const rc_matrix *a, const rc_matrix *b) {
    rc_matrix r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 4; k++) acc += a->m[i][k] * b->m[k][j];
            r.m[i][j] = acc;
        }
    return r;
}

/* asm_purego.go:9-19: left-associated ((m0*x + m1*y) + m2*z) + m3*w. */
void rc_transform_batch(const rc_matrix *m, rc_vec4 *vecs, int64_t n) {
    for (int64_t i = 0; i < n; i++) {
        rc_vec4 v = vecs[i];
        rc_vec4 r;
        r.x = m->m[0][0] * v.x + m->m[0][1] * v.y + m->m[0][2] * v.z + m->m[0][3] * v.w;
        r.y = m->m[1][0] * v.x + m->m[1][1] * v.y + m->m[1][2] * v.z + m->m[1][3] * v.w;
        r.z = m->m[2][0] * v.x + m->m[2][1] * v.y + m->m[2][2] * v.z + m->m[2][3] * v.w;
        r.w = m->m[3][0] * v.x + m->m[3][1] * v.y + m->m[3][2] * v.z + m->m[3][3] * v.w;
        vecs[i] = r;
    }
}

/* The amd64 build's association (c/matrix_amd64.c:6-18 with the transposed
 * matrix of asm_amd64.go:9): (m0*x + m1*y) + (m2*z + m3*w). CPU baseline only. */
void rc_transform_batch_pairwise(const rc_matrix *m, rc_vec4 *vecs, int64_t n) {
    for (int64_t i = 0; i < n; i++) {
        rc_vec4 v = vecs[i];
        float o[4];
        for (int r = 0; r < 4; r++) {
            float p1 = m->m[r][0] * v.x, p2 = m->m[r][1] * v.y;
            float p3 = m->m[r][2] * v.z, p4 = m->m[r][3] * v.w;
            o[r] = (p1 + p2) + (p3 + p4);
        }
        vecs[i].x = o[0];
        vecs[i].y = o[1];
        vecs[i].z = o[2];
        vecs[i].w = o[3];
    }
}

/* matrix.go:167-173 */
void rc_transform_inplace(const rc_matrix *m, rc_vec4 *v) { rc_transform_batch(m, v, 1); }

/* matrix.go:14-21 */
rc_matrix rc_scale_matrix(float x, float y, float z) {
    rc_matrix m = rc_identity_matrix();
    m.m[0][0] = x;
    m.m[1][1] = y;
    m.m[2][2] = z;
    return m;
}

/* matrix.go:23-30 */
rc_matrix rc_translation_matrix(float x, float y, float z) {
    rc_matrix m = rc_identity_matrix();
    m.m[0][3] = x;
    m.m[1][3] = y;
    m.m[2][3] = z;
    return m;
}

/* math32.go:15-25: float64 trig rounded to float32 (libm here, Go math there:
 * parity unpinned beyond exact special values). */
static float sin32(float x) { return (float)sin((double)x); }
static float cos32(float x) { return (float)cos((double)x); }
static float tan32(float x) { return (float)tan((double)x); }

/* matrix.go:32-44 */
rc_matrix rc_rotation_x_matrix(float a) {
    rc_matrix m = rc_identity_matrix();
    if (a == 0) return m;
    float s = sin32(a), c = cos32(a);
    m.m[1][1] = c;
    m.m[1][2] = -s;
    m.m[2][1] = s;
    m.m[2][2] = c;
    return m;
}

/* matrix.go:46-58 */
rc_matrix rc_rotation_y_matrix(float a) {
    rc_matrix m = rc_identity_matrix();
    if (a == 0) return m;
    float s = sin32(a), c = cos32(a);
    m.m[0][0] = c;
    m.m[0][2] = s;
    m.m[2][0] = -s;
    m.m[2][2] = c;
    return m;
}

/* matrix.go:60-72 */
rc_matrix rc_rotation_z_matrix(float a) {
    rc_matrix m = rc_identity_matrix();
    if (a == 0) return m;
    float s = sin32(a), c = cos32(a);
    m.m[0][0] = c;
    m.m[0][1] = -s;
    m.m[1][0] = s;
    m.m[1][1] = c;
    return m;
}

/* matrix.go:74-80 */
rc_matrix rc_rotation_matrix(float x, float y, float z) {
    rc_matrix m = rc_identity_matrix();
    rc_matrix t = rc_rotation_x_matrix(x);
    m = rc_matrix_multiply(&m, &t);
    t = rc_rotation_y_matrix(y);
    m = rc_matrix_multiply(&m, &t);
    t = rc_rotation_z_matrix(z);
    m = rc_matrix_multiply(&m, &t);
    return m;
}

/* matrix.go:82-88 */
rc_matrix rc_world_matrix(rc_vec3 scale, rc_vec3 rotation, rc_vec3 translation) {
    rc_matrix m = rc_identity_matrix();
    rc_matrix t = rc_scale_matrix(scale.x, scale.y, scale.z);
    m = rc_matrix_multiply(&t, &m);
    t = rc_rotation_matrix(rotation.x, rotation.y, rotation.z);
    m = rc_matrix_multiply(&t, &m);
    t = rc_translation_matrix(translation.x, translation.y, translation.z);
    m = rc_matrix_multiply(&t, &m);
    return m;
}

/* matrix.go:92-106 */
rc_matrix rc_perspective_matrix(float fov, float aspect, float znear, float zfar) {
    float tan_half = tan32(fov / 2.0f);
    float m00 = 1.0f / (aspect * tan_half);
    float m11 = 1.0f / tan_half;
    float m22 = (zfar + znear) / (znear - zfar);
    float m23 = (2.0f * zfar * znear) / (znear - zfar);
    rc_matrix m;
    memset(&m, 0, sizeof m);
    m.m[0][0] = m00;
    m.m[1][1] = m11;
    m.m[2][2] = -m22;
    m.m[2][3] = -m23;
    m.m[3][2] = -1.0f;
    return m;
}

/* matrix.go:108-118 */
rc_matrix rc_screen_matrix(int32_t width, int32_t height) {
    float hw = (float)width / 2.0f;
    float hh = (float)height / 2.0f;
    rc_matrix m;
    memset(&m, 0, sizeof m);
    m.m[0][0] = hw;
    m.m[0][3] = hw;
    m.m[1][1] = hh;
    m.m[1][3] = hh;
    m.m[2][2] = 0.5f;
    m.m[2][3] = 0.5f;
    m.m[3][3] = 1.0f;
    return m;
}

/* matrix.go:133-144 */
rc_matrix rc_view_matrix(rc_vec3 eye, rc_vec3 direction, rc_vec3 up) {
    rc_vec3 z = rc_vec3_normalize(direction);
    rc_vec3 x = rc_vec3_normalize(v3_cross(up, z));
    rc_vec3 y = rc_vec3_normalize(v3_cross(z, x));
    rc_matrix m;
    m.m[0][0] = x.x;
    m.m[0][1] = x.y;
    m.m[0][2] = x.z;
    m.m[0][3] = -v3_dot(x, eye);
    m.m[1][0] = y.x;
    m.m[1][1] = y.y;
    m.m[1][2] = y.z;
    m.m[1][3] = -v3_dot(y, eye);
    m.m[2][0] = z.x;
    m.m[2][1] = z.y;
    m.m[2][2] = z.z;
    m.m[2][3] = -v3_dot(z, eye);
    m.m[3][0] = 0.0f;
    m.m[3][1] = 0.0f;
    m.m[3][2] = 0.0f;
    m.m[3][3] = 1.0f;
    return m;
}

/* ------------------------------------------------------------------------- */
/* mesh.go                                                                    */
/* ------------------------------------------------------------------------- */

/* mesh.go:53-60 */
void rc_face_normals(const rc_vec4 *vertices, const int32_t *vi, int32_t num_faces, rc_vec4 *out) {
    for (int32_t i = 0; i < num_faces; i++) {
        rc_vec3 v0 = v4_xyz(vertices[vi[3 * i + 0]]);
        rc_vec3 v1 = v4_xyz(vertices[vi[3 * i + 1]]);
        rc_vec3 v2 = v4_xyz(vertices[vi[3 * i + 2]]);
        rc_vec3 n = rc_vec3_normalize(v3_cross(v3_sub(v1, v0), v3_sub(v2, v0)));
        out[i].x = n.x;
        out[i].y = n.y;
        out[i].z = n.z;
        out[i].w = 1.0f;
    }
}

/* mesh.go:28-51 */
void rc_bounding_box(const rc_vec4 *v, int32_t n, rc_vec4 out[8]) {
    float minx = v[0].x, miny = v[0].y, minz = v[0].z;
    float maxx = minx, maxy = miny, maxz = minz;
    for (int32_t i = 0; i < n; i++) {
        minx = rc_go_min(minx, v[i].x);
        miny = rc_go_min(miny, v[i].y);
        minz = rc_go_min(minz, v[i].z);
        maxx = rc_go_max(maxx, v[i].x);
        maxy = rc_go_max(maxy, v[i].y);
        maxz = rc_go_max(maxz, v[i].z);
    }
    const float xs[2] = {minx, maxx}, ys[2] = {miny, maxy}, zs[2] = {minz, maxz};
    int k = 0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                out[k].x = xs[a];
                out[k].y = ys[b];
                out[k].z = zs[c];
                out[k].w = 1.0f;
                k++;
            }
}

/* ------------------------------------------------------------------------- */
/* clipping.go                                                                */
/* ------------------------------------------------------------------------- */

#define RC_MAX_CLIP 9

typedef struct {
    rc_vec4 point, normal;
} rc_plane;

/* clipping.go:93-129, in the order Left, Right, Top, Bottom, Near, Far. */
static void make_planes(float znear, float zfar, rc_plane p[6]) {
    const rc_plane planes[6] = {
        {{-1, 0, 0, 1}, {1, 0, 0, 1}},
        {{1, 0, 0, 1}, {-1, 0, 0, 1}},
        {{0, -1, 0, 1}, {0, 1, 0, 1}},
        {{0, 1, 0, 1}, {0, -1, 0, 1}},
        {{0, 0, znear, 1}, {0, 0, -1, 0}},
        {{0, 0, zfar, 1}, {0, 0, 1, 0}},
    };
    memcpy(p, planes, sizeof planes);
}

/* clipping.go:69-71 */
static float plane_distance(const rc_plane *p, rc_vec4 v) {
    return v4_dot(p->normal, v) - v4_dot(p->normal, p->point);
}

/* clipping.go:74-76 */
static int plane_inside(const rc_plane *p, rc_vec4 q) {
    return v4_dot(v4_sub(q, p->point), p->normal) <= 0;
}

/* clipping.go:79-86 */
static rc_vec4 plane_intersect(const rc_plane *p, rc_vec4 q0, rc_vec4 q1, float *factor) {
    rc_vec4 u = v4_sub(q1, q0);
    rc_vec4 w = v4_sub(q0, p->point);
    float d = v4_dot(p->normal, u);
    float n = -v4_dot(p->normal, w);
    float f = n / d;
    *factor = f;
    return v4_add(q0, v4_mul(u, f));
}

/* clipping.go:131-154 */
int rc_box_visibility(float znear, float zfar, const rc_vec4 bbox[8]) {
    rc_plane planes[6];
    make_planes(znear, zfar, planes);
    for (int i = 0; i < 6; i++) {
        int outside = 0;
        for (int k = 0; k < 8; k++)
            if (plane_distance(&planes[i], bbox[k]) > 0) outside++;
        if (outside == 8) return RC_BOX_OUTSIDE;
        if (outside > 0) return RC_BOX_INTERSECT;
    }
    return RC_BOX_INSIDE;
}

typedef struct {
    float intensity[RC_MAX_CLIP];
    rc_vec4 points[RC_MAX_CLIP];
    rc_uv uvs[RC_MAX_CLIP];
    int count;
} rc_polygon;

/* clipping.go:35-40; returns 0 when the polygon is full (reference panics). */
static int poly_add(rc_polygon *p, rc_vec4 v, rc_uv uv, float intensity) {
    if (p->count >= RC_MAX_CLIP) return 0;
    p->intensity[p->count] = intensity;
    p->points[p->count] = v;
    p->uvs[p->count] = uv;
    p->count++;
    return 1;
}

/* clipping.go:156-165 */
static float lerp32(float a, float b, float f) { return a + (b - a) * f; }
static rc_uv lerp_uv(rc_uv a, rc_uv b, float f) {
    rc_uv r = {a.u + (b.u - a.u) * f, a.v + (b.v - a.v) * f};
    return r;
}

/* clipping.go:167-236 */
int rc_clip_triangle(float znear, float zfar, const rc_vec4 in_pts[3], const rc_uv in_uvs[3],
                     const float in_int[3], rc_vec4 out_pts[][3], rc_uv out_uvs[][3],
                     float out_int[][3]) {
    rc_plane planes[6];
    make_planes(znear, zfar, planes);
    rc_polygon pa, pb;
    rc_polygon *poly = &pa, *poly2 = &pb;
    poly->count = 0;
    for (int k = 0; k < 3; k++) poly_add(poly, in_pts[k], in_uvs[k], in_int[k]);

    for (int pi = 0; pi < 6; pi++) {
        const rc_plane *plane = &planes[pi];
        poly2->count = 0;
        for (int b = 0; b < poly->count; b++) {
            int a = (b + 1) % poly->count;
            rc_uv uva = poly->uvs[a], uvb = poly->uvs[b];
            rc_vec4 va = poly->points[a], vb = poly->points[b];
            float ia = poly->intensity[a], ib = poly->intensity[b];
            if (plane_inside(plane, va)) {
                if (!plane_inside(plane, vb)) {
                    float f;
                    rc_vec4 x = plane_intersect(plane, va, vb, &f);
                    if (!poly_add(poly2, x, lerp_uv(uva, uvb, f), lerp32(ia, ib, f))) return -1;
                }
                if (!poly_add(poly2, va, uva, ia)) return -1;
            } else if (plane_inside(plane, vb)) {
                float f;
                rc_vec4 x = plane_intersect(plane, va, vb, &f);
                if (!poly_add(poly2, x, lerp_uv(uva, uvb, f), lerp32(ia, ib, f))) return -1;
            }
        }
        if (poly2->count == 0) return 0;
        rc_polygon *t = poly;
        poly = poly2;
        poly2 = t;
    }

    /* clipping.go:42-62 fan triangulation from vertex 0 */
    if (poly->count < 3) return 0;
    int n = 0;
    for (int i = 0; i < poly->count - 2; i++) {
        out_int[n][0] = poly->intensity[0];
        out_int[n][1] = poly->intensity[i + 1];
        out_int[n][2] = poly->intensity[i + 2];
        out_pts[n][0] = poly->points[0];
        out_pts[n][1] = poly->points[i + 1];
        out_pts[n][2] = poly->points[i + 2];
        out_uvs[n][0] = poly->uvs[0];
        out_uvs[n][1] = poly->uvs[i + 1];
        out_uvs[n][2] = poly->uvs[i + 2];
        n++;
    }
    return n;
}

/* renderer.go:246-250: backface test in clip space, before the divide. */
int rc_facing_camera(const rc_vec4 pts[3]) {
    rc_vec3 v0 = v4_xyz(pts[0]), v1 = v4_xyz(pts[1]), v2 = v4_xyz(pts[2]);
    rc_vec3 n = v3_cross(v3_sub(v1, v0), v3_sub(v2, v0));
    rc_vec3 zero = {0.0f, 0.0f, 0.0f};
    return v3_dot(n, v3_sub(zero, v0)) > 0;
}

/* ------------------------------------------------------------------------- */
/* texture.go / rasterizer.go                                                 */
/* ------------------------------------------------------------------------- */

/* texture.go:69-88 */
void rc_sample(const rc_texture *t, float u, float v, uint8_t out[4]) {
    switch (t->type) {
    case RC_TEX_SOLID:
        memcpy(out, t->color, 4);
        return;
    case RC_TEX_IMAGE_FAST: {
        int64_t x = rc_go_int((1 - u) * t->scale * t->width_f) & (int64_t)(t->width - 1);
        int64_t y = rc_go_int(v * t->scale * t->height_f) & (int64_t)(t->height - 1);
        memcpy(out, t->pixels + 4 * (y * t->width + x), 4);
        return;
    }
    case RC_TEX_IMAGE: {
        int64_t x = rc_go_int((1 - u) * t->scale * t->width_f) % (int64_t)t->width;
        int64_t y = rc_go_int(v * t->scale * t->height_f) % (int64_t)t->height;
        int64_t idx = y * t->width + x;
        if (idx < 0) idx = 0;
        memcpy(out, t->pixels + 4 * idx, 4);
        return;
    }
    default:
        out[0] = 255;
        out[1] = 0;
        out[2] = 255;
        out[3] = 255;
        return;
    }
}

/* rasterizer.go:81-88 */
void rc_color_intensity(const uint8_t c[4], float intensity, uint8_t out[4]) {
    out[0] = rc_go_uint8((float)c[0] * intensity);
    out[1] = rc_go_uint8((float)c[1] * intensity);
    out[2] = rc_go_uint8((float)c[2] * intensity);
    out[3] = c[3];
}

/* rasterizer.go:120-125 */
static int64_t edge_adjust(int64_t f, int64_t dx, int64_t dy) {
    if (dy > 0 || (dy == 0 && dx > 0)) return f;
    return wsub(f, 1);
}

/* rasterizer.go:90-183 */
int64_t rc_triangle(int32_t width, int32_t height, uint8_t *rgba, float *depth, int64_t x0,
                    int64_t y0, float z0, float u0, float v0, int64_t x1, int64_t y1, float z1,
                    float u1, float v1, int64_t x2, int64_t y2, float z2, float u2, float v2,
                    int64_t tsx, int64_t tsy, int64_t tex, int64_t tey, float ia, float ib,
                    float ic, const rc_texture *texture, const uint8_t face_color[4],
                    int64_t *passed, uint8_t *written) {
    int64_t covered = 0, npass = 0;
    int64_t minx = i64_min(i64_min(x0, x1), x2), maxx = i64_max(i64_max(x0, x1), x2);
    int64_t miny = i64_min(i64_min(y0, y1), y2), maxy = i64_max(i64_max(y0, y1), y2);
    minx = i64_max(i64_max(minx, tsx), 0);
    maxx = i64_min(i64_min(maxx, tex), (int64_t)width - 1);
    miny = i64_max(i64_max(miny, tsy), 0);
    maxy = i64_min(i64_min(maxy, tey), (int64_t)height - 1);

    int64_t f01 = wadd(wadd(wmul(wsub(y0, y1), minx), wmul(wsub(x1, x0), miny)),
                       wsub(wmul(x0, y1), wmul(x1, y0)));
    int64_t f12 = wadd(wadd(wmul(wsub(y1, y2), minx), wmul(wsub(x2, x1), miny)),
                       wsub(wmul(x1, y2), wmul(x2, y1)));
    int64_t f20 = wadd(wadd(wmul(wsub(y2, y0), minx), wmul(wsub(x0, x2), miny)),
                       wsub(wmul(x2, y0), wmul(x0, y2)));
    int64_t f01dx = wsub(y0, y1), f01dy = wsub(x1, x0);
    int64_t f12dx = wsub(y1, y2), f12dy = wsub(x2, x1);
    int64_t f20dx = wsub(y2, y0), f20dy = wsub(x0, x2);
    f01 = edge_adjust(f01, f01dx, f01dy);
    f12 = edge_adjust(f12, f12dx, f12dy);
    f20 = edge_adjust(f20, f20dx, f20dy);

    float v0z0 = v0 / z0, u0z0 = u0 / z0;
    float u1z1 = u1 / z1, v1z1 = v1 / z1;
    float u2z2 = u2 / z2, v2z2 = v2 / z2;

    for (int64_t y = miny; y <= maxy; y++) {
        int64_t fx01 = f01, fx12 = f12, fx20 = f20;
        for (int64_t x = minx; x <= maxx; x++) {
            if (fx01 < 0 && fx12 < 0 && fx20 < 0) {
                covered++;
                int64_t sum = wadd(wadd(fx12, fx20), fx01);
                float alpha = (float)fx12 / (float)sum;
                float beta = (float)fx20 / (float)sum;
                float gamma = 1 - alpha - beta;
                float zrec = -(alpha / z0 + beta / z1 + gamma / z2);
                int64_t index = y * width + x;
                if (zrec >= depth[index]) {
                    npass++;
                    float u = (alpha * u0z0 + beta * u1z1 + gamma * u2z2) / zrec;
                    float v = (alpha * v0z0 + beta * v1z1 + gamma * v2z2) / zrec;
                    float intensity = alpha * ia + beta * ib + gamma * ic;
                    uint8_t c[4];
                    if (texture)
                        rc_sample(texture, u, v, c);
                    else
                        memcpy(c, face_color, 4);
                    depth[index] = zrec;
                    rc_color_intensity(c, intensity, rgba + 4 * index);
                    if (written) written[index] = 1;
                }
            }
            fx01 = wadd(fx01, f01dx);
            fx12 = wadd(fx12, f12dx);
            fx20 = wadd(fx20, f20dx);
        }
        f01 = wadd(f01, f01dy);
        f12 = wadd(f12, f12dy);
        f20 = wadd(f20, f20dy);
    }
    if (passed) *passed = npass;
    return covered;
}

/* rasterizer.go:36-44: colour fill and z = -1. */
void rc_clear(int32_t width, int32_t height, uint8_t *rgba, float *depth, const uint8_t c[4]) {
    int64_t n = (int64_t)width * height;
    for (int64_t i = 0; i < n; i++) {
        memcpy(rgba + 4 * i, c, 4);
        depth[i] = -1.0f;
    }
}

/* rasterizer.go:46-52 with Pixel's idx>0 guard (rasterizer.go:25-30). A step
 * below 1 never terminates in the reference; here it draws nothing. */
void rc_dot_grid(int32_t width, int32_t height, uint8_t *rgba, const uint8_t c[4], int32_t step) {
    if (step < 1) return;
    for (int64_t y = step; y < height; y += step)
        for (int64_t x = step; x < width; x += step) {
            int64_t idx = y * width + x;
            if (idx > 0 && idx < (int64_t)width * height) memcpy(rgba + 4 * idx, c, 4);
        }
}

/* rasterizer.go:25-30: only the flat index is checked, so x outside the row
 * wraps into a neighbouring row, and index 0 is never written. */
void rc_pixel(int32_t width, int32_t height, uint8_t *rgba, int64_t x, int64_t y,
              const uint8_t c[4]) {
    int64_t idx = wadd(wmul(y, width), x);
    if (idx > 0 && idx < (int64_t)width * height) memcpy(rgba + 4 * idx, c, 4);
}

/* utils.go:11-21: minInt maps to maxInt. */
int64_t rc_go_abs(int64_t v) {
    if (v < 0) v = wsub(0, v);
    if (v < 0) v = wsub(v, 1);
    return v;
}

/* rasterizer.go:54-63 */
void rc_rect(int32_t width, int32_t height, uint8_t *rgba, int64_t x, int64_t y, int64_t w,
             int64_t h, const uint8_t c[4]) {
    if (x >= width || y >= height) return;
    for (int64_t py = y; py < wadd(y, h); py = wadd(py, 1))
        for (int64_t px = x; px < wadd(x, w); px = wadd(px, 1)) rc_pixel(width, height, rgba, px, py, c);
}

/* rasterizer.go:65-79: a float32 DDA accumulated step by step. */
void rc_line(int32_t width, int32_t height, uint8_t *rgba, int64_t x0, int64_t y0, int64_t x1,
             int64_t y1, const uint8_t c[4]) {
    int64_t dx = wsub(x1, x0), dy = wsub(y1, y0);
    int64_t ax = rc_go_abs(dx), ay = rc_go_abs(dy);
    int64_t side = ax > ay ? ax : ay;
    float xs = (float)dx / (float)side, ys = (float)dy / (float)side;
    float cx = (float)x0, cy = (float)y0;
    for (int64_t i = 0; i <= side; i++) {
        rc_pixel(width, height, rgba, rc_go_int(cx), rc_go_int(cy), c);
        cx += xs;
        cy += ys;
    }
}

/* renderer.go:50-76 */
void rc_tile_bounds(int32_t tile, int32_t num_tiles, int32_t width, int32_t height, rc_vec2 *start,
                    rc_vec2 *end) {
    if (num_tiles == 1) {
        start->x = 0;
        start->y = 0;
        end->x = (float)width;
        end->y = (float)height;
        return;
    }
    uint64_t ntx = (uint64_t)sqrt((double)num_tiles);
    uint64_t nty = ((uint64_t)num_tiles + ntx - 1) / ntx;
    uint64_t tw = ((uint64_t)width + ntx - 1) / ntx;
    uint64_t th = ((uint64_t)height + nty - 1) / nty;
    start->x = (float)(((uint64_t)tile % ntx) * tw);
    start->y = (float)(((uint64_t)tile / ntx) * th);
    end->x = start->x + (float)tw;
    end->y = start->y + (float)th;
    if (end->x > (float)width) end->x = (float)width;
    if (end->y > (float)height) end->y = (float)height;
}

/* renderer.go:422-434 (drawTilesBoundaries) and rasterizer.go:209-217
 * (CrossHair), in the order of renderer.go:476-480. int(float32) of the tile
 * bounds, W/2 and H/2 as Go's integer division. */
void rc_frame_decorations(const rc_frame *fr, uint8_t *rgba) {
    const int32_t W = fr->width, H = fr->height;
    if (fr->flags & RC_SHOW_TILE_BOUNDS) {
        const uint8_t red[4] = {255, 0, 0, 255};
        for (int32_t i = 0; i < fr->ref_tiles; i++) {
            rc_vec2 s, e;
            rc_tile_bounds(i, fr->ref_tiles, W, H, &s, &e);
            int64_t sx = rc_go_int(s.x), sy = rc_go_int(s.y);
            int64_t ex = rc_go_int(e.x), ey = rc_go_int(e.y);
            rc_line(W, H, rgba, sx, sy, ex, sy, red);
            rc_line(W, H, rgba, ex, sy, ex, ey, red);
            rc_line(W, H, rgba, ex, ey, sx, ey, red);
            rc_line(W, H, rgba, sx, ey, sx, sy, red);
        }
    }
    if (fr->flags & RC_SHOW_CROSSHAIR) {
        const uint8_t yellow[4] = {255, 255, 0, 255};
        const int64_t size = 5, offset = 3, x = W / 2, y = H / 2;
        rc_line(W, H, rgba, x - size, y, x - offset, y, yellow);
        rc_line(W, H, rgba, x + offset, y, x + size, y, yellow);
        rc_line(W, H, rgba, x, y - size, x, y - offset, yellow);
        rc_line(W, H, rgba, x, y + offset, x, y + size, yellow);
    }
}

/* rasterizer.go:185-191: uint8(float32(a)*(1-f) + float32(b)*f) per channel. */
static uint8_t blend8(uint8_t a, uint8_t b, float f) {
    float t = (float)a * (1 - f);
    float u = (float)b * f;
    return rc_go_uint8(t + u);
}

/* rasterizer.go:193-207 */
void rc_fog(int32_t width, int32_t height, uint8_t *rgba, const float *depth, float fog_start,
            float fog_end, const uint8_t c[4]) {
    for (int64_t i = 0; i < (int64_t)width * height; i++) {
        float d = depth[i];
        uint8_t *p = rgba + 4 * i;
        if (d >= fog_start) {
            /* noop */
        } else if (d <= fog_end) {
            memcpy(p, c, 4);
        } else {
            float f = 1 - ((fog_end - d) / (fog_end - fog_start));
            for (int k = 0; k < 4; k++) p[k] = blend8(p[k], c[k], f);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* renderer.go: geometry stage and the frame                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
    rc_vec4 points[3];
    rc_uv uvs[3];
    float intensity[3];
    int32_t texture;
} rc_tri;

typedef struct {
    rc_tri *v;
    int64_t n, cap;
} rc_trilist;

static int tri_push(rc_trilist *l, const rc_tri *t) {
    if (l->n == l->cap) {
        int64_t nc = l->cap ? 2 * l->cap : 1024;
        rc_tri *nv = (rc_tri *)realloc(l->v, (size_t)nc * sizeof(rc_tri));
        if (!nv) return 0;
        l->v = nv;
        l->cap = nc;
    }
    l->v[l->n++] = *t;
    return 1;
}

/* renderer.go:226-244 with a single tile rect: float bbox (Go min/max) overlap. */
static int tri_overlaps(const rc_vec4 p[3], rc_vec2 start, rc_vec2 end) {
    float minx = rc_go_min(rc_go_min(p[0].x, p[1].x), p[2].x);
    float maxx = rc_go_max(rc_go_max(p[0].x, p[1].x), p[2].x);
    float miny = rc_go_min(rc_go_min(p[0].y, p[1].y), p[2].y);
    float maxy = rc_go_max(rc_go_max(p[0].y, p[1].y), p[2].y);
    return maxx >= start.x && minx <= end.x && maxy >= start.y && miny <= end.y;
}

/* renderer.go:326-335 (one corner: ambientStrength + vn.DotProduct(light) *
 * diffuseStrength) and renderer.go:337-341 (flat: the same two operations on
 * the face normal): the world-space normal is normalised in 4-D with its W
 * (vector.go:115-125), then its X, Y, Z are dotted with the light. */
float rc_light_intensity(rc_vec4 world_normal, rc_vec3 light, float ambient, float diffuse) {
    rc_vec3 n = v4_xyz(rc_vec4_normalize(world_normal));
    float d = v3_dot(n, light) * diffuse;
    return ambient + d;
}

/* renderer.go:361-370: all four components divided by W (vector.go:103-105),
 * the screen matrix applied with all 16 products (matrix.go:167-173), then W
 * restored to the clip-space W. */
rc_vec4 rc_project_point(rc_vec4 clip, const rc_matrix *screen) {
    float orig_w = clip.w;
    rc_vec4 p = v4_div(clip, clip.w);
    rc_transform_inplace(screen, &p);
    p.w = orig_w;
    return p;
}

/* renderer.go:226-244 over the num_tiles reference tiles of
 * calculateTileBoundaries (renderer.go:50-76): the number of tiles whose
 * inclusive float rectangle the triangle's float bounding box overlaps, i.e.
 * what one triangle adds to TPF (renderer.go:436-441). */
int64_t rc_tile_refs(const rc_vec4 points[3], int32_t num_tiles, int32_t width, int32_t height) {
    int64_t n = 0;
    for (int32_t t = 0; t < num_tiles; t++) {
        rc_vec2 s, e;
        rc_tile_bounds(t, num_tiles, width, height, &s, &e);
        if (tri_overlaps(points, s, e)) n++;
    }
    return n;
}

typedef struct {
    const rc_frame *frame;
    const rc_mesh *mesh;
    const rc_object *obj;
    rc_trilist *out;  /* triangles that pass the full-screen bin test */
    rc_trilist *bins; /* optional: per-tile lists of a disjoint tiles_x*tiles_y grid */
    int32_t tiles_x, tiles_y;
    int64_t tpf;      /* refs on the reference tile grid */
    int64_t front;
    int64_t binned;
    int64_t overflow;
    int error;
} rc_project_job;

/* Disjoint-tile binning by the clamped integer bounding box (the pixels
 * fb.Triangle can touch), used by the threaded baseline only. */
static int bin_triangle(rc_project_job *job, const rc_tri *t) {
    const rc_frame *fr = job->frame;
    int64_t xs[3], ys[3];
    for (int k = 0; k < 3; k++) {
        xs[k] = rc_go_int(t->points[k].x);
        ys[k] = rc_go_int(t->points[k].y);
    }
    int64_t mnx = i64_max(i64_min(i64_min(xs[0], xs[1]), xs[2]), 0);
    int64_t mxx = i64_min(i64_max(i64_max(xs[0], xs[1]), xs[2]), (int64_t)fr->width - 1);
    int64_t mny = i64_max(i64_min(i64_min(ys[0], ys[1]), ys[2]), 0);
    int64_t mxy = i64_min(i64_max(i64_max(ys[0], ys[1]), ys[2]), (int64_t)fr->height - 1);
    if (mnx > mxx || mny > mxy) return 1;
    int64_t tw = (fr->width + job->tiles_x - 1) / job->tiles_x;
    int64_t th = (fr->height + job->tiles_y - 1) / job->tiles_y;
    for (int64_t ty = mny / th; ty <= mxy / th; ty++)
        for (int64_t tx = mnx / tw; tx <= mxx / tw; tx++)
            if (!tri_push(&job->bins[ty * job->tiles_x + tx], t)) return 0;
    return 1;
}

/* renderer.go:254-407 (sequential: one tile = whole framebuffer). */
static void project_object(rc_project_job *job) {
    const rc_frame *fr = job->frame;
    const rc_mesh *mesh = job->mesh;
    const rc_object *obj = job->obj;
    uint32_t flags = fr->flags;

    rc_vec4 bbox[8];
    memcpy(bbox, mesh->bbox, sizeof bbox);
    rc_transform_batch(&obj->mvp, bbox, 8);
    int vis = rc_box_visibility(fr->znear, fr->zfar, bbox);
    if (vis == RC_BOX_OUTSIDE) return;

    int32_t V = mesh->num_vertices, N = mesh->num_vertex_normals, F = mesh->num_faces;
    rc_vec4 *tv = (rc_vec4 *)malloc(sizeof(rc_vec4) * (size_t)(V > 0 ? V : 1));
    rc_vec4 *wfn = (rc_vec4 *)malloc(sizeof(rc_vec4) * (size_t)(F > 0 ? F : 1));
    rc_vec4 *wvn = (rc_vec4 *)malloc(sizeof(rc_vec4) * (size_t)(N > 0 ? N : 1));
    if (!tv || !wfn || !wvn) {
        job->error = 1;
        free(tv);
        free(wfn);
        free(wvn);
        return;
    }
    memcpy(tv, mesh->vertices, sizeof(rc_vec4) * (size_t)V);
    rc_transform_batch(&obj->mvp, tv, V);
    memcpy(wfn, mesh->face_normals, sizeof(rc_vec4) * (size_t)F);
    if (N) memcpy(wvn, mesh->vertex_normals, sizeof(rc_vec4) * (size_t)N);
    rc_transform_batch(&obj->world, wfn, F);
    rc_transform_batch(&obj->world, wvn, N);
    int has_vn = N != 0;

    int32_t nt = fr->ref_tiles == 16 ? 16 : 1;
    rc_vec2 full_s, full_e;
    rc_tile_bounds(0, 1, fr->width, fr->height, &full_s, &full_e);

    rc_vec4 cv[RC_MAX_CLIP][3];
    rc_uv cuv[RC_MAX_CLIP][3];
    float cint[RC_MAX_CLIP][3];
    for (int32_t fi = 0; fi < F; fi++) {
        const int32_t *vi = mesh->vertex_indices + 3 * fi;
        rc_vec4 verts[3] = {tv[vi[0]], tv[vi[1]], tv[vi[2]]};
        if ((flags & RC_BACKFACE_CULLING) && !rc_facing_camera(verts)) continue;
        job->front++;

        float vint[3];
        if (flags & RC_LIGHTING) {
            if (has_vn && !(flags & RC_FLAT_SHADING)) {
                const int32_t *ni = mesh->normal_indices + 3 * fi;
                for (int k = 0; k < 3; k++) {
                    if (ni[k] < 0 || ni[k] >= N) {
                        job->error = 2; /* reference: index out of range panic */
                        goto done;
                    }
                    vint[k] = rc_light_intensity(wvn[ni[k]], fr->light, fr->ambient, fr->diffuse);
                }
            } else {
                float intensity = rc_light_intensity(wfn[fi], fr->light, fr->ambient, fr->diffuse);
                vint[0] = vint[1] = vint[2] = intensity;
            }
        } else {
            vint[0] = vint[1] = vint[2] = fr->ambient;
        }

        const rc_uv *fuv = mesh->uvs + 3 * fi;
        int count;
        if ((flags & RC_FRUSTUM_CLIPPING) && vis != RC_BOX_INSIDE) {
            count = rc_clip_triangle(fr->znear, fr->zfar, verts, fuv, vint, cv, cuv, cint);
            if (count < 0) {
                job->overflow++;
                count = 0;
            }
        } else {
            for (int k = 0; k < 3; k++) {
                cv[0][k] = verts[k];
                cuv[0][k] = fuv[k];
                cint[0][k] = vint[k];
            }
            count = 1;
        }

        for (int i = 0; i < count; i++) {
            rc_tri t;
            for (int j = 0; j < 3; j++) {
                t.points[j] = rc_project_point(cv[i][j], &fr->screen);
                t.uvs[j] = cuv[i][j];
                t.intensity[j] = cint[i][j];
            }
            t.texture = mesh->face_textures ? mesh->face_textures[fi] : -1;
            job->tpf += rc_tile_refs(t.points, nt, fr->width, fr->height);
            if (tri_overlaps(t.points, full_s, full_e)) {
                job->binned++;
                if (job->bins) {
                    if (!bin_triangle(job, &t)) {
                        job->error = 1;
                        goto done;
                    }
                } else if (!tri_push(job->out, &t)) {
                    job->error = 1;
                    goto done;
                }
            }
        }
    }
done:
    free(tv);
    free(wfn);
    free(wvn);
}

/* renderer.go:18-19 */
static const uint8_t kEdgeColor[4] = {0, 0, 0, 255};
static const uint8_t kVertexColor[4] = {255, 161, 0, 255};
static const uint8_t kWhite[4] = {255, 255, 255, 255};

/* drawProjection (renderer.go:166-217) over a tile list in order: the face,
 * then the edge and vertex overlays. */
static void raster_list(const rc_frame *fr, const rc_texture *textures, const rc_trilist *l,
                        int64_t tsx, int64_t tsy, int64_t tex, int64_t tey, uint8_t *rgba,
                        float *depth, int64_t *covered, int64_t *passed, uint8_t *written) {
    int show_tex = (fr->flags & RC_SHOW_TEXTURES) != 0;
    int W = fr->width, H = fr->height;
    for (int64_t i = 0; i < l->n; i++) {
        const rc_tri *t = &l->v[i];
        const rc_vec4 *a = &t->points[0], *b = &t->points[1], *c = &t->points[2];
        const rc_texture *tx = (show_tex && t->texture >= 0) ? &textures[t->texture] : NULL;
        int64_t ax = rc_go_int(a->x), ay = rc_go_int(a->y), bx = rc_go_int(b->x),
                by = rc_go_int(b->y), cx = rc_go_int(c->x), cy = rc_go_int(c->y);
        if (fr->flags & RC_SHOW_FACES) {
            int64_t p = 0;
            *covered += rc_triangle(W, H, rgba, depth, ax, ay, a->w, t->uvs[0].u, t->uvs[0].v,
                                    bx, by, b->w, t->uvs[1].u, t->uvs[1].v, cx, cy, c->w,
                                    t->uvs[2].u, t->uvs[2].v, tsx, tsy, tex, tey,
                                    t->intensity[0], t->intensity[1], t->intensity[2], tx,
                                    fr->face_color, &p, written);
            *passed += p;
        }
        if (fr->flags & RC_SHOW_EDGES) {
            const uint8_t *col = (fr->flags & RC_SHOW_FACES) ? kEdgeColor : kWhite;
            rc_line(W, H, rgba, ax, ay, bx, by, col);
            rc_line(W, H, rgba, bx, by, cx, cy, col);
            rc_line(W, H, rgba, cx, cy, ax, ay, col);
            if (fr->flags & RC_SHOW_FACES) {
                float mx = (a->x + b->x + c->x) / 3.0f, my = (a->y + b->y + c->y) / 3.0f;
                rc_rect(W, H, rgba, wsub(rc_go_int(mx), 1), wsub(rc_go_int(my), 1), 3, 3, col);
            }
        }
        if (fr->flags & RC_SHOW_VERTICES) {
            rc_rect(W, H, rgba, wsub(ax, 1), wsub(ay, 1), 3, 3, kVertexColor);
            rc_rect(W, H, rgba, wsub(bx, 1), wsub(by, 1), 3, 3, kVertexColor);
            rc_rect(W, H, rgba, wsub(cx, 1), wsub(cy, 1), 3, 3, kVertexColor);
        }
    }
}

static int validate(const rc_frame *fr, const rc_mesh *meshes, int32_t nm, int32_t ntex,
                    const rc_object *objs, int32_t no) {
    if (fr->width <= 0 || fr->height <= 0) return 3;
    for (int32_t i = 0; i < no; i++) {
        if (objs[i].mesh < 0 || objs[i].mesh >= nm) return 3;
        const rc_mesh *m = &meshes[objs[i].mesh];
        if (m->num_vertices <= 0) return 3;
        for (int64_t k = 0; k < 3 * (int64_t)m->num_faces; k++)
            if (m->vertex_indices[k] < 0 || m->vertex_indices[k] >= m->num_vertices) return 2;
        if (m->face_textures)
            for (int32_t k = 0; k < m->num_faces; k++)
                if (m->face_textures[k] < -1 || m->face_textures[k] >= ntex) return 3;
    }
    return 0;
}

/* renderer.go:443-483, sequential branch (renderer.go:466-474). Returns 0 on
 * success, 1 out of memory, 2 index out of range, 3 invalid argument. */
int rc_draw(const rc_frame *fr, const rc_mesh *meshes, int32_t nm, const rc_texture *textures,
            int32_t ntex, const rc_object *objs, int32_t no, uint8_t *rgba, float *depth,
            rc_stats *st) {
    int err = validate(fr, meshes, nm, ntex, objs, no);
    if (err) return err;
    rc_stats s;
    memset(&s, 0, sizeof s);
    rc_clear(fr->width, fr->height, rgba, depth, fr->clear_color);
    rc_dot_grid(fr->width, fr->height, rgba, fr->dot_color, fr->dot_step);

    rc_trilist list = {0, 0, 0};
    for (int32_t i = 0; i < no; i++) {
        rc_project_job job;
        memset(&job, 0, sizeof job);
        job.frame = fr;
        job.mesh = &meshes[objs[i].mesh];
        job.obj = &objs[i];
        job.out = &list;
        project_object(&job);
        s.tpf += job.tpf;
        s.faces_front += job.front;
        s.clip_overflow += job.overflow;
        if (job.error) {
            free(list.v);
            return job.error;
        }
    }
    s.triangles = list.n;
    if (fr->flags & (RC_SHOW_FACES | RC_SHOW_EDGES | RC_SHOW_VERTICES)) {
        uint8_t *written = (uint8_t *)calloc((size_t)fr->width * fr->height, 1);
        if (!written) {
            free(list.v);
            return 1;
        }
        raster_list(fr, textures, &list, 0, 0, fr->width, fr->height, rgba, depth,
                    &s.fragments_covered, &s.fragments_passed, written);
        for (int64_t i = 0; i < (int64_t)fr->width * fr->height; i++) s.pixels_written += written[i];
        free(written);
    }
    free(list.v);
    rc_frame_decorations(fr, rgba);
    if (st) *st = s;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Threaded CPU baseline                                                      */
/* ------------------------------------------------------------------------- */

typedef struct {
    const rc_frame *frame;
    const rc_texture *textures;
    rc_project_job *jobs; /* per object: bins[tile] in draw order */
    int32_t num_objects;
    int32_t tiles_x, tiles_y;
    int32_t next_tile; /* claimed with __atomic */
    uint8_t *rgba;
    float *depth;
    int64_t covered, passed;
    pthread_mutex_t mu;
} rc_raster_shared;

static void *raster_worker(void *arg) {
    rc_raster_shared *sh = (rc_raster_shared *)arg;
    const rc_frame *fr = sh->frame;
    int64_t tw = (fr->width + sh->tiles_x - 1) / sh->tiles_x;
    int64_t th = (fr->height + sh->tiles_y - 1) / sh->tiles_y;
    int64_t cov = 0, pas = 0;
    for (;;) {
        int32_t tile = __atomic_fetch_add(&sh->next_tile, 1, __ATOMIC_RELAXED);
        if (tile >= sh->tiles_x * sh->tiles_y) break;
        int64_t x0 = (tile % sh->tiles_x) * tw, y0 = (tile / sh->tiles_x) * th;
        int64_t x1 = x0 + tw - 1, y1 = y0 + th - 1; /* disjoint, inclusive */
        for (int32_t i = 0; i < sh->num_objects; i++)
            raster_list(fr, sh->textures, &sh->jobs[i].bins[tile], x0, y0, x1, y1, sh->rgba,
                        sh->depth, &cov, &pas, NULL);
    }
    pthread_mutex_lock(&sh->mu);
    sh->covered += cov;
    sh->passed += pas;
    pthread_mutex_unlock(&sh->mu);
    return NULL;
}

static void *project_worker(void *arg) {
    project_object((rc_project_job *)arg);
    return NULL;
}

/* Same image as rc_draw. Geometry is object-parallel (renderer.go:451-459),
 * binning happens inside the geometry stage (renderer.go:379-394), raster is
 * tile-parallel over disjoint tiles (renderer.go:461-465). */
int rc_draw_parallel(const rc_frame *fr, const rc_mesh *meshes, int32_t nm,
                     const rc_texture *textures, int32_t ntex, const rc_object *objs,
                     int32_t no, uint8_t *rgba, float *depth, rc_stats *st, int32_t threads,
                     int32_t tiles_x, int32_t tiles_y) {
    int err = validate(fr, meshes, nm, ntex, objs, no);
    if (err) return err;
    /* Overlays write outside their tile (Pixel wraps rows): sequential form only. */
    if (fr->flags & (RC_SHOW_EDGES | RC_SHOW_VERTICES)) return 3;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if (tiles_x < 1) tiles_x = 1;
    if (tiles_y < 1) tiles_y = 1;
    int32_t ntiles = tiles_x * tiles_y;
    rc_stats s;
    memset(&s, 0, sizeof s);
    rc_clear(fr->width, fr->height, rgba, depth, fr->clear_color);
    rc_dot_grid(fr->width, fr->height, rgba, fr->dot_color, fr->dot_step);

    rc_project_job *jobs = (rc_project_job *)calloc((size_t)(no > 0 ? no : 1), sizeof *jobs);
    pthread_t tid[256];
    int rc = 0;
    for (int32_t i = 0; i < no; i++) {
        jobs[i].frame = fr;
        jobs[i].mesh = &meshes[objs[i].mesh];
        jobs[i].obj = &objs[i];
        jobs[i].tiles_x = tiles_x;
        jobs[i].tiles_y = tiles_y;
        jobs[i].bins = (rc_trilist *)calloc((size_t)ntiles, sizeof(rc_trilist));
        if (!jobs[i].bins) rc = 1;
    }
    for (int32_t base = 0; base < no && !rc; base += threads) {
        int32_t cnt = no - base < threads ? no - base : threads;
        for (int32_t k = 0; k < cnt; k++)
            pthread_create(&tid[k], NULL, project_worker, &jobs[base + k]);
        for (int32_t k = 0; k < cnt; k++) pthread_join(tid[k], NULL);
    }
    for (int32_t i = 0; i < no; i++) {
        s.tpf += jobs[i].tpf;
        s.faces_front += jobs[i].front;
        s.triangles += jobs[i].binned;
        s.clip_overflow += jobs[i].overflow;
        if (jobs[i].error) rc = jobs[i].error;
    }
    if (!rc && (fr->flags & RC_SHOW_FACES)) {
        rc_raster_shared sh;
        memset(&sh, 0, sizeof sh);
        sh.frame = fr;
        sh.textures = textures;
        sh.jobs = jobs;
        sh.num_objects = no;
        sh.tiles_x = tiles_x;
        sh.tiles_y = tiles_y;
        sh.rgba = rgba;
        sh.depth = depth;
        pthread_mutex_init(&sh.mu, NULL);
        for (int32_t k = 0; k < threads; k++) pthread_create(&tid[k], NULL, raster_worker, &sh);
        for (int32_t k = 0; k < threads; k++) pthread_join(tid[k], NULL);
        pthread_mutex_destroy(&sh.mu);
        s.fragments_covered = sh.covered;
        s.fragments_passed = sh.passed;
    }
    for (int32_t i = 0; i < no; i++) {
        if (jobs[i].bins) {
            for (int32_t t = 0; t < ntiles; t++) free(jobs[i].bins[t].v);
            free(jobs[i].bins);
        }
    }
    free(jobs);
    if (rc) return rc;
    rc_frame_decorations(fr, rgba);
    if (st) *st = s;
    return 0;
}
