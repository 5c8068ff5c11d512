/*
 * refcpu — CPU restatement of gorender's render path, used ONLY as a test oracle
 * and as the CPU baseline timed beside the GPU. Nothing in the product library
 * links, loads or calls this code.
 *
 * It follows the reference's sequential draw order (renderer.go:466-474):
 * objects in order, faces in order, clipped fan triangles in order, one tile
 * covering the whole framebuffer. Every float32 operation is written in the
 * reference's evaluation order and compiled with -ffp-contract=off on x86-64
 * SSE (no FMA, round-to-nearest-even, denormals kept), matching Go gc amd64
 * with GOAMD64=v1.
 *
 * Pinning status of each function is listed in the header comment of
 * refcpu.c and checked by tests/test_oracle_*.py.
 */
#ifndef GORENDER_REFCPU_H
#define GORENDER_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y; } rc_vec2;
typedef struct { float x, y, z; } rc_vec3;
typedef struct { float x, y, z, w; } rc_vec4;
typedef struct { float u, v; } rc_uv;
typedef struct { float m[4][4]; } rc_matrix;

/* texture.go:13-17 */
enum { RC_TEX_SOLID = 0, RC_TEX_IMAGE = 1, RC_TEX_IMAGE_FAST = 2 };

/* texture.go:19-26 */
typedef struct {
    int32_t type;
    int32_t width, height;
    float width_f, height_f;
    float scale;
    uint8_t color[4];
    const uint8_t *pixels; /* RGBA8 row-major, width*height*4 bytes */
} rc_texture;

/* mesh.go:19-26 plus the per-face data of mesh.go:12-17 */
typedef struct {
    int32_t num_vertices;
    const rc_vec4 *vertices;
    int32_t num_vertex_normals;
    const rc_vec4 *vertex_normals;
    int32_t num_faces;
    const int32_t *vertex_indices; /* [F][3] */
    const int32_t *normal_indices; /* [F][3] */
    const rc_uv *uvs;              /* [F][3] */
    const int32_t *face_textures;  /* [F], index into the texture array or -1 for nil */
    const rc_vec4 *face_normals;   /* [F] */
    rc_vec4 bbox[8];
} rc_mesh;

typedef struct {
    int32_t mesh;
    rc_matrix mvp;
    rc_matrix world;
} rc_object;

/* Renderer toggles, renderer.go:90-97 */
enum {
    RC_SHOW_FACES = 1u << 0,
    RC_BACKFACE_CULLING = 1u << 1,
    RC_LIGHTING = 1u << 2,
    RC_FLAT_SHADING = 1u << 3,
    RC_FRUSTUM_CLIPPING = 1u << 4,
    RC_SHOW_TEXTURES = 1u << 5,
    RC_SHOW_EDGES = 1u << 6,
    RC_SHOW_VERTICES = 1u << 7,
    RC_SHOW_TILE_BOUNDS = 1u << 8, /* drawTilesBoundaries (renderer.go:422-434) */
    RC_SHOW_CROSSHAIR = 1u << 9,   /* FrameBuffer.CrossHair (rasterizer.go:209-217) */
};

typedef struct {
    int32_t width, height;
    uint32_t flags;
    int32_t ref_tiles;   /* 1 (sequential) or 16 (parallel): only changes TPF */
    rc_matrix screen;    /* NewScreenMatrix, matrix.go:108-118 */
    rc_vec3 light;       /* already normalised light direction, renderer.go:265 */
    float ambient;       /* renderer.go:13 */
    float diffuse;       /* renderer.go:12 */
    float znear, zfar;   /* frustum planes, clipping.go:93-129 */
    uint8_t face_color[4];
    uint8_t clear_color[4];
    uint8_t dot_color[4];
    int32_t dot_step;
} rc_frame;

typedef struct {
    int64_t tpf;               /* renderer.go:436-441 with ref_tiles tiles */
    int64_t triangles;         /* post-clip triangles that pass the full-screen bin test */
    int64_t fragments_covered; /* pixels inside a triangle (rasterizer.go:147) */
    int64_t fragments_passed;  /* pixels passing the depth test (rasterizer.go:156) */
    int64_t faces_front;       /* faces surviving the backface test */
    int64_t clip_overflow;     /* faces whose clip polygon exceeded 9 vertices */
    int64_t pixels_written;    /* pixels written by at least one fragment (sequential only) */
} rc_stats;

/* ---- Go conversion semantics (Appendix-A helpers) ---- */
int64_t rc_go_int(float x);
uint8_t rc_go_uint8(float x);
float rc_go_min(float a, float b);
float rc_go_max(float a, float b);

/* ---- math (vector.go, matrix.go, asm_purego.go) ---- */
float rc_sqrt32(float x);
rc_vec3 rc_vec3_normalize(rc_vec3 v);
rc_vec4 rc_vec4_normalize(rc_vec4 v);
rc_matrix rc_matrix_multiply(const rc_matrix *a, const rc_matrix *b);
void rc_transform_batch(const rc_matrix *m, rc_vec4 *vecs, int64_t n);
void rc_transform_batch_pairwise(const rc_matrix *m, rc_vec4 *vecs, int64_t n);
void rc_transform_inplace(const rc_matrix *m, rc_vec4 *v);
rc_matrix rc_identity_matrix(void);
rc_matrix rc_scale_matrix(float x, float y, float z);
rc_matrix rc_translation_matrix(float x, float y, float z);
rc_matrix rc_rotation_x_matrix(float a);
rc_matrix rc_rotation_y_matrix(float a);
rc_matrix rc_rotation_z_matrix(float a);
rc_matrix rc_rotation_matrix(float x, float y, float z);
rc_matrix rc_world_matrix(rc_vec3 scale, rc_vec3 rotation, rc_vec3 translation);
rc_matrix rc_perspective_matrix(float fov, float aspect, float znear, float zfar);
rc_matrix rc_screen_matrix(int32_t width, int32_t height);
rc_matrix rc_view_matrix(rc_vec3 eye, rc_vec3 direction, rc_vec3 up);

/* ---- mesh set-up (mesh.go:28-69) ---- */
void rc_face_normals(const rc_vec4 *vertices, const int32_t *vertex_indices, int32_t num_faces,
                     rc_vec4 *out);
void rc_bounding_box(const rc_vec4 *vertices, int32_t n, rc_vec4 out[8]);

/* ---- clipping (clipping.go) ---- */
enum { RC_BOX_OUTSIDE = 0, RC_BOX_INTERSECT = 1, RC_BOX_INSIDE = 2 };
int rc_box_visibility(float znear, float zfar, const rc_vec4 bbox[8]);
/* Returns the number of output triangles (0..7), or -1 when the polygon would
 * exceed 9 vertices (the reference indexes out of range and panics). */
int rc_clip_triangle(float znear, float zfar, const rc_vec4 in_pts[3], const rc_uv in_uvs[3],
                     const float in_int[3], rc_vec4 out_pts[][3], rc_uv out_uvs[][3],
                     float out_int[][3]);
int rc_facing_camera(const rc_vec4 pts[3]);

/* ---- raster (rasterizer.go, texture.go) ---- */
void rc_sample(const rc_texture *t, float u, float v, uint8_t out[4]);
void rc_color_intensity(const uint8_t c[4], float intensity, uint8_t out[4]);
/* fb.Triangle with tile bounds (rasterizer.go:90-183). Returns covered-fragment
 * count; passed-fragment count in *passed. */
int64_t rc_triangle(int32_t width, int32_t height, uint8_t *rgba, float *depth, int64_t x0,
                    int64_t y0, float z0, float u0, float v0, int64_t x1, int64_t y1, float z1,
                    float u1, float v1, int64_t x2, int64_t y2, float z2, float u2, float v2,
                    int64_t tsx, int64_t tsy, int64_t tex, int64_t tey, float ia, float ib,
                    float ic, const rc_texture *texture, const uint8_t face_color[4],
                    int64_t *passed, uint8_t *written);
void rc_clear(int32_t width, int32_t height, uint8_t *rgba, float *depth, const uint8_t c[4]);
void rc_dot_grid(int32_t width, int32_t height, uint8_t *rgba, const uint8_t c[4], int32_t step);
/* FrameBuffer.Pixel / Rect / Line (rasterizer.go:25-30,54-79), abs (utils.go:11-21). */
void rc_pixel(int32_t width, int32_t height, uint8_t *rgba, int64_t x, int64_t y,
              const uint8_t c[4]);
int64_t rc_go_abs(int64_t v);
void rc_rect(int32_t width, int32_t height, uint8_t *rgba, int64_t x, int64_t y, int64_t w,
             int64_t h, const uint8_t c[4]);
void rc_line(int32_t width, int32_t height, uint8_t *rgba, int64_t x0, int64_t y0, int64_t x1,
             int64_t y1, const uint8_t c[4]);
void rc_tile_bounds(int32_t tile, int32_t num_tiles, int32_t width, int32_t height, rc_vec2 *start,
                    rc_vec2 *end);

/* The end of Draw outside demo mode (renderer.go:476-480): the reference
 * tiles' borders in red, then the crosshair in yellow, colour only. */
void rc_frame_decorations(const rc_frame *frame, uint8_t *rgba);
/* FrameBuffer.Fog (rasterizer.go:185-207), over colour and depth. */
void rc_fog(int32_t width, int32_t height, uint8_t *rgba, const float *depth, float fog_start,
            float fog_end, const uint8_t c[4]);

/* ---- geometry stage pieces (renderer.go:226-244,312-370) ---- */
/* Lighting of one corner (Gouraud) or face (flat) from its world-space normal. */
float rc_light_intensity(rc_vec4 world_normal, rc_vec3 light, float ambient, float diffuse);
/* Perspective divide + screen transform of a clip-space point, W restored. */
rc_vec4 rc_project_point(rc_vec4 clip, const rc_matrix *screen);
/* Reference tiles (1 or 16) a screen triangle is binned to: its TPF share. */
int64_t rc_tile_refs(const rc_vec4 points[3], int32_t num_tiles, int32_t width, int32_t height);

/* ---- whole frame (renderer.go:443-483, sequential path) ---- */
int rc_draw(const rc_frame *frame, const rc_mesh *meshes, int32_t num_meshes,
            const rc_texture *textures, int32_t num_textures, const rc_object *objects,
            int32_t num_objects, uint8_t *rgba, float *depth, rc_stats *stats);

/* Same frame, with the reference's parallel structure: geometry of each object on
 * its own thread, raster over a tiles_x*tiles_y grid of disjoint tiles on
 * `threads` threads. Produces the same image as rc_draw (disjoint tiles remove
 * the reference's shared-boundary race). Used only as the timed CPU baseline. */
int rc_draw_parallel(const rc_frame *frame, const rc_mesh *meshes, int32_t num_meshes,
                     const rc_texture *textures, int32_t num_textures,
                     const rc_object *objects, int32_t num_objects, uint8_t *rgba,
                     float *depth, rc_stats *stats, int32_t threads, int32_t tiles_x,
                     int32_t tiles_y);

#ifdef __cplusplus
}
#endif
#endif
