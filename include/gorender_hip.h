/*
 * gorender_hip.h — C ABI of the MI355X (gfx950) render path.
 *
 * This is the drop-in boundary for gorender's render path: vertex transform,
 * frustum clipping, tile binning and the edge-function rasteriser with z-buffer,
 * flat/Gouraud shading and texturing. A Go build with the `hip` tag binds these
 * entry points through cgo (one call per Renderer.Draw); every type here is
 * pointer-free or a plain pointer + size, so Go memory can be passed for the
 * duration of a call and nothing is retained after it returns.
 *
 * Reference interfaces replaced (file:line in maxpoletaev/gorender):
 *   gr_create / gr_destroy      NewRenderer (renderer.go:114-164): owns device state
 *   gr_texture_create           NewColorTexture / NewImageTexture (texture.go:28-63)
 *   gr_mesh_create              NewMesh (mesh.go:53-69): face normals + bounding box
 *   gr_draw                     (*Renderer).Draw (renderer.go:443-483) writing
 *                               FrameBuffer.Pixels / ZBuffer (rasterizer.go:7-23) and TPF
 *   gr_draw_device / gr_draw_batch   the same Draw, asynchronous, device-resident
 *                               outputs (no reference counterpart: batching form)
 *   gr_server_*                 a continuous-batching request loop over resident
 *                               meshes and textures, frames read back to pinned
 *                               host memory (no reference counterpart: serving form;
 *                               replaces the frame loop of main.go:222-227 for
 *                               many independent requests)
 *
 * Output semantics: the image is the reference's sequential draw order
 * (renderer.go:466-474): objects, then faces, then clipped fan triangles.
 * Errors never cross the ABI as exceptions; every call returns a gr_status and
 * gr_last_error() returns a thread-local message.
 */
#ifndef GORENDER_HIP_H
#define GORENDER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GR_ABI_VERSION 5

typedef enum {
    GR_OK = 0,
    GR_ERR_INVALID_ARGUMENT = 1,
    GR_ERR_INVALID_HANDLE = 2,
    GR_ERR_OUT_OF_MEMORY = 3,
    GR_ERR_DEVICE = 4,
    /* A vertex/normal index is out of range: the reference panics with
     * "index out of range" (renderer.go:318-330). */
    GR_ERR_INDEX_OUT_OF_RANGE = 5,
    GR_ERR_UNSUPPORTED = 6,
    /* Bin capacity was exceeded on an asynchronous draw; the next draw on the
     * context allocates the reported size. gr_draw retries by itself. */
    GR_ERR_CAPACITY = 7,
    /* Clip polygon exceeded 9 vertices (reference: index out of range panic,
     * clipping.go:35-40); the face is dropped and counted in gr_stats. */
    GR_ERR_CLIP_OVERFLOW = 8,
    /* gr_server_wait: the request did not finish within the timeout. */
    GR_ERR_TIMEOUT = 9,
    /* Serving: the queue was full at submission, or the request waited past
     * the deadline before admission and was shed. */
    GR_ERR_BUSY = 10,
} gr_status;

/* Renderer toggles (renderer.go:90-97). Defaults of NewRenderer
 * (renderer.go:130-136): SHOW_FACES|BACKFACE_CULLING|LIGHTING|FRUSTUM_CLIPPING|SHOW_TEXTURES. */
enum {
    GR_SHOW_FACES = 1u << 0,
    GR_BACKFACE_CULLING = 1u << 1,
    GR_LIGHTING = 1u << 2,
    GR_FLAT_SHADING = 1u << 3,
    GR_FRUSTUM_CLIPPING = 1u << 4,
    GR_SHOW_TEXTURES = 1u << 5,
    GR_SHOW_EDGES = 1u << 6,    /* debug overlay: triangle edges (renderer.go:191-210) */
    GR_SHOW_VERTICES = 1u << 7, /* debug overlay: vertex dots (renderer.go:212-216); an
                                   overlay line longer than 65536 pixels (unclipped
                                   geometry far off screen) -> GR_ERR_UNSUPPORTED */
    /* The end of Draw outside demo mode (renderer.go:476-480), drawn after the
     * frame, colour only: the reference tiles' borders in red
     * (drawTilesBoundaries, renderer.go:422-434), then the yellow crosshair
     * (FrameBuffer.CrossHair, rasterizer.go:209-217). */
    GR_SHOW_TILE_BOUNDS = 1u << 8,
    GR_SHOW_CROSSHAIR = 1u << 9,
};
#define GR_DEFAULT_FLAGS \
    (GR_SHOW_FACES | GR_BACKFACE_CULLING | GR_LIGHTING | GR_FRUSTUM_CLIPPING | GR_SHOW_TEXTURES)

/* TextureType (texture.go:13-17) */
enum { GR_TEXTURE_SOLID = 0, GR_TEXTURE_IMAGE = 1, GR_TEXTURE_IMAGE_FAST = 2 };

/* Texture fields (texture.go:19-26). IMAGE_FAST requires power-of-two sizes
 * (texture.go:40-42); NewImageTexture picks it by itself, see
 * gr_texture_kind_for_size. */
typedef struct {
    int32_t type;
    int32_t width, height;
    float scale;      /* 1.0 by default (texture.go:53), SetScale (texture.go:65-67) */
    uint8_t color[4]; /* returned by SOLID textures */
} gr_texture_desc;

/* Mesh data in the reference's layout (mesh.go:8-26): Vec4 positions and
 * normals, per-face index triples and per-corner UVs. Pointers are read during
 * the call only. Optional arrays may be NULL. */
typedef struct {
    int32_t num_vertices;         /* >= 1 (boundingBox reads vertices[0], mesh.go:29) */
    const float *vertices;        /* [V][4] X,Y,Z,W */
    int32_t num_vertex_normals;   /* 0: faces are lit by face normals (renderer.go:313) */
    const float *vertex_normals;  /* [N][4] or NULL when N == 0 */
    int32_t num_faces;
    const int32_t *vertex_indices; /* [F][3], each in [0, V) */
    const int32_t *normal_indices; /* [F][3] or NULL (all zero, a Go zero Face) */
    const float *uvs;              /* [F][3][2] or NULL (all zero) */
    const int32_t *face_textures;  /* [F] texture handle or -1 for nil; NULL = all nil */
    const float *face_normals;     /* [F][4] or NULL: computed as NewMesh does (mesh.go:53-60) */
} gr_mesh_desc;

/* One object of a draw: a resident mesh and its matrices, row-major [4][4]
 * float32 as Go builds them: mvp = ((I*P)*V)*W (renderer.go:259-262) and the
 * world matrix W (renderer.go:255). */
typedef struct {
    int32_t mesh;
    int32_t reserved;
    float mvp[16];
    float world[16];
} gr_object;

/* Per-frame parameters. The reference hard-codes most of them
 * (renderer.go:10-20,118-121,265,448-449); gr_frame_defaults fills those values. */
typedef struct {
    int32_t width, height;  /* 1..16384 */
    uint32_t flags;         /* GR_* toggles */
    int32_t ref_tiles;      /* 16 (parallel, main.go:19) or 1 (sequential): TPF grid only */
    float screen[16];       /* NewScreenMatrix (matrix.go:108-118) */
    float light[3];         /* normalised light direction (renderer.go:265) */
    float ambient;          /* ambientStrength 0.5 (renderer.go:13) */
    float diffuse;          /* diffuseStrength 0.5 (renderer.go:12) */
    float znear, zfar;      /* NewFrustum planes (clipping.go:93-129) */
    uint8_t face_color[4];  /* faceColor (renderer.go:17) */
    uint8_t clear_color[4]; /* fb.Clear colour (renderer.go:448) */
    uint8_t dot_color[4];   /* fb.DotGrid colour (renderer.go:449) */
    int32_t dot_step;       /* DotGrid step, 10 (renderer.go:449); < 1 draws no dots */
} gr_frame;

typedef struct {
    int64_t tpf;               /* Renderer.TPF (renderer.go:436-441) on the ref_tiles grid */
    int64_t triangles;         /* post-clip triangles binned on the full-screen tile */
    int64_t fragments_covered; /* (pixel, triangle) pairs inside the edge test */
    int64_t pixels_shaded;     /* pixels whose final colour comes from a triangle */
    int64_t faces_front;       /* faces passing the backface test */
    int64_t clip_overflow;     /* faces dropped by GR_ERR_CLIP_OVERFLOW */
    int64_t refs;              /* tile-list entries used */
    int64_t tiles_split;       /* tiles whose list several raster workgroups shared */
} gr_stats;

/* One frame of an asynchronous batch: outputs are device pointers. */
typedef struct {
    gr_frame frame;
    int32_t first_object, num_objects; /* range in the batch's object array */
    void *d_rgba;                      /* device uint8 [H][W][4] */
    void *d_depth;                     /* device float [H][W] */
} gr_batch_frame;

/* Tunables (autotuned). Zero fields keep the current value. */
typedef struct {
    int32_t tile_w, tile_h;       /* raster tile in pixels: 8, 16, 32 or 64 */
    int64_t refs_capacity;        /* tile-list entries per batch, and triangle records
                                     (at most 1792 per 256 faces); grown on demand */
    int32_t raster_waves;         /* waves per raster workgroup: 4 */
    int32_t tile_split;           /* tile-list entries per raster workgroup, 64..2^24: a
                                     longer list is split into parts that several
                                     workgroups resolve (same result, more parallelism) */
    int32_t pipelines;            /* 1 or 2 (default): with 2, consecutive asynchronous draws
                                     alternate between two sets of buffers and streams, so a
                                     draw's geometry stages overlap the previous draw's raster;
                                     frames are still written in submission order. A stream
                                     set with gr_set_stream runs everything on it (1). */
    int32_t raster_grid;          /* 1 (default): one raster workgroup per screen tile;
                                     2: a persistent grid (every CU at the raster's
                                     occupancy) whose workgroups stride over the tiles.
                                     Same images; slower here, kept for comparison. */
} gr_config;

typedef struct gr_context gr_context;

/* ---- context ---- */
gr_status gr_create(int32_t device, gr_context **out);
void gr_destroy(gr_context *ctx);
const char *gr_last_error(void);
int32_t gr_abi_version(void);
gr_status gr_set_config(gr_context *ctx, const gr_config *cfg);
gr_status gr_get_config(gr_context *ctx, gr_config *cfg);
/* Run all work on `hip_stream` (a hipStream_t; NULL = the context's own stream). */
gr_status gr_set_stream(gr_context *ctx, void *hip_stream);
void *gr_get_stream(gr_context *ctx);

/* ---- resident assets (HBM) ---- */
/* Image textures hold at most 2^28 texels (GR_ERR_INVALID_ARGUMENT beyond). */
gr_status gr_texture_create(gr_context *ctx, const gr_texture_desc *desc, const uint8_t *rgba,
                            int32_t *out_handle);
gr_status gr_texture_destroy(gr_context *ctx, int32_t handle);
/* GR_TEXTURE_IMAGE_FAST when both sizes are powers of two (texture.go:40-42,
 * utils.go:7-9), else GR_TEXTURE_IMAGE. */
int32_t gr_texture_kind_for_size(int32_t width, int32_t height);
gr_status gr_mesh_create(gr_context *ctx, const gr_mesh_desc *desc, int32_t *out_handle);
gr_status gr_mesh_destroy(gr_context *ctx, int32_t handle);
/* Bytes of HBM held by a mesh (for budget-based admission). */
int64_t gr_mesh_device_bytes(gr_context *ctx, int32_t handle);

/* ---- helpers reproducing the reference's host-side set-up ----
 * float32 math in Go's evaluation order; row-major float[16] matrices. */
void gr_frame_defaults(gr_frame *frame, int32_t width, int32_t height); /* renderer.go:114-136,448-449 */
void gr_matrix_multiply(const float *a, const float *b, float *out);     /* matrix.go:155-165 */
void gr_matrix_world(const float scale[3], const float rotation[3], const float translation[3],
                     float *out);                                         /* matrix.go:82-88 */
void gr_matrix_view(const float eye[3], const float direction[3], const float up[3],
                    float *out);                                          /* matrix.go:133-144 */
void gr_matrix_perspective(float fov, float aspect, float znear, float zfar,
                           float *out);                                   /* matrix.go:92-106 */
void gr_matrix_screen(int32_t width, int32_t height, float *out);         /* matrix.go:108-118 */
void gr_matrix_mvp(const float *perspective, const float *view, const float *world,
                   float *out);                                           /* renderer.go:259-262 */
void gr_vec3_normalize(const float in[3], float out[3]);                  /* vector.go:92-94 */
/* Frustum.BoxVisibility of an object's 8 bounding-box corners (bbox [8][4],
 * as NewMesh's boundingBox gives them) transformed by mvp: 0 outside, 1
 * intersect, 2 inside (renderer.go:268-275, clipping.go:131-154). Every draw
 * classifies each object this way on the host; -1 for a null argument. */
int32_t gr_box_visibility(const float *mvp, const float *bbox, float znear, float zfar);

/* ---- drawing ---- */
/* Synchronous Draw into caller-owned host buffers: rgba [H][W][4], depth [H][W]
 * (either may be NULL), stats may be NULL. */
gr_status gr_draw(gr_context *ctx, const gr_frame *frame, const gr_object *objects,
                  int32_t num_objects, uint8_t *rgba, float *depth, gr_stats *stats);
/* Asynchronous Draw on the context stream into device buffers. */
gr_status gr_draw_device(gr_context *ctx, const gr_frame *frame, const gr_object *objects,
                         int32_t num_objects, void *d_rgba, void *d_depth);
/* Many independent frames in one set of launches (batched form). */
gr_status gr_draw_batch(gr_context *ctx, const gr_batch_frame *frames, int32_t num_frames,
                        const gr_object *objects, int32_t num_objects);
/* gr_draw_device that allocates nothing: no counterpart in the reference, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Every per-draw buffer
 * lives in one fixed 32 MiB device array per GPU, held until the process ends;
 * host data reaches it through kernel arguments, and the draw calls no
 * allocation, free or memset. Limits: one frame of at most 16 objects, no edge
 * or vertex overlays (GR_ERR_UNSUPPORTED), and the frame's tile counters, 320
 * triangle records per 256 faces and 2^16 tile-list entries within 32 MiB
 * (about 130k faces at 1080p): larger frames -> GR_ERR_INVALID_ARGUMENT. Tile
 * lists or clipped triangles beyond what is left of the array ->
 * GR_ERR_CAPACITY from gr_synchronize. The array is
 * shared by every call on the GPU and keeps the last call's contents: calls
 * that overlap on one GPU (two contexts or streams at once) race; calls on
 * one context's stream are ordered. The image equals gr_draw_device's. */
gr_status gr_draw_device_noalloc(gr_context *ctx, const gr_frame *frame,
                                 const gr_object *objects, int32_t num_objects, void *d_rgba,
                                 void *d_depth);
/* ---- the batched form captured in HIP graphs (no reference counterpart) ----
 * The launch sequence of one gr_draw_batch (counter zeroing, geometry and
 * clipping, binning, raster) captured once, with device buffers of its own,
 * in two HIP graphs (geometry + binning, then raster) that
 * gr_batch_graph_launch replays: two graph launches per batch, no host work
 * per frame. Meshes, textures and output buffers are bound at creation
 * (destroying one the graph uses invalidates it); gr_batch_graph_update
 * gives the same objects new matrices. Replays are ordered like gr_draw_batch
 * calls: frames are written in submission order, and with two pipes a
 * replay's geometry overlaps the previous batch's raster. A replay whose tile
 * lists overflow makes gr_synchronize return GR_ERR_CAPACITY; the graph's next
 * launch regrows them (draw it again). gr_last_stats then describes the
 * graph's frames. */
typedef struct gr_batch_graph gr_batch_graph;
gr_status gr_batch_graph_create(gr_context *ctx, const gr_batch_frame *frames,
                                int32_t num_frames, const gr_object *objects,
                                int32_t num_objects, gr_batch_graph **out);
/* New mvp/world matrices for the graph's objects (same count and meshes). */
gr_status gr_batch_graph_update(gr_batch_graph *graph, const gr_object *objects,
                                int32_t num_objects);
gr_status gr_batch_graph_launch(gr_batch_graph *graph);
void gr_batch_graph_destroy(gr_batch_graph *graph);

/* ---- autotuner (no reference counterpart) ----
 * Draws one batch under each candidate configuration and keeps the fastest in
 * the context's config: raster tiles {8,16,32,64} x {8,16,32,64}, tile_split
 * {256, 512, 1024, 4096}, with the tile-list capacity sized to what the batch
 * needs plus a quarter. Each candidate is timed over `reps` launches (HIP
 * events, one pipe) after one untimed launch, and its frames' colour and depth
 * are checksummed on the device: the image may not depend on the
 * configuration, so a candidate whose checksum differs from the first's fails
 * the call (GR_ERR_DEVICE). results (may be NULL) receives up to max_results
 * candidates in the order tried. */
typedef struct {
    int32_t tile_w, tile_h;
    int32_t tile_split;
    int32_t raster_waves;
    double ms_per_launch;  /* mean over the timed launches */
    uint64_t checksum;     /* of every frame's colour and depth buffers */
} gr_tune_result;
gr_status gr_autotune(gr_context *ctx, const gr_batch_frame *frames, int32_t num_frames,
                      const gr_object *objects, int32_t num_objects, int32_t reps,
                      gr_tune_result *results, int32_t max_results, int32_t *num_results,
                      gr_config *best);

/* Wait for queued draws; returns GR_ERR_CAPACITY if one of them overflowed its
 * bins (its output is then incomplete and must be drawn again). */
gr_status gr_synchronize(gr_context *ctx);
/* The present path of the reference's window loop (main.go:201-227 with
 * FrameBuffer.SwapBuffers, rasterizer.go:32-34): gr_present_draw queues a
 * Draw and the copy of its colour (and depth when with_depth) to pinned host
 * memory the context owns, and returns at once; gr_present_wait waits for the
 * oldest queued frame and points rgba/depth (depth NULL without with_depth) at
 * its pixels, which stay valid until the second gr_present_draw after it. Two
 * frames are in flight at most (Pixels / Pixels2): the draw of frame N+1
 * overlaps the copy-out of frame N. A frame whose tile lists overflow is drawn
 * again inside gr_present_wait. */
gr_status gr_present_draw(gr_context *ctx, const gr_frame *frame, const gr_object *objects,
                          int32_t num_objects, int32_t with_depth);
gr_status gr_present_wait(gr_context *ctx, const uint8_t **rgba, const float **depth,
                          gr_stats *stats);
/* Stats of the frames of the last completed draw call (after gr_synchronize). */
gr_status gr_last_stats(gr_context *ctx, gr_stats *stats, int32_t max_frames,
                        int32_t *num_frames);

/* ---- device memory for the asynchronous forms (no reference counterpart:
 * a Go caller of gr_draw_device / gr_draw_batch owns its output buffers) ---- */
gr_status gr_device_alloc(gr_context *ctx, size_t bytes, void **out);
gr_status gr_device_free(gr_context *ctx, void *ptr);
/* Fill `bytes` bytes with `value` on the context stream (asynchronous). */
gr_status gr_device_fill(gr_context *ctx, void *ptr, int32_t value, size_t bytes);
/* Device -> device copy of `bytes` bytes on the context stream (asynchronous);
 * the regions must not overlap. */
gr_status gr_device_copy(gr_context *ctx, void *dst, const void *src, size_t bytes);
/* FrameBuffer.Fog (rasterizer.go:185-207) over a drawn frame in device
 * memory, on the context stream (asynchronous): depth >= fog_start keeps the
 * pixel, depth <= fog_end takes `color`, depths between blend toward it. */
gr_status gr_fog(gr_context *ctx, void *rgba, const void *depth, int32_t width, int32_t height,
                 float fog_start, float fog_end, const uint8_t color[4]);
/* Adds an order-independent 64-bit digest of `bytes` (a multiple of 4) device
 * bytes at ptr to *digest_dev (device memory, 8 bytes), on the context stream
 * after the draws queued before it: equal buffers give equal digests, so two
 * forms of the same frames can be compared without copying them out. */
gr_status gr_device_checksum(gr_context *ctx, const void *ptr, size_t bytes, void *digest_dev);
/* Wait for the context stream, then copy device -> host. */
gr_status gr_copy_to_host(gr_context *ctx, void *dst, const void *src, size_t bytes);

/* ---- asset ingestion (host side; no device needed to parse) ---- */
typedef struct gr_obj_file gr_obj_file;
/* LoadObjFile (obj.go:196-309): parse an OBJ file, the MTL libraries it names
 * (parseMtlLibFile, obj.go:156-193) and their PNG map_Kd images into host
 * memory, one mesh per "o" group unless single_mesh. Line, face and number
 * rules are the reference's (fmt.Sscanf formats of obj.go:42-151, including the
 * "f v//vn" normal-index bug of obj.go:77-82). Malformed input ->
 * GR_ERR_INVALID_ARGUMENT; indexes the reference would panic on while parsing
 * -> GR_ERR_INDEX_OUT_OF_RANGE. */
gr_status gr_obj_open(const char *path, int32_t single_mesh, gr_obj_file **out);
void gr_obj_close(gr_obj_file *obj);
int32_t gr_obj_num_meshes(const gr_obj_file *obj);
int32_t gr_obj_num_textures(const gr_obj_file *obj);
/* Arrays of mesh i, valid until gr_obj_close; face_textures index the file's
 * texture table (gr_obj_texture), -1 = nil. Face normals are left to
 * gr_mesh_create (NewMesh, mesh.go:53-69). */
gr_status gr_obj_mesh(const gr_obj_file *obj, int32_t i, gr_mesh_desc *desc);
/* Texture i of the file: a decoded map_Kd image, or the solid magenta default
 * the reference gives materials without one (obj.go:219,231-234). */
gr_status gr_obj_texture(const gr_obj_file *obj, int32_t i, gr_texture_desc *desc,
                         const uint8_t **rgba);
/* Upload every texture and mesh of the file to the resident store; one mesh
 * handle per mesh (GR_ERR_CAPACITY, uploading nothing, if max_meshes is short). */
gr_status gr_obj_upload(gr_context *ctx, const gr_obj_file *obj, int32_t *mesh_handles,
                        int32_t max_meshes, int32_t *num_meshes);
/* image/png decoding as LoadTextureFile uses it (texture.go:91-103), pixels
 * converted by color.RGBAModel (texture.go:57: alpha premultiplied); *rgba is
 * allocated here, release it with gr_host_free. JPEG -> GR_ERR_UNSUPPORTED. */
gr_status gr_png_decode(const char *path, int32_t *width, int32_t *height, uint8_t **rgba);
void gr_host_free(void *ptr);
/* LoadTextureFile + gr_texture_create (type from gr_texture_kind_for_size). */
gr_status gr_texture_load(gr_context *ctx, const char *path, int32_t *out_handle);

/* ---- serving (no reference counterpart) ----
 * A request loop on one context: requests are admitted in arrival order while
 * their colour + depth targets (width * height * 8 bytes each) fit the budget,
 * every admitted request joins the next batched launch, and each frame is read
 * back into pinned host memory on a second stream while later batches draw.
 * A batch that overflows its tile lists is drawn again with more capacity.
 * Frames are bit-identical to gr_draw of the same request alone. One worker
 * thread drives the context: upload meshes and textures first, and do not use
 * the context directly while the server runs. */
typedef struct gr_server gr_server;
typedef struct {
    int64_t target_budget; /* bytes of device colour + depth targets admitted at once */
    int32_t max_batch;     /* requests per batched launch */
    int32_t max_inflight;  /* launches queued on the device at once */
    int32_t colour_only;   /* 1: read back colour only (result depth = NULL), as a
                              window's present needs (main.go:296-298); 0: both */
    int32_t max_queue;     /* requests waiting for admission at most (0: no bound);
                              gr_server_submit beyond it -> GR_ERR_BUSY */
    int32_t deadline_ms;   /* a request that waited longer for admission is shed:
                              its wait returns GR_ERR_BUSY (0: never) */
    int32_t reserved;      /* 0 */
} gr_server_config;
typedef struct {
    int32_t status;        /* gr_status of the request */
    int32_t width, height;
    int32_t batch_size;    /* requests in the launch that drew it */
    const uint8_t *rgba;   /* [H][W][4], pinned host memory valid until gr_server_release */
    const float *depth;    /* [H][W] */
    gr_stats stats;
    int64_t submit_ns;     /* steady clock: submitted */
    int64_t start_ns;      /* its batch was launched */
    int64_t done_ns;       /* colour and depth are in host memory */
} gr_serve_result;
gr_status gr_server_create(gr_context *ctx, const gr_server_config *cfg, gr_server **out);
/* Finishes every submitted request, then frees the server (results included). */
void gr_server_destroy(gr_server *server);
/* Queue a request (frame and objects are copied). It is checked here as
 * gr_draw_batch would check it, so a bad request is refused alone: a frame
 * larger than the budget or with bad fields -> GR_ERR_INVALID_ARGUMENT, a dead
 * mesh or texture handle -> GR_ERR_INVALID_HANDLE, Gouraud lighting over normal
 * indices out of range -> GR_ERR_INDEX_OUT_OF_RANGE. */
gr_status gr_server_submit(gr_server *server, const gr_frame *frame, const gr_object *objects,
                           int32_t num_objects, int64_t *ticket);
/* Wait for a request (timeout_ms < 0: no limit; GR_ERR_TIMEOUT when it passes)
 * and describe its result; returns the request's status, with its message in
 * gr_last_error. A frame whose overlay lines are too long fails alone
 * (GR_ERR_UNSUPPORTED); the other frames of its batch complete. */
gr_status gr_server_wait(gr_server *server, int64_t ticket, int32_t timeout_ms,
                         gr_serve_result *out);
/* Hand a finished request's host buffers back to the server. */
gr_status gr_server_release(gr_server *server, int64_t ticket);
/* Open-loop load for measurement: submits request i % num_distinct (frame d
 * with objects[first_object[d] .. + num_objects[d]]) at arrival_s[i] seconds
 * after the call starts (busy-waiting on the clock), while a second thread
 * waits for and releases results in submission order; latency_ns[i]
 * (submit -> colour and depth in host memory; -1 refused by a full queue, -2
 * shed past the deadline) and batch_size[i] may be NULL. */
gr_status gr_server_replay(gr_server *server, const gr_frame *frames, const gr_object *objects,
                           const int32_t *first_object, const int32_t *num_objects,
                           int32_t num_distinct, const double *arrival_s, int32_t n,
                           int64_t *latency_ns, int32_t *batch_size);

/* ---- measurement ---- */
/* When enabled, each kernel launch is bracketed by HIP events on the context
 * stream and its duration accumulated per kernel name. */
gr_status gr_set_kernel_timing(gr_context *ctx, int32_t enabled);
/* Copies up to `max` entries; names are static strings. */
int32_t gr_kernel_times(gr_context *ctx, const char **names, double *total_ms, int64_t *launches,
                        int32_t max);
void gr_reset_kernel_times(gr_context *ctx);

/* ---- self-check (no reference counterpart) ----
 * The raster stage divides by per-triangle divisors through their correctly
 * rounded reciprocals and two residual corrections, which equals IEEE float32
 * division on the divisor/numerator ranges it admits. This runs `count` pairs
 * drawn from `seed` over those ranges on the device, compares both quotients
 * bit for bit, and returns the mismatch count; first_mismatch (4 floats, may
 * be NULL) receives numerator, divisor, IEEE quotient, tested quotient. */
gr_status gr_check_exact_division(gr_context *ctx, uint64_t count, uint64_t seed,
                                  uint64_t *mismatches, float *first_mismatch);

#ifdef __cplusplus
}
#endif
#endif
