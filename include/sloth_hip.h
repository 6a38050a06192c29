/*
 * sloth_hip.h -- C-ABI of the MI355X rasterization library (libsloth_hip.so).
 *
 * Drop-in boundary for the hot path of ecumene/rust-sloth: the reference's
 * `Context` (src/context.rs:12-19), `SimpleMesh` (src/geometry.rs:78-81) and
 * `draw_mesh` (src/rasterizer.rs:39-46).  Plain pointers and sizes only; every
 * entry point returns an int status (SLOTH_OK = 0) and never throws or aborts
 * across the boundary.  sloth_last_error() describes the last failure on the
 * calling thread.
 *
 * Data formats
 *   positions  float[n][3][3]  triangle soup, vertex v1,v2,v3 as x,y,z; w = 1
 *              (every constructor in the reference sets w = 1:
 *              geometry.rs:26-32, 93-98, 166-170)
 *   colours    uint8[n][3]     the per-triangle `color` (geometry.rs:18)
 *   matrices   float[16]       4x4, column-major (nalgebra's storage order)
 *   cells      uint32[W*H (+H in image mode)]  byte0 = glyph, byte1..3 = R,G,B;
 *              the (char, (u8,u8,u8)) of Context.frame_buffer (context.rs:16)
 *   z          float[W*H]      Context.z_buffer (context.rs:17)
 *   winners    int32[W*H]      frame-global index of the triangle that owns
 *              each z slot, -1 if none (not in the reference; for tests)
 *
 * Threading: a handle must not be used from two threads at once.  Different
 * handles are independent.  All device work of a handle runs on one HIP
 * stream; functions that return host data synchronise it.
 */
#ifndef SLOTH_HIP_H
#define SLOTH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLOTH_OK 0
#define SLOTH_EINVAL 1     /* bad argument (null handle, W or H outside 1..65535, ...) */
#define SLOTH_ENOMEM 2     /* device or host allocation failed */
#define SLOTH_EHIP 3       /* HIP runtime error (message in sloth_last_error) */
#define SLOTH_ENODEV 4     /* no usable gfx950 device */
#define SLOTH_ESTATE 5     /* call not valid in the handle's current state */

#define SLOTH_MAX_RAMP 16

/* A step shader.  The reference passes a closure `Fn(f32) -> char`
 * (rasterizer.rs:39-42); a closure cannot cross the ABI, so the shader is a
 * table: the first i with shade <= thresholds[i] gives glyphs[i], otherwise
 * `fallback` (NaN included).  default_shader (rasterizer.rs:5-27) is
 * n = 9, thresholds {0.2,...,1.0}, glyphs ".:-=+*#%@", fallback ' '.
 * Glyphs are 8-bit characters. */
typedef struct sloth_ramp {
    uint32_t n;
    float thresholds[SLOTH_MAX_RAMP];
    uint8_t glyphs[SLOTH_MAX_RAMP];
    uint8_t fallback;
    uint8_t reserved[3];
} sloth_ramp;

typedef struct sloth_mesh sloth_mesh;
typedef struct sloth_context sloth_context;
typedef struct sloth_batch sloth_batch;

/* ---- library ---------------------------------------------------------- */

/* Selects the HIP device for handles created afterwards (default 0) and
 * checks that it is a gfx950 part.  No reference counterpart (the reference
 * has no device). */
int sloth_init(int device);

/* Thread-local description of the last failed call ("" if none). */
const char *sloth_last_error(void);

/* Fills `out` with default_shader's table (rasterizer.rs:5-27). */
int sloth_default_ramp(sloth_ramp *out);

/* ---- SimpleMesh (geometry.rs:78-81) ----------------------------------- */

/* Uploads a triangle soup to device memory.  n may be 0.  The bounding box
 * follows the OBJ loader (geometry.rs:85-88,129-135): it starts at the origin
 * and grows by every triangle's aabb. */
int sloth_mesh_create(const float *positions, const uint8_t *colors, uint32_t n, sloth_mesh **out);

/* Wraps positions/colours that already live in device memory (no copy; the
 * caller keeps them alive and unchanged while the mesh is used).  The
 * bounding box is computed from a copy read back to the host. */
int sloth_mesh_create_device(const float *d_positions, const uint8_t *d_colors, uint32_t n,
                             sloth_mesh **out);

/* An indexed device mesh: vertex rows float[n_vertices][3] and per triangle
 * its three vertex ids faces uint32[n][3] and its colour uint8[n][3].  It
 * renders exactly as the soup positions[t][k] = vertices[faces[t][k]] (the
 * kernels read the same floats through the ids), in 12 bytes per vertex and
 * 15 per triangle instead of 39 per triangle.  Face ids must be below
 * n_vertices (else SLOTH_EINVAL).  The bounding box follows the OBJ loader
 * over the triangles.  No reference counterpart: SimpleMesh is a soup
 * (geometry.rs:78-81). */
int sloth_mesh_create_indexed(const float *vertices, uint32_t n_vertices, const uint32_t *faces,
                              const uint8_t *colors, uint32_t n, sloth_mesh **out);

/* 1 if the mesh is indexed (and its vertex count in *n_vertices), else 0. */
int sloth_mesh_is_indexed(const sloth_mesh *mesh, uint32_t *n_vertices);

/* SimpleMesh.bounding_box (geometry.rs:79): min[3], max[3]. */
int sloth_mesh_bbox(const sloth_mesh *mesh, float *min3, float *max3);
uint32_t sloth_mesh_triangles(const sloth_mesh *mesh);
int sloth_mesh_destroy(sloth_mesh *mesh);

/* ---- model files ----------------------------------------------------------
 * The mesh queue of match_meshes (inputs.rs:93-129) for one file, parsed on
 * the host: ".obj" (any case) as tobj 3.2.2 loads it with GPU_LOAD_OPTIONS
 * (inputs.rs:108) and to_simple_mesh_with_materials converts it
 * (geometry.rs:83-138): one mesh per model, faces fan-triangulated, colour =
 * material Kd * 255 as u8, or the first corner's vertex colour, or (1,1,1)
 * without materials; bounding box from the origin.  ".stl" as stl_io 0.4.2's
 * read_stl (inputs.rs:117) and to_simple_mesh (geometry.rs:146-181): ASCII or
 * binary, colour (255,255,0), bounding box from (f32::MAX, f32::MIN).  Files
 * on which the reference panics (a named material library that cannot be
 * read, a model without a material while materials exist, bad indices)
 * return SLOTH_EINVAL. */
typedef struct sloth_model sloth_model;
int sloth_model_load(const char *path, sloth_model **out);
int sloth_model_mesh_count(const sloth_model *model, uint32_t *n_meshes);
/* Mesh i's soup (positions [n][3][3], colours [n][3]; valid until the model
 * is destroyed), triangle count and bounding box; any output may be NULL. */
int sloth_model_mesh(const sloth_model *model, uint32_t i, const float **positions, const uint8_t **colors,
                     uint32_t *n_tris, float *bbox_min3, float *bbox_max3);
/* Mesh i uploaded as a device mesh, indexed: positions the soup repeats bit
 * for bit are stored once (it renders as the soup does); sloth_mesh_bbox
 * then reports the loader's bounding box. */
int sloth_model_upload(const sloth_model *model, uint32_t i, sloth_mesh **out);
int sloth_model_destroy(sloth_model *model);

/* ---- Context (context.rs:12-142) -------------------------------------- */

/* Context::blank (context.rs:22-34) followed by a fixed size, as image mode's
 * match_dimensions (inputs.rs:159-169) sets it.  width, height in 1..65535
 * (Context::update passes them through u16, context.rs:99).  image != 0 adds
 * one newline cell per row (context.rs:38-40).  utransform starts as the
 * identity. */
int sloth_context_create(uint32_t width, uint32_t height, int image, sloth_context **out);
int sloth_context_destroy(sloth_context *ctx);

/* Context::update (context.rs:93-141): utransform from the size and the
 * largest bounding-box maximum of the given meshes.  bbox_max: [n][3]. */
int sloth_context_update(sloth_context *ctx, const float *bbox_max, uint32_t n_meshes);

/* Context.utransform (context.rs:13), column-major. */
int sloth_context_set_utransform(sloth_context *ctx, const float *m16);
int sloth_context_get_utransform(const sloth_context *ctx, float *m16);

/* Context::clear (context.rs:35-45): all cells (' ', (0,0,0)), z = f32::MAX,
 * and the list of draws since the last clear is emptied. */
int sloth_context_clear(sloth_context *ctx);

/* rasterizer.rs:39-46 draw_mesh(context, mesh, transform, shader).  Records
 * the draw; the frame is rasterized when its buffers are read (or by
 * sloth_context_render).  Triangle order across draws since the last clear is
 * the z-tie order, as in the reference.  The mesh must stay alive until the
 * next clear.  shader = NULL means default_shader. */
int sloth_draw_mesh(sloth_context *ctx, const sloth_mesh *mesh, const float *transform16,
                    const sloth_ramp *shader);

/* Rasterizes every draw recorded since the last clear into the context's
 * device buffers, on `stream` (a hipStream_t; NULL = the context's own
 * stream).  Returns without waiting for the device.  On a caller's stream
 * the launch waits for the context's pending uploads, and the context's
 * stream waits for the launch, so sloth_context_read waits for it too. */
int sloth_context_render(sloth_context *ctx, void *stream);

/* Renders if needed, then copies the frame buffer (W*H (+H) cells), z buffer
 * and winning triangle ids (W*H each) to the host (any pointer may be NULL)
 * and waits. */
int sloth_context_read(sloth_context *ctx, uint32_t *cells, float *z, int32_t *winners);

/* Device addresses of the context's cell and z buffers (valid until the
 * context is destroyed), and the frame's element counts: W*H (+H) cells and
 * W*H depths. */
int sloth_context_device_buffers(sloth_context *ctx, uint32_t **d_cells, float **d_z, uint64_t *n_cells,
                                 uint64_t *n_z);

/* Context::flush's byte streams (context.rs:50-92), encoded on the device:
 *   SLOTH_ENCODE_TEXT  the no-colour flush (context.rs:58-61): every cell's
 *                      char, then println!'s '\n'
 *   SLOTH_ENCODE_ANSI  the colour terminal flush (context.rs:62-78): per
 *                      cell crossterm 0.18.0's PrintStyledContent of the char
 *                      with its RGB on RGB (25,25,25): ESC "[48;2;25;25;25m",
 *                      ESC "[38;2;R;G;Bm", the char, ESC "[0m"
 *   SLOTH_ENCODE_HTML  the -j export (context.rs:79-87): per cell
 *                      `<span style="color:rgb(R,G,B)">` then the char
 * A terminal-mode frame (image = 0) starts with crossterm's MoveTo(0, 0),
 * ESC "[1;1H" (context.rs:53-55).  Chars are written as UTF-8 (a glyph byte is
 * a code point below 256).  With out = NULL only *len is set; otherwise out
 * must hold cap >= *len bytes.  Renders first if needed. */
#define SLOTH_ENCODE_TEXT 0
#define SLOTH_ENCODE_ANSI 1
#define SLOTH_ENCODE_HTML 2
int sloth_context_encode(sloth_context *ctx, int mode, char *out, uint64_t cap, uint64_t *len);

/* ---- batches (many frames in one launch sequence; no reference
 * counterpart: the reference renders one frame per loop iteration,
 * main.rs:60-111) ---------------------------------------------------- */

typedef struct sloth_frame_desc {
    uint32_t width, height;   /* 1..65535 */
    int32_t image;
    float utransform[16];     /* column-major; e.g. from sloth_screen_transform */
} sloth_frame_desc;

typedef struct sloth_draw_desc {
    uint32_t frame;           /* index into the frame array */
    const sloth_mesh *mesh;
    float transform[16];      /* column-major */
} sloth_draw_desc;

/* Builds a batch: every frame cleared, then its draws in array order (draws
 * of one frame need not be contiguous; their relative order is the frame's
 * draw order).  Output buffers for all frames are allocated in device memory.
 * shader = NULL means default_shader.  Frames of any widths render on the
 * tiled path; a batch with a frame of 2^22 - 1 or more triangles takes the
 * general path (per-slot 64-bit keys in device memory), which has no tiles.
 * Results are identical either way. */
int sloth_batch_create(const sloth_frame_desc *frames, uint32_t n_frames, const sloth_draw_desc *draws,
                       uint32_t n_draws, const sloth_ramp *shader, sloth_batch **out);

/* Replaces the per-draw transforms without rebuilding: transforms16 holds
 * exactly n_draws column-major matrices, float[n_draws][16], in the order of
 * the draw array given to sloth_batch_create.  Waits for the batch's stream.
 * The frames on the device keep the old transforms' images until the next
 * sloth_batch_run / _run_per_frame; reading before that returns SLOTH_ESTATE.
 * If the new transforms need longer tile lists than the ones sized so far,
 * the next synchronising call (sloth_batch_sync, _read_frame, _encode_frame)
 * re-sizes them, renders again and re-captures the graph at the next run. */
int sloth_batch_set_transforms(sloth_batch *b, const float *transforms16);

/* Renders all frames on `stream` (NULL = the batch's own stream).  The first
 * call captures the launch sequence into a HIP graph; later calls replay it.
 * Returns without waiting.  On a caller's stream the launch waits for the
 * batch's own pending work, and the batch's stream waits for the launch, so
 * the synchronising calls below also wait for it. */
int sloth_batch_run(sloth_batch *b, void *stream);

/* Renders every frame with one launch sequence per frame and a device
 * synchronisation after each (the plain, unbatched form).  Waits. */
int sloth_batch_run_per_frame(sloth_batch *b, void *stream);

/* Waits for the batch's stream. */
int sloth_batch_sync(sloth_batch *b);

/* Copies one frame's cells / z / winners to the host (NULL = skip); waits. */
int sloth_batch_read_frame(sloth_batch *b, uint32_t frame, uint32_t *cells, float *z, int32_t *winners);

/* sloth_context_encode for one frame of a batch that has been run. */
int sloth_batch_encode_frame(sloth_batch *b, uint32_t frame, int mode, char *out, uint64_t cap, uint64_t *len);

/* The `-j` export (main.rs:55-58, 85-106) of frames [first, first + count)
 * of a run batch: "let frames = [\n", then for each frame "`\n", the frame's
 * flush bytes and "`,\n" ("`];\n" after the last).  mode is
 * SLOTH_ENCODE_HTML (the colour export's spans, context.rs:79-87) or
 * SLOTH_ENCODE_TEXT (with --no-color the export writes the plain flush,
 * context.rs:58-61, newline included).  Which frames the export holds is the
 * caller's schedule: main.rs stops after the requested count or once the
 * accumulated pitch exceeds 9.42477.  With out = NULL only *len is set;
 * otherwise out must hold cap >= *len bytes. */
int sloth_batch_encode_export(sloth_batch *b, uint32_t first, uint32_t count, int mode, char *out, uint64_t cap,
                              uint64_t *len);

/* Device addresses of one frame's cells (W*H (+H) of them, starting on a
 * 256-byte boundary) and z (W*H). */
int sloth_batch_device_frame(sloth_batch *b, uint32_t frame, uint32_t **d_cells, float **d_z);

/* Times each stage of one (non-graph) run with HIP events on the batch's
 * stream: out_ms[i] = milliseconds of stage i, names[i] its kernel name;
 * *n_stages in: capacity, out: count.  On the tiled path the first stage
 * (k_tile_clear) runs only for batches with frames of odd width; otherwise
 * it times two adjacent events. */
int sloth_batch_profile(sloth_batch *b, float *out_ms, const char **names, uint32_t *n_stages);

/* Runs `warmup` untimed renders, then `steps` timed renders bracketed by HIP
 * events on the batch's stream, and waits.  mode 0: graph replay
 * (sloth_batch_run); mode 1: one launch sequence per frame with a device
 * synchronisation after each (sloth_batch_run_per_frame).  *ms = elapsed
 * milliseconds of the timed renders. */
int sloth_batch_time(sloth_batch *b, uint32_t steps, uint32_t warmup, int mode, float *ms);

/* Tile size of the batch's frames, in cells (width a power of two in 8..512,
 * height in 4..64).  Takes effect at the next render, which fails with
 * SLOTH_EINVAL when a tile does not fit the raster's workgroup: at most 64 KiB
 * of LDS (12 KiB of triangle rows, 8 bytes a z slot, 12 for odd widths) and
 * at most 256 z slots a tile row (a slot covers two cells of an even-width
 * frame and one of an odd-width frame). */
int sloth_batch_set_tile(sloth_batch *b, uint32_t tile_w, uint32_t tile_h);

/* Tile raster variant of the batch's next renders: 1 = wavefronts walk the
 * boxes of more than 128 pixels in a tile together, 0 = every lane walks its
 * own box, -1 = the choice of the sizing run (boxes of more than 128 pixels
 * in at least 1 triangle of 512).  Results are identical either way. */
int sloth_batch_set_walk(sloth_batch *b, int mode);

/* Triangles per thread of the binning kernel (1, 2 or 4) for the batch's
 * next renders: more keeps more loads and atomics in flight per barrier,
 * fewer spreads small draws over more workgroups.  Results are identical
 * either way. */
int sloth_batch_set_bin(sloth_batch *b, uint32_t tris_per_thread);

/* Renders the batch with each candidate tile size that fits -- 16, 32, 64,
 * 128, 256 or 512 cells (8..256 logical pixels) wide by 8, 16, 32 or 64 rows
 * -- timed with HIP events as the least of three runs of 3 graph renders
 * each, keeps the fastest, then times the tile raster with and without the
 * wavefront walk of big boxes at that size and keeps the faster, then the
 * binning kernel with 4, 2 and 1 triangles per thread.  The choice is remembered for batches of the same
 * class (frame sizes, cell count and log2 of the triangle count) in this
 * process and reused without timing.  Writes the chosen size to tile_w /
 * tile_h (may be NULL; 0 for batches on the general path, which has no
 * tiles). */
int sloth_batch_autotune(sloth_batch *b, uint32_t *tile_w, uint32_t *tile_h);

/* The batch's current raster plan: tile size in cells (0 x 0 on the general
 * path), whether the tile raster walks big boxes by wavefront (1), not (0),
 * or -1 on the general path, and the binning kernel's triangles per thread
 * (0 on the general path).  Any pointer may be NULL. */
int sloth_batch_plan(const sloth_batch *b, uint32_t *tile_w, uint32_t *tile_h, int *big_walk,
                     uint32_t *bin_tris_per_thread);

int sloth_batch_destroy(sloth_batch *b);

/* ---- differentiable shading (no reference counterpart: the reference shades
 * with a flat glyph, rasterizer.rs:76; this is the build's added form) ---- */

/* Shades the visible fragments of a batch with per-vertex colours, per-vertex
 * normals and a light direction, with visibility fixed by the raster.  For
 * each covered z slot (winning triangle t, pixel p) with barycentrics
 * lambda_k = w_k(p) / orient(v1,v2,v3) (rasterizer.rs:67-74):
 *     rgb = (L . sum_k lambda_k G n[f_k]) * (sum_k lambda_k c[f_k])
 * where f = faces[t] and G = cof(M3) / |det M3|^(2/3) maps model-space normals
 * to the screen space of Triangle::normal (M3 = upper-left 3x3 of
 * utransform * transform).  With n = the face normal and L = (0,0,1) the
 * intensity is the reference's flat shade.  The loss is the sum over covered
 * slots and channels of (rgb - target)^2.
 *
 * faces: [n_tris][3] vertex ids shared by every draw of the batch (every draw
 * must draw a mesh of n_tris triangles).  Creating the shading renders the
 * batch once to fix the visibility. */
typedef struct sloth_shading sloth_shading;
int sloth_shading_create(sloth_batch *b, const uint32_t *faces, uint32_t n_tris, uint32_t n_verts,
                         sloth_shading **out);
/* Re-renders the visibility (after sloth_batch_set_transforms). */
int sloth_shading_visibility(sloth_shading *s);
/* Number of z slots over all frames (the rows of target and out_rgb). */
int sloth_shading_slots(const sloth_shading *s, uint64_t *n_slots);
/* target: float[n_slots][3] on the host (copied); NULL removes it. */
int sloth_shading_set_target(sloth_shading *s, const float *target);
/* Forward: vcolor, vnormal float[n_verts][3], light float[3].  out_rgb
 * (float[n_slots][3], zero where no triangle is visible) and loss (needs a
 * target) may be NULL. */
int sloth_shading_forward(sloth_shading *s, const float *vcolor, const float *vnormal, const float *light,
                          float *out_rgb, double *loss);
/* Backward of the loss at the parameters of the last forward call: gradients
 * float[n_verts][3], float[n_verts][3], float[3] (any may be NULL). */
int sloth_shading_backward(sloth_shading *s, float *d_vcolor, float *d_vnormal, float *d_light);
/* One training step in a single pass over the fragments: the loss (as
 * sloth_shading_forward gives it) and its gradients (as
 * sloth_shading_backward gives them) at vcolor, vnormal, light.  Needs a
 * target; any output may be NULL. */
int sloth_shading_step(sloth_shading *s, const float *vcolor, const float *vnormal, const float *light,
                       double *loss, float *d_vcolor, float *d_vnormal, float *d_light);
/* Times `steps` training steps (loss and gradients, as sloth_shading_step) on
 * the device-resident parameters and target after `warmup` untimed ones;
 * *ms = elapsed milliseconds. */
int sloth_shading_time(sloth_shading *s, uint32_t steps, uint32_t warmup, float *ms);
int sloth_shading_destroy(sloth_shading *s);

/* ---- rendering without allocation -------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Renders one frame of one draw -- Context::clear,
 * draw_mesh(mesh, transform, shader) with the given utransform
 * (rasterizer.rs:39-93), as a one-draw context renders it -- from device
 * positions / colours into the caller's device buffers d_cells (W*H (+H in
 * image mode) cells) and d_z (W*H floats), on `stream` (NULL = the null
 * stream), without waiting.  It calls no allocator: its working memory is a
 * fixed device array of the library, one per GPU, 33 MiB, held until the
 * process ends.  W*H is limited to 4,194,304 z slots (2560 x 1440 fits;
 * larger frames return SLOTH_EINVAL).  Every call on a GPU uses that GPU's
 * one array: calls that overlap on one GPU (two streams, two threads) race
 * and must be serialised by the caller.  Any width, odd included. */
int sloth_render_frame_noalloc(const float *d_positions, const uint8_t *d_colors, uint32_t n, uint32_t width,
                               uint32_t height, int image, const float *utransform16, const float *transform16,
                               const sloth_ramp *shader, uint32_t *d_cells, float *d_z, void *stream);

/* ---- helpers ---------------------------------------------------------- */

/* Context::update's utransform for a size and bounding-box maxima
 * (context.rs:104-133), without a context. */
int sloth_screen_transform(uint32_t width, uint32_t height, const float *bbox_max, uint32_t n_meshes,
                           float *out16);

/* Rotation3::from_euler_angles(roll, pitch, yaw).to_homogeneous()
 * (main.rs:76-77), column-major, with the C library's sinf / cosf as Rust's
 * f32::sin_cos calls them. */
int sloth_rotation_euler(float roll, float pitch, float yaw, float *out16);

/* utransform * transform in nalgebra's Matrix4 product order
 * (rasterizer.rs:57). */
int sloth_mat4_mul(const float *a16, const float *b16, float *out16);

#ifdef __cplusplus
}
#endif
#endif /* SLOTH_HIP_H */
