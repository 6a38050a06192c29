/*
 * CPU oracle for sloth's rasterization path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the arithmetic of the Rust reference
 * (ecumene/rust-sloth) for the per-triangle setup and the per-pixel
 * coverage / depth / shade work of `draw_triangle`.  It exists to check the
 * HIP library and to time the CPU baseline; nothing in the product links it.
 * Only the test suite, the smoke check and the benchmark's CPU-baseline leg
 * load it.
 *
 * Floating point: every operation below is one IEEE-754 binary32 rounding in
 * the order the reference evaluates it (Rust never contracts a*b+c into an
 * FMA).  Build with -O2 -ffp-contract=off on x86-64 (SSE, FLT_EVAL_METHOD 0),
 * without -ffast-math.
 *
 * Matrices are 4x4, column-major (element (r,c) at m[c*4+r]), which is the
 * storage order of nalgebra's Matrix4<f32>.
 *
 * Cell packing: one uint32 per frame-buffer cell,
 *   byte0 = glyph (an 8-bit character), byte1 = R, byte2 = G, byte3 = B.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include "sloth_oracle.h"

#define CELL(g, r, gg, b) \
    ((uint32_t)(uint8_t)(g) | ((uint32_t)(r) << 8) | ((uint32_t)(gg) << 16) | ((uint32_t)(b) << 24))

/* ------------------------------------------------------------------ */
/* nalgebra 0.22.1 semantics (Cargo.lock:241-257; source not vendored)  */
/* ------------------------------------------------------------------ */

/* Matrix4 * Matrix4 (used at rasterizer.rs:57).  nalgebra's gemm walks the
 * result column by column and calls gemv, which accumulates columns of the
 * left operand with axpy: out(i,j) = (((A(i,0)B(0,j) + A(i,1)B(1,j))
 * + A(i,2)B(2,j)) + A(i,3)B(3,j)), each product rounded before the add. */
void oracle_mat4_mul(const float *A, const float *B, float *out)
{
    float r[16];
    for (int j = 0; j < 4; ++j) {
        for (int i = 0; i < 4; ++i) {
            float acc = A[0 * 4 + i] * B[j * 4 + 0];
            acc = A[1 * 4 + i] * B[j * 4 + 1] + acc;
            acc = A[2 * 4 + i] * B[j * 4 + 2] + acc;
            acc = A[3 * 4 + i] * B[j * 4 + 3] + acc;
            r[j * 4 + i] = acc;
        }
    }
    memcpy(out, r, sizeof r);
}

/* Alternative operation orders, for bounding what the unpinned nalgebra
 * order could change (tests only; per calling thread, 0 = the restatement
 * above and below):
 *   ORACLE_NORM_SEQUENTIAL  norm_squared as ((x*x + y*y) + z*z) + w*w
 *                           instead of the 4-wide dot's (x*x + z*z) + (y*y + w*w)
 *   ORACLE_MATVEC_PAIRWISE  Matrix4 * Vector4 rows as the 4-wide dot's
 *                           (M(i,0) x + M(i,2) z) + (M(i,1) y + M(i,3) w)
 *                           instead of gemv's ((M(i,0) x + M(i,1) y) + M(i,2) z) + M(i,3) w */
static __thread int g_variant = 0;

void oracle_set_arith_variant(int v) { g_variant = v; }

/* Matrix4 * Vector4 (geometry.rs:44-46), same gemv order. */
static inline void mat4_vec(const float *M, const float *v, float *out)
{
    float r[4];
    for (int i = 0; i < 4; ++i) {
        if (g_variant & ORACLE_MATVEC_PAIRWISE) {
            float a = M[0 * 4 + i] * v[0], b = M[1 * 4 + i] * v[1];
            float c = M[2 * 4 + i] * v[2], d = M[3 * 4 + i] * v[3];
            r[i] = (a + c) + (b + d);
            continue;
        }
        float acc = M[0 * 4 + i] * v[0];
        acc = M[1 * 4 + i] * v[1] + acc;
        acc = M[2 * 4 + i] * v[2] + acc;
        acc = M[3 * 4 + i] * v[3] + acc;
        r[i] = acc;
    }
    memcpy(out, r, sizeof r);
}

/* Rotation3::from_euler_angles(roll, pitch, yaw).to_homogeneous()
 * (main.rs:76-77).  sin/cos come from the C library, as Rust's
 * f32::sin_cos does on Linux. */
void oracle_rotation_euler(float roll, float pitch, float yaw, float *out)
{
    float sr = sinf(roll), cr = cosf(roll);
    float sp = sinf(pitch), cp = cosf(pitch);
    float sy = sinf(yaw), cy = cosf(yaw);
    float m00 = cy * cp;
    float m01 = cy * sp * sr - sy * cr;
    float m02 = cy * sp * cr + sy * sr;
    float m10 = sy * cp;
    float m11 = sy * sp * sr + cy * cr;
    float m12 = sy * sp * cr - cy * sr;
    float m20 = -sp;
    float m21 = cp * sr;
    float m22 = cp * cr;
    float M[16] = {m00, m10, m20, 0.0f, m01, m11, m21, 0.0f,
                   m02, m12, m22, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    memcpy(out, M, sizeof M);
}

/* ------------------------------------------------------------------ */
/* geometry.rs                                                          */
/* ------------------------------------------------------------------ */

/* Triangle::mul (geometry.rs:43-48): v <- M * v for the three vertices.
 * tri holds 3 Vector4 (x, y, z, w) = 12 floats. */
void oracle_tri_mul(const float *M, float *tri)
{
    mat4_vec(M, tri + 0, tri + 0);
    mat4_vec(M, tri + 4, tri + 4);
    mat4_vec(M, tri + 8, tri + 8);
}

/* Triangle::aabb (geometry.rs:37-42): per component v1.min(v2.min(v3)) with
 * Rust's f32::min / f32::max (IEEE minNum / maxNum: a NaN operand yields the
 * other one).  out = min[4], max[4]. */
void oracle_tri_aabb(const float *tri, float *out)
{
    for (int c = 0; c < 4; ++c) {
        out[c] = fminf(tri[c], fminf(tri[4 + c], tri[8 + c]));
        out[4 + c] = fmaxf(tri[c], fmaxf(tri[4 + c], tri[8 + c]));
    }
}

/* Triangle::normal (geometry.rs:49-56) followed by Unit::new_normalize.
 * Cross product with each product rounded; nalgebra's norm_squared on a
 * Vector4 uses the 4-wide dot special case (a+=c; b+=d; a+b), i.e.
 * ((x*x + z*z) + (y*y + w*w)) with w = 0, then sqrt, then each component is
 * divided by the norm.  out = unit normal (x, y, z, w). */
void oracle_tri_normal(const float *tri, float *out)
{
    float e1[4], e2[4];
    for (int c = 0; c < 4; ++c) {
        e1[c] = tri[4 + c] - tri[c];
        e2[c] = tri[8 + c] - tri[c];
    }
    float x = (e1[1] * e2[2]) - (e1[2] * e2[1]);
    float y = (e1[2] * e2[0]) - (e1[0] * e2[2]);
    float z = (e1[0] * e2[1]) - (e1[1] * e2[0]);
    float w = 0.0f;
    float a = x * x, b = y * y, c = z * z, d = w * w;
    float n2;
    if (g_variant & ORACLE_NORM_SEQUENTIAL) {
        n2 = ((a + b) + c) + d;
    } else {
        a += c;
        b += d;
        n2 = 0.0f + (a + b);
    }
    float n = sqrtf(n2);
    out[0] = x / n;
    out[1] = y / n;
    out[2] = z / n;
    out[3] = w / n;
}

/* ------------------------------------------------------------------ */
/* rasterizer.rs                                                        */
/* ------------------------------------------------------------------ */

/* default_shader (rasterizer.rs:5-27).  The thresholds are f32 literals. */
uint8_t oracle_default_shader(float shade)
{
    if (shade <= 0.20f) return '.';
    else if (shade <= 0.30f) return ':';
    else if (shade <= 0.40f) return '-';
    else if (shade <= 0.50f) return '=';
    else if (shade <= 0.60f) return '+';
    else if (shade <= 0.70f) return '*';
    else if (shade <= 0.80f) return '#';
    else if (shade <= 0.90f) return '%';
    else if (shade <= 1.0f) return '@';
    else return ' ';
}

/* A step shader: the first i with shade <= thresholds[i] gives glyphs[i],
 * otherwise the fallback.  default_shader is the 9-step table
 * {0.2 .. 1.0} -> ".:-=+*#%@" with fallback ' '. */
uint8_t oracle_ramp_shade(const oracle_ramp *r, float shade)
{
    for (uint32_t i = 0; i < r->n; ++i)
        if (shade <= r->thresholds[i]) return r->glyphs[i];
    return r->fallback;
}

void oracle_default_ramp(oracle_ramp *r)
{
    static const float t[9] = {0.20f, 0.30f, 0.40f, 0.50f, 0.60f, 0.70f, 0.80f, 0.90f, 1.0f};
    static const char g[9] = {'.', ':', '-', '=', '+', '*', '#', '%', '@'};
    memset(r, 0, sizeof *r);
    r->n = 9;
    for (int i = 0; i < 9; ++i) {
        r->thresholds[i] = t[i];
        r->glyphs[i] = (uint8_t)g[i];
    }
    r->fallback = ' ';
}

/* orient (rasterizer.rs:30-32). */
static inline float orient(const float *a, const float *b, const float *c)
{
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
}

/* `f as usize` in Rust: saturating, NaN -> 0. */
static inline uint64_t as_usize(float f)
{
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)f;
}

/* draw_triangle (rasterizer.rs:48-93).  `tri` is the untransformed
 * triangle (3 Vector4), `M` = context.utransform * transform computed by the
 * caller exactly as rasterizer.rs:57 does (the product is identical for every
 * triangle, the reference merely recomputes it).  Updates z, cells and, if
 * given, win (winning global triangle index per z slot, -1 = none). */
static void draw_triangle(uint32_t W, uint32_t H, int image, const float *UR,
                          float *zb, uint32_t *cells, int32_t *win, int32_t *winxy, const float *src,
                          uint8_t r, uint8_t g, uint8_t b, const oracle_ramp *ramp,
                          int32_t tri_index, uint64_t *stats)
{
    float t[12];
    memcpy(t, src, sizeof t);
    oracle_tri_mul(UR, t);
    float bb[8];
    oracle_tri_aabb(t, bb);
    uint64_t x0 = as_usize(ceilf(fmaxf(bb[0], 1.0f)));
    uint64_t y0 = as_usize(ceilf(fmaxf(bb[1], 1.0f)));
    uint64_t x1 = as_usize(ceilf(fminf(bb[4] * 2.0f, (float)(W - 1))));
    uint64_t y1 = as_usize(ceilf(fminf(bb[5], (float)(H - 1))));
    float a = 1.0f / orient(t + 0, t + 4, t + 8);
    const float *v1 = t + 0, *v2 = t + 4, *v3 = t + 8;

    for (uint64_t y = y0; y < y1; ++y) {
        for (uint64_t x = x0; x < x1; ++x) {
            float p[4] = {(float)x, (float)y, 0.0f, 0.0f};
            float w0 = orient(v2, v3, p);
            float w1 = orient(v3, v1, p);
            float w2 = orient(v1, v2, p);
            if (stats) stats[0]++;
            if (w0 >= 0.0f && w1 >= 0.0f && w2 >= 0.0f) {
                float n[4];
                oracle_tri_normal(t, n); /* recomputed per pixel, as the reference does */
                float pixel_shade = n[2] * a * (w0 + w1 + w2);
                float z = v1[2] + a * (w1 * (v2[2] - v1[2]) + w2 * (v3[2] - v1[2]));
                uint64_t id = y * (uint64_t)W + x * 2;
                if (stats) stats[1]++;
                if (z < zb[id]) {
                    zb[id] = z;
                    uint32_t cell = CELL(oracle_ramp_shade(ramp, pixel_shade), r, g, b);
                    cells[id] = cell;
                    cells[id + 1] = cell;
                    if (win) win[id] = tri_index;
                    if (winxy) {
                        winxy[2 * id] = (int32_t)x;
                        winxy[2 * id + 1] = (int32_t)y;
                    }
                    if (stats) stats[2]++;
                }
            }
        }
        if (image) cells[y * (uint64_t)W + 1] = CELL('\n', 0, 0, 0);
    }
}

/* draw_mesh (rasterizer.rs:39-46): triangles in mesh order.  xyz holds
 * [n][3][3] floats (w = 1, as every loader in the reference sets it), rgb
 * holds [n][3] bytes.  tri_base is the number of triangles drawn earlier in
 * the same frame (global draw order = z-tie order). */
void oracle_draw_mesh(uint32_t W, uint32_t H, int image, const float *U, uint32_t *cells,
                      float *zb, int32_t *win, int32_t *winxy, const float *xyz, const uint8_t *rgb,
                      uint32_t n, const float *R, const oracle_ramp *ramp, uint32_t tri_base,
                      uint64_t *stats)
{
    float UR[16];
    for (uint32_t i = 0; i < n; ++i) {
        /* rasterizer.rs:57 recomputes utransform * transform per triangle. */
        oracle_mat4_mul(U, R, UR);
        float tri[12];
        for (int k = 0; k < 3; ++k) {
            tri[4 * k + 0] = xyz[(size_t)i * 9 + 3 * k + 0];
            tri[4 * k + 1] = xyz[(size_t)i * 9 + 3 * k + 1];
            tri[4 * k + 2] = xyz[(size_t)i * 9 + 3 * k + 2];
            tri[4 * k + 3] = 1.0f;
        }
        draw_triangle(W, H, image, UR, zb, cells, win, winxy, tri, rgb[(size_t)i * 3 + 0],
                      rgb[(size_t)i * 3 + 1], rgb[(size_t)i * 3 + 2], ramp,
                      (int32_t)(tri_base + i), stats);
    }
}

/* ------------------------------------------------------------------ */
/* context.rs                                                           */
/* ------------------------------------------------------------------ */

/* Context::update (context.rs:93-141) for a given size: scale = max of the
 * meshes' bounding-box maxima (x, y, z), starting from 0; then
 * s = min(H, W/2) / scale / 2 and utransform = [[s,0,0,W/4],[0,-s,0,H/2],
 * [0,0,s,0],[0,0,0,1]].  bbox_max holds [n_meshes][3]. */
void oracle_update(uint32_t W, uint32_t H, const float *bbox_max, uint32_t n_meshes, float *U)
{
    float scale = 0.0f;
    for (uint32_t m = 0; m < n_meshes; ++m) {
        scale = fmaxf(scale, bbox_max[3 * m + 0]);
        scale = fmaxf(scale, bbox_max[3 * m + 1]);
        scale = fmaxf(scale, bbox_max[3 * m + 2]);
    }
    float fw = (float)(uint16_t)W, fh = (float)(uint16_t)H;
    scale = fminf(fh, fw / 2.0f) / scale / 2.0f;
    float M[16] = {scale, 0.0f, 0.0f, 0.0f, 0.0f, -scale, 0.0f, 0.0f,
                   0.0f, 0.0f, scale, 0.0f, fw / 4.0f, fh / 2.0f, 0.0f, 1.0f};
    memcpy(U, M, sizeof M);
}

/* Bounding box of an OBJ-loaded SimpleMesh (geometry.rs:85-88,129-135):
 * starts at the origin and grows by each triangle's aabb.  out = min[3],
 * max[3]. */
void oracle_mesh_bbox(const float *xyz, uint32_t n, float *out)
{
    float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < n; ++i) {
        float tri[12];
        for (int k = 0; k < 3; ++k) {
            tri[4 * k + 0] = xyz[(size_t)i * 9 + 3 * k + 0];
            tri[4 * k + 1] = xyz[(size_t)i * 9 + 3 * k + 1];
            tri[4 * k + 2] = xyz[(size_t)i * 9 + 3 * k + 2];
            tri[4 * k + 3] = 1.0f;
        }
        float bb[8];
        oracle_tri_aabb(tri, bb);
        for (int c = 0; c < 3; ++c) {
            mn[c] = fminf(bb[c], mn[c]);
            mx[c] = fmaxf(bb[4 + c], mx[c]);
        }
    }
    memcpy(out, mn, sizeof mn);
    memcpy(out + 3, mx, sizeof mx);
}

/* Context::clear (context.rs:35-45): cells = (' ', (0,0,0)) for W*H (+H in
 * image mode) cells, z = f32::MAX for W*H slots. */
void oracle_clear(uint32_t W, uint32_t H, int image, uint32_t *cells, float *zb)
{
    uint64_t nc = (uint64_t)W * H + (image ? H : 0);
    uint64_t nz = (uint64_t)W * H;
    for (uint64_t i = 0; i < nc; ++i) cells[i] = CELL(' ', 0, 0, 0);
    for (uint64_t i = 0; i < nz; ++i) zb[i] = 3.40282347e+38f;
    (void)0;
}

/* ------------------------------------------------------------------ */
/* Multi-frame driver for the CPU baseline                              */
/* ------------------------------------------------------------------ */

/* One frame of the main loop (main.rs:76-83) in image mode: clear, then
 * draw_mesh with transform R.  U is the frame's utransform. */
typedef struct {
    uint32_t W, H;
    int image;
    const float *U, *R;
    uint32_t F;
    const float *xyz;
    const uint8_t *rgb;
    uint32_t n;
    uint32_t *cells_out;
    float *z_out;
    const oracle_ramp *ramp;
    uint32_t next;
    pthread_mutex_t mu;
} frames_job;

static void *frames_worker(void *arg)
{
    frames_job *j = (frames_job *)arg;
    uint64_t nc = (uint64_t)j->W * j->H + (j->image ? j->H : 0);
    uint64_t nz = (uint64_t)j->W * j->H;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        uint32_t f = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (f >= j->F) break;
        /* Context::clear allocates fresh buffers every frame. */
        uint32_t *cells = (uint32_t *)malloc(nc * sizeof(uint32_t));
        float *zb = (float *)malloc(nz * sizeof(float));
        if (!cells || !zb) {
            free(cells);
            free(zb);
            break;
        }
        oracle_clear(j->W, j->H, j->image, cells, zb);
        oracle_draw_mesh(j->W, j->H, j->image, j->U + 16 * (size_t)f, cells, zb, NULL, NULL,
                         j->xyz, j->rgb, j->n, j->R + 16 * (size_t)f, j->ramp, 0, NULL);
        if (j->cells_out) memcpy(j->cells_out + nc * f, cells, nc * sizeof(uint32_t));
        if (j->z_out) memcpy(j->z_out + nz * f, zb, nz * sizeof(float));
        free(cells);
        free(zb);
    }
    return NULL;
}

int oracle_render_frames(uint32_t W, uint32_t H, int image, const float *U, const float *R,
                         uint32_t F, const float *xyz, const uint8_t *rgb, uint32_t n,
                         uint32_t *cells_out, float *z_out, const oracle_ramp *ramp, int nthreads)
{
    frames_job j = {W, H, image, U, R, F, xyz, rgb, n, cells_out, z_out, ramp, 0,
                    PTHREAD_MUTEX_INITIALIZER};
    if (nthreads < 1) nthreads = 1;
    if (nthreads == 1) {
        frames_worker(&j);
        return 0;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    if (!th) return -1;
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, frames_worker, &j);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    return 0;
}

/* Exhaustive check that the step table built by oracle_default_ramp agrees
 * with the literal if-chain of default_shader (rasterizer.rs:5-27) for every
 * binary32 bit pattern whose upper 16 bits lie in [lo_hi16, hi_hi16).
 * Returns the number of mismatching patterns. */
uint32_t oracle_shader_exhaustive_mismatch(uint32_t lo_hi16, uint32_t hi_hi16)
{
    oracle_ramp r;
    oracle_default_ramp(&r);
    uint32_t bad = 0;
    for (uint32_t h = lo_hi16; h < hi_hi16 && h < 65536u; ++h) {
        for (uint32_t l = 0; l < 65536u; ++l) {
            uint32_t bits = (h << 16) | l;
            float f;
            memcpy(&f, &bits, 4);
            if (oracle_default_shader(f) != oracle_ramp_shade(&r, f)) ++bad;
        }
    }
    return bad;
}

/* ------------------------------------------------------------------ */
/* Differentiable shading, FP64 (no reference counterpart)              */
/* ------------------------------------------------------------------ */

/* The shading model of the build's added differentiable form.  Visibility
 * is fixed: win[id] / winxy[id] give the triangle and the pixel that own z
 * slot id (from oracle_draw_mesh).  For that pixel, with the screen triangle
 * of draw_triangle (binary32 transform, rasterizer.rs:56-57):
 *   lambda_k = w_k / orient(v1,v2,v3)        (w_k as rasterizer.rs:72-74)
 *   G        = cof(M3) / |det M3|^(2/3)      (maps model-space normals to
 *              the screen-space normal of Triangle::normal, scale-free)
 *   I        = L . sum_k lambda_k G n[v_k]
 *   C        = sum_k lambda_k c[v_k]
 *   out      = I * C                         (RGB)
 *   loss     = sum over covered slots, channels of (out - target)^2
 * With n[v_k] the face's unit normal and L = (0,0,1), I is the reference's
 * flat shade nz_unit * a * (w0 + w1 + w2).  Everything below is float64.
 * Gradients (if non-NULL) are added into dcolor [nv][3], dnormal [nv][3],
 * dlight [3].  Returns the loss (0 without a target). */
double oracle_shade_f64(uint32_t W, uint32_t H, const float *M, const float *xyz, const int32_t *faces,
                        uint32_t nv, const double *vcolor, const double *vnormal, const double *light,
                        const int32_t *win, const int32_t *winxy, const float *target, double *out,
                        double *dcolor, double *dnormal, double *dlight)
{
    (void)nv;
    /* G = cof(M3) / |det M3|^(2/3), M3 = upper-left 3x3 of M (column-major) */
    double m[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r][c] = (double)M[c * 4 + r];
    double cof[3][3];
    cof[0][0] = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    cof[0][1] = -(m[1][0] * m[2][2] - m[1][2] * m[2][0]);
    cof[0][2] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    cof[1][0] = -(m[0][1] * m[2][2] - m[0][2] * m[2][1]);
    cof[1][1] = m[0][0] * m[2][2] - m[0][2] * m[2][0];
    cof[1][2] = -(m[0][0] * m[2][1] - m[0][1] * m[2][0]);
    cof[2][0] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    cof[2][1] = -(m[0][0] * m[1][2] - m[0][2] * m[1][0]);
    cof[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
    double det = m[0][0] * cof[0][0] + m[0][1] * cof[0][1] + m[0][2] * cof[0][2];
    double k = pow(fabs(det), 2.0 / 3.0);
    double G[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) G[r][c] = cof[r][c] / k;

    double loss = 0.0;
    uint64_t nz = (uint64_t)W * H;
    for (uint64_t id = 0; id < nz; ++id) {
        int32_t T = win[id];
        if (T < 0) continue;
        double px = winxy[2 * id], py = winxy[2 * id + 1];
        float t[12];
        for (int v = 0; v < 3; ++v) {
            t[4 * v + 0] = xyz[(size_t)T * 9 + 3 * v + 0];
            t[4 * v + 1] = xyz[(size_t)T * 9 + 3 * v + 1];
            t[4 * v + 2] = xyz[(size_t)T * 9 + 3 * v + 2];
            t[4 * v + 3] = 1.0f;
        }
        oracle_tri_mul(M, t);
        double x1 = t[0], y1 = t[1], x2 = t[4], y2 = t[5], x3 = t[8], y3 = t[9];
        double w[3];
        w[0] = (x3 - x2) * (py - y2) - (y3 - y2) * (px - x2);
        w[1] = (x1 - x3) * (py - y3) - (y1 - y3) * (px - x3);
        w[2] = (x2 - x1) * (py - y1) - (y2 - y1) * (px - x1);
        double o = (x2 - x1) * (y3 - y1) - (y2 - y1) * (x3 - x1);
        double lam[3] = {w[0] / o, w[1] / o, w[2] / o};
        double nt[3][3], n[3] = {0, 0, 0}, C[3] = {0, 0, 0};
        for (int v = 0; v < 3; ++v) {
            int32_t vi = faces[(size_t)T * 3 + v];
            for (int r = 0; r < 3; ++r)
                nt[v][r] = G[r][0] * vnormal[3 * vi + 0] + G[r][1] * vnormal[3 * vi + 1] + G[r][2] * vnormal[3 * vi + 2];
            for (int r = 0; r < 3; ++r) {
                n[r] += lam[v] * nt[v][r];
                C[r] += lam[v] * vcolor[3 * vi + r];
            }
        }
        double I = light[0] * n[0] + light[1] * n[1] + light[2] * n[2];
        double O[3] = {I * C[0], I * C[1], I * C[2]};
        if (out)
            for (int r = 0; r < 3; ++r) out[3 * id + r] = O[r];
        if (!target) continue;
        double g[3];
        for (int r = 0; r < 3; ++r) {
            double d = O[r] - (double)target[3 * id + r];
            loss += d * d;
            g[r] = 2.0 * d;
        }
        if (!dcolor && !dnormal && !dlight) continue;
        double dI = g[0] * C[0] + g[1] * C[1] + g[2] * C[2];
        for (int v = 0; v < 3; ++v) {
            int32_t vi = faces[(size_t)T * 3 + v];
            if (dcolor)
                for (int r = 0; r < 3; ++r) dcolor[3 * vi + r] += lam[v] * g[r] * I;
            if (dnormal) {
                /* d/dn[v] of L . (lam G n[v]) = lam G^T L */
                for (int c = 0; c < 3; ++c)
                    dnormal[3 * vi + c] += dI * lam[v] * (G[0][c] * light[0] + G[1][c] * light[1] + G[2][c] * light[2]);
            }
        }
        if (dlight)
            for (int r = 0; r < 3; ++r) dlight[r] += dI * n[r];
    }
    return loss;
}

/* ------------------------------------------------------------------ */
/* Context::flush byte streams (context.rs:50-92)                       */
/* ------------------------------------------------------------------ */

static size_t put_char_utf8(char *p, uint32_t g)
{
    if (g < 0x80u) {
        if (p) p[0] = (char)g;
        return 1;
    }
    if (p) {
        p[0] = (char)(0xC0u | (g >> 6));
        p[1] = (char)(0x80u | (g & 0x3Fu));
    }
    return 2;
}

/* Context::flush (context.rs:50-92).  A terminal-mode context (image = 0)
 * first queues crossterm's MoveTo(0, 0) = CSI "1;1H" (context.rs:53-55).
 * mode 0: `let frame: String = frame_buffer.iter().map(|p| p.0).collect();
 * println!("{}", frame);` (context.rs:58-61).  mode 1: per cell crossterm
 * 0.18.0's PrintStyledContent of style(char).with(Rgb fg).on(Rgb 25,25,25)
 * (context.rs:62-78): background CSI "48;2;25;25;25m", foreground CSI
 * "38;2;R;G;Bm", the char, then ResetColor CSI "0m" (crossterm's
 * StyledContent Display; the crate is not vendored, no fixture pins these
 * bytes).  mode 2: for every cell
 * print!("<span style=\"color:rgb({},{},{})\">{}", r, g, b, char)
 * (context.rs:79-87).  Writes into out (if non-NULL, cap bytes) and returns
 * the full length. */
size_t oracle_encode(const uint32_t *cells, uint64_t n, int mode, int image, char *out, size_t cap)
{
    size_t len = 0;
    char tmp[96];
    if (!image) {
        const char *mv = "\x1b[1;1H";
        if (out && len + 6 <= cap) memcpy(out + len, mv, 6);
        len += 6;
    }
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t c = cells[i];
        int k = 0;
        if (mode == 1)
            k = snprintf(tmp, sizeof tmp, "\x1b[48;2;25;25;25m\x1b[38;2;%u;%u;%um", (c >> 8) & 0xFFu,
                         (c >> 16) & 0xFFu, c >> 24);
        else if (mode == 2)
            k = snprintf(tmp, sizeof tmp, "<span style=\"color:rgb(%u,%u,%u)\">", (c >> 8) & 0xFFu,
                         (c >> 16) & 0xFFu, c >> 24);
        if (out && len + (size_t)k <= cap) memcpy(out + len, tmp, (size_t)k);
        len += (size_t)k;
        char u[2];
        size_t ku = put_char_utf8(u, c & 0xFFu);
        if (out && len + ku <= cap) memcpy(out + len, u, ku);
        len += ku;
        if (mode == 1) {
            if (out && len + 4 <= cap) memcpy(out + len, "\x1b[0m", 4);
            len += 4;
        }
    }
    if (mode == 0) {
        if (out && len < cap) out[len] = '\n';
        len += 1;
    }
    return len;
}
