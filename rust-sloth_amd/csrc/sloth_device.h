// Device-side data layout and per-triangle arithmetic shared by every
// raster kernel.  All floating-point expressions keep the reference's
// binary32 operation order (rasterizer.rs, geometry.rs of ecumene/rust-sloth);
// the library is compiled with -ffp-contract=off so none of them fuses.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sloth {

constexpr uint64_t KEY_EMPTY = ~0ull;
constexpr uint32_t CELL_BLANK = 0x20u;     // (' ', (0,0,0))   context.rs:36-37
constexpr uint32_t CELL_NEWLINE = 0x0Au;   // ('\n', (0,0,0))  rasterizer.rs:90
constexpr float F32_MAX = 3.40282347e+38f;
constexpr int MAX_RAMP = 16;

// A step shader: the first i with shade <= thr[i] gives glyph[i], else
// fallback.  default_shader (rasterizer.rs:5-27) is the 9-step table.
struct Ramp {
    uint32_t n;
    float thr[MAX_RAMP];
    uint8_t glyph[MAX_RAMP];
    uint8_t fallback;
    uint8_t sorted;   // thresholds non-decreasing and none NaN (set by the host,
                      // which also sets thr[n..MAX_RAMP) to +inf)
    uint8_t pad[2];
};

// One draw_mesh call (rasterizer.rs:39-46) inside one frame.
struct Draw {
    const float* xyz;     // [n][3][3] positions, w = 1 (a soup); or [V][3] vertex rows when faces is set
    const uint32_t* faces;   // [n][3] vertex ids of an indexed mesh, or null for a soup
    const uint8_t* rgb;   // [n][3]
    uint32_t n;           // triangles
    uint32_t tri_begin;   // frame-global index of the first triangle (tie order)
    uint32_t frame;
    uint32_t ramp;
    uint64_t item_begin;  // index of the first triangle among all draws of the launch
    float M[16];          // utransform * transform (rasterizer.rs:57), column-major
};

// One Context (context.rs:12-19) being rendered.
struct Frame {
    uint32_t W, H;
    uint32_t image;
    uint32_t draw_begin, draw_end;   // [begin, end) into the draw table
    uint32_t n_tris;
    uint64_t cell_off;    // into the cell buffer (a multiple of 64), W*H (+H) cells
    uint64_t z_off;       // into the z / key / winner buffers, W*H slots
    uint64_t nl_off;      // into the per-row newline table, H rows
    uint64_t cell_begin;  // index of this frame's first cell among all frames
    uint32_t tiles_x, tiles_y;   // tiled path: tile grid over the W x H cells
    uint64_t tile_begin;         // tiled path: index of the frame's first tile
};

// Cells of a frame: W*H, plus one newline cell per row in image mode
// (context.rs:38-40).  Frames are laid out further apart than this.
__host__ __device__ inline uint64_t frame_cells(const Frame& f)
{
    return (uint64_t)f.W * f.H + (f.image ? f.H : 0);
}

// Transformed triangle + the quantities draw_triangle derives from it.
struct Tri {
    float x1, y1, z1, x2, y2, z2, x3, y3, z3;
    float A0, B0, A1, B1, A2, B2;   // edge deltas of orient(v2,v3,.), orient(v3,v1,.), orient(v1,v2,.)
    float a;                        // 1 / orient(v1,v2,v3)      rasterizer.rs:67
    float minx, maxx;               // the vertices' x extent
    int x0, x1d, y0, y1d;           // pixel domain [x0,x1d) x [y0,y1d)  rasterizer.rs:58-66
};

// Matrix4 * Vector4 with w = 1 in nalgebra's gemv order (geometry.rs:44-46):
// row i = (((M(i,0) x + M(i,1) y) + M(i,2) z) + M(i,3) * 1).
__device__ __forceinline__ float xform_row(const float* M, int i, float x, float y, float z)
{
    float acc = M[0 + i] * x;
    acc = M[4 + i] * y + acc;
    acc = M[8 + i] * z + acc;
    acc = M[12 + i] + acc;
    return acc;
}

// `f as usize` (saturating, NaN -> 0), clamped to 2^30: every domain bound is
// later compared against W-1 or H-1 < 2^16, so the clamp keeps empty ranges
// empty.
__device__ __forceinline__ int sat_index(float f)
{
    // fmaxf returns 0 for NaN (IEEE maxNum), and the clamp keeps the
    // conversion in range
    return (int)fminf(fmaxf(f, 0.0f), 1073741824.0f);
}

// Triangle::clone + mul + aabb + pixel domain + a (rasterizer.rs:56-67).
// Mesh arrays are device-accessible global memory, but their addresses come
// from tables in memory, so the compiler cannot tell them from LDS and emits
// flat loads (which wait on both memory counters).  Reading through an
// explicit global pointer gives global loads.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* global_ptr(const T* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ void tri_finish(uint32_t W, uint32_t H, Tri& t);

// Transformed triangle from its 9 model-space coordinates p (v1, v2, v3).
template <class P>
__device__ __forceinline__ void setup_tri_from(const float* M, const P& p, uint32_t W, uint32_t H, Tri& t)
{
    t.x1 = xform_row(M, 0, p[0], p[1], p[2]);
    t.y1 = xform_row(M, 1, p[0], p[1], p[2]);
    t.z1 = xform_row(M, 2, p[0], p[1], p[2]);
    t.x2 = xform_row(M, 0, p[3], p[4], p[5]);
    t.y2 = xform_row(M, 1, p[3], p[4], p[5]);
    t.z2 = xform_row(M, 2, p[3], p[4], p[5]);
    t.x3 = xform_row(M, 0, p[6], p[7], p[8]);
    t.y3 = xform_row(M, 1, p[6], p[7], p[8]);
    t.z3 = xform_row(M, 2, p[6], p[7], p[8]);
    tri_finish(W, H, t);
}

// The rest of Triangle set-up from the screen vertices t.x*, t.y*: extent,
// pixel domain, edge deltas and a (rasterizer.rs:58-67).
__device__ __forceinline__ void tri_finish(uint32_t W, uint32_t H, Tri& t)
{
    t.minx = fminf(t.x1, fminf(t.x2, t.x3));
    t.maxx = fmaxf(t.x1, fmaxf(t.x2, t.x3));
    const float miny = fminf(t.y1, fminf(t.y2, t.y3));
    const float maxy = fmaxf(t.y1, fmaxf(t.y2, t.y3));
    t.x0 = sat_index(ceilf(fmaxf(t.minx, 1.0f)));
    t.y0 = sat_index(ceilf(fmaxf(miny, 1.0f)));
    t.x1d = sat_index(ceilf(fminf(t.maxx * 2.0f, (float)(W - 1))));
    t.y1d = sat_index(ceilf(fminf(maxy, (float)(H - 1))));
    t.A0 = t.x3 - t.x2;
    t.B0 = t.y3 - t.y2;
    t.A1 = t.x1 - t.x3;
    t.B1 = t.y1 - t.y3;
    t.A2 = t.x2 - t.x1;
    t.B2 = t.y2 - t.y1;
    float o = t.A2 * (t.y3 - t.y1) - t.B2 * (t.x3 - t.x1);
    t.a = 1.0f / o;
}

__device__ __forceinline__ void setup_tri(const float* M, const float* pv, uint32_t W, uint32_t H, Tri& t)
{
    setup_tri_from(M, global_ptr(pv), W, H, t);
}

// Triangle `local` of a draw: a soup's 9 consecutive floats, or the vertex
// rows an indexed mesh's face names (the same floats, so the same results).
__device__ __forceinline__ void setup_draw_tri(const Draw& dr, uint32_t local, uint32_t W, uint32_t H, Tri& t)
{
    if (dr.faces) {
        const auto* f = global_ptr(dr.faces + (size_t)local * 3);
        const auto* vp = global_ptr(dr.xyz);
        const uint32_t a = 3u * f[0], b = 3u * f[1], c = 3u * f[2];
        const float v[9] = {vp[a], vp[a + 1], vp[a + 2], vp[b], vp[b + 1], vp[b + 2], vp[c], vp[c + 1], vp[c + 2]};
        setup_tri_from(dr.M, v, W, H, t);
    } else {
        setup_tri(dr.M, dr.xyz + (size_t)local * 9, W, H, t);
    }
}

// The three edge functions at integer pixel (x, y) (rasterizer.rs:71-74).
__device__ __forceinline__ void edges(const Tri& t, float px, float py, float& w0, float& w1, float& w2)
{
    w0 = t.A0 * (py - t.y2) - t.B0 * (px - t.x2);
    w1 = t.A1 * (py - t.y3) - t.B1 * (px - t.x3);
    w2 = t.A2 * (py - t.y1) - t.B2 * (px - t.x1);
}

// w0 >= 0 && w1 >= 0 && w2 >= 0 (rasterizer.rs:80), as one v_minimum3_f32 and
// one compare: IEEE 754-2019 minimum returns NaN when any operand is NaN (then
// the compare fails, as `NaN >= 0` does) and -0 for (-0, +0) (which passes, as
// `-0 >= 0` does), so the two forms agree on every input.
__device__ __forceinline__ bool inside(float w0, float w1, float w2)
{
    return __builtin_elementwise_minimum(__builtin_elementwise_minimum(w0, w1), w2) >= 0.0f;
}

// Depth (rasterizer.rs:77-79).
__device__ __forceinline__ float depth(const Tri& t, float w1, float w2)
{
    return t.z1 + t.a * (w1 * (t.z2 - t.z1) + w2 * (t.z3 - t.z1));
}

// z component of Triangle::normal() after Unit::new_normalize
// (geometry.rs:49-56): products rounded before the subtraction, the squared
// norm summed as ((x^2 + z^2) + (y^2 + 0^2)), then z / sqrt.
// normal_parts gives nz and the squared norm n2, unit_normal_z = nz / sqrt(n2).
__device__ __forceinline__ void normal_parts(const Tri& t, float& nz, float& n2)
{
    float e1x = t.x2 - t.x1, e1y = t.y2 - t.y1, e1z = t.z2 - t.z1;
    float e2x = t.x3 - t.x1, e2y = t.y3 - t.y1, e2z = t.z3 - t.z1;
    float nx = (e1y * e2z) - (e1z * e2y);
    float ny = (e1z * e2x) - (e1x * e2z);
    nz = (e1x * e2y) - (e1y * e2x);
    float a = nx * nx, b = ny * ny, c = nz * nz;
    float d = 0.0f;
    a += c;
    b += d;
    n2 = 0.0f + (a + b);
}

__device__ __forceinline__ float unit_normal_z(const Tri& t)
{
    float nz, n2;
    normal_parts(t, nz, n2);
    return nz / sqrtf(n2);
}

// Flat shade value at a covered pixel (rasterizer.rs:76).
__device__ __forceinline__ float shade_value(float nzu, float a, float w0, float w1, float w2)
{
    return nzu * a * (w0 + w1 + w2);
}

// Step index of shade s under a shader with non-decreasing thresholds: the
// first i with s <= thr[i], or n when there is none or s is NaN (the
// fallback).  The count is the one ramp_glyph uses below.
__device__ __forceinline__ uint32_t ramp_index_sorted(const Ramp& r, float s)
{
    const uint32_t c = (uint32_t)(r.thr[3] < s) + (uint32_t)(r.thr[7] < s) + (uint32_t)(r.thr[11] < s) +
                       (uint32_t)(r.thr[15] < s);
    const uint32_t b = 4u * min(c, 3u);
    const uint32_t i = b + (uint32_t)(r.thr[b] < s) + (uint32_t)(r.thr[b + 1] < s) + (uint32_t)(r.thr[b + 2] < s) +
                       (uint32_t)(r.thr[b + 3] < s);
    return (i >= r.n || s != s) ? r.n : i;
}

__device__ __forceinline__ uint8_t ramp_glyph(const Ramp& r, float s)
{
    if (r.sorted) {
        // Non-decreasing thresholds: {k : thr[k] < s} is a prefix [0, i), so
        // i is the first index with s <= thr[i] (none when i == n); a NaN
        // shade passes no threshold.  The host pads thr[n..16) with +inf,
        // which no shade exceeds, so i is counted over all 16 entries in two
        // steps of independent reads: c blocks of four lie wholly below s
        // (their last entries are), and the block b = min(c, 3) adds its own
        // entries below s (four when c = 4, giving 16).
        static_assert(MAX_RAMP == 16, "the count below covers 16 thresholds");
        const uint32_t c = (uint32_t)(r.thr[3] < s) + (uint32_t)(r.thr[7] < s) + (uint32_t)(r.thr[11] < s) +
                           (uint32_t)(r.thr[15] < s);
        const uint32_t b = 4u * min(c, 3u);
        const uint32_t i = b + (uint32_t)(r.thr[b] < s) + (uint32_t)(r.thr[b + 1] < s) + (uint32_t)(r.thr[b + 2] < s) +
                           (uint32_t)(r.thr[b + 3] < s);
        return (i >= r.n || s != s) ? r.fallback : r.glyph[i];
    }
    for (uint32_t i = 0; i < r.n; ++i)
        if (s <= r.thr[i]) return r.glyph[i];
    return r.fallback;
}

// Order-preserving map of a depth that passed `z < f32::MAX` to 32 bits.
// -0 and +0 compare equal in the reference's `<`, so both map to +0's code.
__device__ __forceinline__ uint32_t depth_code(float z)
{
    // z + 0 turns -0 into +0 (and changes no other value), so that -0 == +0;
    // then negative values have every bit flipped and the others the sign bit
    const uint32_t u = __float_as_uint(z + 0.0f);
    return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}

// Inverse of depth_code for every depth but -0 (whose code is +0's).
__device__ __forceinline__ float depth_of_code(uint32_t code)
{
    return __uint_as_float((code & 0x80000000u) ? (code & 0x7FFFFFFFu) : ~code);
}
constexpr uint32_t DEPTH_CODE_ZERO = 0x80000000u;   // depth_code(+0) = depth_code(-0)

// Packed z-test key: (depth code, frame triangle index, sub).  `sub` orders the
// two logical pixels a z slot can receive from one triangle when 2x >= W wraps
// id = y*W + 2x into the next row: the wrapped pixel (row y-1) is drawn first.
__device__ __forceinline__ uint64_t make_key(float z, uint32_t tri, uint32_t sub)
{
    return ((uint64_t)depth_code(z) << 32) | (uint64_t)((tri << 1) | sub);
}

// Logical pixel (x, y) that a key with `sub` denotes at z slot `id`.
__device__ __forceinline__ void key_pixel(uint32_t id, uint32_t sub, uint32_t W, uint32_t& x, uint32_t& y)
{
    uint32_t row = id / W, col = id - row * W;
    if (sub) {
        y = row;
        x = col >> 1;
    } else {
        y = row - 1;
        x = (col + W) >> 1;
    }
}

// Classification of a triangle's candidate pixels (see setup_class).
enum : uint32_t { CLS_EMPTY = 0, CLS_BOX = 1, CLS_SPANS = 2, CLS_FULL = 3, CLS_TIGHT = 4 };

// An upper bound of sum_k |A_k| |py - ay_k| + |B_k| |px - ax_k| (edge k
// anchored at vertex (ax_k, ay_k)) over the pixels of [X0, X1] x [Y0, Y1],
// X0 <= X1 and Y0 <= Y1: for a vertex x in [minx, maxx], |px - x| <=
// DX = max(|X1 - minx|, |X0 - maxx|) (|X0 - x| <= X1 - minx when x <= X0,
// <= maxx - X0 otherwise; likewise for X1), so the sum is at most
// (sum |A_k|) DY + (sum |B_k|) DX.  Its few roundings (relative error 2^-24
// each) are within the margins of the bounds that use it.  miny, maxy: the
// vertices' y extent.
__device__ __forceinline__ float edge_terms_bound(const Tri& t, float X0, float X1, float Y0, float Y1, float miny,
                                                  float maxy)
{
    const float DX = fmaxf(fabsf(X1 - t.minx), fabsf(X0 - t.maxx));
    const float DY = fmaxf(fabsf(Y1 - miny), fabsf(Y0 - maxy));
    return (fabsf(t.A0) + fabsf(t.A1) + fabsf(t.A2)) * DY + (fabsf(t.B0) + fabsf(t.B1) + fabsf(t.B2)) * DX;
}

// Conservative proof that the reference's FP32 edge test cannot pass outside
// the box [floor(minx), ceil(maxx)] (CLS_BOX), outside the tight box
// [ceil(minx - 1/32), floor(maxx + 1/32)] (CLS_TIGHT), or anywhere at all (back
// face, CLS_EMPTY).  Uses |w_fp - w_exact| <= 4.0001 u (|A| |py-ay| + |B|
// |px-ax|) per edge over the whole pixel domain (summed over the edges by
// edge_terms_bound), and that the exact edge functions sum to the exact
// orient(v1,v2,v3) = o.
//
// Box argument: a pixel d to the left of minx (or right of maxx) has some
// exact barycentric lambda_i <= -d / (2 (maxx - minx)) (at most two of the
// three can be negative, and they carry the x offset), so its exact w_i <=
// -o d / (2 width) and its FP32 w_i < 0 once o d / (2 width) > bound.  With
// bound * width < lo / 2 that holds for d >= 1 (CLS_BOX); with bound * width
// < lo / 128 for d >= 1/64 (CLS_TIGHT), and every integer x >= minx - 1/64 is
// >= fl(minx - 1/32) (the subtraction rounds by at most 1/64 when ulp(minx) <=
// 1/32, and when ulp(minx) >= 1/16 minx sits on a 1/16 grid, so x >= minx).
//
// Triangles the bound cannot decide, and those with coordinates outside
// +-2^22, fall to CLS_SPANS (exact per-row spans found from the monotonicity of
// each FP32 edge function along a row) or CLS_FULL (every pixel of the
// domain).  u = 2^-24; the factor 8 covers the bound's own rounding.
//
// Every test is evaluated and the class picked by selects, lowest precedence
// first (no branches: a wavefront's lanes take different classes).  Where a
// higher test decides, the values of the lower ones (NaN or inf included) are
// not used.
__device__ __forceinline__ uint32_t setup_class(const Tri& t)
{
    const bool no_pixels = t.x0 >= t.x1d || t.y0 >= t.y1d;
    const float LIM = 4194304.0f;
    // all six |coordinates| < 2^22 and none NaN: the IEEE 754-2019 maximum
    // propagates NaN, so one compare of the largest decides
    const float mx = __builtin_elementwise_maximum(__builtin_elementwise_maximum(fabsf(t.x1), fabsf(t.x2)), fabsf(t.x3));
    const float my = __builtin_elementwise_maximum(__builtin_elementwise_maximum(fabsf(t.y1), fabsf(t.y2)), fabsf(t.y3));
    const bool finite = __builtin_elementwise_maximum(mx, my) < LIM;
    const float K = 8.0f * 5.9604645e-08f;
    // a non-empty domain has X0 <= X1 and Y0 <= Y1 (an empty one is CLS_EMPTY)
    const float X0 = (float)t.x0, X1 = (float)(t.x1d - 1), Y0 = (float)t.y0, Y1 = (float)(t.y1d - 1);
    const float bound = K * edge_terms_bound(t, X0, X1, Y0, Y1, fminf(t.y1, fminf(t.y2, t.y3)),
                                             fmaxf(t.y1, fmaxf(t.y2, t.y3)));
    float ey = t.y3 - t.y1, ex = t.x3 - t.x1;
    float o = t.A2 * ey - t.B2 * ex;
    float eo = K * (fabsf(t.A2) * fabsf(ey) + fabsf(t.B2) * fabsf(ex));
    const float lo = o - eo;
    const float bw = bound * (t.maxx - t.minx);
    uint32_t c = CLS_SPANS;
    c = lo > 0.0f && bw < 0.5f * lo ? CLS_BOX : c;
    c = lo > 0.0f && bw < 0.0078125f * lo ? CLS_TIGHT : c;
    c = o + eo < -bound ? CLS_EMPTY : c;           // back face, provably uncovered
    c = finite ? c : CLS_FULL;
    return no_pixels ? CLS_EMPTY : c;
}

// Per-row candidate spans of a CLS_BOX / CLS_TIGHT triangle, for walking a
// box [bx0, bx1] x [by0, by1] row by row from row y_first over at most 64 rows.
//
// Edge k (anchor (xk, yk), coefficients A, B as in Tri) passes in FP32 only
// where its exact value E = A (py - yk) - B (px - xk) is >= -err, with err <=
// 3.0001 u (|A| |py - yk| + |B| |px - xk|) (three roundings of the
// reference's expression, u = 2^-24) <= e = 4u (|A| Dy + |B| Dx) over the box
// (Dx, Dy the largest distances from the anchor).  So a passing pixel has
// px <= T(py) = xk + (A (py - yk) + e) / B when B > 0, and px >= T(py) when
// B < 0; B = 0 gives no limit.  T is linear in py (slope s = A / B): it is
// evaluated at y_first and stepped by s per row.  The float evaluation (a
// hardware reciprocal within 1 ulp, a few roundings, up to 64 steps) errs by
// well under 2^-17 of the magnitudes involved (M below); the limits are moved
// outward by d = 2^-13 M + 2^-12 pixel, so every pixel that can pass lies
// within [ceil(max L), floor(min H)] of its row.  Overflow or NaN in any of
// these values only removes the limit (min / max ignore NaN; +-inf limits are
// no limit), so the span is never narrower than the passing set; the walk
// still evaluates the reference's test on every pixel of it.
struct RowSpans {
    float H[3], L[3], s[3];   // per edge: upper / lower limit at the current row (or +-inf), step per row
};

__device__ __forceinline__ void row_spans_begin(const Tri& t, int bx0, int bx1, int by0, int by1, int y_first,
                                                RowSpans& r)
{
    const float ax[3] = {t.x2, t.x3, t.x1}, ay[3] = {t.y2, t.y3, t.y1};
    const float A[3] = {t.A0, t.A1, t.A2}, B[3] = {t.B0, t.B1, t.B2};
    const float X0 = (float)bx0, X1 = (float)bx1, Y0 = (float)by0, Y1 = (float)by1, Yf = (float)y_first;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float rB = __builtin_amdgcn_rcpf(B[k]);
        const float sk = A[k] * rB;
        const float Dy = fmaxf(fabsf(Y0 - ay[k]), fabsf(Y1 - ay[k]));
        const float Dx = fmaxf(fabsf(X0 - ax[k]), fabsf(X1 - ax[k]));
        const float e = 0x1p-22f * (fabsf(A[k]) * Dy + fabsf(B[k]) * Dx) + 0x1p-120f;
        const float T = ax[k] + (A[k] * (Yf - ay[k]) + e) * rB;
        const float M = fabsf(ax[k]) + (fabsf(A[k]) * Dy + e) * fabsf(rB) + fabsf(sk) * 64.0f;
        const float d = 0x1p-13f * M + 0x1p-12f;
        r.s[k] = sk;
        r.H[k] = B[k] > 0.0f ? T + d : INFINITY;
        r.L[k] = B[k] < 0.0f ? T - d : -INFINITY;
    }
}

// No limits (careful entries, whose boxes k_careful computed).
__device__ __forceinline__ void row_spans_none(RowSpans& r)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.s[k] = 0.0f;
        r.H[k] = INFINITY;
        r.L[k] = -INFINITY;
    }
}

// x range of the candidate box of a CLS_BOX / CLS_TIGHT triangle, clipped to
// its domain (empty when bx0 > bx1).  |minx|, |maxx| < 2^22 for both classes.
__device__ __forceinline__ void box_x(const Tri& t, bool tight, int& bx0, int& bx1)
{
    if (tight) {
        bx0 = max(t.x0, (int)ceilf(t.minx - 0.03125f));
        bx1 = min(t.x1d - 1, (int)floorf(t.maxx + 0.03125f));
    } else {
        bx0 = max(t.x0, (int)floorf(t.minx));
        bx1 = min(t.x1d - 1, (int)ceilf(t.maxx));
    }
}

// Exact covered x-range on row y for CLS_SPANS triangles: each edge function
// w_i(x) = fl(P_i - fl(B_i * fl(x - a_ix))) is monotone in x (rounding is
// monotone), so {x : w_i(x) >= 0} is a prefix (B_i > 0), a suffix (B_i < 0) or
// all/nothing (B_i == 0) of the row; binary search finds each boundary.
// Returns false when the row has no covered pixel.
__device__ __forceinline__ bool edge_bound(float A, float B, float ax, float ay, float py, int xa, int xb,
                                           int& lo, int& hi)
{
    float P = A * (py - ay);
    if (B == 0.0f) {
        float w = P - B * ((float)xa - ax);
        return w >= 0.0f;
    }
    if (B > 0.0f) {
        // non-increasing: largest x in [lo,hi] with w >= 0
        int l = lo, h = hi;
        if (!(P - B * ((float)l - ax) >= 0.0f)) return false;
        while (l < h) {
            int m = l + ((h - l + 1) >> 1);
            if (P - B * ((float)m - ax) >= 0.0f) l = m; else h = m - 1;
        }
        hi = l;
        return true;
    }
    if (B < 0.0f) {
        // non-decreasing: smallest x in [lo,hi] with w >= 0
        int l = lo, h = hi;
        if (!(P - B * ((float)h - ax) >= 0.0f)) return false;
        while (l < h) {
            int m = l + ((h - l) >> 1);
            if (P - B * ((float)m - ax) >= 0.0f) h = m; else l = m + 1;
        }
        lo = l;
        return true;
    }
    return false;   // NaN slope: never covered
}

__device__ __forceinline__ bool row_span(const Tri& t, int y, int& lo, int& hi)
{
    float py = (float)y;
    lo = t.x0;
    hi = t.x1d - 1;
    if (!edge_bound(t.A0, t.B0, t.x2, t.y2, py, t.x0, t.x1d - 1, lo, hi)) return false;
    if (lo > hi) return false;
    if (!edge_bound(t.A1, t.B1, t.x3, t.y3, py, t.x0, t.x1d - 1, lo, hi)) return false;
    if (lo > hi) return false;
    if (!edge_bound(t.A2, t.B2, t.x1, t.y1, py, t.x0, t.x1d - 1, lo, hi)) return false;
    return lo <= hi;
}

// Candidate box [bx0,bx1] x [by0,by1] (logical pixels, inclusive) that holds
// every pixel of the triangle's domain the FP32 edge test can pass; false when
// there is none.  CLS_BOX / CLS_TIGHT use the proof in setup_class, CLS_SPANS the union of
// the exact row spans, CLS_FULL the whole domain.
__device__ __forceinline__ bool tri_box(const Tri& t, uint32_t cls, int& bx0, int& bx1, int& by0, int& by1)
{
    if (cls == CLS_EMPTY) return false;
    by0 = t.y0;
    by1 = t.y1d - 1;
    if (cls == CLS_BOX || cls == CLS_TIGHT) {
        box_x(t, cls == CLS_TIGHT, bx0, bx1);
        return bx0 <= bx1;
    }
    if (cls == CLS_FULL) {
        bx0 = t.x0;
        bx1 = t.x1d - 1;
        return true;
    }
    int lo_all = 1 << 30, hi_all = -1, first = -1, last = -1;
    for (int y = t.y0; y < t.y1d; ++y) {
        int lo, hi;
        if (!row_span(t, y, lo, hi)) continue;
        lo_all = min(lo_all, lo);
        hi_all = max(hi_all, hi);
        if (first < 0) first = y;
        last = y;
    }
    if (first < 0) return false;
    bx0 = lo_all;
    bx1 = hi_all;
    by0 = first;
    by1 = last;
    return true;
}

}  // namespace sloth
