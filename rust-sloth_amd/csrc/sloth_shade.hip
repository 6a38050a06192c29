// Differentiable shading of the visible fragments (added form; the reference
// shades with a flat glyph only, rasterizer.rs:76).
//
// Visibility is fixed: for every z slot, the raster's winner (triangle and
// pixel) is kept as (triangle << 1 | sub).  For that pixel, with the
// triangle's screen vertices and edge functions exactly as draw_triangle
// computes them (rasterizer.rs:56-74),
//   lambda_k = a * w_k                      (a = 1/orient, rasterizer.rs:67)
//   I        = L . sum_k lambda_k G n[v_k]  (G = cof(M3) / |det M3|^(2/3): the
//                                            model-to-screen normal map, scale-free)
//   C        = sum_k lambda_k c[v_k]
//   out      = I * C                         (RGB, float)
//   loss     = sum (out - target)^2          (accumulated in double)
// With n[v_k] = the face's unit normal and L = (0,0,1), I equals the flat
// shade nz_unit * a * (w0 + w1 + w2) up to rounding order.
//
// Backward: per fragment, with g = 2 (out - target) and dI = g . C,
//   d c[v_k] += lambda_k I g,  d n[v_k] += dI lambda_k G^T L,  d L += dI n.
// Per-vertex sums are formed per tile of z entries in LDS and added to the
// vertices' binary64 accumulators with hardware atomics, then converted to
// binary32.
#include "sloth_shade.h"

namespace sloth {

constexpr int SHADE_BLOCK = 256;

__device__ __forceinline__ const Draw& shade_draw(const Draw* __restrict__ draws, const Frame& fr, uint32_t tri)
{
    uint32_t lo = fr.draw_begin, hi = fr.draw_end - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (draws[mid].tri_begin <= tri) lo = mid; else hi = mid - 1;
    }
    return draws[lo];
}

// What every fragment of triangle T (frame fr) uses: its screen vertices,
// edge deltas and 1/orient as draw_triangle computes them (rasterizer.rs:56-67),
// its vertex ids, the normal map G applied to each vertex normal, the vertex
// colours, and G^T L (the light in model space for this view).
struct ShadeTri {
    float x1, y1, x2, y2, x3, y3, A0, B0, A1, B1, A2, B2, a;
    float gn[3][3];    // G n[v_k]
    float col[3][3];   // c[v_k]
    float gl[3];       // G^T L
    uint32_t vid[3];
};

__device__ __forceinline__ void shade_tri(const ShadeArgs& p, const Frame& fr, uint32_t T, ShadeTri& s)
{
    const uint32_t d = (uint32_t)(&shade_draw(p.draws, fr, T) - p.draws);
    const Draw& dr = p.draws[d];
    const uint32_t li = T - dr.tri_begin;
    Tri t;
    setup_draw_tri(dr, li, fr.W, fr.H, t);
    s.x1 = t.x1;
    s.y1 = t.y1;
    s.x2 = t.x2;
    s.y2 = t.y2;
    s.x3 = t.x3;
    s.y3 = t.y3;
    s.A0 = t.A0;
    s.B0 = t.B0;
    s.A1 = t.A1;
    s.B1 = t.B1;
    s.A2 = t.A2;
    s.B2 = t.B2;
    s.a = t.a;
    const float* G = p.G + (size_t)d * 9;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t vid = p.faces[(size_t)li * 3 + k];
        s.vid[k] = vid;
        const float* nk = p.vnormal + (size_t)vid * 3;
        const float* ck = p.vcolor + (size_t)vid * 3;
        const float a0 = nk[0], a1 = nk[1], a2 = nk[2];
        s.gn[k][0] = G[0] * a0 + G[1] * a1 + G[2] * a2;
        s.gn[k][1] = G[3] * a0 + G[4] * a1 + G[5] * a2;
        s.gn[k][2] = G[6] * a0 + G[7] * a1 + G[8] * a2;
        s.col[k][0] = ck[0];
        s.col[k][1] = ck[1];
        s.col[k][2] = ck[2];
    }
    const float L0 = p.light[0], L1 = p.light[1], L2 = p.light[2];
    s.gl[0] = G[0] * L0 + G[3] * L1 + G[6] * L2;
    s.gl[1] = G[1] * L0 + G[4] * L1 + G[7] * L2;
    s.gl[2] = G[2] * L0 + G[5] * L1 + G[8] * L2;
}

// One visible fragment of triangle s at pixel (px, py): the barycentric
// weights, the interpolated normal (mapped by G) and colour, and the shade I.
struct Fragment {
    float lam[3];
    float n0, n1, n2, c0, c1, c2, I;
};

__device__ __forceinline__ Fragment shade_at(const ShadeArgs& p, const ShadeTri& s, float px, float py)
{
    Fragment f;
    // the expressions of edges() (rasterizer.rs:71-74)
    const float w0 = s.A0 * (py - s.y2) - s.B0 * (px - s.x2);
    const float w1 = s.A1 * (py - s.y3) - s.B1 * (px - s.x3);
    const float w2 = s.A2 * (py - s.y1) - s.B2 * (px - s.x1);
    f.lam[0] = s.a * w0;
    f.lam[1] = s.a * w1;
    f.lam[2] = s.a * w2;
    f.n0 = f.n1 = f.n2 = f.c0 = f.c1 = f.c2 = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        f.n0 += f.lam[k] * s.gn[k][0];
        f.n1 += f.lam[k] * s.gn[k][1];
        f.n2 += f.lam[k] * s.gn[k][2];
        f.c0 += f.lam[k] * s.col[k][0];
        f.c1 += f.lam[k] * s.col[k][1];
        f.c2 += f.lam[k] * s.col[k][2];
    }
    f.I = p.light[0] * f.n0 + p.light[1] * f.n1 + p.light[2] * f.n2;
    return f;
}

// The pixel of z entry (row, col) of a frame of width W whose winner has
// `sub` (key_pixel without the division).
__device__ __forceinline__ void entry_pixel(uint32_t row, uint32_t col, uint32_t sub, uint32_t W, float& px, float& py)
{
    px = (float)(sub ? col >> 1 : (col + W) >> 1);
    py = (float)(sub ? row : row - 1);
}

// Block sums of up to four doubles in a fixed order, written to partial[block * 4 + k].
__device__ __forceinline__ void block_partials(double v0, double v1, double v2, double v3, double* partial,
                                               uint32_t block)
{
    __shared__ double s_red[4][SHADE_BLOCK / 64];
    for (int sh = 32; sh >= 1; sh >>= 1) {
        v0 += __shfl_xor(v0, sh);
        v1 += __shfl_xor(v1, sh);
        v2 += __shfl_xor(v2, sh);
        v3 += __shfl_xor(v3, sh);
    }
    const int wv = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        s_red[0][wv] = v0;
        s_red[1][wv] = v1;
        s_red[2][wv] = v2;
        s_red[3][wv] = v3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int w = 0; w < SHADE_BLOCK / 64; ++w) s += s_red[threadIdx.x][w];
        partial[(size_t)block * 4 + threadIdx.x] = s;
    }
}

// Forward: one thread per z slot; output colours and per-block loss sums.
__global__ void __launch_bounds__(SHADE_BLOCK) k_shade_fwd(ShadeArgs p)
{
    const uint64_t g = (uint64_t)blockIdx.x * SHADE_BLOCK + threadIdx.x;
    double loss = 0.0;
    if (g < p.n_slots) {
        uint32_t lo = 0, hi = p.n_frames - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (p.frames[mid].z_off <= g) lo = mid; else hi = mid - 1;
        }
        const int32_t v = p.vis[g];
        if (v >= 0) {
            const Frame& fr = p.frames[lo];
            const uint32_t id = (uint32_t)(g - fr.z_off), row = id / fr.W;
            ShadeTri st;
            shade_tri(p, fr, (uint32_t)v >> 1, st);
            float px, py;
            entry_pixel(row, id - row * fr.W, (uint32_t)v & 1u, fr.W, px, py);
            const Fragment f = shade_at(p, st, px, py);
            const float o0 = f.I * f.c0, o1 = f.I * f.c1, o2 = f.I * f.c2;
            if (p.out) {
                p.out[g * 3 + 0] = o0;
                p.out[g * 3 + 1] = o1;
                p.out[g * 3 + 2] = o2;
            }
            if (p.target) {
                const float r0 = o0 - p.target[g * 3 + 0];
                const float r1 = o1 - p.target[g * 3 + 1];
                const float r2 = o2 - p.target[g * 3 + 2];
                loss = (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
            }
        }
    }
    block_partials(loss, 0.0, 0.0, 0.0, p.partial, blockIdx.x);
}

constexpr uint32_t BWD_TILE_W = 64, BWD_TILE_H = 16;   // z entries per backward tile
constexpr uint32_t BWD_TABLE = 256;     // LDS rows for the tile's distinct vertices
constexpr uint32_t BWD_PROBES = 32;     // a vertex that finds no row adds to HBM directly
constexpr uint64_t BWD_EMPTY = ~0ull;
constexpr uint32_t BWD_RUN = 4;         // consecutive z entries of a row per thread
constexpr uint32_t BWD_GROUP = 2;       // run entries whose loads are issued together
constexpr uint32_t BWD_ROW = 7;         // 64-bit LDS words per vertex row (6 terms, padded)
static_assert(BWD_TILE_W % BWD_RUN == 0 && SHADE_BLOCK / (BWD_TILE_W / BWD_RUN) == BWD_TILE_H,
              "the block's runs cover one tile in one pass");
static_assert(BWD_TABLE == SHADE_BLOCK, "one thread per table row at the end of a tile");

// Vertex rows hold their 6 sums in fixed point: term t is added as the
// integer round(t 2^(FIX_BITS - e)), with e fixed per row when the row is
// claimed (the claiming sums' largest binary exponent + FIX_HEADROOM), so
// that |integer| < 2^FIX_BITS for every |t| < 2^e and 2^(63 - FIX_BITS)
// adds cannot overflow (a row takes at most one add per z entry of the
// tile, 1024).  A term outside the row's range (or not finite) goes to the
// binary64 accumulator directly.  The rounding error of an add is at most
// 2^(e - FIX_BITS - 1), i.e. 2^-39 of the claiming sums' magnitude.  LDS has
// integer atomics at full rate, while a float atomic add takes several
// cycles per lane.
constexpr int FIX_BITS = 50;
constexpr int FIX_HEADROOM = 12;
static_assert(BWD_TILE_W * BWD_TILE_H <= (1u << (63 - FIX_BITS)), "a row's adds stay within int64");

__device__ __forceinline__ int fix_exponent(const float* t)
{
    int e = -100;
#pragma unroll
    for (int m = 0; m < 6; ++m)
        if (t[m] != 0.0f) e = max(e, __builtin_amdgcn_frexp_expf(t[m]));
    return min(e, 100) + FIX_HEADROOM;
}

// Backward: a block takes BWD_TILE_W x BWD_TILE_H tiles of z entries (tile j of the
// job; frames hold tile_begin[f] .. tile_begin[f+1]), grid-striding over the
// tiles, each thread a run of BWD_RUN entries of a row (every other entry
// for even widths, whose odd entries hold no winner).  The run's visibility
// words, then the targets of its fragments, are loaded BWD_GROUP entries at
// a time with their loads in flight together.  A fragment's 6 colour /
// normal terms for each of its three vertices are summed over the run's
// consecutive fragments of one triangle, then go into an LDS table keyed by
// vertex, in fixed point (above); after the tile, each table row adds its
// sums to the vertex's binary64 accumulators with one global atomic per
// term.  Several pixels of a tile share each vertex, so this takes most
// same-address atomics off the memory system.  A vertex that finds no free
// row adds its terms to the accumulators directly.  The loss and the light
// gradient are summed per block.
__global__ void __launch_bounds__(SHADE_BLOCK) k_shade_bwd(ShadeArgs p, const uint64_t* __restrict__ tile_begin,
                                                           uint64_t n_tiles)
{
    // key: exponent << 32 | vertex id (set together by the claiming CAS)
    __shared__ uint64_t s_key[BWD_TABLE];
    __shared__ uint64_t s_acc[BWD_TABLE * BWD_ROW];
    double loss = 0.0, dl0 = 0.0, dl1 = 0.0, dl2 = 0.0;
    for (uint64_t j = blockIdx.x; j < n_tiles; j += gridDim.x) {
        s_key[threadIdx.x] = BWD_EMPTY;
        for (uint32_t i = threadIdx.x; i < BWD_TABLE * BWD_ROW; i += SHADE_BLOCK) s_acc[i] = 0;
        __syncthreads();
        uint32_t lo = 0, hi = p.n_frames - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (tile_begin[mid] <= j) lo = mid; else hi = mid - 1;
        }
        const Frame& fr = p.frames[lo];
        const uint32_t W = fr.W;
        const uint32_t tiles_x = (W + BWD_TILE_W - 1) / BWD_TILE_W;
        const uint32_t local = (uint32_t)(j - tile_begin[lo]);
        const uint32_t ty = local / tiles_x, tx = local - ty * tiles_x;
        // Each thread takes BWD_RUN consecutive z entries of a row.  Their
        // terms are summed in registers while they share a triangle and go
        // to the table when the triangle changes or the run ends: adjacent
        // pixels mostly share a triangle.
        constexpr uint32_t RUNS = BWD_TILE_W / BWD_RUN;   // threads per tile row
        const uint32_t col0 = tx * BWD_TILE_W + (threadIdx.x % RUNS) * BWD_RUN;
        const uint32_t cstep = (W & 1u) ? 1u : 2u;        // runs start on even columns
        int32_t cur = -1;   // triangle of the sums below and of st, or none
        ShadeTri st;
        float acc[3][6];
        // adds the run's sums for triangle `cur` to the table: the three
        // vertices' first probes are issued together
        auto flush = [&]() {
            uint64_t mine[3], old[3];
            uint32_t h[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                mine[k] = ((uint64_t)(uint32_t)fix_exponent(acc[k]) << 32) | st.vid[k];
                h[k] = (st.vid[k] * 0x9E3779B1u) >> 24;   // BWD_TABLE = 256 rows
                old[k] = atomicCAS((unsigned long long*)&s_key[h[k]], (unsigned long long)BWD_EMPTY,
                                   (unsigned long long)mine[k]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const uint32_t vid = st.vid[k];
                bool found = old[k] == BWD_EMPTY || (uint32_t)old[k] == vid;
                for (uint32_t probe = 1; !found && probe < BWD_PROBES; ++probe) {
                    h[k] = (h[k] + 1) & (BWD_TABLE - 1);
                    old[k] = atomicCAS((unsigned long long*)&s_key[h[k]], (unsigned long long)BWD_EMPTY,
                                       (unsigned long long)mine[k]);
                    found = old[k] == BWD_EMPTY || (uint32_t)old[k] == vid;
                }
                // the row's exponent: the claimer's (this lane's when it claimed the row)
                const int e = (int)(uint32_t)((old[k] == BWD_EMPTY ? mine[k] : old[k]) >> 32);
                double* dst = p.vert_acc + (size_t)vid * 6;
                const float lim = __builtin_ldexpf(1.0f, e);
                const double sc = __builtin_ldexp(1.0, FIX_BITS - e);
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    const float t = acc[k][m];
                    if (found && fabsf(t) < lim)
                        atomicAdd((unsigned long long*)&s_acc[h[k] * BWD_ROW + m],
                                  (unsigned long long)__double2ll_rn((double)t * sc));
                    else
                        unsafeAtomicAdd(dst + m, (double)t);
                }
            }
        };
        for (uint32_t row = ty * BWD_TILE_H + threadIdx.x / RUNS; row < min(fr.H, (ty + 1) * BWD_TILE_H);
             row += SHADE_BLOCK / RUNS) {
            const uint64_t g_row = fr.z_off + (uint64_t)row * W;
            for (uint32_t b = 0; b < BWD_RUN; b += BWD_GROUP * cstep) {
                int32_t vv[BWD_GROUP];
#pragma unroll
                for (uint32_t i = 0; i < BWD_GROUP; ++i) {
                    const uint32_t col = col0 + b + i * cstep;
                    vv[i] = col < W ? p.vis[g_row + col] : -1;
                }
                float tg[BWD_GROUP][3];
#pragma unroll
                for (uint32_t i = 0; i < BWD_GROUP; ++i) {
                    const float* t = p.target + (g_row + col0 + b + i * cstep) * 3;
                    tg[i][0] = vv[i] >= 0 ? t[0] : 0.0f;
                    tg[i][1] = vv[i] >= 0 ? t[1] : 0.0f;
                    tg[i][2] = vv[i] >= 0 ? t[2] : 0.0f;
                }
#pragma unroll
                for (uint32_t i = 0; i < BWD_GROUP; ++i) {
                    const int32_t v = vv[i];
                    if (v < 0) continue;
                    // a thread's run mostly stays on one triangle: its
                    // set-up, vertex loads and the normal map are redone
                    // only when the triangle changes
                    const int32_t T = v >> 1;
                    if (T != cur) {
                        if (cur >= 0) flush();
                        cur = T;
                        shade_tri(p, fr, (uint32_t)T, st);
#pragma unroll
                        for (int k = 0; k < 3; ++k)
#pragma unroll
                            for (int m = 0; m < 6; ++m) acc[k][m] = 0.0f;
                    }
                    float px, py;
                    entry_pixel(row, col0 + b + i * cstep, (uint32_t)v & 1u, W, px, py);
                    const Fragment f = shade_at(p, st, px, py);
                    const float o0 = f.I * f.c0, o1 = f.I * f.c1, o2 = f.I * f.c2;
                    const float r0 = o0 - tg[i][0];
                    const float r1 = o1 - tg[i][1];
                    const float r2 = o2 - tg[i][2];
                    loss += (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;   // as the forward pass sums it
                    const float g0 = 2.f * r0, g1 = 2.f * r1, g2 = 2.f * r2;
                    const float dI = g0 * f.c0 + g1 * f.c1 + g2 * f.c2;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float lk = f.lam[k];
                        acc[k][0] += lk * f.I * g0;
                        acc[k][1] += lk * f.I * g1;
                        acc[k][2] += lk * f.I * g2;
                        acc[k][3] += dI * lk * st.gl[0];
                        acc[k][4] += dI * lk * st.gl[1];
                        acc[k][5] += dI * lk * st.gl[2];
                    }
                    dl0 += (double)(dI * f.n0);
                    dl1 += (double)(dI * f.n1);
                    dl2 += (double)(dI * f.n2);
                }
            }
            if (cur >= 0) flush();
            cur = -1;
        }
        __syncthreads();
        {
            const uint64_t key = s_key[threadIdx.x];
            if (key != BWD_EMPTY) {
                const uint32_t vid = (uint32_t)key;
                const double sc = __builtin_ldexp(1.0, (int)(uint32_t)(key >> 32) - FIX_BITS);
                double* dst = p.vert_acc + (size_t)vid * 6;
#pragma unroll
                for (int m = 0; m < 6; ++m)
                    unsafeAtomicAdd(dst + m, (double)(int64_t)s_acc[threadIdx.x * BWD_ROW + m] * sc);
            }
        }
        __syncthreads();   // the table is cleared for the next tile
    }
    block_partials(loss, dl0, dl1, dl2, p.partial, blockIdx.x);
}

__global__ void k_fill_f64(double* __restrict__ a, uint64_t n)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = 0.0;
}

// The vertices' binary64 sums as the binary32 gradients.
__global__ void k_shade_finish(const double* __restrict__ vert_acc, uint32_t n_verts, float* __restrict__ d_color,
                               float* __restrict__ d_normal)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_verts) return;
    const double* a = vert_acc + (size_t)v * 6;
    d_color[v * 3 + 0] = (float)a[0];
    d_color[v * 3 + 1] = (float)a[1];
    d_color[v * 3 + 2] = (float)a[2];
    d_normal[v * 3 + 0] = (float)a[3];
    d_normal[v * 3 + 1] = (float)a[4];
    d_normal[v * 3 + 2] = (float)a[5];
}

uint32_t shade_blocks(uint64_t n_slots) { return (uint32_t)((n_slots + SHADE_BLOCK - 1) / SHADE_BLOCK); }

// the backward's workgroups: one per tile, at most BWD_MAX_BLOCKS (which then
// stride over the tiles), never more than the forward's partial sums
constexpr uint32_t BWD_MAX_BLOCKS = 8192;
uint32_t bwd_blocks(uint64_t n_slots, uint64_t n_bwd_tiles)
{
    return (uint32_t)std::min<uint64_t>({n_bwd_tiles, BWD_MAX_BLOCKS, shade_blocks(n_slots)});
}

void bwd_tiles(const Frame* frames, uint32_t n_frames, uint64_t* begin)
{
    begin[0] = 0;
    for (uint32_t f = 0; f < n_frames; ++f) {
        const uint64_t tx = (frames[f].W + BWD_TILE_W - 1) / BWD_TILE_W, ty = (frames[f].H + BWD_TILE_H - 1) / BWD_TILE_H;
        begin[f + 1] = begin[f] + tx * ty;
    }
}

int launch_shade(const ShadeArgs& a, bool backward, hipStream_t s)
{
    const uint32_t nb = shade_blocks(a.n_slots);
    if (nb == 0) return SLOTH_OK;
    if (backward) {
        const uint64_t n_acc = (uint64_t)a.n_verts * 6;
        hipLaunchKernelGGL(k_fill_f64, dim3(std::max(1u, std::min(4096u, (uint32_t)((n_acc + 255) / 256)))),
                           dim3(256), 0, s, a.vert_acc, n_acc);
        HIP_TRY(hipGetLastError());
        const uint32_t g = bwd_blocks(a.n_slots, a.n_bwd_tiles);
        if (g) hipLaunchKernelGGL(k_shade_bwd, dim3(g), dim3(SHADE_BLOCK), 0, s, a, a.bwd_tile_begin, a.n_bwd_tiles);
    } else {
        hipLaunchKernelGGL(k_shade_fwd, dim3(nb), dim3(SHADE_BLOCK), 0, s, a);
    }
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

int launch_shade_finish(const ShadeArgs& a, float* d_color, float* d_normal, hipStream_t s)
{
    if (a.n_verts == 0) return SLOTH_OK;
    hipLaunchKernelGGL(k_shade_finish, dim3((a.n_verts + 255) / 256), dim3(256), 0, s, a.vert_acc, a.n_verts, d_color,
                       d_normal);
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

}  // namespace sloth
