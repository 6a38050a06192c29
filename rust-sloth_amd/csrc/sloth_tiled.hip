// Tiled raster path (even widths): bin-to-tile and LDS tile raster.
// The data flow is described in sloth_tiled.h.
#include <algorithm>
#include <cstring>

#include "sloth_tiled.h"

namespace sloth {

constexpr int BIN_BLOCK = 256;
constexpr int RASTER_BLOCK = 64;

// ---------------------------------------------------------------------------
// k_tile_clear
// ---------------------------------------------------------------------------
__global__ void k_tile_clear(uint32_t* __restrict__ cursor, uint64_t n_tiles, int32_t* __restrict__ nl,
                             uint64_t n_rows, uint32_t* __restrict__ overflow, uint32_t n_frames)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = i0; i < n_tiles; i += stride) cursor[i] = 0;
    for (uint64_t i = i0; i < n_rows; i += stride) nl[i] = -1;
    for (uint64_t i = i0; i < n_frames; i += stride) overflow[i] = 0;
}

// ---------------------------------------------------------------------------
// k_bin
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t find_draw_t(const Draw* __restrict__ draws, uint32_t d_lo, uint32_t d_hi,
                                                uint64_t item)
{
    uint32_t lo = d_lo, hi = d_hi - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (draws[mid].item_begin <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

// Tile rectangles (inclusive, in tile units) that a candidate box reaches:
// the direct part (2x < W, cell (2x, y)) and the wrapped part (2x >= W, cell
// (2x - W, y + 1)).  An empty rectangle has x0 > x1.
struct TileRects {
    int dx0, dx1, dy0, dy1;
    int wx0, wx1, wy0, wy1;
};

__device__ __forceinline__ void tile_rects(int bx0, int bx1, int by0, int by1, uint32_t W, uint32_t TW,
                                           uint32_t TH, TileRects& r)
{
    const int half = (int)(W >> 1);
    int dxe = min(bx1, half - 1);
    if (bx0 <= dxe) {
        r.dx0 = (2 * bx0) / (int)TW;
        r.dx1 = (2 * dxe) / (int)TW;
        r.dy0 = by0 / (int)TH;
        r.dy1 = by1 / (int)TH;
    } else {
        r.dx0 = 1;
        r.dx1 = 0;
        r.dy0 = 1;
        r.dy1 = 0;
    }
    int wxs = max(bx0, half);
    if (wxs <= bx1) {
        r.wx0 = (2 * wxs - (int)W) / (int)TW;
        r.wx1 = (2 * bx1 - (int)W) / (int)TW;
        r.wy0 = (by0 + 1) / (int)TH;
        r.wy1 = (by1 + 1) / (int)TH;
    } else {
        r.wx0 = 1;
        r.wx1 = 0;
        r.wy0 = 1;
        r.wy1 = 0;
    }
}

__device__ __forceinline__ int rect_count(int x0, int x1, int y0, int y1)
{
    return (x0 <= x1 && y0 <= y1) ? (x1 - x0 + 1) * (y1 - y0 + 1) : 0;
}

// k-th tile of the lane's set (direct rectangle first, then the wrapped one);
// returns false for a wrapped tile that the direct rectangle already holds.
__device__ __forceinline__ bool tile_at(const TileRects& r, int k, int& tx, int& ty)
{
    int nd = rect_count(r.dx0, r.dx1, r.dy0, r.dy1);
    if (k < nd) {
        int w = r.dx1 - r.dx0 + 1;
        tx = r.dx0 + k % w;
        ty = r.dy0 + k / w;
        return true;
    }
    k -= nd;
    int w = r.wx1 - r.wx0 + 1;
    tx = r.wx0 + k % w;
    ty = r.wy0 + k / w;
    return !(nd > 0 && tx >= r.dx0 && tx <= r.dx1 && ty >= r.dy0 && ty <= r.dy1);
}

// COUNT = true is the sizing run: it only counts each tile's entries.
template <bool COUNT>
__global__ void __launch_bounds__(BIN_BLOCK) k_bin(const Draw* __restrict__ draws, uint32_t d_lo, uint32_t d_hi,
                                                   const Frame* __restrict__ frames, uint64_t item_base,
                                                   uint64_t n_items, uint32_t TW, uint32_t TH,
                                                   uint32_t* __restrict__ cursor, const uint64_t* __restrict__ off,
                                                   uint32_t* __restrict__ pool, int32_t* __restrict__ nl,
                                                   uint32_t* __restrict__ overflow)
{
    const uint64_t t_id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t_id < n_items;
    const int lane = (int)__lane_id();

    uint32_t tri = 0, frame = 0;
    uint64_t tile_begin = 0;
    uint32_t tiles_x = 0;
    int nt = 0;
    TileRects r;
    r.dx0 = r.wx0 = 1;
    r.dx1 = r.wx1 = 0;
    r.dy0 = r.wy0 = 1;
    r.dy1 = r.wy1 = 0;

    // Blocks rarely straddle two draws: if this one does not, the draw index is
    // uniform and its fields (the 4x4 matrix included) come in scalar loads.
    const uint64_t blk0 = item_base + (uint64_t)blockIdx.x * blockDim.x;
    const uint64_t blk1 = item_base + min((uint64_t)blockIdx.x * blockDim.x + blockDim.x, n_items) - 1;
    const uint32_t d_first = (uint32_t)__builtin_amdgcn_readfirstlane((int)find_draw_t(draws, d_lo, d_hi, blk0));
    const uint32_t d_last = (uint32_t)__builtin_amdgcn_readfirstlane((int)find_draw_t(draws, d_lo, d_hi, blk1));
    int y0 = 0, y1d = 0;
    if (valid) {
        uint64_t item = item_base + t_id;
        uint32_t d = d_first == d_last ? d_first : find_draw_t(draws, d_lo, d_hi, item);
        const Draw& dr = draws[d];
        uint32_t local = (uint32_t)(item - dr.item_begin);
        frame = dr.frame;
        const Frame& fr = frames[frame];
        Tri t;
        setup_tri(dr.M, dr.xyz + (size_t)local * 9, fr.W, fr.H, t);
        tri = dr.tri_begin + local;
        if (fr.image) {
            y0 = t.y0;
            y1d = t.y1d;
        }
        uint32_t cls = setup_class(t);
        int bx0, bx1, by0, by1;
        if (tri_box(t, cls, bx0, bx1, by0, by1)) {
            tile_rects(bx0, bx1, by0, by1, fr.W, TW, TH, r);
            nt = rect_count(r.dx0, r.dx1, r.dy0, r.dy1) + rect_count(r.wx0, r.wx1, r.wy0, r.wy1);
        }
        tile_begin = fr.tile_begin;
        tiles_x = fr.tiles_x;
    }

    // rasterizer.rs:89-91: a newline cell after every row of the domain; the
    // last triangle (in draw order) whose rows include r decides row r.  When
    // the wave is in one frame and its rows span < 64, one ballot per row finds
    // the highest lane (= latest triangle) covering it and one atomicMax
    // publishes it; otherwise every lane publishes its own rows.
    {
        const bool has_rows = valid && y0 < y1d;
        const uint64_t rmask = __ballot(has_rows);
        if (rmask) {
            const int l0 = __ffsll((unsigned long long)rmask) - 1;
            const uint32_t f0 = (uint32_t)__builtin_amdgcn_readlane((int)frame, l0);
            const bool same = __ballot(has_rows && frame != f0) == 0;
            int lo_r = has_rows ? y0 : (1 << 30), hi_r = has_rows ? y1d - 1 : -1;
            for (int sh = 32; sh >= 1; sh >>= 1) {
                lo_r = min(lo_r, __shfl_xor(lo_r, sh));
                hi_r = max(hi_r, __shfl_xor(hi_r, sh));
            }
            lo_r = __builtin_amdgcn_readfirstlane(lo_r);
            hi_r = __builtin_amdgcn_readfirstlane(hi_r);
            int32_t* fnl = nl + frames[f0].nl_off;
            if (same && hi_r - lo_r < 64) {
                for (int row = lo_r; row <= hi_r; ++row) {
                    uint64_t m = __ballot(has_rows && y0 <= row && row < y1d);
                    if (m) {
                        int top = 63 - __clzll((long long)m);
                        if (lane == top) atomicMax(&fnl[row], (int32_t)tri);
                    }
                }
            } else if (has_rows) {
                int32_t* mine = nl + frames[frame].nl_off;
                for (int y = y0; y < y1d; ++y) atomicMax(&mine[y], (int32_t)tri);
            }
        }
    }

    // Append `tri` to each of its tiles; lanes that share a tile share one
    // atomicAdd on that tile's cursor.
    int k = 0;
    for (;;) {
        int tx = 0, ty = 0;
        bool active = false;
        while (k < nt) {
            if (tile_at(r, k, tx, ty)) {
                active = true;
                break;
            }
            ++k;
        }
        uint64_t amask = __ballot(active);
        if (amask == 0) break;
        uint64_t key = active ? tile_begin + (uint64_t)ty * tiles_x + (uint64_t)tx : ~0ull;
        int leader = __ffsll((unsigned long long)amask) - 1;
        uint64_t lk = readlane64(key, leader);
        bool match = active && key == lk;
        uint64_t mmask = __ballot(match);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&cursor[lk], (uint32_t)__popcll(mmask));
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
        if (match) {
            uint32_t pos = base + (uint32_t)__popcll(mmask & __lanemask_lt());
            if (!COUNT) {
                uint64_t o = off[lk], cap = off[lk + 1] - o;
                if (pos < cap) pool[o + pos] = tri;
                else overflow[frame] = 1u;
            }
            ++k;
        }
    }
}

// ---------------------------------------------------------------------------
// k_tile_raster
// ---------------------------------------------------------------------------
__device__ __forceinline__ const Draw& draw_of_t(const Draw* __restrict__ draws, const Frame& fr, uint32_t tri)
{
    uint32_t lo = fr.draw_begin, hi = fr.draw_end - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (draws[mid].tri_begin <= tri) lo = mid; else hi = mid - 1;
    }
    return draws[lo];
}

struct TileGeom {
    int tc0, tc1, tr0, tr1;   // cell columns [tc0,tc1), rows [tr0,tr1)
    int sw;                   // slots per tile row (TW/2)
    int W;
};

// Calls f(x, y, slot, sub) for every logical pixel of the box that lands in
// the tile: the direct region (2x in [tc0,tc1), row y) and the wrapped region
// (2x - W in [tc0,tc1), row y + 1).
template <class Fn>
__device__ __forceinline__ void walk_tile(const TileGeom& g, int bx0, int bx1, int by0, int by1, Fn&& f)
{
    int xa = max(bx0, g.tc0 >> 1), xb = min(bx1, (g.tc1 >> 1) - 1);
    int ya = max(by0, g.tr0), yb = min(by1, g.tr1 - 1);
    for (int y = ya; y <= yb; ++y)
        for (int x = xa; x <= xb; ++x) f(x, y, (y - g.tr0) * g.sw + (x - (g.tc0 >> 1)), 1u);
    int wx0 = (g.tc0 + g.W) >> 1;
    xa = max(bx0, wx0);
    xb = min(bx1, ((g.tc1 + g.W) >> 1) - 1);
    ya = max(by0, g.tr0 - 1);
    yb = min(by1, g.tr1 - 2);
    for (int y = ya; y <= yb; ++y)
        for (int x = xa; x <= xb; ++x) f(x, y, (y + 1 - g.tr0) * g.sw + (x - wx0), 0u);
}

__global__ void __launch_bounds__(RASTER_BLOCK) k_tile_raster(
    const Draw* __restrict__ draws, const Ramp* __restrict__ ramps, const Frame* __restrict__ frames, uint32_t f_lo,
    uint32_t f_hi, uint64_t tile_base, uint32_t TW, uint32_t TH, const uint32_t* __restrict__ cursor,
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ pool, const int32_t* __restrict__ nl,
    uint32_t* __restrict__ cells, float* __restrict__ zbuf, int32_t* __restrict__ win)
{
    extern __shared__ uint64_t smem[];
    const uint32_t S = (TW >> 1) * TH;
    uint64_t* keys = smem;
    uint32_t* st_cell = (uint32_t*)(smem + S);
    float* st_z = (float*)(st_cell + S);

    const uint64_t gt = tile_base + blockIdx.x;
    uint32_t lo = f_lo, hi = f_hi - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (frames[mid].tile_begin <= gt) lo = mid; else hi = mid - 1;
    }
    const Frame& fr = frames[lo];
    const uint32_t W = fr.W, H = fr.H;
    const uint32_t local = (uint32_t)(gt - fr.tile_begin);
    const uint32_t tx = local % fr.tiles_x, ty = local / fr.tiles_x;
    TileGeom g;
    g.tc0 = (int)(tx * TW);
    g.tc1 = min((int)W, g.tc0 + (int)TW);
    g.tr0 = (int)(ty * TH);
    g.tr1 = min((int)H, g.tr0 + (int)TH);
    g.sw = (int)(TW >> 1);
    g.W = (int)W;

    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) keys[s] = KEY_EMPTY;
    const uint64_t o = off[gt];
    const uint32_t n = min(cursor[gt], (uint32_t)(off[gt + 1] - o));
    const uint32_t* list = pool + o;
    __syncthreads();

    const bool one_draw = fr.draw_end - fr.draw_begin == 1;
    auto draw_for = [&](uint32_t T) -> const Draw& {
        return one_draw ? draws[fr.draw_begin] : draw_of_t(draws, fr, T);
    };

    // pass 1: depth test of every candidate fragment
    auto test = [&](const Tri& t, uint32_t T, int bx0, int bx1, int by0, int by1) {
        walk_tile(g, bx0, bx1, by0, by1, [&](int x, int y, int slot, uint32_t sub) {
            float w0, w1, w2;
            edges(t, (float)x, (float)y, w0, w1, w2);
            if (!inside(w0, w1, w2)) return;
            float z = depth(t, w1, w2);
            if (!(z < F32_MAX)) return;
            atomicMin((unsigned long long*)&keys[slot], (unsigned long long)make_key(z, T, sub));
        });
    };
    // pass 2: the fragment whose key won stages its cell and depth
    auto stage = [&](const Tri& t, uint32_t T, const Draw& dr, int bx0, int bx1, int by0, int by1) {
        const uint32_t li = T - dr.tri_begin;
        bool have_nz = false;
        float nzu = 0.0f;
        uint32_t color = 0;
        walk_tile(g, bx0, bx1, by0, by1, [&](int x, int y, int slot, uint32_t sub) {
            float w0, w1, w2;
            edges(t, (float)x, (float)y, w0, w1, w2);
            if (!inside(w0, w1, w2)) return;
            float z = depth(t, w1, w2);
            if (!(z < F32_MAX)) return;
            if (keys[slot] != make_key(z, T, sub)) return;
            if (!have_nz) {
                nzu = unit_normal_z(t);
                const uint8_t* c = dr.rgb + (size_t)li * 3;
                color = ((uint32_t)c[0] << 8) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 24);
                have_nz = true;
            }
            float s = shade_value(nzu, t.a, w0, w1, w2);
            st_cell[slot] = (uint32_t)ramp_glyph(ramps[dr.ramp], s) | color;
            st_z[slot] = z;
        });
    };

    // The first entry of each thread stays in registers for both passes.
    Tri t0;
    uint32_t T0 = 0;
    int c0x0 = 0, c0x1 = -1, c0y0 = 0, c0y1 = -1;
    bool ok0 = false;
    if (threadIdx.x < n) {
        T0 = list[threadIdx.x];
        const Draw& dr = draw_for(T0);
        setup_tri(dr.M, dr.xyz + (size_t)(T0 - dr.tri_begin) * 9, W, H, t0);
        ok0 = tri_box(t0, setup_class(t0), c0x0, c0x1, c0y0, c0y1);
    }
    if (ok0) test(t0, T0, c0x0, c0x1, c0y0, c0y1);
    for (uint32_t e = threadIdx.x + blockDim.x; e < n; e += blockDim.x) {
        const uint32_t T = list[e];
        const Draw& dr = draw_for(T);
        Tri t;
        setup_tri(dr.M, dr.xyz + (size_t)(T - dr.tri_begin) * 9, W, H, t);
        int bx0, bx1, by0, by1;
        if (tri_box(t, setup_class(t), bx0, bx1, by0, by1)) test(t, T, bx0, bx1, by0, by1);
    }
    __syncthreads();
    if (ok0) stage(t0, T0, draw_for(T0), c0x0, c0x1, c0y0, c0y1);
    for (uint32_t e = threadIdx.x + blockDim.x; e < n; e += blockDim.x) {
        const uint32_t T = list[e];
        const Draw& dr = draw_for(T);
        Tri t;
        setup_tri(dr.M, dr.xyz + (size_t)(T - dr.tri_begin) * 9, W, H, t);
        int bx0, bx1, by0, by1;
        if (tri_box(t, setup_class(t), bx0, bx1, by0, by1)) stage(t, T, dr, bx0, bx1, by0, by1);
    }
    __syncthreads();

    // write-out: every slot owns cells (id, id+1) and z slots (id, id+1)
    const int32_t* fnl = nl + fr.nl_off;
    uint32_t* fc = cells + fr.cell_off;
    float* fz = zbuf + fr.z_off;
    for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
        int sr = (int)s / g.sw, sc = (int)s - sr * g.sw;
        int row = g.tr0 + sr, col = g.tc0 + 2 * sc;
        if (row >= g.tr1 || col >= g.tc1) continue;
        uint64_t k = keys[s];
        uint32_t c0 = CELL_BLANK, c1 = CELL_BLANK;
        float z0 = F32_MAX;
        int32_t wt = -1;
        if (k != KEY_EMPTY) {
            c0 = c1 = st_cell[s];
            z0 = st_z[s];
            wt = (int32_t)((uint32_t)k >> 1);
        }
        if (fr.image && col == 0) {
            // cell row*W + 1: the newline of row `row` is written after every
            // pixel of its triangle's row and before its next row, so it
            // overwrites a wrapped pixel (drawn in row-1) of the same or an
            // earlier triangle
            int32_t nlr = fnl[row];
            if (nlr >= 0 && (k == KEY_EMPTY || nlr >= wt)) c1 = CELL_NEWLINE;
        }
        uint64_t id = (uint64_t)row * W + (uint64_t)col;
        *(uint2*)(fc + id) = make_uint2(c0, c1);
        *(float2*)(fz + id) = make_float2(z0, F32_MAX);
        if (win) *(int2*)(win + fr.z_off + id) = make_int2(wt, -1);
    }
    // image mode: the H cells after the W*H grid are never written (context.rs:38-40)
    if (fr.image && local == 0)
        for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) fc[(uint64_t)W * H + i] = CELL_BLANK;
}

// ---------------------------------------------------------------------------
// TiledPlan
// ---------------------------------------------------------------------------
static unsigned blocks(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

int TiledPlan::plan(std::vector<Frame>& frames, uint64_t rows)
{
    uint64_t tiles = 0;
    for (Frame& f : frames) {
        f.tiles_x = (f.W + TW - 1) / TW;
        f.tiles_y = (f.H + TH - 1) / TH;
        f.tile_begin = tiles;
        tiles += (uint64_t)f.tiles_x * f.tiles_y;
    }
    if (tiles != n_tiles) sized = false;
    n_tiles = tiles;
    n_rows = rows;
    int rc;
    if ((rc = d_cursor.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, n_tiles)))) return rc;
    if ((rc = d_off.reserve(sizeof(uint64_t) * (n_tiles + 1)))) return rc;
    if ((rc = d_nl.reserve(sizeof(int32_t) * std::max<uint64_t>(1, n_rows)))) return rc;
    if ((rc = d_overflow.reserve(sizeof(uint32_t) * std::max<size_t>(1, frames.size())))) return rc;
    return SLOTH_OK;
}

int TiledPlan::launch_clear(const std::vector<Frame>& frames, const Frame* d_frames, uint32_t f_lo, uint32_t f_hi,
                            uint32_t* cells, hipStream_t s)
{
    (void)d_frames;
    (void)cells;
    const Frame& a = frames[f_lo];
    const Frame& b = frames[f_hi - 1];
    uint64_t t_lo = a.tile_begin, t_hi = b.tile_begin + (uint64_t)b.tiles_x * b.tiles_y;
    uint64_t r_lo = a.nl_off, r_hi = b.nl_off + b.H;
    uint64_t n = std::max(t_hi - t_lo, r_hi - r_lo);
    unsigned g = std::max(1u, std::min(blocks(n, 256), 4096u));
    hipLaunchKernelGGL(k_tile_clear, dim3(g), dim3(256), 0, s, d_cursor.as<uint32_t>() + t_lo, t_hi - t_lo,
                       d_nl.as<int32_t>() + r_lo, r_hi - r_lo, d_overflow.as<uint32_t>() + f_lo, f_hi - f_lo);
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

int TiledPlan::launch_bin(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
                          const Draw* d_draws, uint32_t f_lo, uint32_t f_hi, uint64_t n_items, bool count_only,
                          hipStream_t s)
{
    uint32_t d_lo = frames[f_lo].draw_begin, d_hi = frames[f_hi - 1].draw_end;
    uint64_t it_lo = d_lo < draws.size() ? draws[d_lo].item_begin : n_items;
    uint64_t it_hi = d_hi < draws.size() ? draws[d_hi].item_begin : n_items;
    if (it_hi <= it_lo) return SLOTH_OK;
    if (count_only)
        hipLaunchKernelGGL(k_bin<true>, dim3(blocks(it_hi - it_lo, BIN_BLOCK)), dim3(BIN_BLOCK), 0, s, d_draws, d_lo,
                           d_hi, d_frames, it_lo, it_hi - it_lo, TW, TH, d_cursor.as<uint32_t>(), d_off.as<uint64_t>(),
                           d_pool.as<uint32_t>(), d_nl.as<int32_t>(), d_overflow.as<uint32_t>());
    else
        hipLaunchKernelGGL(k_bin<false>, dim3(blocks(it_hi - it_lo, BIN_BLOCK)), dim3(BIN_BLOCK), 0, s, d_draws, d_lo,
                           d_hi, d_frames, it_lo, it_hi - it_lo, TW, TH, d_cursor.as<uint32_t>(), d_off.as<uint64_t>(),
                           d_pool.as<uint32_t>(), d_nl.as<int32_t>(), d_overflow.as<uint32_t>());
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

int TiledPlan::size(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
                    const Draw* d_draws, uint64_t n_items, hipStream_t s)
{
    if (frames.empty()) return SLOTH_OK;
    const uint32_t F = (uint32_t)frames.size();
    int rc;
    if ((rc = launch_clear(frames, d_frames, 0, F, nullptr, s))) return rc;
    if ((rc = launch_bin(frames, draws, d_frames, d_draws, 0, F, n_items, true, s))) return rc;
    std::vector<uint32_t> counts(n_tiles);
    if (n_tiles)
        HIP_TRY(hipMemcpyAsync(counts.data(), d_cursor.p, sizeof(uint32_t) * n_tiles, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    off.assign(n_tiles + 1, 0);
    for (uint64_t t = 0; t < n_tiles; ++t) {
        uint64_t c = counts[t];
        uint64_t cap = c ? c + c / 8 + 16 : 0;
        off[t + 1] = off[t] + cap;
    }
    if ((rc = d_pool.reserve(sizeof(uint32_t) * std::max<uint64_t>(1, off[n_tiles])))) return rc;
    HIP_TRY(hipMemcpyAsync(d_off.p, off.data(), sizeof(uint64_t) * (n_tiles + 1), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    sized = true;
    ++generation;
    return SLOTH_OK;
}

int TiledPlan::launch(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
                      const Draw* d_draws, const Ramp* d_ramps, uint32_t f_lo, uint32_t f_hi, uint32_t* cells,
                      float* z, int32_t* win, hipStream_t s, std::vector<hipEvent_t>* ev)
{
    if (!sized) return fail(SLOTH_ESTATE, "tile lists not sized");
    uint64_t n_items = draws.empty() ? 0 : draws.back().item_begin + draws.back().n;
    int rc;
    if (ev) HIP_TRY(hipEventRecord((*ev)[0], s));
    if ((rc = launch_clear(frames, d_frames, f_lo, f_hi, cells, s))) return rc;
    if (ev) HIP_TRY(hipEventRecord((*ev)[1], s));
    if ((rc = launch_bin(frames, draws, d_frames, d_draws, f_lo, f_hi, n_items, false, s))) return rc;
    if (ev) HIP_TRY(hipEventRecord((*ev)[2], s));
    const Frame& a = frames[f_lo];
    const Frame& b = frames[f_hi - 1];
    uint64_t t_lo = a.tile_begin, t_hi = b.tile_begin + (uint64_t)b.tiles_x * b.tiles_y;
    size_t lds = (size_t)(TW / 2) * TH * 16;
    if (t_hi > t_lo) {
        hipLaunchKernelGGL(k_tile_raster, dim3((unsigned)(t_hi - t_lo)), dim3(RASTER_BLOCK), lds, s, d_draws, d_ramps,
                           d_frames, f_lo, f_hi, t_lo, TW, TH, d_cursor.as<uint32_t>(), d_off.as<uint64_t>(),
                           d_pool.as<uint32_t>(), d_nl.as<int32_t>(), cells, z, win);
        HIP_TRY(hipGetLastError());
    }
    if (ev) HIP_TRY(hipEventRecord((*ev)[3], s));
    return SLOTH_OK;
}

int TiledPlan::overflowed(uint32_t nf, hipStream_t s, bool& ovf)
{
    ovf = false;
    std::vector<uint32_t> h(nf);
    if (nf) HIP_TRY(hipMemcpyAsync(h.data(), d_overflow.p, sizeof(uint32_t) * nf, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (uint32_t v : h) ovf |= (v != 0);
    return SLOTH_OK;
}

}  // namespace sloth
