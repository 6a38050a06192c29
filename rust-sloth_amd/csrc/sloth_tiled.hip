// Tiled raster path (even widths): bin-to-tile and LDS tile raster.
// The data flow is described in sloth_tiled.h.
#include <algorithm>
#include <cstring>

#include "sloth_tiled.h"

namespace sloth {

constexpr int BIN_BLOCK = 256;
constexpr int RASTER_BLOCK = 256;

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

// Pool entries: frame triangle index, plus this bit when the triangle's
// candidate box is the cheap CLS_BOX one (the raster then skips the proof).
constexpr uint32_t BOX_FLAG = 0x80000000u;

// Tile rectangles (inclusive, in tile units) that a candidate box reaches:
// the direct part (2x < W, cell (2x, y)) and the wrapped part (2x >= W, cell
// (2x - W, y + 1)).  An empty rectangle has x0 > x1.
struct TileRects {
    int dx0, dx1, dy0, dy1;
    int wx0, wx1, wy0, wy1;
};

__device__ __forceinline__ void tile_rects(int bx0, int bx1, int by0, int by1, uint32_t W, int tws, int ths,
                                           TileRects& r)
{
    const int half = (int)(W >> 1);
    int dxe = min(bx1, half - 1);
    if (bx0 <= dxe) {
        r.dx0 = (2 * bx0) >> tws;
        r.dx1 = (2 * dxe) >> tws;
        r.dy0 = by0 >> ths;
        r.dy1 = by1 >> ths;
    } else {
        r.dx0 = r.dy0 = 1;
        r.dx1 = r.dy1 = 0;
    }
    int wxs = max(bx0, half);
    if (wxs <= bx1) {
        r.wx0 = (2 * wxs - (int)W) >> tws;
        r.wx1 = (2 * bx1 - (int)W) >> tws;
        r.wy0 = (by0 + 1) >> ths;
        r.wy1 = (by1 + 1) >> ths;
    } else {
        r.wx0 = r.wy0 = 1;
        r.wx1 = r.wy1 = 0;
    }
}

// Walks the lane's tiles: the direct rectangle, then the wrapped one minus the
// tiles the direct rectangle already holds.  phase 2 = done.
struct TileIter {
    int tx, ty, phase;
};

__device__ __forceinline__ bool in_direct(const TileRects& r, int tx, int ty)
{
    return tx >= r.dx0 && tx <= r.dx1 && ty >= r.dy0 && ty <= r.dy1;
}

__device__ __forceinline__ void iter_settle(const TileRects& r, TileIter& it)
{
    for (;;) {
        if (it.phase == 0) {
            if (r.dx0 <= r.dx1 && it.ty <= r.dy1) return;
            it.phase = 1;
            it.tx = r.wx0;
            it.ty = r.wy0;
        }
        if (it.phase == 1) {
            if (!(r.wx0 <= r.wx1) || it.ty > r.wy1) {
                it.phase = 2;
                return;
            }
            if (!in_direct(r, it.tx, it.ty)) return;
            if (++it.tx > r.wx1) {
                it.tx = r.wx0;
                ++it.ty;
            }
            continue;
        }
        return;
    }
}

__device__ __forceinline__ void iter_next(const TileRects& r, TileIter& it)
{
    if (it.phase == 0) {
        if (++it.tx > r.dx1) {
            it.tx = r.dx0;
            ++it.ty;
        }
    } else if (it.phase == 1) {
        if (++it.tx > r.wx1) {
            it.tx = r.wx0;
            ++it.ty;
        }
    }
    iter_settle(r, it);
}

// Grid: x = blocks of one draw's triangles, y = draw index - d_lo, so the draw,
// its frame and its matrix are uniform per block (scalar loads).
template <bool COUNT>
__global__ void __launch_bounds__(BIN_BLOCK) k_bin(const Draw* __restrict__ draws, uint32_t d_lo,
                                                   const Frame* __restrict__ frames, int tws, int ths,
                                                   uint32_t* __restrict__ cursor, const uint64_t* __restrict__ off,
                                                   uint32_t* __restrict__ pool, int32_t* __restrict__ nl,
                                                   uint32_t* __restrict__ overflow)
{
    __shared__ int32_t s_nl[BIN_BLOCK];
    __shared__ int s_lo, s_hi;

    const Draw& dr = draws[d_lo + blockIdx.y];
    const uint32_t base = blockIdx.x * BIN_BLOCK;
    if (base >= dr.n) return;
    const uint32_t local = base + threadIdx.x;
    const bool valid = local < dr.n;
    const uint32_t frame = dr.frame;
    const Frame& fr = frames[frame];
    const uint32_t W = fr.W, H = fr.H;
    const uint32_t tri = dr.tri_begin + local;
    const int lane = (int)__lane_id();

    int y0 = 0, y1d = 0;
    uint32_t entry = tri;
    TileRects r;
    r.dx0 = r.dy0 = r.wx0 = r.wy0 = 1;
    r.dx1 = r.dy1 = r.wx1 = r.wy1 = 0;
    if (valid) {
        Tri t;
        setup_tri(dr.M, dr.xyz + (size_t)local * 9, W, H, t);
        y0 = t.y0;
        y1d = t.y1d;
        uint32_t cls = setup_class(t);
        if (cls == CLS_BOX) entry |= BOX_FLAG;
        int bx0, bx1, by0, by1;
        if (tri_box(t, cls, bx0, bx1, by0, by1)) tile_rects(bx0, bx1, by0, by1, W, tws, ths, r);
    }

    // rasterizer.rs:89-91: a newline cell after every row of the domain; the
    // last triangle (in draw order) whose rows include r decides row r.  The
    // block reduces its rows in LDS and publishes one atomicMax per row.
    if (fr.image) {
        if (threadIdx.x == 0) {
            s_lo = 1 << 30;
            s_hi = -1;
        }
        s_nl[threadIdx.x] = -1;
        __syncthreads();
        const bool has_rows = valid && y0 < y1d;
        if (has_rows) {
            atomicMin(&s_lo, y0);
            atomicMax(&s_hi, y1d - 1);
        }
        __syncthreads();
        const int lo = s_lo, hi = s_hi;
        int32_t* fnl = nl + fr.nl_off;
        if (hi >= lo) {
            if (hi - lo < BIN_BLOCK) {
                if (has_rows)
                    for (int y = y0; y < y1d; ++y) atomicMax(&s_nl[y - lo], (int32_t)tri);
                __syncthreads();
                const int32_t v = s_nl[threadIdx.x];
                if ((int)threadIdx.x <= hi - lo && v >= 0) atomicMax(&fnl[lo + (int)threadIdx.x], v);
            } else if (has_rows) {
                for (int y = y0; y < y1d; ++y) atomicMax(&fnl[y], (int32_t)tri);
            }
        }
    }

    // Append the entry to each of its tiles; lanes that share a tile share one
    // atomicAdd on that tile's cursor (ballot + prefix popcount).
    const uint64_t tile_begin = fr.tile_begin;
    const uint32_t tiles_x = fr.tiles_x;
    TileIter it;
    it.phase = 0;
    it.tx = r.dx0;
    it.ty = r.dy0;
    iter_settle(r, it);
    for (;;) {
        const bool active = it.phase < 2;
        const uint64_t amask = __ballot(active);
        if (amask == 0) break;
        const uint32_t key = active ? (uint32_t)it.ty * tiles_x + (uint32_t)it.tx : 0xFFFFFFFFu;
        const int leader = __ffsll((unsigned long long)amask) - 1;
        const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
        const bool match = active && key == lk;
        const uint64_t mmask = __ballot(match);
        const uint64_t gt = tile_begin + lk;
        uint32_t pos0 = 0;
        if (lane == leader) pos0 = atomicAdd(&cursor[gt], (uint32_t)__popcll(mmask));
        pos0 = (uint32_t)__builtin_amdgcn_readlane((int)pos0, leader);
        if (match) {
            const uint32_t pos = pos0 + (uint32_t)__popcll(mmask & __lanemask_lt());
            if (!COUNT) {
                const uint64_t o = off[gt], cap = off[gt + 1] - o;
                if (pos < cap) pool[o + pos] = entry;
                else overflow[frame] = 1u;
            }
            iter_next(r, it);
        }
    }
}

// ---------------------------------------------------------------------------
// k_tile_raster
// ---------------------------------------------------------------------------
constexpr uint32_t CACHE_CAP = 256;   // list length up to which the fast path runs
constexpr uint32_t HASH_CAP = 512;    // open-addressing table: triangle -> cache row
constexpr uint32_t CACHE_WORDS = 12;
constexpr uint32_t EMPTY32 = 0xFFFFFFFFu;

__device__ __forceinline__ const Draw& draw_of_t(const Draw* __restrict__ draws, const Frame& fr, uint32_t tri)
{
    uint32_t lo = fr.draw_begin, hi = fr.draw_end - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (draws[mid].tri_begin <= tri) lo = mid; else hi = mid - 1;
    }
    return draws[lo];
}

__device__ __forceinline__ uint32_t hash_tri(uint32_t T) { return (T * 0x9E3779B1u) >> 23; }   // 9 bits

struct TileGeom {
    int tc0, tc1, tr0, tr1;   // cell columns [tc0,tc1), rows [tr0,tr1)
    int sws;                  // log2(slots per tile row) = log2(TW/2)
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
        for (int x = xa; x <= xb; ++x) f(x, y, ((y - g.tr0) << g.sws) + (x - (g.tc0 >> 1)), 1u);
    int wx0 = (g.tc0 + g.W) >> 1;
    xa = max(bx0, wx0);
    xb = min(bx1, ((g.tc1 + g.W) >> 1) - 1);
    ya = max(by0, g.tr0 - 1);
    yb = min(by1, g.tr1 - 2);
    for (int y = ya; y <= yb; ++y)
        for (int x = xa; x <= xb; ++x) f(x, y, ((y + 1 - g.tr0) << g.sws) + (x - wx0), 0u);
}

// Candidate box of a pool entry.
__device__ __forceinline__ bool entry_box(const Tri& t, uint32_t e, int& bx0, int& bx1, int& by0, int& by1)
{
    if (e & BOX_FLAG) {
        bx0 = max(t.x0, (int)floorf(t.minx));
        bx1 = min(t.x1d - 1, (int)ceilf(t.maxx));
        by0 = t.y0;
        by1 = t.y1d - 1;
        return bx0 <= bx1 && by0 <= by1;
    }
    return tri_box(t, setup_class(t), bx0, bx1, by0, by1);
}

__global__ void __launch_bounds__(RASTER_BLOCK) k_tile_raster(
    const Draw* __restrict__ draws, const Ramp* __restrict__ ramps, const Frame* __restrict__ frames, uint32_t f_lo,
    uint32_t f_hi, uint64_t tile_base, int tws, int ths, const uint32_t* __restrict__ cursor,
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ pool, const int32_t* __restrict__ nl,
    uint32_t* __restrict__ cells, float* __restrict__ zbuf, int32_t* __restrict__ win)
{
    extern __shared__ uint64_t smem[];
    const int sws = tws - 1;
    const uint32_t S = 1u << (sws + ths);
    uint64_t* keys = smem;

    const uint64_t gt = tile_base + blockIdx.x;
    uint32_t lo = f_lo, hi = f_hi - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (frames[mid].tile_begin <= gt) lo = mid; else hi = mid - 1;
    }
    const Frame& fr = frames[lo];
    const uint32_t W = fr.W, H = fr.H;
    const uint32_t local = (uint32_t)(gt - fr.tile_begin);
    const uint32_t ty = local / fr.tiles_x, tx = local - ty * fr.tiles_x;
    TileGeom g;
    g.tc0 = (int)(tx << tws);
    g.tc1 = min((int)W, g.tc0 + (1 << tws));
    g.tr0 = (int)(ty << ths);
    g.tr1 = min((int)H, g.tr0 + (1 << ths));
    g.sws = sws;
    g.W = (int)W;

    const uint64_t o = off[gt];
    const uint32_t n = min(cursor[gt], (uint32_t)(off[gt + 1] - o));
    const uint32_t* list = pool + o;
    const bool one_draw = fr.draw_end - fr.draw_begin == 1;
    auto draw_for = [&](uint32_t T) -> const Draw& {
        return one_draw ? draws[fr.draw_begin] : draw_of_t(draws, fr, T);
    };
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

    const int32_t* fnl = nl + fr.nl_off;
    uint32_t* fc = cells + fr.cell_off;
    float* fz = zbuf + fr.z_off;
    // Stores the cells and depths of slot s from its winning cell (or blank).
    auto store = [&](uint32_t s, uint64_t k, uint32_t cell, float z) {
        const int sr = (int)(s >> sws), sc = (int)(s & ((1u << sws) - 1));
        const int row = g.tr0 + sr, col = g.tc0 + 2 * sc;
        uint32_t c0 = CELL_BLANK, c1 = CELL_BLANK;
        float z0 = F32_MAX;
        int32_t wt = -1;
        if (k != KEY_EMPTY) {
            c0 = c1 = cell;
            z0 = z;
            wt = (int32_t)((uint32_t)k >> 1);
        }
        if (fr.image && col == 0) {
            // cell row*W + 1: the newline of row `row` comes after every pixel
            // of its triangle's row and before its next row, so it overwrites
            // a wrapped pixel (drawn in row-1) of the same or an earlier
            // triangle
            const int32_t nlr = fnl[row];
            if (nlr >= 0 && (k == KEY_EMPTY || nlr >= wt)) c1 = CELL_NEWLINE;
        }
        const uint64_t id = (uint64_t)row * W + (uint64_t)col;
        *(uint2*)(fc + id) = make_uint2(c0, c1);
        *(float2*)(fz + id) = make_float2(z0, F32_MAX);
        if (win) *(int2*)(win + fr.z_off + id) = make_int2(wt, -1);
    };
    auto slot_valid = [&](uint32_t s) {
        return g.tr0 + (int)(s >> sws) < g.tr1 && g.tc0 + 2 * (int)(s & ((1u << sws) - 1)) < g.tc1;
    };

    if (n <= CACHE_CAP) {
        // Fast path: one walk.  Each entry's transformed triangle goes to an
        // LDS row found through a hash of its triangle index; the winner of a
        // slot is re-evaluated from that row.
        uint32_t* htag = (uint32_t*)(keys + S);
        uint32_t* hrow = htag + HASH_CAP;
        float* cache = (float*)(hrow + HASH_CAP);
        for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) keys[s] = KEY_EMPTY;
        for (uint32_t h = threadIdx.x; h < HASH_CAP; h += blockDim.x) htag[h] = EMPTY32;
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < n; p += blockDim.x) {
            const uint32_t e = list[p];
            const uint32_t T = e & ~BOX_FLAG;
            const Draw& dr = draw_for(T);
            const uint32_t li = T - dr.tri_begin;
            Tri t;
            setup_tri(dr.M, dr.xyz + (size_t)li * 9, W, H, t);
            const float na = unit_normal_z(t) * t.a;   // rasterizer.rs:76 multiplies these first
            const uint8_t* c = dr.rgb + (size_t)li * 3;
            const uint32_t cw = ((uint32_t)c[0] << 8) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 24) | dr.ramp;
            float4* row4 = (float4*)(cache + p * CACHE_WORDS);
            row4[0] = make_float4(t.x1, t.y1, t.x2, t.y2);
            row4[1] = make_float4(t.x3, t.y3, t.z1, t.z2 - t.z1);
            row4[2] = make_float4(t.z3 - t.z1, t.a, na, __uint_as_float(cw));
            uint32_t h = hash_tri(T);
            for (uint32_t probe = 0; probe < HASH_CAP; ++probe) {
                if (atomicCAS(&htag[h], EMPTY32, T) == EMPTY32) {
                    hrow[h] = p;
                    break;
                }
                h = (h + 1) & (HASH_CAP - 1);
            }
            int bx0, bx1, by0, by1;
            if (entry_box(t, e, bx0, bx1, by0, by1)) test(t, T, bx0, bx1, by0, by1);
        }
        __syncthreads();
        for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
            if (!slot_valid(s)) continue;
            const uint64_t k = keys[s];
            uint32_t cell = CELL_BLANK;
            float z = F32_MAX;
            if (k != KEY_EMPTY) {
                const uint32_t T = (uint32_t)k >> 1, sub = (uint32_t)k & 1u;
                uint32_t h = hash_tri(T);
                for (uint32_t probe = 0; probe < HASH_CAP && htag[h] != T; ++probe) h = (h + 1) & (HASH_CAP - 1);
                const float4* row4 = (const float4*)(cache + hrow[h] * CACHE_WORDS);
                const float4 q0 = row4[0], q1 = row4[1], q2 = row4[2];
                const int row = g.tr0 + (int)(s >> sws), col = g.tc0 + 2 * (int)(s & ((1u << sws) - 1));
                const float px = (float)(sub ? (col >> 1) : ((col + g.W) >> 1));
                const float py = (float)(sub ? row : row - 1);
                // the same expressions as edges(), depth() and shade_value()
                const float x1 = q0.x, y1 = q0.y, x2 = q0.z, y2 = q0.w, x3 = q1.x, y3 = q1.y;
                const float w0 = (x3 - x2) * (py - y2) - (y3 - y2) * (px - x2);
                const float w1 = (x1 - x3) * (py - y3) - (y1 - y3) * (px - x3);
                const float w2 = (x2 - x1) * (py - y1) - (y2 - y1) * (px - x1);
                z = q1.z + q2.y * (w1 * q1.w + w2 * q2.x);
                const float sh = q2.z * (w0 + w1 + w2);
                const uint32_t cw = __float_as_uint(q2.w);
                cell = (uint32_t)ramp_glyph(ramps[cw & 0xFFu], sh) | (cw & 0xFFFFFF00u);
            }
            store(s, k, cell, z);
        }
    } else {
        // Long lists: two walks; the second stages the winner's cell and depth.
        uint32_t* st_cell = (uint32_t*)(keys + S);
        float* st_z = (float*)(st_cell + S);
        for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) keys[s] = KEY_EMPTY;
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < n; p += blockDim.x) {
            const uint32_t e = list[p];
            const uint32_t T = e & ~BOX_FLAG;
            const Draw& dr = draw_for(T);
            Tri t;
            setup_tri(dr.M, dr.xyz + (size_t)(T - dr.tri_begin) * 9, W, H, t);
            int bx0, bx1, by0, by1;
            if (entry_box(t, e, bx0, bx1, by0, by1)) test(t, T, bx0, bx1, by0, by1);
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < n; p += blockDim.x) {
            const uint32_t e = list[p];
            const uint32_t T = e & ~BOX_FLAG;
            const Draw& dr = draw_for(T);
            const uint32_t li = T - dr.tri_begin;
            Tri t;
            setup_tri(dr.M, dr.xyz + (size_t)li * 9, W, H, t);
            int bx0, bx1, by0, by1;
            if (!entry_box(t, e, bx0, bx1, by0, by1)) continue;
            bool have_nz = false;
            float na = 0.0f;
            uint32_t color = 0;
            walk_tile(g, bx0, bx1, by0, by1, [&](int x, int y, int slot, uint32_t sub) {
                float w0, w1, w2;
                edges(t, (float)x, (float)y, w0, w1, w2);
                if (!inside(w0, w1, w2)) return;
                float z = depth(t, w1, w2);
                if (!(z < F32_MAX)) return;
                if (keys[slot] != make_key(z, T, sub)) return;
                if (!have_nz) {
                    na = unit_normal_z(t) * t.a;
                    const uint8_t* c = dr.rgb + (size_t)li * 3;
                    color = ((uint32_t)c[0] << 8) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 24);
                    have_nz = true;
                }
                st_cell[slot] = (uint32_t)ramp_glyph(ramps[dr.ramp], na * (w0 + w1 + w2)) | color;
                st_z[slot] = z;
            });
        }
        __syncthreads();
        for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) {
            if (!slot_valid(s)) continue;
            const uint64_t k = keys[s];
            store(s, k, k != KEY_EMPTY ? st_cell[s] : CELL_BLANK, k != KEY_EMPTY ? st_z[s] : F32_MAX);
        }
    }
    // image mode: the H cells after the W*H grid are never written (context.rs:38-40)
    if (fr.image && local == 0)
        for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) fc[(uint64_t)W * H + i] = CELL_BLANK;
}

// ---------------------------------------------------------------------------
// TiledPlan
// ---------------------------------------------------------------------------
static unsigned blocks(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

static int log2u(uint32_t v)
{
    int s = 0;
    while ((1u << s) < v) ++s;
    return s;
}

int TiledPlan::set_tile(uint32_t tw, uint32_t th)
{
    if (tw < 8 || tw > 128 || th < 4 || th > 64 || (tw & (tw - 1)) || (th & (th - 1)))
        return fail(SLOTH_EINVAL, "tile size %ux%u: width must be a power of two in 8..128, height in 4..64", tw, th);
    if (tw != TW || th != TH) sized = false;
    TW = tw;
    TH = th;
    return SLOTH_OK;
}

size_t TiledPlan::raster_lds() const
{
    size_t S = (size_t)(TW / 2) * TH;
    size_t fast = (size_t)HASH_CAP * 8 + (size_t)CACHE_CAP * CACHE_WORDS * 4;
    return S * 8 + std::max(fast, S * 8);
}

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

int TiledPlan::launch_clear(const std::vector<Frame>& frames, uint32_t f_lo, uint32_t f_hi, hipStream_t s)
{
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
                          const Draw* d_draws, uint32_t f_lo, uint32_t f_hi, bool count_only, hipStream_t s)
{
    const uint32_t d_lo = frames[f_lo].draw_begin, d_hi = frames[f_hi - 1].draw_end;
    const int tws = log2u(TW), ths = log2u(TH);
    for (uint32_t d0 = d_lo; d0 < d_hi; d0 += 65535u) {
        const uint32_t d1 = std::min(d_hi, d0 + 65535u);
        uint32_t max_n = 0;
        for (uint32_t d = d0; d < d1; ++d) max_n = std::max(max_n, draws[d].n);
        if (max_n == 0) continue;
        dim3 grid(blocks(max_n, BIN_BLOCK), d1 - d0);
        if (count_only)
            hipLaunchKernelGGL(k_bin<true>, grid, dim3(BIN_BLOCK), 0, s, d_draws, d0, d_frames, tws, ths,
                               d_cursor.as<uint32_t>(), d_off.as<uint64_t>(), d_pool.as<uint32_t>(),
                               d_nl.as<int32_t>(), d_overflow.as<uint32_t>());
        else
            hipLaunchKernelGGL(k_bin<false>, grid, dim3(BIN_BLOCK), 0, s, d_draws, d0, d_frames, tws, ths,
                               d_cursor.as<uint32_t>(), d_off.as<uint64_t>(), d_pool.as<uint32_t>(),
                               d_nl.as<int32_t>(), d_overflow.as<uint32_t>());
        HIP_TRY(hipGetLastError());
    }
    return SLOTH_OK;
}

int TiledPlan::size(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
                    const Draw* d_draws, hipStream_t s)
{
    if (frames.empty()) return SLOTH_OK;
    const uint32_t F = (uint32_t)frames.size();
    int rc;
    if ((rc = launch_clear(frames, 0, F, s))) return rc;
    if ((rc = launch_bin(frames, draws, d_frames, d_draws, 0, F, true, s))) return rc;
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
    int rc;
    if (ev) HIP_TRY(hipEventRecord((*ev)[0], s));
    if ((rc = launch_clear(frames, f_lo, f_hi, s))) return rc;
    if (ev) HIP_TRY(hipEventRecord((*ev)[1], s));
    if ((rc = launch_bin(frames, draws, d_frames, d_draws, f_lo, f_hi, false, s))) return rc;
    if (ev) HIP_TRY(hipEventRecord((*ev)[2], s));
    const Frame& a = frames[f_lo];
    const Frame& b = frames[f_hi - 1];
    uint64_t t_lo = a.tile_begin, t_hi = b.tile_begin + (uint64_t)b.tiles_x * b.tiles_y;
    if (t_hi > t_lo) {
        hipLaunchKernelGGL(k_tile_raster, dim3((unsigned)(t_hi - t_lo)), dim3(RASTER_BLOCK), raster_lds(), s, d_draws,
                           d_ramps, d_frames, f_lo, f_hi, t_lo, log2u(TW), log2u(TH), d_cursor.as<uint32_t>(),
                           d_off.as<uint64_t>(), d_pool.as<uint32_t>(), d_nl.as<int32_t>(), cells, z, win);
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
