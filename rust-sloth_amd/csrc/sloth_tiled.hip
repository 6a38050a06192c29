// Tiled raster path (even widths): bin-to-tile and LDS tile raster.
// The data flow is described in sloth_tiled.h.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "sloth_tiled.h"

namespace sloth {

constexpr int BIN_BLOCK = 256;
constexpr int RASTER_BLOCK = 256;

// k_tile_raster<true> walks a box with more than BIG_BOX pixels in a tile by
// the whole wavefront.  careful_cnt[BIG_CNT] is where the counting run of
// k_bin counts the triangles whose (unclipped) box has more.
constexpr uint32_t BIG_BOX = 128;
constexpr uint32_t BIG_CNT = 2 + MAX_GROUPS;
// careful_cnt[TALL_CNT]: triangles the counting run finds spanning more than
// TALL_ROWS rows (TiledPlan::bin_tree)
constexpr uint32_t TALL_CNT = BIG_CNT + 1;
constexpr int TALL_ROWS = 16;

// ---------------------------------------------------------------------------
// k_tile_clear
// ---------------------------------------------------------------------------
__global__ void k_tile_clear(uint32_t* __restrict__ cursor, uint64_t n_tiles, int32_t* __restrict__ nl,
                             uint64_t n_rows, uint32_t* __restrict__ overflow, uint32_t n_frames,
                             uint32_t* __restrict__ careful_cnt)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = i0; i < n_tiles; i += stride) cursor[i] = 0;
    for (uint64_t i = i0; i < n_rows; i += stride) nl[i] = -1;
    for (uint64_t i = i0; i < n_frames; i += stride) overflow[i] = 0;
    if (i0 == 0) careful_cnt[0] = careful_cnt[1] = careful_cnt[BIG_CNT] = careful_cnt[TALL_CNT] = 0;   // list length, first group's start
}

// ---------------------------------------------------------------------------
// tile geometry shared by the binning kernels
// ---------------------------------------------------------------------------

// Pool entries: the frame triangle index of a CLS_TIGHT triangle, LOOSE_FLAG |
// the index of a CLS_BOX one (the raster recomputes either box from the
// triangle's x extent), or CAREFUL_FLAG | row of the careful-box table for
// triangles whose box k_careful computed.  Frames on this path hold fewer than
// TILED_MAX_TRIS triangles.
constexpr uint32_t CAREFUL_FLAG = 0x80000000u;
constexpr uint32_t LOOSE_FLAG = 0x40000000u;


// Tile rectangles (inclusive, in tile units) that a candidate box reaches:
// the direct part (2x < W, cell (2x, y)) and the wrapped part (2x >= W, cell
// (2x - W, y + 1)).  An empty rectangle has x0 > x1.
struct TileRects {
    int dx0, dx1, dy0, dy1;
    int wx0, wx1, wy0, wy1;
};

__device__ __forceinline__ void empty_rects(TileRects& r)
{
    r.dx0 = r.dy0 = r.wx0 = r.wy0 = 1;
    r.dx1 = r.dy1 = r.wx1 = r.wy1 = 0;
}

__device__ __forceinline__ void tile_rects(int bx0, int bx1, int by0, int by1, uint32_t W, int tws, int ths,
                                           TileRects& r)
{
    // direct pixels: 2x < W, i.e. x <= (W - 1) / 2; wrapped: x >= (W + 1) / 2
    empty_rects(r);
    int dxe = min(bx1, (int)((W - 1) >> 1));
    if (bx0 <= dxe && by0 <= by1) {
        r.dx0 = (2 * bx0) >> tws;
        r.dx1 = (2 * dxe) >> tws;
        r.dy0 = by0 >> ths;
        r.dy1 = by1 >> ths;
    }
    int wxs = max(bx0, (int)((W + 1) >> 1));
    if (wxs <= bx1 && by0 <= by1) {
        r.wx0 = (2 * wxs - (int)W) >> tws;
        r.wx1 = (2 * bx1 - (int)W) >> tws;
        r.wy0 = (by0 + 1) >> ths;
        r.wy1 = (by1 + 1) >> ths;
    }
}

// Walks a lane's tiles: the direct rectangle, then the wrapped one minus the
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

__device__ __forceinline__ void iter_begin(const TileRects& r, TileIter& it)
{
    it.phase = 0;
    it.tx = r.dx0;
    it.ty = r.dy0;
    iter_settle(r, it);
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

// One entry appended to tile `gt` at `pos` (when the pool has room).
template <bool COUNT>
__device__ __forceinline__ void put_entry(uint64_t gt, uint32_t pos, uint32_t entry, const uint64_t* __restrict__ off,
                                          uint32_t* __restrict__ pool, uint32_t* __restrict__ overflow, uint32_t frame)
{
    if (COUNT) return;
    const uint64_t o = off[gt], cap = off[gt + 1] - o;
    if (pos < cap) pool[o + pos] = entry;
    else overflow[frame] = 1u;
}

// ---------------------------------------------------------------------------
// k_bin
// ---------------------------------------------------------------------------
constexpr int BIN_WIN = 16;   // the block aggregates appends in a 16 x 16 window of tiles

// Wavefront-wide min / max (all 64 lanes must be active), returned uniform.
// A block reduces through these first so that each wavefront issues one LDS
// atomic: LDS serialises lanes that hit the same address.  Within each row of
// 16 lanes a DPP butterfly (quad_perm [1,0,3,2] and [2,3,0,1], then row_ror 4
// and 8) leaves the row's extreme in every lane; the four rows are then read
// by readlane.  (Shuffles through ds_bpermute cost six LDS round trips.)
template <class Op>
__device__ __forceinline__ int wave_reduce(int v, Op op)
{
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x124, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false));
    return op(op(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
              op(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
__device__ __forceinline__ int wave_min(int v)
{
    return wave_reduce(v, [](int a, int b) { return min(a, b); });
}
__device__ __forceinline__ int wave_max(int v)
{
    return wave_reduce(v, [](int a, int b) { return max(a, b); });
}

// Grid: x = blocks of TPT * BIN_BLOCK consecutive triangles of one draw,
// y = draw index - d_lo, so the draw, its frame and its matrix are uniform
// per block (scalar loads).  CLS_BOX / CLS_TIGHT triangles are appended to
// their tiles here; CLS_SPANS / CLS_FULL ones go to the careful list for
// k_careful.
//
// Each thread takes TPT triangles (k * BIN_BLOCK + thread), set up one after
// another with their loads in flight together, so that the block's barriers
// and its global atomic round trip serve TPT times as many triangles (the
// kernel is bound by latency, not by issue).  TPT is 1, 2 or 4
// (TiledPlan::bin_tpt; the autotuner picks it).  Per triangle a thread
// keeps two words between the phases (its tile rectangle and flags, and its
// rows) and two more of window counts between the last two.
//
// Appends: a block's consecutive triangles lie close together on screen.
// Triangles whose box touches at most 2 x 2 tiles, all in the direct region
// or all in the wrapped one (nearly all), count themselves into LDS counters
// of a 16 x 16 tile window anchored at the block's smallest tile coordinates;
// one global atomicAdd per window tile then reserves that tile's run (the
// thread that reserves it also reads the tile's pool range), and each
// triangle writes its entry at run + its LDS count.  Others append with one
// global atomicAdd per tile.
//
// geo word: qx0 (bits 0-13), qy0 (14-27) of the tile rectangle, wide (28),
// tall (29), CLS_BOX (30); tile coordinates stay below 2^14 (frames of at
// most 65535 cells a side, tiles of at least 8 x 4 cells).
constexpr uint32_t GEO_WIDE = 1u << 28, GEO_TALL = 1u << 29, GEO_LOOSE = 1u << 30;

// Newline rows over a window of BIN_BLOCK rows: a max segment tree in LDS
// (node 1 spans the window, node i has children 2i and 2i + 1, row r's leaf
// is BIN_BLOCK + r).  A triangle raises the nodes that tile its rows [a, b)
// of the window, at most 2 log2(BIN_BLOCK) of them (a range of up to 4 rows
// takes its leaves); row r's value is the maximum on its leaf's path to the
// root.
__device__ __forceinline__ void nl_raise(int32_t* tree, int a, int b, int32_t v)
{
    if (b - a <= 4) {
        for (int y = a; y < b; ++y) atomicMax(&tree[BIN_BLOCK + y], v);
        return;
    }
    for (int l = a + BIN_BLOCK, r = b + BIN_BLOCK; l < r; l >>= 1, r >>= 1) {
        if (l & 1) atomicMax(&tree[l++], v);
        if (r & 1) atomicMax(&tree[--r], v);
    }
}
__device__ __forceinline__ int32_t nl_value(const int32_t* tree, int r)
{
    int32_t v = -1;
    for (int i = BIN_BLOCK + r; i >= 1; i >>= 1) v = max(v, tree[i]);
    return v;
}

template <bool COUNT, int TPT, bool TREE>
__global__ void __launch_bounds__(BIN_BLOCK) k_bin(const Draw* __restrict__ draws, uint32_t d_lo,
                                                   const Frame* __restrict__ frames, int tws, int ths,
                                                   uint32_t* __restrict__ cursor, const uint64_t* __restrict__ off,
                                                   uint32_t* __restrict__ pool, int32_t* __restrict__ nl,
                                                   uint32_t* __restrict__ overflow, uint32_t* __restrict__ careful_cnt,
                                                   uint4* __restrict__ careful_in, uint32_t careful_cap,
                                                   const uint2* __restrict__ flat_blocks)
{
    __shared__ uint32_t w_cnt[BIN_WIN * BIN_WIN];
    __shared__ uint64_t w_base[BIN_WIN * BIN_WIN];   // pool index of the tile's reserved run
    __shared__ uint64_t w_end[BIN_WIN * BIN_WIN];    // end of the tile's pool range
    // newline rows: a window of rows, or (TREE) its segment tree (nl_raise)
    __shared__ int32_t s_nl[TREE ? 2 * BIN_BLOCK : BIN_BLOCK];
    constexpr int LEAF = TREE ? BIN_BLOCK : 0;   // row r of the window at s_nl[LEAF + r]
    __shared__ int s_tx0, s_ty0, s_lo, s_hi;

    // the block's draw and first triangle: from the grid, or (flat grids of
    // draws of different sizes) from its table entry
    uint32_t d, base;
    if (flat_blocks) {
        const uint2 e = flat_blocks[blockIdx.x];
        d = e.x;
        base = e.y;
    } else {
        d = d_lo + blockIdx.y;
        base = blockIdx.x * (BIN_BLOCK * TPT);
    }
    const Draw& dr = draws[d];
    if (base >= dr.n) return;
    const uint32_t frame = dr.frame;
    const Frame& fr = frames[frame];
    const uint32_t W = fr.W, H = fr.H;
    const uint64_t tile_begin = fr.tile_begin;
    const uint32_t tiles_x = fr.tiles_x;

    w_cnt[threadIdx.x] = 0;
    s_nl[threadIdx.x] = -1;
    if (TREE) s_nl[LEAF + threadIdx.x] = -1;
    if (threadIdx.x == 0) {
        s_tx0 = s_ty0 = s_lo = 1 << 30;
        s_hi = -1;
    }

    // Phase 1: set-up and classes.  Lanes past the draw's end set up its
    // last triangle and discard it, so that the set-up has no branch around
    // it (the lanes of a wavefront take different classes; selects cost less
    // than divergent paths).
    uint32_t geo[TPT], rows[TPT];
    uint32_t tiled = 0, small = 0;   // bit k: triangle k has tiles here / has a small rectangle
    int acc_lo = 1 << 30, acc_hi = -1, acc_tx = 1 << 30, acc_ty = 1 << 30;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const uint32_t local = base + (uint32_t)k * BIN_BLOCK + threadIdx.x;
        const bool valid = local < dr.n;
        Tri t;
        setup_draw_tri(dr, valid ? local : dr.n - 1, W, H, t);
        const uint32_t cls = valid ? setup_class(t) : CLS_EMPTY;
        const bool boxed = cls == CLS_BOX || cls == CLS_TIGHT;
        const bool careful = cls == CLS_SPANS || cls == CLS_FULL;
        uint32_t g = cls == CLS_BOX ? GEO_LOOSE : 0u;
        // A wavefront whose triangles are all back faces or empty (half of
        // a closed mesh's, in runs) has no tiles to find: a uniform branch.
        if (__ballot(boxed)) {
            int bx0 = 1, bx1 = 0;
            if (boxed) box_x(t, cls == CLS_TIGHT, bx0, bx1);   // empty when not boxed
            if (COUNT) {
                const bool big = boxed && (int64_t)(bx1 - bx0 + 1) * (int64_t)(t.y1d - t.y0) > (int64_t)BIG_BOX;
                const uint64_t m = __ballot(big);
                if (__lane_id() == 0 && m) atomicAdd(&careful_cnt[BIG_CNT], (uint32_t)__popcll(m));
            }
            // A box in one region appends through that region's tile
            // rectangle [qx0, qx1] x [qy0, qy1]; a box that wraps in part
            // (it spans 2x = W) takes the per-tile path.  (Meshes wider than
            // the frame's logical width put whole patches of triangles in
            // the wrapped region, whose per-tile atomics would meet on few
            // addresses.)
            bool has_direct, has_wrap;
            int qx0, qx1, qy0, qy1;
            if (!__ballot(bx1 >= (int)((W + 1) >> 1))) {
                // no box of the wavefront reaches the wrapped region (x >=
                // (W + 1) / 2): tile_rects' direct rectangle with x <= bx1
                has_direct = bx0 <= bx1 && t.y0 < t.y1d;
                has_wrap = false;
                qx0 = (2 * bx0) >> tws;
                qx1 = (2 * bx1) >> tws;
                qy0 = t.y0 >> ths;
                qy1 = (t.y1d - 1) >> ths;
            } else {
                TileRects r;
                tile_rects(bx0, bx1, t.y0, t.y1d - 1, W, tws, ths, r);   // empty when bx0 > bx1
                has_direct = r.dx0 <= r.dx1 && r.dy0 <= r.dy1;
                has_wrap = r.wx0 <= r.wx1 && r.wy0 <= r.wy1;
                const bool wrap_only = has_wrap && !has_direct;
                qx0 = wrap_only ? r.wx0 : r.dx0;
                qx1 = wrap_only ? r.wx1 : r.dx1;
                qy0 = wrap_only ? r.wy0 : r.dy0;
                qy1 = wrap_only ? r.wy1 : r.dy1;
            }
            const bool sm = has_direct != has_wrap && qx1 - qx0 <= 1 && qy1 - qy0 <= 1;
            if (has_direct || has_wrap) tiled |= 1u << k;
            if (sm) {
                small |= 1u << k;
                g |= (uint32_t)qx0 | ((uint32_t)qy0 << 14) | (qx1 > qx0 ? GEO_WIDE : 0u) | (qy1 > qy0 ? GEO_TALL : 0u);
                acc_tx = min(acc_tx, qx0);
                acc_ty = min(acc_ty, qy0);
            }
        }
        geo[k] = g;
        if (careful) {
            const uint32_t idx = atomicAdd(careful_cnt, 1u);
            if (idx < careful_cap) {
                // the screen vertices come along, so that k_careful starts
                // from them instead of the draw record and the mesh
                careful_in[2 * idx] = make_uint4(frame, dr.tri_begin + local, __float_as_uint(t.x1), __float_as_uint(t.y1));
                careful_in[2 * idx + 1] = make_uint4(__float_as_uint(t.x2), __float_as_uint(t.y2),
                                                     __float_as_uint(t.x3), __float_as_uint(t.y3));
            }
            else if (!COUNT) overflow[frame] = 1u;
        }
        const bool has_rows = fr.image && valid && t.y0 < t.y1d;
        rows[k] = has_rows ? (uint32_t)t.y0 | ((uint32_t)t.y1d << 16) : 0u;
        if (has_rows) {
            acc_lo = min(acc_lo, t.y0);
            acc_hi = max(acc_hi, t.y1d - 1);
        }
        if (COUNT) {
            const uint64_t m = __ballot(has_rows && t.y1d - t.y0 > TALL_ROWS);
            if (__lane_id() == 0 && m) atomicAdd(&careful_cnt[TALL_CNT], (uint32_t)__popcll(m));
        }
    }
    __syncthreads();   // LDS initialised

    {
        const int wlo = wave_min(acc_lo), whi = wave_max(acc_hi);
        if (__lane_id() == 0 && whi >= 0) {
            atomicMin(&s_lo, wlo);
            atomicMax(&s_hi, whi);
        }
        if (__ballot(small != 0)) {
            const int wtx = wave_min(acc_tx), wty = wave_min(acc_ty);
            if (__lane_id() == 0) {
                atomicMin(&s_tx0, wtx);
                atomicMin(&s_ty0, wty);
            }
        }
    }
    __syncthreads();

    // Phase 2: window counts and newline rows in LDS.
    // rasterizer.rs:89-91: a newline cell after every row of the domain; the
    // last triangle (in draw order) whose rows include r decides row r.
    // The block reduces its rows in LDS, one global atomicMax per row: here
    // when they span fewer than BIN_BLOCK rows, else in windows of BIN_BLOCK
    // rows at the end (tall triangles put every row of a tall frame under
    // the atomics of each triangle that covers it).
    const int ox = s_tx0, oy = s_ty0, lo = s_lo, hi = s_hi;
    const bool nl_lds = hi >= lo && hi - lo < BIN_BLOCK;
    auto in_window = [&](uint32_t g, int& c) {
        const int qx0 = (int)(g & 0x3FFFu), qy0 = (int)((g >> 14) & 0x3FFFu);
        c = (qy0 - oy) * BIN_WIN + (qx0 - ox);
        return qx0 + (int)((g >> 28) & 1u) - ox < BIN_WIN && qy0 + (int)((g >> 29) & 1u) - oy < BIN_WIN;
    };
    const uint32_t tri0 = dr.tri_begin + base + threadIdx.x;
    uint32_t lpa[TPT], lpb[TPT];   // LDS counts: tiles (0,0) | (1,0) << 16, (0,1) | (1,1) << 16
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        lpa[k] = lpb[k] = 0;
        int c;
        if (((small >> k) & 1u) && in_window(geo[k], c)) {
            const bool wide = geo[k] & GEO_WIDE, tall = geo[k] & GEO_TALL;
            uint32_t l0 = atomicAdd(&w_cnt[c], 1u), l1 = 0, l2 = 0, l3 = 0;
            if (wide) l1 = atomicAdd(&w_cnt[c + 1], 1u);
            if (tall) l2 = atomicAdd(&w_cnt[c + BIN_WIN], 1u);
            if (wide && tall) l3 = atomicAdd(&w_cnt[c + BIN_WIN + 1], 1u);
            lpa[k] = l0 | (l1 << 16);
            lpb[k] = l2 | (l3 << 16);
        }
        if (nl_lds && rows[k]) {
            const int y0 = (int)(rows[k] & 0xFFFFu), y1d = (int)(rows[k] >> 16);
            if (TREE) nl_raise(s_nl, y0 - lo, y1d - lo, (int32_t)(tri0 + (uint32_t)k * BIN_BLOCK));
            else
                for (int y = y0; y < y1d; ++y) atomicMax(&s_nl[y - lo], (int32_t)(tri0 + (uint32_t)k * BIN_BLOCK));
        }
    }
    __syncthreads();

    int32_t* fnl = nl + fr.nl_off;
    {
        const uint32_t n = w_cnt[threadIdx.x];
        if (n) {
            const uint64_t gt = tile_begin + (uint64_t)(oy + (int)(threadIdx.x / BIN_WIN)) * tiles_x +
                                (uint64_t)(ox + (int)(threadIdx.x % BIN_WIN));
            const uint32_t pos = atomicAdd(&cursor[gt], n);
            if (!COUNT) {
                w_base[threadIdx.x] = off[gt] + pos;
                w_end[threadIdx.x] = off[gt + 1];
            }
        }
    }
    if (nl_lds && (int)threadIdx.x <= hi - lo) {
        const int32_t v = TREE ? nl_value(s_nl, (int)threadIdx.x) : s_nl[threadIdx.x];
        if (v >= 0) atomicMax(&fnl[lo + (int)threadIdx.x], v);
    }
    __syncthreads();

    // Phase 3: entries.
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        if (!((tiled >> k) & 1u)) continue;
        const uint32_t local = base + (uint32_t)k * BIN_BLOCK + threadIdx.x;
        const uint32_t entry = (dr.tri_begin + local) | ((geo[k] & GEO_LOOSE) ? LOOSE_FLAG : 0u);
        int c;
        if ((small >> k) & 1u) {
            const bool wide = geo[k] & GEO_WIDE, tall = geo[k] & GEO_TALL;
            if (in_window(geo[k], c)) {
                if (!COUNT) {
                    auto put = [&](int cc, uint32_t l) {
                        const uint64_t p = w_base[cc] + l;
                        if (p < w_end[cc]) pool[p] = entry;
                        else overflow[frame] = 1u;
                    };
                    put(c, lpa[k] & 0xFFFFu);
                    if (wide) put(c + 1, lpa[k] >> 16);
                    if (tall) put(c + BIN_WIN, lpb[k] & 0xFFFFu);
                    if (wide && tall) put(c + BIN_WIN + 1, lpb[k] >> 16);
                }
            } else {
                // a small rectangle outside the window: one global atomic per tile
                const int qx0 = (int)(geo[k] & 0x3FFFu), qy0 = (int)((geo[k] >> 14) & 0x3FFFu);
                for (int j = 0; j <= (int)tall; ++j)
                    for (int i = 0; i <= (int)wide; ++i) {
                        const uint64_t gt = tile_begin + (uint64_t)(qy0 + j) * tiles_x + (uint64_t)(qx0 + i);
                        const uint32_t pos = atomicAdd(&cursor[gt], 1u);
                        put_entry<COUNT>(gt, pos, entry, off, pool, overflow, frame);
                    }
            }
        } else {
            // large or part-wrapping boxes (few): set up again, one global
            // atomic per tile
            Tri t;
            setup_draw_tri(dr, local, W, H, t);
            int bx0, bx1;
            box_x(t, !(geo[k] & GEO_LOOSE), bx0, bx1);
            TileRects r;
            tile_rects(bx0, bx1, t.y0, t.y1d - 1, W, tws, ths, r);
            TileIter it;
            iter_begin(r, it);
            while (it.phase < 2) {
                const uint64_t gt = tile_begin + (uint64_t)it.ty * tiles_x + (uint64_t)it.tx;
                const uint32_t pos = atomicAdd(&cursor[gt], 1u);
                put_entry<COUNT>(gt, pos, entry, off, pool, overflow, frame);
                iter_next(r, it);
            }
        }
    }
    if (hi >= lo && !nl_lds) {   // uniform per block
        for (int ws = lo; ws <= hi; ws += BIN_BLOCK) {
            __syncthreads();   // the previous window's reads are done
            s_nl[threadIdx.x] = -1;
            if (TREE) s_nl[LEAF + threadIdx.x] = -1;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                if (!rows[k]) continue;
                const int y0 = max((int)(rows[k] & 0xFFFFu), ws), y1d = min((int)(rows[k] >> 16), ws + BIN_BLOCK);
                if (TREE) {
                    if (y0 < y1d) nl_raise(s_nl, y0 - ws, y1d - ws, (int32_t)(tri0 + (uint32_t)k * BIN_BLOCK));
                } else {
                    for (int y = y0; y < y1d; ++y) atomicMax(&s_nl[y - ws], (int32_t)(tri0 + (uint32_t)k * BIN_BLOCK));
                }
            }
            __syncthreads();
            if (ws + (int)threadIdx.x <= hi) {
                const int32_t v = TREE ? nl_value(s_nl, (int)threadIdx.x) : s_nl[threadIdx.x];
                if (v >= 0) atomicMax(&fnl[ws + (int)threadIdx.x], v);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_careful: exact boxes of the triangles the bound could not decide
// ---------------------------------------------------------------------------
// One wavefront per careful triangle.  CLS_SPANS: the lanes split the rows of
// the domain, each row's exact span comes from row_span(), and the box is the
// wave's min/max; CLS_FULL: the whole domain.  The lanes then append the
// triangle to every tile of the box, and lane 0 stores the box for the raster.
template <bool COUNT>
__device__ __forceinline__ void careful_one(const Frame* __restrict__ frames, int tws,
                                            int ths, uint32_t* __restrict__ cursor, const uint64_t* __restrict__ off,
                                            uint32_t* __restrict__ pool, uint32_t* __restrict__ overflow,
                                            const uint4* __restrict__ careful_in, uint4* __restrict__ careful_box,
                                            uint32_t w, int lane)
{
    // (frame, frame triangle index, screen vertices as k_bin transformed them)
    const uint4 in = careful_in[2 * w], in2 = careful_in[2 * w + 1];
    const uint32_t frame = in.x;
    const Frame& fr = frames[frame];
    Tri t;
    t.x1 = __uint_as_float(in.z);
    t.y1 = __uint_as_float(in.w);
    t.x2 = __uint_as_float(in2.x);
    t.y2 = __uint_as_float(in2.y);
    t.x3 = __uint_as_float(in2.z);
    t.y3 = __uint_as_float(in2.w);
    tri_finish(fr.W, fr.H, t);
    const uint32_t T = in.y;
    const uint32_t cls = setup_class(t);
    int bx0 = t.x0, bx1 = t.x1d - 1, by0 = t.y0, by1 = t.y1d - 1;
    if (cls == CLS_SPANS) {
        int lo_x = 1 << 30, hi_x = -1, first = 1 << 30, last = -1;
        for (int y = t.y0 + lane; y < t.y1d; y += 64) {
            int l, h;
            if (row_span(t, y, l, h)) {
                lo_x = min(lo_x, l);
                hi_x = max(hi_x, h);
                first = min(first, y);
                last = max(last, y);
            }
        }
        bx0 = wave_min(lo_x);
        bx1 = wave_max(hi_x);
        by0 = wave_min(first);
        by1 = wave_max(last);
    } else if (cls != CLS_FULL) {
        bx1 = -1;   // decided after all (cannot happen for entries k_bin sent here)
    }
    if (lane == 0 && !COUNT)
        careful_box[w] = make_uint4(T, (uint32_t)max(bx0, 0), (uint32_t)max(bx1, 0),
                                    (uint32_t)max(by0, 0) | ((uint32_t)max(by1, 0) << 16));
    if (bx0 > bx1 || by0 > by1) return;
    TileRects r;
    tile_rects(bx0, bx1, by0, by1, fr.W, tws, ths, r);
    const int dw = r.dx1 - r.dx0 + 1, dh = r.dy1 - r.dy0 + 1;
    const int nd = (r.dx0 <= r.dx1 && r.dy0 <= r.dy1) ? dw * dh : 0;
    const int ww = r.wx1 - r.wx0 + 1, wh = r.wy1 - r.wy0 + 1;
    const int nw = (r.wx0 <= r.wx1 && r.wy0 <= r.wy1) ? ww * wh : 0;
    for (int k = lane; k < nd + nw; k += 64) {
        int tx, ty;
        if (k < nd) {
            tx = r.dx0 + k % dw;
            ty = r.dy0 + k / dw;
        } else {
            tx = r.wx0 + (k - nd) % ww;
            ty = r.wy0 + (k - nd) / ww;
            if (nd > 0 && in_direct(r, tx, ty)) continue;
        }
        const uint64_t gt = fr.tile_begin + (uint64_t)ty * fr.tiles_x + (uint64_t)tx;
        const uint32_t pos = atomicAdd(&cursor[gt], 1u);
        put_entry<COUNT>(gt, pos, CAREFUL_FLAG | w, off, pool, overflow, frame);
    }
}

// careful_cnt[0] is the list length; careful_cnt[1 + g] the index where group
// g's triangles start (the bins of group g ran after those of group g - 1).
// Group g takes [careful_cnt[1 + g], length) and records the length as group
// g + 1's start.
template <bool COUNT>
__global__ void __launch_bounds__(256) k_careful(const Frame* __restrict__ frames, int tws, int ths, uint32_t* __restrict__ cursor,
                                                 const uint64_t* __restrict__ off, uint32_t* __restrict__ pool,
                                                 uint32_t* __restrict__ overflow, uint32_t* __restrict__ careful_cnt,
                                                 uint32_t group, const uint4* __restrict__ careful_in,
                                                 uint4* __restrict__ careful_box, uint32_t careful_cap)
{
    const int lane = (int)__lane_id();
    const uint32_t n = min(careful_cnt[0], careful_cap);
    const uint32_t first = careful_cnt[1 + group];
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    if (blockIdx.x == 0 && threadIdx.x == 0) careful_cnt[2 + group] = n;
    for (uint32_t w = first + ((blockIdx.x * blockDim.x + threadIdx.x) >> 6); w < n; w += n_waves)
        careful_one<COUNT>(frames, tws, ths, cursor, off, pool, overflow, careful_in, careful_box, w, lane);
}


// ---------------------------------------------------------------------------
// k_tile_raster
// ---------------------------------------------------------------------------
constexpr uint32_t CHUNK = RASTER_BLOCK;   // list entries whose triangles sit in LDS at once (one per thread)
constexpr uint32_t CACHE_WORDS = 12;
constexpr uint32_t LDS_RAMPS = 4;          // shaders staged in LDS per tile
constexpr uint32_t SPAN_MIN_W = 6;         // box width (pixels) from which a wavefront walks row spans
static_assert(CHUNK <= 256, "tile keys hold a cache row index in one byte");

// A tile's z-test key: (depth code, frame triangle, sub) decide the order
// exactly as make_key's do; below them a `stored` flag and the cache row p
// of the entry that wrote the key (its lane in the chunk).  Two fragments
// with equal (depth code, triangle, sub) are the same fragment, so the low
// bits never change a comparison, and the resolve finds the winner's
// transformed triangle from the key alone.  `stored` is set on keys whose
// winner a multi-chunk tile has already stored.
constexpr uint64_t TILE_KEY_STORED = 1u << 8;
__device__ __forceinline__ uint64_t make_tile_key(float z, uint32_t tri, uint32_t sub, uint32_t p)
{
    return ((uint64_t)depth_code(z) << 32) | (uint64_t)((tri << 10) | (sub << 9) | p);
}
__device__ __forceinline__ uint32_t tile_key_tri(uint64_t k) { return (uint32_t)k >> 10; }
__device__ __forceinline__ uint32_t tile_key_sub(uint64_t k) { return ((uint32_t)k >> 9) & 1u; }
__device__ __forceinline__ uint32_t tile_key_row(uint64_t k) { return (uint32_t)k & 0xFFu; }

__device__ __forceinline__ const Draw& draw_of_t(const Draw* __restrict__ draws, const Frame& fr, uint32_t tri)
{
    uint32_t lo = fr.draw_begin, hi = fr.draw_end - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (draws[mid].tri_begin <= tri) lo = mid; else hi = mid - 1;
    }
    return draws[lo];
}

// A tile's cells and z slots.  With W even a z slot id = y*W + 2x sits on an
// even column and owns its two cells: a tile row of TW cells holds TW/2
// slots (cs = 1).  With W odd, direct pixels land on even columns and
// wrapped ones (2x >= W, row y + 1) on odd columns, so every column holds a
// slot (cs = 0).  A pixel whose slot is cell column c of the tile's row r
// takes slot ((r - tr0) << sws) + ((c - tc0) >> cs).
struct TileGeom {
    int tc0, tc1, tr0, tr1;   // cell columns [tc0,tc1), rows [tr0,tr1)
    int sws;                  // log2(slots per tile row): log2(TW >> cs)
    int cs;                   // 1 for even W, 0 for odd W
    int W;
};

// Calls f(px, py, slot, sub) for the pixels (x, y) in [xa, xb] x [ya, yb],
// px = x and py = y as floats, slot = row0 + ((y - ya) << sws) + (x - xa) * xs.
// Pixel coordinates are below 2^16, so stepping px by 1.0f is exact (one add
// instead of a conversion per pixel).
template <class Fn>
__device__ __forceinline__ void walk_rect(int xa, int xb, int ya, int yb, int row0, int sws, int xs, uint32_t sub,
                                          Fn&& f)
{
    const float fxb = (float)xb;
    for (int y = ya; y <= yb; ++y, row0 += 1 << sws) {
        const float py = (float)y;
        int slot = row0;
        for (float px = (float)xa; px <= fxb; px += 1.0f, slot += xs) f(px, py, slot, sub);
    }
}

// walk_rect over the direct region with each row cut to its candidate span
// (row_spans_begin, started at row ya): slot = row0 + ((y - ya) << sws) +
// (x - xa) * xs as there.
template <class Fn>
__device__ __forceinline__ void walk_rect_spans(int xa, int xb, int ya, int yb, int row0, int sws, int xs,
                                                RowSpans& r, Fn&& f)
{
    const float fxa = (float)xa, fxb = (float)xb, fxe = fxb + 1.0f;
    for (int y = ya; y <= yb; ++y, row0 += 1 << sws) {
        // lo within [xa, xb + 1] (an empty row when above hi), so that its
        // integer conversion is exact
        const float lo = fminf(fmaxf(fmaxf(fmaxf(r.L[0], r.L[1]), r.L[2]), fxa), fxe);
        const float hi = fminf(fminf(fminf(r.H[0], r.H[1]), r.H[2]), fxb);
        const float py = (float)y, px0 = ceilf(lo), px1 = floorf(hi);
        int slot = row0 + ((int)px0 - xa) * xs;
        for (float px = px0; px <= px1; px += 1.0f, slot += xs) f(px, py, slot, 1u);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            r.H[k] += r.s[k];
            r.L[k] += r.s[k];
        }
    }
}

// Calls f(px, py, slot, sub) for every logical pixel of the box that lands in
// the tile's wrapped region (column 2x - W in [tc0,tc1), row y + 1).
template <class Fn>
__device__ __forceinline__ void walk_wrapped(const TileGeom& g, int bx0, int bx1, int by0, int by1, Fn&& f)
{
    const int xa = max(bx0, (g.tc0 + g.W + 1) >> 1), xb = min(bx1, (g.tc1 - 1 + g.W) >> 1);
    const int ya = max(by0, g.tr0 - 1), yb = min(by1, g.tr1 - 2);
    walk_rect(xa, xb, ya, yb, ((ya + 1 - g.tr0) << g.sws) + ((2 * xa - g.W - g.tc0) >> g.cs), g.sws, 2 >> g.cs, 0u,
              f);
}

// Calls f(px, py, slot, sub) for the candidate pixels of the box that land in
// the tile: the direct region (column 2x in [tc0,tc1), row y) with each row
// cut to its span (spans started at row max(by0, tr0)), then the wrapped one.
template <class Fn>
__device__ __forceinline__ void walk_tile(const TileGeom& g, int bx0, int bx1, int by0, int by1, bool use_spans,
                                          RowSpans& r, Fn&& f)
{
    const int xa = max(bx0, (g.tc0 + 1) >> 1), xb = min(bx1, (g.tc1 - 1) >> 1);
    const int ya = max(by0, g.tr0), yb = min(by1, g.tr1 - 1);
    const int row0 = ((ya - g.tr0) << g.sws) + ((2 * xa - g.tc0) >> g.cs);
    if (use_spans)
        walk_rect_spans(xa, xb, ya, yb, row0, g.sws, 2 >> g.cs, r, f);
    else
        walk_rect(xa, xb, ya, yb, row0, g.sws, 2 >> g.cs, 1u, f);
    walk_wrapped(g, bx0, bx1, by0, by1, f);
}

// The direct part of the walk: the box clipped to the tile's logical pixels,
// as tile-relative [xa, xb] x [ya, yb] (empty when xa > xb or ya > yb).
__device__ __forceinline__ void direct_rect(const TileGeom& g, int bx0, int bx1, int by0, int by1, int& xa, int& xb,
                                            int& ya, int& yb)
{
    xa = max(bx0, g.tc0 >> 1) - (g.tc0 >> 1);
    xb = min(bx1, (g.tc1 >> 1) - 1) - (g.tc0 >> 1);
    ya = max(by0, g.tr0) - g.tr0;
    yb = min(by1, g.tr1 - 1) - g.tr0;
}

// Triangle index of a pool entry; the candidate box of a careful entry comes
// from its table row (given = true), the others' from box_x.
__device__ __forceinline__ uint32_t entry_tri(uint32_t e, const uint4* __restrict__ careful_box, int& bx0, int& bx1,
                                              int& by0, int& by1, bool& given, bool& tight)
{
    given = (e & CAREFUL_FLAG) != 0;
    tight = (e & LOOSE_FLAG) == 0;
    if (!given) return e & ~LOOSE_FLAG;
    const uint4 b = careful_box[e & ~CAREFUL_FLAG];
    bx0 = (int)b.y;
    bx1 = (int)b.z;
    by0 = (int)(b.w & 0xFFFFu);
    by1 = (int)(b.w >> 16);
    return b.x;
}

// One workgroup per tile; grid (tile x, tile y, frame - f_lo).  Blocks past a
// smaller frame's tile grid return at once.
//
// The tile's list is taken CHUNK entries at a time: each entry's transformed
// triangle goes to the LDS row of its lane, and its fragments take a 64-bit
// LDS atomicMin of packed (depth, triangle, sub, row) keys (make_tile_key).
// With one chunk, every slot then re-evaluates its winner from the row its
// key names and stores it.  With several, after each chunk the slots whose
// key was written in that chunk store that winner and flag the key as stored
// (the last store of a slot is its final winner's, since a later winner
// comes from a later chunk), and the empty slots are stored at the end.
//
// BIG: an entry whose box covers more than BIG_BOX of the tile's pixels in
// its direct region is not walked by its own lane; after its own walk, each
// wavefront takes its lanes' big boxes one at a time and all 64 lanes stride
// over the box.  The variant costs registers on every path, so it is launched
// only for lists with many big boxes (TiledPlan::big_walk).
//
// ONE: every frame of the launch has exactly one draw, so the draw is the
// same for the whole workgroup and its fields come through scalar loads that
// do not wait for the list entries.
//
// ODD: frames of odd width.  Every cell column holds a z slot (direct pixels
// on even columns, wrapped ones on odd columns) and a cell's final value is
// the later write of two slots' winners: its own slot's and the one to its
// left (rasterizer.rs:82-85 writes cells id and id + 1), and the row's
// newline for cell 1.  The tile first derives every slot's winner (depth
// and winner out, the cell into LDS), then resolves each cell whose two
// slots it holds; a first-column cell takes its own slot's winner here and
// is finished by k_tile_edges, from the winners this tile and its left
// neighbour leave in `edge_out` (first and last column, per row).
// Tile (tx, ty) of frame fr, global tile index gt, whose list is pool[o, o + n).
// FLAT is that of the calling kernel and changes nothing here: it gives each
// kernel its own copy of the lambdas below, which the compiler then inlines
// as it does for a single caller (shared by the two kernels, they cost the
// 3-D instances two VGPRs spilled to scratch).
template <bool BIG, bool ONE, bool ODD, bool FLAT>
__device__ __forceinline__ void raster_tile(const Draw* __restrict__ draws, const Ramp* __restrict__ ramps,
                                            const Ramp* s_ramp, const Frame& fr, uint32_t tx, uint32_t ty, uint64_t gt,
                                            uint64_t o, uint32_t n, int tws, int ths, const uint32_t* __restrict__ pool,
                                            const int32_t* __restrict__ nl, const uint4* __restrict__ careful_box,
                                            uint32_t* __restrict__ cells, float* __restrict__ zbuf,
                                            int32_t* __restrict__ win, uint32_t win_raw, uint4* __restrict__ edge_out,
                                            uint64_t* smem)
{
    static_assert(!(BIG && ODD), "the big-box walk maps slots of even widths only");
    const uint32_t W = fr.W, H = fr.H;
    const int sws = tws - (ODD ? 0 : 1);
    const uint32_t S = 1u << (sws + ths);
    TileGeom g;
    g.tc0 = (int)(tx << tws);
    g.tc1 = min((int)W, g.tc0 + (1 << tws));
    g.tr0 = (int)(ty << ths);
    g.tr1 = min((int)H, g.tr0 + (1 << ths));
    g.sws = sws;
    g.cs = ODD ? 0 : 1;
    g.W = (int)W;

    const uint32_t* list = pool + o;

    const int32_t* fnl = nl + fr.nl_off;
    uint32_t* fc = cells + fr.cell_off;
    float* fz = zbuf + fr.z_off;
    // A thread owns the slots of one column of the tile, every row_step-th
    // row (the block is at least one tile row of slots wide).
    const uint32_t my_sc = threadIdx.x & ((1u << sws) - 1);
    const int my_col = g.tc0 + ((int)my_sc << g.cs);
    const int row_step = (int)(blockDim.x >> sws);
    const uint64_t id_step = (uint64_t)row_step * W;
    // fn(slot, row, id of the slot's first cell) over this thread's slots
    // inside the frame
    auto for_my_slots = [&](auto&& fn) {
        if (my_col >= g.tc1) return;
        int row = g.tr0 + (int)(threadIdx.x >> sws);
        uint64_t id = (uint64_t)row * W + (uint64_t)my_col;
        for (; row < g.tr1; row += row_step, id += id_step)
            fn(((uint32_t)(row - g.tr0) << sws) | my_sc, row, id);
    };
    // ODD: the winners' cells of the tile's slots, for the per-cell resolve
    uint32_t* wcell = (uint32_t*)(smem + S) + CHUNK * CACHE_WORDS;
    // Stores the cells and depths of slot s from its winning cell (or
    // blank); ODD: the depth and winner, and the cell into wcell.
    auto store = [&](uint32_t s, int row, uint64_t id, uint64_t k, uint32_t cell, float z) {
        if constexpr (ODD) {
            const int32_t wt = (k == KEY_EMPTY) ? -1 : (int32_t)tile_key_tri(k);
            __builtin_nontemporal_store(k == KEY_EMPTY ? F32_MAX : z, fz + id);
            if (win) {
                const int32_t wv = (k == KEY_EMPTY) ? -1 : (win_raw ? (int32_t)((uint32_t)wt << 1 | tile_key_sub(k)) : wt);
                __builtin_nontemporal_store(wv, win + fr.z_off + id);
            }
            if (k != KEY_EMPTY) wcell[s] = cell;
            return;
        }
        uint32_t c0 = CELL_BLANK, c1 = CELL_BLANK;
        float z0 = F32_MAX;
        int32_t wt = -1;
        if (k != KEY_EMPTY) {
            c0 = c1 = cell;
            z0 = z;
            wt = (int32_t)tile_key_tri(k);
        }
        if (fr.image && my_col == 0) {
            // cell row*W + 1: the newline of row `row` comes after every pixel
            // of its triangle's row and before its next row, so it overwrites
            // a wrapped pixel (drawn in row-1) of the same or an earlier
            // triangle
            const int32_t nlr = fnl[row];
            if (nlr >= 0 && (k == KEY_EMPTY || nlr >= wt)) c1 = CELL_NEWLINE;
        }
        // streaming stores: the frame is written once and not read back here
        __builtin_nontemporal_store((uint64_t)c0 | ((uint64_t)c1 << 32), (uint64_t*)(fc + id));
        __builtin_nontemporal_store((uint64_t)__float_as_uint(z0) | ((uint64_t)__float_as_uint(F32_MAX) << 32),
                                    (uint64_t*)(fz + id));
        if (win) {
            // win_raw gives triangle << 1 | sub for the shading pass
            const int32_t wv = (k == KEY_EMPTY) ? -1 : (win_raw ? (int32_t)((uint32_t)wt << 1 | tile_key_sub(k)) : wt);
            __builtin_nontemporal_store((uint64_t)(uint32_t)wv | ((uint64_t)0xFFFFFFFFu << 32),
                                        (uint64_t*)(win + fr.z_off + id));
        }
    };
    // Even widths, final stores: a thread takes the slot pairs (2q, 2q + 1) of
    // a row, q = thread mod (slots a row / 2), every (512 >> sws)-th row, and
    // writes a pair's 4 cells, 2 depth pairs and winners with one 16-byte
    // store each where the address is 16-byte aligned (two 8-byte ones
    // otherwise).  fn(first slot, row, id of its first cell, that cell's
    // column, the second slot lies in the frame).
    auto for_my_pairs = [&](auto&& fn) {
        const uint32_t q = threadIdx.x & ((1u << (sws - 1)) - 1);
        const int col = g.tc0 + ((int)q << 2);
        if (col >= g.tc1) return;
        const int step = (int)(blockDim.x >> (sws - 1));
        int row = g.tr0 + (int)(threadIdx.x >> (sws - 1));
        uint64_t id = (uint64_t)row * W + (uint64_t)col;
        for (; row < g.tr1; row += step, id += (uint64_t)step * W)
            fn(((uint32_t)(row - g.tr0) << sws) | (q << 1), row, id, col, col + 2 < g.tc1);
    };
    auto put_pair = [](uint32_t* p, uint32_t a, uint32_t b, uint32_t c, uint32_t d, bool first, bool second) {
        typedef unsigned int v4u __attribute__((ext_vector_type(4)));
        if (first && second && ((uintptr_t)p & 15u) == 0) {
            const v4u v = {a, b, c, d};
            __builtin_nontemporal_store(v, (v4u*)p);
            return;
        }
        if (first) __builtin_nontemporal_store((uint64_t)a | ((uint64_t)b << 32), (uint64_t*)p);
        if (second) __builtin_nontemporal_store((uint64_t)c | ((uint64_t)d << 32), (uint64_t*)(p + 2));
    };
    // Stores the slots of a pair that `need0` / `need1` name, as store()
    // stores one slot (winning cell or blank, depth or f32::MAX, winner).
    auto store_pair = [&](int row, uint64_t id, int col, uint64_t k0, uint32_t c0, float z0, bool need0, uint64_t k1,
                          uint32_t c1, float z1, bool need1) {
        const int32_t t0 = (k0 == KEY_EMPTY) ? -1 : (int32_t)tile_key_tri(k0);
        const int32_t t1 = (k1 == KEY_EMPTY) ? -1 : (int32_t)tile_key_tri(k1);
        uint32_t c0b = c0;
        if (fr.image && col == 0) {
            // cell row*W + 1, as in store()
            const int32_t nlr = fnl[row];
            if (nlr >= 0 && (k0 == KEY_EMPTY || nlr >= t0)) c0b = CELL_NEWLINE;
        }
        put_pair(fc + id, c0, c0b, c1, c1, need0, need1);
        put_pair((uint32_t*)(fz + id), __float_as_uint(z0), __float_as_uint(F32_MAX), __float_as_uint(z1),
                 __float_as_uint(F32_MAX), need0, need1);
        if (win) {
            auto wv = [&](uint64_t k, int32_t t) -> uint32_t {
                return (uint32_t)((k == KEY_EMPTY) ? -1 : (win_raw ? (int32_t)((uint32_t)t << 1 | tile_key_sub(k)) : t));
            };
            put_pair((uint32_t*)(win + fr.z_off + id), wv(k0, t0), 0xFFFFFFFFu, wv(k1, t1), 0xFFFFFFFFu, need0, need1);
        }
    };

    if (n == 0) {
        // no triangle reaches the tile: blank cells, f32::MAX depths, and the
        // newline cells of column 1
        if constexpr (ODD)
            for_my_slots([&](uint32_t s, int row, uint64_t id) { store(s, row, id, KEY_EMPTY, CELL_BLANK, F32_MAX); });
        else
            for_my_pairs([&](uint32_t, int row, uint64_t id, int col, bool two) {
                store_pair(row, id, col, KEY_EMPTY, CELL_BLANK, F32_MAX, true, KEY_EMPTY, CELL_BLANK, F32_MAX, two);
            });
    } else {
        // glyph of shade sh under shader r: the first LDS_RAMPS shaders from
        // LDS, the others from global memory (two calls, so that each reads
        // its own address space rather than through generic pointers)
        auto glyph_of = [&](uint32_t r, float sh) -> uint32_t {
            if (r < LDS_RAMPS) return ramp_glyph(s_ramp[r], sh);
            return ramp_glyph(ramps[r], sh);
        };
        // The glyph a triangle gives every pixel of its box [X0,X1] x [Y0,Y1]
        // when that is provably one glyph, else ~0u.  The pixel's shade is
        // fl(na * S) with S = fl(fl(w0 + w1) + w2) (rasterizer.rs:76).  Each
        // FP32 edge function errs from the exact one by at most
        // 4.0003u (|A| |py - ay| + |B| |px - ax|) (u = 2^-24), the exact ones
        // sum to the exact orient, the two additions add at most 2.0001u of
        // their inputs' magnitudes, and the FP32 orient o errs from the exact
        // one by at most 4.0002u (|A2| |y3 - y1| + |B2| |x3 - x1|).  So every
        // S lies within E = 10u (Sb + Q) of o, with Sb the edge terms' sum
        // over the box and Q o's terms; 10u also covers the rounding of E
        // and of o -+ E (Sb >= |o|).  fl(na * .) and the step index are
        // monotone, so equal indices at the two ends mean one glyph for the
        // whole box.  Shaders that are not non-decreasing, and non-finite
        // values, take the per-pixel evaluation.
        //
        // na comes here from the hardware reciprocal square root: na_h =
        // fl(fl(nz rsq(n2)) a), rsq within 1 ulp, so for n2 in [2^-100,
        // 2^100] na_h lies within 2^-21 |na| of the reference's fl(fl(nz /
        // sqrt(n2)) a) (two roundings each side).  Every fl(na S) then lies
        // between fl(na_h (o - E')) and fl(na_h (o + E')) with E' = E +
        // 2^-19 (|o| + E).  The lower end's step index i is counted, and the
        // upper end compared with threshold i: both ends have index i when
        // it is at most thr[i] (thresholds are non-decreasing).
        auto box_glyph = [&](const Ramp& rp, const Tri& t, float na, int bx0, int bx1, int by0, int by1) -> uint32_t {
            if (!rp.sorted) return ~0u;
            // the edge terms' exact maximum over the box, edge by edge (an
            // empty box has no pixels, so its glyph is never used); the
            // cheaper edge_terms_bound would also do, but its schedule costs
            // this kernel a register spill
            const float X0 = (float)bx0, X1 = (float)bx1, Y0 = (float)by0, Y1 = (float)by1;
            const float Sb =
                fabsf(t.A0) * fmaxf(fabsf(Y0 - t.y2), fabsf(Y1 - t.y2)) + fabsf(t.B0) * fmaxf(fabsf(X0 - t.x2), fabsf(X1 - t.x2)) +
                fabsf(t.A1) * fmaxf(fabsf(Y0 - t.y3), fabsf(Y1 - t.y3)) + fabsf(t.B1) * fmaxf(fabsf(X0 - t.x3), fabsf(X1 - t.x3)) +
                fabsf(t.A2) * fmaxf(fabsf(Y0 - t.y1), fabsf(Y1 - t.y1)) + fabsf(t.B2) * fmaxf(fabsf(X0 - t.x1), fabsf(X1 - t.x1));
            const float ey = t.y3 - t.y1, ex = t.x3 - t.x1;
            const float o = t.A2 * ey - t.B2 * ex;
            const float E = 5.9604645e-07f * (Sb + fabsf(t.A2) * fabsf(ey) + fabsf(t.B2) * fabsf(ex));
            const float E2 = E + 0x1p-19f * (fabsf(o) + E);
            const float s1 = na * (o - E2), s2 = na * (o + E2);
            // NaN or infinity anywhere fails the finite test (v_maximum3 propagates NaN)
            const float mag = __builtin_elementwise_maximum(__builtin_elementwise_maximum(fabsf(na), fabsf(s1)), fabsf(s2));
            if (!(mag < INFINITY)) return ~0u;
            const float lo = fminf(s1, s2), hi = fmaxf(s1, s2);
            const uint32_t i = ramp_index_sorted(rp, lo);
            if (i < rp.n && !(hi <= rp.thr[i])) return ~0u;
            return i < rp.n ? rp.glyph[i] : rp.fallback;
        };
        uint64_t* keys = smem;
        float* cache = (float*)(keys + S);
        const bool one_draw = fr.draw_end - fr.draw_begin == 1;
        const bool multi = n > CHUNK;

        // Winner of a slot with key k (written in the current chunk), from
        // the cache row the key names.  A row whose triangle has one glyph
        // over its box holds the finished cell (low byte of its last word
        // 0xFF); the depth then comes from the key's code, except for a
        // zero depth, whose sign the code does not keep.  Other winners are
        // re-evaluated from the transformed triangle.
        auto winner = [&](int col, int row, uint64_t k, uint32_t& cell, float& z) {
            const uint32_t sub = tile_key_sub(k);
            const float4* row4 = (const float4*)(cache + tile_key_row(k) * CACHE_WORDS);
            const float4 q2 = row4[2];
            const uint32_t cw = __float_as_uint(q2.w);
            const uint32_t code = (uint32_t)(k >> 32);
            const bool fast = (cw & 0xFFu) == 0xFFu;
            if (fast && code != DEPTH_CODE_ZERO) {
                cell = __float_as_uint(q2.z);
                z = depth_of_code(code);
                return;
            }
            const float4 q0 = row4[0], q1 = row4[1];
            const float px = (float)(sub ? (col >> 1) : ((col + g.W) >> 1));
            const float py = (float)(sub ? row : row - 1);
            // the same expressions as edges(), depth() and shade_value()
            const float x1 = q0.x, y1 = q0.y, x2 = q0.z, y2 = q0.w, x3 = q1.x, y3 = q1.y;
            const float w0 = (x3 - x2) * (py - y2) - (y3 - y2) * (px - x2);
            const float w1 = (x1 - x3) * (py - y3) - (y1 - y3) * (px - x3);
            const float w2 = (x2 - x1) * (py - y1) - (y2 - y1) * (px - x1);
            z = q1.z + q2.y * (w1 * q1.w + w2 * q2.x);
            const float sh = q2.z * (w0 + w1 + w2);
            cell = fast ? __float_as_uint(q2.z) : glyph_of(cw & 0xFFu, sh) | (cw & 0xFFFFFF00u);
        };

        for (uint32_t s = threadIdx.x; s < S; s += blockDim.x) keys[s] = KEY_EMPTY;
        for (uint32_t base = 0; base < n; base += CHUNK) {
            __syncthreads();   // keys cleared / the previous chunk's winner stores no longer read its rows
            const uint32_t p = threadIdx.x;
            bool big = false;              // BIG: this lane's box is walked by its wavefront
            uint32_t big_T = 0, big_rect = 0;
            if (base + p < n) {
                int bx0, bx1, by0, by1;
                bool given, tight;
                const uint32_t T = entry_tri(list[base + p], careful_box, bx0, bx1, by0, by1, given, tight);
                const Draw& dr = ONE ? draws[fr.draw_begin] : one_draw ? draws[fr.draw_begin] : draw_of_t(draws, fr, T);
                const uint32_t li = T - dr.tri_begin;
                Tri t;
                setup_draw_tri(dr, li, W, H, t);
                float nz, n2;
                normal_parts(t, nz, n2);
                const auto* c = global_ptr(dr.rgb + (size_t)li * 3);
                const uint32_t rgb = ((uint32_t)c[0] << 8) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 24);
                if (!given) {
                    box_x(t, tight, bx0, bx1);
                    by0 = t.y0;
                    by1 = t.y1d - 1;
                }
                // a triangle with one glyph over its box keeps its finished
                // cell where the shade factor would be, flagged by 0xFF
                uint32_t bg = ~0u;
                if (n2 >= 0x1p-100f && n2 <= 0x1p100f) {
                    const float na_h = nz * __builtin_amdgcn_rsqf(n2) * t.a;
                    const uint32_t r = dr.ramp;
                    bg = r < LDS_RAMPS ? box_glyph(s_ramp[r], t, na_h, bx0, bx1, by0, by1)
                                       : box_glyph(ramps[r], t, na_h, bx0, bx1, by0, by1);
                }
                float w2 = __uint_as_float(bg | rgb);
                // the others keep the shade factor (rasterizer.rs:76
                // multiplies the unit normal's z by a first)
                if (bg == ~0u) w2 = nz / sqrtf(n2) * t.a;
                const uint32_t w3 = rgb | (bg != ~0u ? 0xFFu : dr.ramp);
                float4* row4 = (float4*)(cache + p * CACHE_WORDS);
                row4[0] = make_float4(t.x1, t.y1, t.x2, t.y2);
                row4[1] = make_float4(t.x3, t.y3, t.z1, t.z2 - t.z1);
                row4[2] = make_float4(t.z3 - t.z1, t.a, w2, __uint_as_float(w3));
                RowSpans spans;
                // row spans cost set-up and per-row work that only boxes of
                // some width repay; a wavefront walks as long as its widest
                // box, so it takes them when any lane's box is that wide
                const bool use_spans = __ballot(bx1 - bx0 + 1 >= (int)SPAN_MIN_W) != 0;
                if (use_spans) {
                    if (!given) row_spans_begin(t, bx0, bx1, by0, by1, max(by0, g.tr0), spans);
                    else row_spans_none(spans);
                }
                if (bx0 <= bx1 && by0 <= by1) {
                    auto test = [&](float px, float py, int slot, uint32_t sub) {
                        float w0, w1, w2;
                        edges(t, px, py, w0, w1, w2);
                        const float z = depth(t, w1, w2);
                        if (inside(w0, w1, w2) && z < F32_MAX)
                            atomicMin((unsigned long long*)&keys[slot], (unsigned long long)make_tile_key(z, T, sub, p));
                    };
                    if constexpr (BIG) {
                        int xa, xb, ya, yb;
                        direct_rect(g, bx0, bx1, by0, by1, xa, xb, ya, yb);
                        if (xa <= xb && ya <= yb && (xb - xa + 1) * (yb - ya + 1) > (int)BIG_BOX) {
                            big = true;
                            big_T = T;
                            big_rect =
                                (uint32_t)xa | ((uint32_t)xb << 8) | ((uint32_t)ya << 16) | ((uint32_t)yb << 24);
                        } else {
                            if (use_spans)
                                walk_rect_spans(xa + (g.tc0 >> 1), xb + (g.tc0 >> 1), ya + g.tr0, yb + g.tr0,
                                                (ya << sws) + xa, sws, 1, spans, test);
                            else
                                walk_rect(xa + (g.tc0 >> 1), xb + (g.tc0 >> 1), ya + g.tr0, yb + g.tr0,
                                          (ya << sws) + xa, sws, 1, 1u, test);
                        }
                        walk_wrapped(g, bx0, bx1, by0, by1, test);
                    } else {
                        walk_tile(g, bx0, bx1, by0, by1, use_spans, spans, test);
                    }
                }
            }
            if constexpr (BIG) {
                // The wavefront's big boxes, one at a time: lane l of the box
                // takes pixels i = l, l + 64, ... (row i / nx, column i % nx),
                // with the expressions of edges() and depth() on the box's
                // cache row (row p of the lane that holds it, written by this
                // wavefront above).
                uint64_t m = __ballot(big);
                while (m) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1;
                    const uint32_t T = __builtin_amdgcn_readlane(big_T, l), rc = __builtin_amdgcn_readlane(big_rect, l);
                    const uint32_t prow = (threadIdx.x & ~63u) + (uint32_t)l;   // the owner lane's cache row
                    const float4* row4 = (const float4*)(cache + prow * CACHE_WORDS);
                    const float4 q0 = row4[0], q1 = row4[1], q2 = row4[2];
                    const float x1 = q0.x, y1 = q0.y, x2 = q0.z, y2 = q0.w, x3 = q1.x, y3 = q1.y;
                    const float A0 = x3 - x2, B0 = y3 - y2, A1 = x1 - x3, B1 = y1 - y3, A2 = x2 - x1, B2 = y2 - y1;
                    const float z1 = q1.z, dz2 = q1.w, dz3 = q2.x, a = q2.y;
                    const int xa = (int)(rc & 0xFFu), nx = (int)((rc >> 8) & 0xFFu) - xa + 1;
                    const int ya = (int)((rc >> 16) & 0xFFu), ny = (int)(rc >> 24) - ya + 1;
                    const int dy = 64 / nx, dx = 64 - dy * nx;
                    int i = (int)__lane_id(), yy = i / nx, xx = i - yy * nx;
                    for (; i < nx * ny; i += 64) {
                        const int x = xa + xx, y = ya + yy;
                        const float px = (float)(x + (g.tc0 >> 1)), py = (float)(y + g.tr0);
                        const float w0 = A0 * (py - y2) - B0 * (px - x2);
                        const float w1 = A1 * (py - y3) - B1 * (px - x3);
                        const float w2 = A2 * (py - y1) - B2 * (px - x1);
                        const float z = z1 + a * (w1 * dz2 + w2 * dz3);
                        if (inside(w0, w1, w2) && z < F32_MAX)
                            atomicMin((unsigned long long*)&keys[(y << sws) + x],
                                      (unsigned long long)make_tile_key(z, T, 1u, prow));
                        xx += dx;
                        yy += dy;
                        if (xx >= nx) {
                            xx -= nx;
                            ++yy;
                        }
                    }
                }
            }
            __syncthreads();
            if (multi)
                // slots whose key was written in this chunk (no stored flag)
                // store that winner and flag the key; a later chunk that
                // changes the key stores again.  The same thread owns slot s
                // throughout, so its last store lands last.
                for_my_slots([&](uint32_t s, int row, uint64_t id) {
                    const uint64_t k = keys[s];
                    if (k == KEY_EMPTY || (k & TILE_KEY_STORED)) return;
                    uint32_t cell;
                    float z;
                    winner(my_col, row, k, cell, z);
                    store(s, row, id, k, cell, z);
                    keys[s] = k | TILE_KEY_STORED;
                });
        }
        if constexpr (ODD) {
            for_my_slots([&](uint32_t s, int row, uint64_t id) {
                const uint64_t k = keys[s];
                if (multi && k != KEY_EMPTY) return;   // stored after its chunk
                uint32_t cell = CELL_BLANK;
                float z = F32_MAX;
                if (k != KEY_EMPTY) winner(my_col, row, k, cell, z);
                store(s, row, id, k, cell, z);
            });
        } else {
            for_my_pairs([&](uint32_t s, int row, uint64_t id, int col, bool two) {
                const uint64_t k0 = keys[s], k1 = two ? keys[s + 1] : KEY_EMPTY;
                const bool need0 = !(multi && k0 != KEY_EMPTY);   // else stored after its chunk
                const bool need1 = two && !(multi && k1 != KEY_EMPTY);
                if (!need0 && !need1) return;
                uint32_t c0 = CELL_BLANK, c1 = CELL_BLANK;
                float z0 = F32_MAX, z1 = F32_MAX;
                if (need0 && k0 != KEY_EMPTY) winner(col, row, k0, c0, z0);
                if (need1 && k1 != KEY_EMPTY) winner(col + 2, row, k1, c1, z1);
                store_pair(row, id, col, k0, c0, z0, need0, k1, c1, z1, need1);
            });
        }
    }
    if constexpr (ODD) {
        // Each cell: the later write of its own slot's winner and the left
        // slot's (write time = triangle, then the pixel's row and column,
        // as the sequential loop orders them), and for cell 1 the row's
        // newline, which follows its triangle's pixels of that row.
        const uint64_t* keys = smem;
        const bool any = n > 0;
        if (any) __syncthreads();   // every slot's winner cell is in wcell
        auto slot_time = [&](uint64_t k, int row, int col) -> uint64_t {
            const uint32_t sub = tile_key_sub(k);
            const uint32_t py = sub ? (uint32_t)row : (uint32_t)row - 1u;
            const uint32_t px = sub ? (uint32_t)col >> 1 : (uint32_t)(col + (int)W) >> 1;
            return (((uint64_t)tile_key_tri(k) << 32) | ((uint64_t)py << 16) | px) + 1;
        };
        for_my_slots([&](uint32_t s, int row, uint64_t id) {
            uint64_t t_own = 0, t_best = 0;
            uint32_t c_own = CELL_BLANK, best = CELL_BLANK;
            if (any) {
                const uint64_t k = keys[s];
                if (k != KEY_EMPTY) {
                    t_own = slot_time(k, row, my_col);
                    c_own = wcell[s];
                }
                t_best = t_own;
                best = c_own;
                if (my_col > g.tc0) {
                    const uint64_t kl = keys[s - 1];
                    if (kl != KEY_EMPTY) {
                        const uint64_t tl = slot_time(kl, row, my_col - 1);
                        if (tl > t_best) {
                            t_best = tl;
                            best = wcell[s - 1];
                        }
                    }
                }
            }
            if (fr.image && my_col == 1) {
                const int32_t nlr = fnl[row];
                if (nlr >= 0) {
                    const uint64_t tn = (((uint64_t)(uint32_t)nlr << 32) | ((uint64_t)row << 16) | 0xFFFFu) + 1;
                    if (tn > t_best) best = CELL_NEWLINE;
                }
            }
            __builtin_nontemporal_store(best, fc + id);
            uint4* e = edge_out + (gt * (1u << ths) + (uint64_t)(row - g.tr0)) * 2;
            if (my_col == g.tc0) e[0] = make_uint4(c_own, 0u, (uint32_t)t_own, (uint32_t)(t_own >> 32));
            if (my_col == g.tc1 - 1) e[1] = make_uint4(c_own, 0u, (uint32_t)t_own, (uint32_t)(t_own >> 32));
        });
    }
    // image mode: the H cells after the W*H grid are never written (context.rs:38-40)
    if (fr.image && tx == 0 && ty == 0)
        for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) __builtin_nontemporal_store(CELL_BLANK, fc + (uint64_t)W * H + i);
}

// One workgroup per tile; grid (tile x, tile y, frame - f_lo).  Blocks past a
// smaller frame's tile grid return at once.
// Occupancy: LDS allows 7 workgroups of 4 wavefronts per CU.  Left to itself
// the compiler gives the instances 76-80 VGPRs (6 wavefronts per SIMD); held
// to 7 they fit 72 and run ~3% faster.  Most fit with SGPR spills to VGPR
// lanes only; the big-box instance for one draw per frame and the plain one
// for several spill 2 VGPRs to scratch, still faster than at 6 (a batch of
// uneven single-draw frames: 0.559 against 0.587 ms).  The big-box instance
// for several draws per frame would spill more at 72 and keeps 6.
template <bool BIG, bool ONE, bool ODD, bool FLAT>
__global__ void __launch_bounds__(RASTER_BLOCK) __attribute__((amdgpu_waves_per_eu(BIG && !ONE ? 6 : 7, 8))) k_tile_raster(
    const Draw* __restrict__ draws, const Ramp* __restrict__ ramps, const Frame* __restrict__ frames, uint32_t f_lo,
    int tws, int ths, uint32_t* __restrict__ cursor, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ pool, int32_t* __restrict__ nl, const uint4* __restrict__ careful_box,
    uint32_t* __restrict__ cells, float* __restrict__ zbuf, int32_t* __restrict__ win, uint32_t win_raw,
    uint32_t n_ramps, uint4* __restrict__ edge_out, uint32_t* __restrict__ careful_reset,
    const uint2* __restrict__ tile_frame, uint32_t t_lo)
{
    extern __shared__ uint64_t smem[];
    // the job's first LDS_RAMPS shaders, so that a glyph lookup reads LDS
    // (read only after the tile's first barrier)
    __shared__ Ramp s_ramp[LDS_RAMPS];

    // the careful list was read by k_careful, which ran before this launch
    if (careful_reset && (blockIdx.x | blockIdx.y | blockIdx.z | threadIdx.x) == 0) careful_reset[0] = 0;
    // the block's tile: (x, y, frame) from a 3-D grid, or (FLAT: frames of
    // different tile grids) the (tile, frame) of entry t_lo + x of a flat one
    uint32_t f, tx, ty;
    if constexpr (FLAT) {
        const uint2 e = tile_frame[t_lo + blockIdx.x];
        f = e.y;
        const uint32_t local = e.x - (uint32_t)frames[f].tile_begin;
        ty = local / frames[f].tiles_x;
        tx = local - ty * frames[f].tiles_x;
    } else {
        f = f_lo + blockIdx.z;
        tx = blockIdx.x;
        ty = blockIdx.y;
    }
    const Frame& fr = frames[f];
    if (tx >= fr.tiles_x || ty >= fr.tiles_y || ((fr.W & 1u) != 0) != ODD) return;
    const uint64_t gt = fr.tile_begin + (uint64_t)ty * fr.tiles_x + tx;
    const uint64_t o = off[gt];
    const uint32_t c = cursor[gt];
    const uint32_t n = min(c, (uint32_t)(off[gt + 1] - o));
    if (n > 0)
        for (uint32_t i = threadIdx.x; i < min(n_ramps, LDS_RAMPS) * (uint32_t)sizeof(Ramp) / 4; i += blockDim.x)
            ((uint32_t*)s_ramp)[i] = ((const uint32_t*)ramps)[i];
    raster_tile<BIG, ONE, ODD, FLAT>(draws, ramps, s_ramp, fr, tx, ty, gt, o, n, tws, ths, pool, nl, careful_box, cells, zbuf,
                               win, win_raw, edge_out, smem);
    // The tile leaves its binning state as k_tile_clear would for the next
    // render: its cursor at 0 and, for even widths (k_tile_edges reads odd
    // frames' rows after this), the newline rows of its tile row back at -1
    // (the first tile of the row is the one that read them).  Every wave
    // has read the cursor and the rows once the block has met here.
    const bool reset_nl = !ODD && fr.image && tx == 0;
    if (c != 0 || reset_nl) {
        __syncthreads();
        if (c != 0 && threadIdx.x == 0) cursor[gt] = 0;
        if (reset_nl) {
            const uint32_t r0 = ty << ths, r1 = min(fr.H, r0 + (1u << ths));
            for (uint32_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) nl[fr.nl_off + r] = -1;
        }
    }
}

// Frames of odd width: the first-column cells of the tiles.  Cell (r, c)
// with c the first column of a tile is the later write of slot (r, c)'s
// winner and slot (r, c - 1)'s, which the tile to the left holds; cell
// (r, 0) takes the last slot of row r - 1 instead (id + 1 of the row's last
// direct pixel wraps to the next row).  The tiles left each first and last
// column slot's winning cell and write time in `edges` (time 0: none).  A
// frame one column wide has no pixels (the domain's x range is empty), and
// its cell (r, 0) is row r - 1's newline cell (r - 1) * W + 1.
// Grid (rows / 256, tile columns, frame - f_lo).
__global__ void __launch_bounds__(256) k_tile_edges(const Frame* __restrict__ frames, uint32_t f_lo, int tws,
                                                    int ths, const uint4* __restrict__ edges,
                                                    const int32_t* __restrict__ nl, uint32_t* __restrict__ cells)
{
    const Frame& fr = frames[f_lo + blockIdx.z];
    const uint32_t b = blockIdx.y;
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (!(fr.W & 1u) || b >= fr.tiles_x || r >= fr.H || (b == 0 && r == 0)) return;
    const uint32_t TH = 1u << ths;
    const uint64_t own = fr.tile_begin + (uint64_t)(r >> ths) * fr.tiles_x + b;
    const uint4 e_own = edges[(own * TH + (r & (TH - 1))) * 2];
    uint4 e_left;
    if (b > 0) {
        e_left = edges[((own - 1) * TH + (r & (TH - 1))) * 2 + 1];
    } else {
        const uint32_t rl = r - 1;
        const uint64_t left = fr.tile_begin + (uint64_t)(rl >> ths) * fr.tiles_x + (fr.tiles_x - 1);
        e_left = edges[(left * TH + (rl & (TH - 1))) * 2 + 1];
    }
    const uint64_t t_own = ((uint64_t)e_own.w << 32) | e_own.z, t_left = ((uint64_t)e_left.w << 32) | e_left.z;
    uint32_t cell = CELL_BLANK;
    if (t_own | t_left) cell = t_own > t_left ? e_own.x : e_left.x;
    if (fr.W == 1 && fr.image && nl[fr.nl_off + r - 1] >= 0) cell = CELL_NEWLINE;
    cells[fr.cell_off + (uint64_t)r * fr.W + ((uint64_t)b << tws)] = cell;
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
    if (tw < 8 || tw > 512 || th < 4 || th > 64 || (tw & (tw - 1)) || (th & (th - 1)))
        return fail(SLOTH_EINVAL, "tile size %ux%u: width must be a power of two in 8..512, height in 4..64", tw, th);
    if (tw != TW || th != TH) sized = false;
    TW = tw;
    TH = th;
    user_tile = true;
    return SLOTH_OK;
}

size_t TiledPlan::raster_lds(bool odd) const
{
    // keys (8 B per slot), triangle rows; odd widths: a slot per cell and
    // the slots' winning cells (4 B per slot)
    size_t S = (size_t)(odd ? TW : TW / 2) * TH;
    return S * 8 + (size_t)CHUNK * CACHE_WORDS * 4 + (odd ? S * 4 : 0);
}

bool TiledPlan::tile_fits() const
{
    // the raster's threads own whole columns of slots: a tile row holds at
    // most one workgroup's worth (odd widths: a slot per cell)
    auto fits = [&](bool odd) { return raster_lds(odd) <= 65536 && (odd ? TW : TW / 2) <= (uint32_t)RASTER_BLOCK; };
    return (!has_even || fits(false)) && (!has_odd || fits(true));
}

int TiledPlan::plan(std::vector<Frame>& frames, uint64_t rows)
{
    bool ev = false, od = false;
    for (const Frame& f : frames) ((f.W & 1u) ? od : ev) = true;
    if (ev != has_even || od != has_odd) sized = false;
    has_even = ev;
    has_odd = od;
    // odd widths hold twice the slots per tile: half the default height
    if (!user_tile) TH = has_odd ? 16 : 32;
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
    if ((rc = d_careful_cnt.reserve(sizeof(uint32_t) * (TALL_CNT + 1)))) return rc;
    if (has_odd && (rc = d_edges.reserve(sizeof(uint4) * 2 * std::max<uint64_t>(1, n_tiles) * TH))) return rc;
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
                       d_nl.as<int32_t>() + r_lo, r_hi - r_lo, d_overflow.as<uint32_t>() + f_lo, f_hi - f_lo,
                       d_careful_cnt.as<uint32_t>());
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

int TiledPlan::launch_bin(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
                          const Draw* d_draws, uint32_t f_lo, uint32_t f_hi, bool count_only, hipStream_t s)
{
    const uint32_t d_lo = frames[f_lo].draw_begin, d_hi = frames[f_hi - 1].draw_end;
    const int tws = log2u(TW), ths = log2u(TH);
    auto go = [&](dim3 grid, uint32_t d0, const uint2* flat) {
        auto run = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid, dim3(BIN_BLOCK), 0, s, d_draws, d0, d_frames, tws, ths,
                               d_cursor.as<uint32_t>(), d_off.as<uint64_t>(), d_pool.as<uint32_t>(),
                               d_nl.as<int32_t>(), d_overflow.as<uint32_t>(), d_careful_cnt.as<uint32_t>(),
                               d_careful_in.as<uint4>(), careful_cap, flat);
        };
        // the counting run takes the tree variant (it counts the tall triangles)
        if (bin_tpt == 1)
            count_only ? run(k_bin<true, 1, true>) : bin_tree ? run(k_bin<false, 1, true>) : run(k_bin<false, 1, false>);
        else if (bin_tpt == 2)
            count_only ? run(k_bin<true, 2, true>) : bin_tree ? run(k_bin<false, 2, true>) : run(k_bin<false, 2, false>);
        else
            count_only ? run(k_bin<true, 4, true>) : bin_tree ? run(k_bin<false, 4, true>) : run(k_bin<false, 4, false>);
        HIP_TRY(hipGetLastError());
        return SLOTH_OK;
    };
    if (flat_bin) {
        // one block per (draw, BIN_BLOCK * TPT triangles) of the range
        const int ti = bin_tpt == 1 ? 0 : bin_tpt == 2 ? 1 : 2;
        const uint64_t b0 = bin_block_begin[ti][d_lo], b1 = bin_block_begin[ti][d_hi];
        if (b1 - b0 > 0x7FFFFFFFull) return fail(SLOTH_EINVAL, "too many binning blocks in one launch");
        if (b1 > b0) return go(dim3((unsigned)(b1 - b0)), d_lo, d_bin_blocks[ti].as<uint2>() + b0);
        return SLOTH_OK;
    }
    int rc;
    for (uint32_t d0 = d_lo; d0 < d_hi; d0 += 65535u) {
        const uint32_t d1 = std::min(d_hi, d0 + 65535u);
        uint32_t max_n = 0;
        for (uint32_t d = d0; d < d1; ++d) max_n = std::max(max_n, draws[d].n);
        if (max_n == 0) continue;
        if ((rc = go(dim3(blocks(max_n, BIN_BLOCK * (uint32_t)bin_tpt), d1 - d0), d0, nullptr))) return rc;
    }
    return SLOTH_OK;
}

int TiledPlan::launch_careful(const Frame* d_frames, const Draw* d_draws, bool count_only, uint32_t group,
                              hipStream_t s)
{
    if (careful_cap == 0) return SLOTH_OK;
    const int tws = log2u(TW), ths = log2u(TH);
    // a grid of up to 16,384 wavefronts (the chip holds about 8,192 of
    // these at once) strides over the careful list, one triangle per
    // wavefront at a time (the sweep's pass 0.015 -> 0.012 ms against 4,096)
    dim3 grid(std::min(blocks(careful_cap, 4), 4096u));
    if (count_only)
        hipLaunchKernelGGL(k_careful<true>, grid, dim3(256), 0, s, d_frames, tws, ths,
                           d_cursor.as<uint32_t>(), d_off.as<uint64_t>(), d_pool.as<uint32_t>(),
                           d_overflow.as<uint32_t>(), d_careful_cnt.as<uint32_t>(), group, d_careful_in.as<uint4>(),
                           d_careful_box.as<uint4>(), careful_cap);
    else
        hipLaunchKernelGGL(k_careful<false>, grid, dim3(256), 0, s, d_frames, tws, ths,
                           d_cursor.as<uint32_t>(), d_off.as<uint64_t>(), d_pool.as<uint32_t>(),
                           d_overflow.as<uint32_t>(), d_careful_cnt.as<uint32_t>(), group, d_careful_in.as<uint4>(),
                           d_careful_box.as<uint4>(), careful_cap);
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

int TiledPlan::size(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
                    const Draw* d_draws, hipStream_t s)
{
    if (frames.empty()) return SLOTH_OK;
    const uint32_t F = (uint32_t)frames.size();
    int rc;
    // 0. flat k_bin grids (draws of different sizes): each block's (draw,
    // first triangle) for TPT 1, 2, 4
    flat_bin = false;
    for (const Draw& d : draws) flat_bin = flat_bin || d.n != draws[0].n;
    if (flat_bin) {
        for (int ti = 0; ti < 3; ++ti) {
            const uint32_t per = (uint32_t)BIN_BLOCK << ti;
            std::vector<uint2> tbl;
            std::vector<uint64_t>& beg = bin_block_begin[ti];
            beg.assign(draws.size() + 1, 0);
            for (size_t d = 0; d < draws.size(); ++d) {
                beg[d] = tbl.size();
                for (uint32_t b = 0; b < draws[d].n; b += per) tbl.push_back(make_uint2((uint32_t)d, b));
            }
            beg[draws.size()] = tbl.size();
            if ((rc = d_bin_blocks[ti].reserve(sizeof(uint2) * std::max<size_t>(1, tbl.size())))) return rc;
            if (!tbl.empty())
                HIP_TRY(hipMemcpyAsync(d_bin_blocks[ti].p, tbl.data(), sizeof(uint2) * tbl.size(),
                                       hipMemcpyHostToDevice, s));
            HIP_TRY(hipStreamSynchronize(s));   // before tbl goes
        }
    }
    // 1. count tile entries of CLS_BOX triangles and the careful triangles
    if ((rc = launch_clear(frames, 0, F, s))) return rc;
    if ((rc = launch_bin(frames, draws, d_frames, d_draws, 0, F, true, s))) return rc;
    uint32_t cnt[TALL_CNT + 1];
    HIP_TRY(hipMemcpyAsync(cnt, d_careful_cnt.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const uint32_t n_careful = cnt[0];
    // the wavefront walk of big boxes pays once about one wavefront in ten
    // holds one (a big box costs its lane's wavefront 10x the common walk)
    uint64_t n_tri = 0;
    for (const Draw& d : draws) n_tri += d.n;
    big_walk = (uint64_t)cnt[BIG_CNT] * 512 >= n_tri && cnt[BIG_CNT] > 0;
    if (const char* e = getenv("SLOTH_BIG_WALK")) big_walk = e[0] == '1';
    // the binning kernel's segment-tree newline rows pay once one triangle
    // in 256 spans more than TALL_ROWS rows (a wavefront's row loop lasts as
    // long as its tallest triangle's)
    const bool tree = (uint64_t)cnt[TALL_CNT] * 256 >= n_tri && cnt[TALL_CNT] > 0;
    bool want_tree = tree;
    if (const char* e = getenv("SLOTH_BIN_TREE")) want_tree = e[0] == '1';
    if (want_tree != bin_tree) ++generation;   // captured graphs hold the other variant
    bin_tree = want_tree;
    if (n_careful > careful_cap) {
        // 2. room for the careful list, then count again so that it is filled
        careful_cap = n_careful + n_careful / 4 + 256;
        if ((rc = d_careful_in.reserve(2 * sizeof(uint4) * careful_cap))) return rc;
        if ((rc = d_careful_box.reserve(sizeof(uint4) * careful_cap))) return rc;
        if ((rc = launch_clear(frames, 0, F, s))) return rc;
        if ((rc = launch_bin(frames, draws, d_frames, d_draws, 0, F, true, s))) return rc;
    }
    // 3. the careful triangles' tile entries
    if ((rc = launch_careful(d_frames, d_draws, true, 0, s))) return rc;
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
    // 4. flat raster grids (frames of different tile grids): each tile's
    // (tile, frame) in tile order, and the same with the tiles of more than
    // four chunks of entries moved to the front, longest first (the rest
    // keep tile order, whose neighbouring tiles share cached vertices and
    // lists: sorting every tile by length made the uneven batch slower)
    flat_raster = false;
    for (const Frame& f : frames)
        flat_raster = flat_raster || f.tiles_x != frames[0].tiles_x || f.tiles_y != frames[0].tiles_y;
    std::vector<uint2> tf, heavy;
    if (flat_raster) {
        tf.resize(n_tiles);
        for (uint32_t f = 0; f < F; ++f)
            for (uint64_t t = 0; t < (uint64_t)frames[f].tiles_x * frames[f].tiles_y; ++t)
                tf[frames[f].tile_begin + t] = make_uint2((uint32_t)(frames[f].tile_begin + t), f);
        heavy = tf;
        const auto split = std::stable_partition(heavy.begin(), heavy.end(),
                                                 [&](const uint2& a) { return counts[a.x] > 4 * CHUNK; });
        std::stable_sort(heavy.begin(), split,
                         [&](const uint2& a, const uint2& b) { return counts[a.x] > counts[b.x]; });
        if ((rc = d_tile_frame.reserve(sizeof(uint2) * std::max<uint64_t>(1, n_tiles))) ||
            (rc = d_tile_heavy.reserve(sizeof(uint2) * std::max<uint64_t>(1, n_tiles))))
            return rc;
        if (n_tiles) {
            HIP_TRY(hipMemcpyAsync(d_tile_frame.p, tf.data(), sizeof(uint2) * n_tiles, hipMemcpyHostToDevice, s));
            HIP_TRY(hipMemcpyAsync(d_tile_heavy.p, heavy.data(), sizeof(uint2) * n_tiles, hipMemcpyHostToDevice, s));
        }
    }
    // renders start from cleared cursors, rows and careful list (each leaves them so)
    if ((rc = launch_clear(frames, 0, F, s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    sized = true;
    ++generation;
    return SLOTH_OK;
}

// The raster instance for a width parity: big-box walk or not, one draw per
// frame or not, flat grid (frames of different tile grids) or 3-D.  Each
// instance is compiled on its own, so the 3-D instances carry no flat-grid
// mapping and its registers.
template <bool ODD, bool FLAT>
static auto raster_kernel(bool walk, bool one)
{
    if constexpr (ODD)
        return one ? k_tile_raster<false, true, true, FLAT> : k_tile_raster<false, false, true, FLAT>;
    else
        return walk ? (one ? k_tile_raster<true, true, false, FLAT> : k_tile_raster<true, false, false, FLAT>)
                    : (one ? k_tile_raster<false, true, false, FLAT> : k_tile_raster<false, false, false, FLAT>);
}

int TiledPlan::launch_raster(const std::vector<Frame>& frames, const Frame* d_frames, const Draw* d_draws,
                             const Ramp* d_ramps, uint32_t f_lo, uint32_t f_hi, uint32_t* cells, float* z, int32_t* win,
                             bool win_raw, uint32_t n_ramps, bool reset_careful, hipStream_t s)
{
    // the first raster launch resets the careful list's length (k_careful
    // has read it) when asked
    uint32_t* careful_reset = reset_careful ? d_careful_cnt.as<uint32_t>() : nullptr;
    // grid (tile x, tile y, frame) over the largest tile grid of the range,
    // at most 65535 frames per launch
    for (uint32_t f0 = f_lo; f0 < f_hi; f0 += 65535u) {
        const uint32_t f1 = std::min(f_hi, f0 + 65535u);
        uint32_t mx = 0, my = 0;
        for (uint32_t f = f0; f < f1; ++f) {
            mx = std::max(mx, frames[f].tiles_x);
            my = std::max(my, frames[f].tiles_y);
        }
        if (mx == 0 || my == 0) continue;
        bool one = true;   // every frame of the range has exactly one draw
        bool ev = false, od = false;
        for (uint32_t f = f0; f < f1; ++f) {
            one = one && frames[f].draw_end - frames[f].draw_begin == 1;
            ((frames[f].W & 1u) ? od : ev) = true;
        }
        // one launch per width parity in the range; each skips the other's
        // frames.  The grid: (tile x, tile y, frame) over the largest tile
        // grid, or one workgroup per tile when the frames' tile grids differ
        const uint64_t t_lo = frames[f0].tile_begin;
        const uint64_t t_hi = frames[f1 - 1].tile_begin + (uint64_t)frames[f1 - 1].tiles_x * frames[f1 - 1].tiles_y;
        if (flat_raster && t_hi > 0x7FFFFFFFull) return fail(SLOTH_EINVAL, "too many tiles for a flat raster grid");
        const dim3 grid = flat_raster ? dim3((unsigned)(t_hi - t_lo)) : dim3(mx, my, f1 - f0);
        const uint2* tf = !flat_raster ? nullptr
                          : f0 == 0 && f1 == (uint32_t)frames.size() ? d_tile_heavy.as<uint2>()
                                                                      : d_tile_frame.as<uint2>();
        if (ev) {
            const auto kern = flat_raster ? raster_kernel<false, true>(walk(), one) : raster_kernel<false, false>(walk(), one);
            hipLaunchKernelGGL(kern, grid, dim3(RASTER_BLOCK), raster_lds(false), s, d_draws,
                               d_ramps, d_frames, f0, log2u(TW), log2u(TH), d_cursor.as<uint32_t>(),
                               d_off.as<uint64_t>(), d_pool.as<uint32_t>(), d_nl.as<int32_t>(),
                               d_careful_box.as<uint4>(), cells, z, win, win_raw ? 1u : 0u, n_ramps, (uint4*)nullptr,
                               careful_reset, tf, (uint32_t)t_lo);
            HIP_TRY(hipGetLastError());
            careful_reset = nullptr;
        }
        if (od) {
            const auto kern = flat_raster ? raster_kernel<true, true>(false, one) : raster_kernel<true, false>(false, one);
            hipLaunchKernelGGL(kern, grid, dim3(RASTER_BLOCK), raster_lds(true), s, d_draws,
                               d_ramps, d_frames, f0, log2u(TW), log2u(TH), d_cursor.as<uint32_t>(),
                               d_off.as<uint64_t>(), d_pool.as<uint32_t>(), d_nl.as<int32_t>(),
                               d_careful_box.as<uint4>(), cells, z, win, win_raw ? 1u : 0u, n_ramps,
                               d_edges.as<uint4>(), careful_reset, tf, (uint32_t)t_lo);
            HIP_TRY(hipGetLastError());
            careful_reset = nullptr;
            uint32_t mh = 0;
            for (uint32_t f = f0; f < f1; ++f) mh = std::max(mh, frames[f].H);
            hipLaunchKernelGGL(k_tile_edges, dim3(blocks(mh, 256), mx, f1 - f0), dim3(256), 0, s, d_frames, f0,
                               log2u(TW), log2u(TH), (const uint4*)d_edges.as<uint4>(),
                               (const int32_t*)d_nl.as<int32_t>(), cells);
            HIP_TRY(hipGetLastError());
        }
    }
    // no raster block ran (frames without tiles): reset the length here
    if (careful_reset) return launch_clear(frames, f_lo, f_hi, s);
    return SLOTH_OK;
}

int TiledPlan::launch(const std::vector<Frame>& frames, const std::vector<Draw>& draws, const Frame* d_frames,
                      const Draw* d_draws, const Ramp* d_ramps, uint32_t f_lo, uint32_t f_hi, uint32_t* cells,
                      float* z, int32_t* win, bool win_raw, uint32_t n_ramps, hipStream_t s,
                      std::vector<hipEvent_t>* ev, hipStream_t s2, const std::vector<hipEvent_t>* join)
{
    if (!sized) return fail(SLOTH_ESTATE, "tile lists not sized");
    if (!tile_fits())
        return fail(SLOTH_EINVAL,
                    "tile %ux%u does not fit a workgroup: %zu bytes of LDS (at most 65536), %u slots a row "
                    "(at most %d)", TW, TH, raster_lds(has_odd), has_odd ? TW : TW / 2, RASTER_BLOCK);
    int rc;
    const uint32_t nf = f_hi - f_lo;
    uint32_t groups = 1;
    if (s2 && join && !ev) groups = std::min<uint32_t>({MAX_GROUPS, nf / PIPE_MIN_FRAMES, (uint32_t)join->size() - 1});
    if (groups < 2) {
        // Every render leaves the tile cursors, the even frames' newline
        // rows and the careful list's length as k_tile_clear sets them
        // (k_tile_raster resets what it has read; size() ends with a
        // clear), so only ranges with odd widths, whose newline rows
        // k_tile_edges reads after the raster, are cleared first.
        bool odd = false;
        for (uint32_t f = f_lo; f < f_hi; ++f) odd = odd || (frames[f].W & 1u);
        if (ev) HIP_TRY(hipEventRecord((*ev)[0], s));
        if (odd && (rc = launch_clear(frames, f_lo, f_hi, s))) return rc;
        if (ev) HIP_TRY(hipEventRecord((*ev)[1], s));
        if ((rc = launch_bin(frames, draws, d_frames, d_draws, f_lo, f_hi, false, s))) return rc;
        if (ev) HIP_TRY(hipEventRecord((*ev)[2], s));
        if ((rc = launch_careful(d_frames, d_draws, false, 0, s))) return rc;
        if (ev) HIP_TRY(hipEventRecord((*ev)[3], s));
        if ((rc = launch_raster(frames, d_frames, d_draws, d_ramps, f_lo, f_hi, cells, z, win, win_raw, n_ramps, true,
                                s)))
            return rc;
        if (ev) HIP_TRY(hipEventRecord((*ev)[4], s));
        return SLOTH_OK;
    }
    // Pipeline over frame groups: s bins group g + 1 while s2 rasters group
    // g.  Group g's raster waits for its bins and careful boxes (event g); s
    // joins s2 at the end.
    if ((rc = launch_clear(frames, f_lo, f_hi, s))) return rc;
    for (uint32_t g = 0; g < groups; ++g) {
        const uint32_t a = f_lo + (uint32_t)((uint64_t)nf * g / groups);
        const uint32_t b = f_lo + (uint32_t)((uint64_t)nf * (g + 1) / groups);
        if ((rc = launch_bin(frames, draws, d_frames, d_draws, a, b, false, s))) return rc;
        if ((rc = launch_careful(d_frames, d_draws, false, g, s))) return rc;
        HIP_TRY(hipEventRecord((*join)[g], s));
        HIP_TRY(hipStreamWaitEvent(s2, (*join)[g], 0));
        // the last group's raster runs after every group's careful pass
        if ((rc = launch_raster(frames, d_frames, d_draws, d_ramps, a, b, cells, z, win, win_raw, n_ramps,
                                g + 1 == groups, s2)))
            return rc;
    }
    HIP_TRY(hipEventRecord((*join)[groups], s2));
    HIP_TRY(hipStreamWaitEvent(s, (*join)[groups], 0));
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
