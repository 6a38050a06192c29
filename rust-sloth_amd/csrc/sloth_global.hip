// General raster path: any width (odd widths included), multi-draw frames.
//
//   k_clear_global    z-slot keys <- empty, newline table <- -1
//   k_raster_global   one thread per triangle: setup, classify, walk the
//                     candidate pixels, 64-bit atomicMin of the packed key
//                     into the frame's z-slot keys; newline rows by atomicMax,
//                     reduced per block in LDS first
//   k_winner_global   one thread per z slot: the winning fragment's cell and
//                     depth, derived once from the slot's key; the depth and
//                     winner go out, and the key's depth half is replaced by
//                     the cell (the triangle and sub half stays)
//   k_resolve_global  one thread per cell: last writer among the winners of
//                     z slots c and c-1 and the row's newline
//
// The order-independent packed key makes the result deterministic and equal
// to the sequential loop of rasterizer.rs:43-45, 69-92.
#include "sloth_kernels.h"

namespace sloth {

__global__ void k_clear_global(uint64_t* __restrict__ keys, uint64_t n_keys, int32_t* __restrict__ nl,
                               uint64_t n_rows)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_keys; i += stride)
        keys[i] = KEY_EMPTY;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rows; i += stride)
        nl[i] = -1;
}

__device__ __forceinline__ void plot(uint64_t* __restrict__ keys, const Tri& t, uint32_t W, uint32_t tri, int x,
                                     int y)
{
    float w0, w1, w2;
    edges(t, (float)x, (float)y, w0, w1, w2);
    if (!inside(w0, w1, w2)) return;
    float z = depth(t, w1, w2);
    if (!(z < F32_MAX)) return;   // the initial z-buffer value never loses; NaN never wins
    uint32_t id = (uint32_t)y * W + 2u * (uint32_t)x;
    uint32_t sub = (2u * (uint32_t)x < W) ? 1u : 0u;
    atomicMin((unsigned long long*)&keys[id], (unsigned long long)make_key(z, tri, sub));
}

// Grid: x = blocks of one draw's triangles, y = draw index - d_lo, so the
// draw, its frame and its matrix are uniform per block (scalar loads).
__global__ void __launch_bounds__(256) k_raster_global(const Draw* __restrict__ draws, uint32_t d_lo,
                                                       const Frame* __restrict__ frames, uint64_t* __restrict__ keys,
                                                       int32_t* __restrict__ nl)
{
    __shared__ int32_t s_nl[256];
    __shared__ int s_lo, s_hi;
    const Draw& dr = draws[d_lo + blockIdx.y];
    const uint32_t base = blockIdx.x * blockDim.x;
    if (base >= dr.n) return;   // the whole block: past this draw's triangles
    const bool valid = base + threadIdx.x < dr.n;
    // threads past the end take the draw's last triangle and do nothing with
    // it, so that every thread reaches the block's barriers
    const uint32_t local = valid ? base + threadIdx.x : dr.n - 1;
    const Frame& fr = frames[dr.frame];
    const uint32_t W = fr.W, H = fr.H;
    Tri t;
    setup_draw_tri(dr, local, W, H, t);
    const uint32_t tri = dr.tri_begin + local;
    uint64_t* fk = keys + fr.z_off;

    // rasterizer.rs:89-91: every row of the domain ends with a newline cell;
    // the last triangle (in draw order) whose rows include r decides row r.
    // The block reduces its rows in LDS, 256 rows at a time, with one global
    // atomicMax per row and window.
    const bool has_rows = valid && fr.image && t.y0 < t.y1d;
    int32_t* fnl = nl + fr.nl_off;
    s_nl[threadIdx.x] = -1;
    if (threadIdx.x == 0) {
        s_lo = 1 << 30;
        s_hi = -1;
    }
    __syncthreads();
    if (has_rows) {
        atomicMin(&s_lo, t.y0);
        atomicMax(&s_hi, t.y1d - 1);
    }
    __syncthreads();
    const int lo = s_lo, hi = s_hi;   // uniform per block
    for (int ws = lo; ws <= hi; ws += (int)blockDim.x) {
        if (ws > lo) {
            __syncthreads();   // the previous window's reads are done
            s_nl[threadIdx.x] = -1;
            __syncthreads();
        }
        if (has_rows)
            for (int y = max(t.y0, ws); y < min(t.y1d, ws + (int)blockDim.x); ++y)
                atomicMax(&s_nl[y - ws], (int32_t)tri);
        __syncthreads();
        const int32_t v = s_nl[threadIdx.x];
        if (ws + (int)threadIdx.x <= hi && v >= 0) atomicMax(&fnl[ws + (int)threadIdx.x], v);
    }
    if (!valid) return;

    uint32_t cls = setup_class(t);
    if (cls == CLS_EMPTY) return;
    if (cls == CLS_BOX || cls == CLS_TIGHT) {
        int xa, xb;
        box_x(t, cls == CLS_TIGHT, xa, xb);
        for (int y = t.y0; y < t.y1d; ++y)
            for (int x = xa; x <= xb; ++x) plot(fk, t, W, tri, x, y);
    } else if (cls == CLS_SPANS) {
        for (int y = t.y0; y < t.y1d; ++y) {
            int lo, hi;
            if (!row_span(t, y, lo, hi)) continue;
            for (int x = lo; x <= hi; ++x) plot(fk, t, W, tri, x, y);
        }
    } else {
        for (int y = t.y0; y < t.y1d; ++y)
            for (int x = t.x0; x < t.x1d; ++x) plot(fk, t, W, tri, x, y);
    }
}

// Draw in frame `fr` that owns frame-global triangle `tri`.
__device__ __forceinline__ const Draw& draw_of(const Draw* __restrict__ draws, const Frame& fr, uint32_t tri)
{
    uint32_t lo = fr.draw_begin, hi = fr.draw_end - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (draws[mid].tri_begin <= tri) lo = mid; else hi = mid - 1;
    }
    return draws[lo];
}

// Recomputes the winning fragment's cell and depth from its key.
__device__ __forceinline__ void winner_fragment(const Draw* __restrict__ draws, const Ramp* __restrict__ ramps,
                                                const Frame& fr, uint32_t id, uint64_t key, uint32_t& cell,
                                                float& z, uint32_t& tri, uint32_t& px, uint32_t& py)
{
    uint32_t lowbits = (uint32_t)key;
    tri = lowbits >> 1;
    key_pixel(id, lowbits & 1u, fr.W, px, py);
    auto derive = [&](const Draw& dr) {
        uint32_t local = tri - dr.tri_begin;
        Tri t;
        setup_draw_tri(dr, local, fr.W, fr.H, t);
        float w0, w1, w2;
        edges(t, (float)px, (float)py, w0, w1, w2);
        z = depth(t, w1, w2);
        float s = shade_value(unit_normal_z(t), t.a, w0, w1, w2);
        const auto* c = global_ptr(dr.rgb + (size_t)local * 3);
        cell = (uint32_t)ramp_glyph(ramps[dr.ramp], s) | ((uint32_t)c[0] << 8) | ((uint32_t)c[1] << 16) |
               ((uint32_t)c[2] << 24);
    };
    // a frame of one draw (the usual case) reads it at a uniform address
    // (scalar loads of the matrix and the shader table): two copies
    if (fr.draw_end - fr.draw_begin == 1)
        derive(draws[fr.draw_begin]);
    else
        derive(draw_of(draws, fr, tri));
}

// Keys of slots that k_winner_global has resolved: the winning cell in the
// high word, the key's (triangle << 1 | sub) in the low word.  Frames hold
// fewer than 2^31 - 1 triangles, so the low word is never all ones and a
// resolved slot never reads as KEY_EMPTY.
//
// Both kernels: grid (cells of the largest frame / 256, frame - f_lo).
__global__ void __launch_bounds__(256) k_winner_global(const Draw* __restrict__ draws, const Ramp* __restrict__ ramps,
                                                       const Frame* __restrict__ frames, uint32_t f_lo,
                                                       uint64_t* __restrict__ keys, float* __restrict__ zbuf,
                                                       int32_t* __restrict__ win, uint32_t win_raw)
{
    const Frame& fr = frames[f_lo + blockIdx.y];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;   // W, H <= 65535: cell ids fit 32 bits
    if (c >= fr.W * fr.H) return;   // the frame's z slots are its first W*H cell ids
    uint64_t* fk = keys + fr.z_off;
    const uint64_t k = fk[c];
    float z = F32_MAX;
    int32_t wt = -1;
    if (k != KEY_EMPTY) {
        uint32_t cell, tri, px, py;
        winner_fragment(draws, ramps, fr, (uint32_t)c, k, cell, z, tri, px, py);
        wt = win_raw ? (int32_t)(uint32_t)k : (int32_t)tri;   // raw: triangle << 1 | sub
        fk[c] = ((uint64_t)cell << 32) | (uint32_t)k;
    }
    // streaming stores: outputs are written once
    __builtin_nontemporal_store(z, zbuf + fr.z_off + c);
    if (win) __builtin_nontemporal_store(wt, win + fr.z_off + c);
}

// Write time of a resolved slot's winner in the sequential loop: triangle,
// then the pixel's row and column (plus one; 0 = no write).  (row, col) is
// the slot's cell; the pixel is as in key_pixel.
__device__ __forceinline__ uint64_t slot_time(uint64_t k, uint32_t row, uint32_t col, uint32_t W)
{
    const uint32_t lowbits = (uint32_t)k;
    const uint32_t px = (lowbits & 1u) ? col >> 1 : (col + W) >> 1;
    const uint32_t py = (lowbits & 1u) ? row : row - 1;
    return (((uint64_t)(lowbits >> 1) << 32) | ((uint64_t)py << 16) | px) + 1;
}

__global__ void __launch_bounds__(256) k_resolve_global(const Frame* __restrict__ frames, uint32_t f_lo,
                                                        const uint64_t* __restrict__ keys,
                                                        const int32_t* __restrict__ nl, uint32_t* __restrict__ cells)
{
    const Frame& fr = frames[f_lo + blockIdx.y];
    const uint32_t W = fr.W, H = fr.H;
    const uint32_t WH = W * H;   // W, H <= 65535: cell ids fit 32 bits
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= frame_cells(fr)) return;
    const uint64_t* fk = keys + fr.z_off;
    // cell c = (row, col); its left neighbour is (row, col - 1), or the last
    // cell of the row above when col = 0 (cells past W*H are the newline rows'
    // tail, whose row may be H)
    const uint32_t row = c / W, col = c - row * W;
    const uint32_t lrow = col ? row : row - 1, lcol = col ? col - 1 : W - 1;

    uint64_t best_time = 0;   // 0 = no event; event times are stored +1
    uint32_t best_cell = CELL_BLANK;

    if (c < WH) {
        const uint64_t k = fk[c];
        if (k != KEY_EMPTY) {
            best_time = slot_time(k, row, col, W);
            best_cell = (uint32_t)(k >> 32);
        }
    }
    if (c >= 1 && c - 1 < WH) {
        const uint64_t k = fk[c - 1];
        if (k != KEY_EMPTY) {
            const uint64_t tm = slot_time(k, lrow, lcol, W);
            if (tm > best_time) {
                best_time = tm;
                best_cell = (uint32_t)(k >> 32);
            }
        }
    }
    if (fr.image && c >= 1 && lcol == 0) {
        const uint32_t r = lrow;
        if (r < H) {
            int32_t n = nl[fr.nl_off + r];
            if (n >= 0) {
                uint64_t tm = (((uint64_t)(uint32_t)n << 32) | ((uint64_t)r << 16) | 0xFFFFu) + 1;
                if (tm > best_time) {
                    best_time = tm;
                    best_cell = CELL_NEWLINE;
                }
            }
        }
    }
    __builtin_nontemporal_store(best_cell, cells + fr.cell_off + c);
}

}  // namespace sloth
