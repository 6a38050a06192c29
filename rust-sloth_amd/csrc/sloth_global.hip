// General raster path: any width (odd widths included), multi-draw frames.
//
//   k_clear_global    z-slot keys <- empty, newline table <- -1
//   k_raster_global   one thread per triangle: setup, classify, walk the
//                     candidate pixels, 64-bit atomicMin of the packed key
//                     into the frame's z-slot keys; newline rows via atomicMax
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

// Index of the draw that owns launch-wide triangle `item`.
__device__ __forceinline__ uint32_t find_draw(const Draw* __restrict__ draws, uint32_t d_lo, uint32_t d_hi,
                                              uint64_t item)
{
    uint32_t lo = d_lo, hi = d_hi - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (draws[mid].item_begin <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
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

__global__ void __launch_bounds__(256) k_raster_global(const Draw* __restrict__ draws, uint32_t d_lo, uint32_t d_hi,
                                                       const Frame* __restrict__ frames, uint64_t item_base,
                                                       uint64_t n_items, uint64_t* __restrict__ keys,
                                                       int32_t* __restrict__ nl)
{
    uint64_t t_id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t_id >= n_items) return;
    uint64_t item = item_base + t_id;
    uint32_t d = find_draw(draws, d_lo, d_hi, item);
    const Draw& dr = draws[d];
    uint32_t local = (uint32_t)(item - dr.item_begin);
    const Frame& fr = frames[dr.frame];
    const uint32_t W = fr.W, H = fr.H;
    Tri t;
    setup_tri(dr.M, dr.xyz + (size_t)local * 9, W, H, t);
    const uint32_t tri = dr.tri_begin + local;
    uint64_t* fk = keys + fr.z_off;

    if (fr.image && t.y0 < t.y1d) {
        // rasterizer.rs:89-91: every row of the domain ends with a newline cell.
        int32_t* fnl = nl + fr.nl_off;
        for (int y = t.y0; y < t.y1d; ++y) atomicMax(&fnl[y], (int32_t)tri);
    }

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
    const Draw& dr = draw_of(draws, fr, tri);
    uint32_t local = tri - dr.tri_begin;
    Tri t;
    setup_tri(dr.M, dr.xyz + (size_t)local * 9, fr.W, fr.H, t);
    float w0, w1, w2;
    edges(t, (float)px, (float)py, w0, w1, w2);
    z = depth(t, w1, w2);
    float s = shade_value(unit_normal_z(t), t.a, w0, w1, w2);
    const auto* c = global_ptr(dr.rgb + (size_t)local * 3);
    cell = (uint32_t)ramp_glyph(ramps[dr.ramp], s) | ((uint32_t)c[0] << 8) | ((uint32_t)c[1] << 16) |
           ((uint32_t)c[2] << 24);
}

__global__ void __launch_bounds__(256) k_resolve_global(const Draw* __restrict__ draws, const Ramp* __restrict__ ramps,
                                                        const Frame* __restrict__ frames, uint32_t f_lo,
                                                        uint32_t f_hi, uint64_t cell_base, uint64_t n_cells,
                                                        const uint64_t* __restrict__ keys,
                                                        const int32_t* __restrict__ nl, uint32_t* __restrict__ cells,
                                                        float* __restrict__ zbuf, int32_t* __restrict__ win,
                                                        uint32_t win_raw)
{
    uint64_t t_id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t_id >= n_cells) return;
    uint64_t g = cell_base + t_id;
    uint32_t lo = f_lo, hi = f_hi - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (frames[mid].cell_begin <= g) lo = mid; else hi = mid - 1;
    }
    const Frame& fr = frames[lo];
    const uint32_t W = fr.W, H = fr.H;
    const uint64_t WH = (uint64_t)W * H;
    const uint64_t c = g - fr.cell_begin;
    if (c >= frame_cells(fr)) return;   // padding between frames
    const uint64_t* fk = keys + fr.z_off;

    uint64_t best_time = 0;   // 0 = no event; event times are stored +1
    uint32_t best_cell = CELL_BLANK;

    if (c < WH) {
        uint64_t k = fk[c];
        float z = F32_MAX;
        int32_t wt = -1;
        if (k != KEY_EMPTY) {
            uint32_t cell, tri, px, py;
            winner_fragment(draws, ramps, fr, (uint32_t)c, k, cell, z, tri, px, py);
            uint64_t tm = (((uint64_t)tri << 32) | ((uint64_t)py << 16) | px) + 1;
            best_time = tm;
            best_cell = cell;
            wt = win_raw ? (int32_t)(uint32_t)k : (int32_t)tri;   // raw: triangle << 1 | sub
        }
        // streaming stores: outputs are written once
        __builtin_nontemporal_store(z, zbuf + fr.z_off + c);
        if (win) __builtin_nontemporal_store(wt, win + fr.z_off + c);
    }
    if (c >= 1 && c - 1 < WH) {
        uint64_t k = fk[c - 1];
        if (k != KEY_EMPTY) {
            uint32_t cell, tri, px, py;
            float z;
            winner_fragment(draws, ramps, fr, (uint32_t)(c - 1), k, cell, z, tri, px, py);
            uint64_t tm = (((uint64_t)tri << 32) | ((uint64_t)py << 16) | px) + 1;
            if (tm > best_time) {
                best_time = tm;
                best_cell = cell;
            }
        }
    }
    if (fr.image && c >= 1 && (c - 1) % W == 0) {
        uint64_t r = (c - 1) / W;
        if (r < H) {
            int32_t n = nl[fr.nl_off + r];
            if (n >= 0) {
                uint64_t tm = (((uint64_t)(uint32_t)n << 32) | (r << 16) | 0xFFFFu) + 1;
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
