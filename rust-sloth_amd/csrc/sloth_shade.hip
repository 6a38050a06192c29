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
// Per-(triangle, corner) sums are added in double with hardware atomics; a
// gather kernel then sums each vertex's corners in a fixed (CSR) order.
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

template <bool BWD>
__global__ void __launch_bounds__(SHADE_BLOCK) k_shade(ShadeArgs p)
{
    __shared__ double s_red[4][SHADE_BLOCK / 64];
    const uint64_t g = (uint64_t)blockIdx.x * SHADE_BLOCK + threadIdx.x;
    double loss = 0.0, dl0 = 0.0, dl1 = 0.0, dl2 = 0.0;
    if (g < p.n_slots) {
        uint32_t lo = 0, hi = p.n_frames - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (p.frames[mid].z_off <= g) lo = mid; else hi = mid - 1;
        }
        const Frame& fr = p.frames[lo];
        const int32_t v = p.vis[g];
        if (v >= 0) {
            const uint32_t T = (uint32_t)v >> 1, sub = (uint32_t)v & 1u;
            uint32_t px, py;
            key_pixel((uint32_t)(g - fr.z_off), sub, fr.W, px, py);
            const uint32_t d = (uint32_t)(&shade_draw(p.draws, fr, T) - p.draws);
            const Draw& dr = p.draws[d];
            const uint32_t li = T - dr.tri_begin;
            Tri t;
            setup_tri(dr.M, dr.xyz + (size_t)li * 9, fr.W, fr.H, t);
            float w0, w1, w2;
            edges(t, (float)px, (float)py, w0, w1, w2);
            const float lam[3] = {t.a * w0, t.a * w1, t.a * w2};
            const float* G = p.G + (size_t)d * 9;
            const float L0 = p.light[0], L1 = p.light[1], L2 = p.light[2];
            float n0 = 0.f, n1 = 0.f, n2 = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
            uint32_t vid[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vid[k] = p.faces[(size_t)li * 3 + k];
                const float* nk = p.vnormal + (size_t)vid[k] * 3;
                const float* ck = p.vcolor + (size_t)vid[k] * 3;
                const float a0 = nk[0], a1 = nk[1], a2 = nk[2];
                n0 += lam[k] * (G[0] * a0 + G[1] * a1 + G[2] * a2);
                n1 += lam[k] * (G[3] * a0 + G[4] * a1 + G[5] * a2);
                n2 += lam[k] * (G[6] * a0 + G[7] * a1 + G[8] * a2);
                c0 += lam[k] * ck[0];
                c1 += lam[k] * ck[1];
                c2 += lam[k] * ck[2];
            }
            const float I = L0 * n0 + L1 * n1 + L2 * n2;
            const float o0 = I * c0, o1 = I * c1, o2 = I * c2;
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
                if (BWD) {
                    const float g0 = 2.f * r0, g1 = 2.f * r1, g2 = 2.f * r2;
                    const float dI = g0 * c0 + g1 * c1 + g2 * c2;
                    // G^T L (model-space direction of the light for this view)
                    const float gl0 = G[0] * L0 + G[3] * L1 + G[6] * L2;
                    const float gl1 = G[1] * L0 + G[4] * L1 + G[7] * L2;
                    const float gl2 = G[2] * L0 + G[5] * L1 + G[8] * L2;
                    double* acc = p.tri_acc + (size_t)li * 18;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float lk = lam[k];
                        unsafeAtomicAdd(acc + k * 6 + 0, (double)(lk * I * g0));
                        unsafeAtomicAdd(acc + k * 6 + 1, (double)(lk * I * g1));
                        unsafeAtomicAdd(acc + k * 6 + 2, (double)(lk * I * g2));
                        unsafeAtomicAdd(acc + k * 6 + 3, (double)(dI * lk * gl0));
                        unsafeAtomicAdd(acc + k * 6 + 4, (double)(dI * lk * gl1));
                        unsafeAtomicAdd(acc + k * 6 + 5, (double)(dI * lk * gl2));
                    }
                    dl0 = (double)(dI * n0);
                    dl1 = (double)(dI * n1);
                    dl2 = (double)(dI * n2);
                }
            }
        }
    }
    // block sums of the loss and the light gradient (fixed order)
    for (int sh = 32; sh >= 1; sh >>= 1) {
        loss += __shfl_xor(loss, sh);
        dl0 += __shfl_xor(dl0, sh);
        dl1 += __shfl_xor(dl1, sh);
        dl2 += __shfl_xor(dl2, sh);
    }
    const int wv = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        s_red[0][wv] = loss;
        s_red[1][wv] = dl0;
        s_red[2][wv] = dl1;
        s_red[3][wv] = dl2;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int w = 0; w < SHADE_BLOCK / 64; ++w) s += s_red[threadIdx.x][w];
        p.partial[(size_t)blockIdx.x * 4 + threadIdx.x] = s;
    }
}

__global__ void k_fill_f64(double* __restrict__ a, uint64_t n)
{
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = 0.0;
}

// Per-vertex sums of the (triangle, corner) accumulators, in CSR order.
__global__ void k_shade_gather(const double* __restrict__ tri_acc, const uint32_t* __restrict__ adj_off,
                               const uint32_t* __restrict__ adj, uint32_t n_verts, float* __restrict__ d_color,
                               float* __restrict__ d_normal)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_verts) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t e = adj_off[v]; e < adj_off[v + 1]; ++e) {
        const double* a = tri_acc + (size_t)adj[e] * 6;
#pragma unroll
        for (int j = 0; j < 6; ++j) s[j] += a[j];
    }
    d_color[v * 3 + 0] = (float)s[0];
    d_color[v * 3 + 1] = (float)s[1];
    d_color[v * 3 + 2] = (float)s[2];
    d_normal[v * 3 + 0] = (float)s[3];
    d_normal[v * 3 + 1] = (float)s[4];
    d_normal[v * 3 + 2] = (float)s[5];
}

uint32_t shade_blocks(uint64_t n_slots) { return (uint32_t)((n_slots + SHADE_BLOCK - 1) / SHADE_BLOCK); }

int launch_shade(const ShadeArgs& a, bool backward, hipStream_t s)
{
    const uint32_t nb = shade_blocks(a.n_slots);
    if (nb == 0) return SLOTH_OK;
    if (backward) {
        const uint64_t n_acc = (uint64_t)a.n_tris * 18;
        hipLaunchKernelGGL(k_fill_f64, dim3(std::max(1u, std::min(4096u, (uint32_t)((n_acc + 255) / 256)))),
                           dim3(256), 0, s, a.tri_acc, n_acc);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_shade<true>, dim3(nb), dim3(SHADE_BLOCK), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_shade<false>, dim3(nb), dim3(SHADE_BLOCK), 0, s, a);
    }
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

int launch_shade_gather(const ShadeArgs& a, const uint32_t* adj_off, const uint32_t* adj, float* d_color,
                        float* d_normal, hipStream_t s)
{
    if (a.n_verts == 0) return SLOTH_OK;
    hipLaunchKernelGGL(k_shade_gather, dim3((a.n_verts + 255) / 256), dim3(256), 0, s, a.tri_acc, adj_off, adj,
                       a.n_verts, d_color, d_normal);
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

}  // namespace sloth
