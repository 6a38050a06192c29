// Output encoders of Context::flush (context.rs:50-92) on the device.
//
//   plain  (context.rs:58-61): every cell's char, then println!'s '\n'
//   ansi   (context.rs:62-78): per cell crossterm 0.18.0's PrintStyledContent
//          of the char with RGB foreground on RGB (25,25,25): CSI
//          "48;2;25;25;25m", CSI "38;2;R;G;Bm", the char, CSI "0m"
//   html   (context.rs:79-87): per cell `<span style="color:rgb(R,G,B)">` and
//          the char (the `-j` export format)
//
// A cell's glyph byte is a Unicode code point below 256 (`char` in the
// reference), written as UTF-8: one byte below 0x80, two bytes otherwise.
// Encoding is three passes over the cells: byte length per cell, an
// exclusive scan of the lengths, then each cell writes its bytes at its
// offset.
#include "sloth_encode.h"

namespace sloth {

constexpr int ENC_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;   // cells per thread in the block-level scan

__device__ __forceinline__ uint32_t utf8_len(uint32_t g) { return g < 0x80u ? 1u : 2u; }

__device__ __forceinline__ uint32_t dec_len(uint32_t v) { return v >= 100u ? 3u : (v >= 10u ? 2u : 1u); }

// `<span style="color:rgb(` = 23 bytes, `)">` = 3 bytes, plus 2 commas
constexpr uint32_t HTML_FIXED = 23 + 3 + 2;
// ESC "[48;2;25;25;25m" = 16 bytes, ESC "[38;2;" = 7, two ';' and 'm' = 3,
// ESC "[0m" = 4
constexpr uint32_t ANSI_FIXED = 16 + 7 + 3 + 4;

__device__ __forceinline__ uint32_t cell_len(uint32_t c, int mode)
{
    const uint32_t g = c & 0xFFu;
    if (mode == 0) return utf8_len(g);
    const uint32_t digits = dec_len((c >> 8) & 0xFFu) + dec_len((c >> 16) & 0xFFu) + dec_len(c >> 24);
    return (mode == 1 ? ANSI_FIXED : HTML_FIXED) + digits + utf8_len(g);
}

// Per block: the exclusive scan of its cells' lengths (written to len) and
// the block total (written to block_sum).
__global__ void __launch_bounds__(ENC_BLOCK) k_enc_scan_local(const uint32_t* __restrict__ cells, uint64_t n,
                                                              int mode, uint64_t* __restrict__ off,
                                                              uint64_t* __restrict__ block_sum)
{
    __shared__ uint64_t s[ENC_BLOCK];
    const uint64_t base = ((uint64_t)blockIdx.x * ENC_BLOCK + threadIdx.x) * SCAN_ITEMS;
    uint64_t l[SCAN_ITEMS];
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        l[i] = (base + i < n) ? cell_len(cells[base + i], mode) : 0;
        t += l[i];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int d = 1; d < ENC_BLOCK; d <<= 1) {   // inclusive Hillis-Steele scan
        uint64_t v = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = s[threadIdx.x] - t;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) off[base + i] = run;
        run += l[i];
    }
    if (threadIdx.x == ENC_BLOCK - 1) block_sum[blockIdx.x] = s[threadIdx.x];
}

// Exclusive scan of the block totals (one block, looping over chunks).
__global__ void __launch_bounds__(ENC_BLOCK) k_enc_scan_blocks(uint64_t* __restrict__ block_sum, uint64_t nb,
                                                               uint64_t* __restrict__ total)
{
    __shared__ uint64_t s[ENC_BLOCK];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t c0 = 0; c0 < nb; c0 += ENC_BLOCK) {
        const uint64_t i = c0 + threadIdx.x;
        const uint64_t v = i < nb ? block_sum[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < ENC_BLOCK; d <<= 1) {
            uint64_t w = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
            __syncthreads();
            s[threadIdx.x] += w;
            __syncthreads();
        }
        if (i < nb) block_sum[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == ENC_BLOCK - 1) carry += s[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__device__ __forceinline__ uint32_t put_dec(uint8_t* p, uint32_t v)
{
    if (v >= 100u) {
        p[0] = (uint8_t)('0' + v / 100u);
        p[1] = (uint8_t)('0' + (v / 10u) % 10u);
        p[2] = (uint8_t)('0' + v % 10u);
        return 3;
    }
    if (v >= 10u) {
        p[0] = (uint8_t)('0' + v / 10u);
        p[1] = (uint8_t)('0' + v % 10u);
        return 2;
    }
    p[0] = (uint8_t)('0' + v);
    return 1;
}

__device__ __forceinline__ uint32_t put_utf8(uint8_t* p, uint32_t g)
{
    if (g < 0x80u) {
        p[0] = (uint8_t)g;
        return 1;
    }
    p[0] = (uint8_t)(0xC0u | (g >> 6));
    p[1] = (uint8_t)(0x80u | (g & 0x3Fu));
    return 2;
}

__global__ void __launch_bounds__(ENC_BLOCK) k_enc_write(const uint32_t* __restrict__ cells, uint64_t n, int mode,
                                                         const uint64_t* __restrict__ off,
                                                         const uint64_t* __restrict__ block_off,
                                                         uint8_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * ENC_BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = cells[i];
    uint8_t* p = out + block_off[i / ((uint64_t)ENC_BLOCK * SCAN_ITEMS)] + off[i];
    if (mode == 0) {
        put_utf8(p, c & 0xFFu);
        return;
    }
    if (mode == 1) {
        const char* bg = "\x1b[48;2;25;25;25m\x1b[38;2;";
        uint32_t k = 0;
        for (; k < 23; ++k) p[k] = (uint8_t)bg[k];
        k += put_dec(p + k, (c >> 8) & 0xFFu);
        p[k++] = ';';
        k += put_dec(p + k, (c >> 16) & 0xFFu);
        p[k++] = ';';
        k += put_dec(p + k, c >> 24);
        p[k++] = 'm';
        k += put_utf8(p + k, c & 0xFFu);
        p[k++] = 0x1b;
        p[k++] = '[';
        p[k++] = '0';
        p[k++] = 'm';
        return;
    }
    const char* head = "<span style=\"color:rgb(";
    uint32_t k = 0;
    for (; k < 23; ++k) p[k] = (uint8_t)head[k];
    k += put_dec(p + k, (c >> 8) & 0xFFu);
    p[k++] = ',';
    k += put_dec(p + k, (c >> 16) & 0xFFu);
    p[k++] = ',';
    k += put_dec(p + k, c >> 24);
    p[k++] = ')';
    p[k++] = '"';
    p[k++] = '>';
    put_utf8(p + k, c & 0xFFu);
}

uint64_t encode_blocks(uint64_t n_cells) { return (n_cells + (uint64_t)ENC_BLOCK * SCAN_ITEMS - 1) / ((uint64_t)ENC_BLOCK * SCAN_ITEMS); }

int launch_encode_size(const uint32_t* cells, uint64_t n, int mode, uint64_t* off, uint64_t* block_sum,
                       uint64_t* total, hipStream_t s)
{
    const uint64_t nb = encode_blocks(n);
    if (nb == 0) return SLOTH_OK;
    hipLaunchKernelGGL(k_enc_scan_local, dim3((unsigned)nb), dim3(ENC_BLOCK), 0, s, cells, n, mode, off, block_sum);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_enc_scan_blocks, dim3(1), dim3(ENC_BLOCK), 0, s, block_sum, nb, total);
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

int launch_encode_write(const uint32_t* cells, uint64_t n, int mode, const uint64_t* off, const uint64_t* block_sum,
                        uint8_t* out, hipStream_t s)
{
    if (n == 0) return SLOTH_OK;
    hipLaunchKernelGGL(k_enc_write, dim3((unsigned)((n + ENC_BLOCK - 1) / ENC_BLOCK)), dim3(ENC_BLOCK), 0, s, cells,
                       n, mode, off, block_sum, out);
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

}  // namespace sloth
