// Host side of libsloth_hip.so: the C-ABI declared in include/sloth_hip.h.
//
// A "job" is a set of frames (Contexts) and their draw_mesh calls, laid out in
// device tables (Frame, Draw, Ramp) plus the output buffers of all frames.
// A context is a job with one frame; a batch is a job with many, whose launch
// sequence is captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <unordered_map>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "sloth_hip.h"
#include "sloth_host.h"
#include "sloth_encode.h"
#include "sloth_kernels.h"
#include "sloth_models.h"
#include "sloth_shade.h"
#include "sloth_tiled.h"

using namespace sloth;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
static int g_device = 0;

int sloth::fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// ---------------------------------------------------------------------------
// host arithmetic in the reference's order (compiled with -ffp-contract=off)
// ---------------------------------------------------------------------------

// Matrix4 * Matrix4 as nalgebra's gemm computes it (rasterizer.rs:57): each
// result column is a gemv, i.e. an axpy over the columns of the left operand.
static void mat4_mul(const float* A, const float* B, float* out)
{
    float r[16];
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            float acc = A[0 * 4 + i] * B[j * 4 + 0];
            acc = A[1 * 4 + i] * B[j * 4 + 1] + acc;
            acc = A[2 * 4 + i] * B[j * 4 + 2] + acc;
            acc = A[3 * 4 + i] * B[j * 4 + 3] + acc;
            r[j * 4 + i] = acc;
        }
    memcpy(out, r, sizeof r);
}

// Context::update (context.rs:104-133).
static void screen_transform(uint32_t W, uint32_t H, const float* bbox_max, uint32_t n, float* U)
{
    float scale = 0.0f;
    for (uint32_t m = 0; m < n; ++m) {
        scale = fmaxf(scale, bbox_max[3 * m + 0]);
        scale = fmaxf(scale, bbox_max[3 * m + 1]);
        scale = fmaxf(scale, bbox_max[3 * m + 2]);
    }
    float fw = (float)(uint16_t)W, fh = (float)(uint16_t)H;
    scale = fminf(fh, fw / 2.0f) / scale / 2.0f;
    float M[16] = {scale, 0.0f, 0.0f, 0.0f, 0.0f, -scale, 0.0f, 0.0f,
                   0.0f,  0.0f, scale, 0.0f, fw / 4.0f, fh / 2.0f, 0.0f, 1.0f};
    memcpy(U, M, sizeof M);
}

static const float IDENTITY[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

// Marks tables whose thresholds are non-decreasing with no NaN (the device
// then counts thresholds below the shade instead of scanning).
static void finish_ramp(Ramp& r)
{
    bool ok = true;
    for (uint32_t i = 0; i < r.n; ++i) {
        if (r.thr[i] != r.thr[i]) ok = false;
        if (i + 1 < r.n && r.thr[i] > r.thr[i + 1]) ok = false;
    }
    r.sorted = ok ? 1 : 0;
    // unused steps: +inf, which no shade exceeds (the device's binary search
    // over all MAX_RAMP thresholds relies on it)
    for (uint32_t i = r.n; i < MAX_RAMP; ++i) r.thr[i] = INFINITY;
}

static Ramp default_ramp()
{
    static const float t[9] = {0.20f, 0.30f, 0.40f, 0.50f, 0.60f, 0.70f, 0.80f, 0.90f, 1.0f};
    static const char g[9] = {'.', ':', '-', '=', '+', '*', '#', '%', '@'};
    Ramp r;
    memset(&r, 0, sizeof r);
    r.n = 9;
    for (int i = 0; i < 9; ++i) {
        r.thr[i] = t[i];
        r.glyph[i] = (uint8_t)g[i];
    }
    r.fallback = ' ';
    finish_ramp(r);
    return r;
}

static int to_ramp(const sloth_ramp* s, Ramp& r)
{
    if (!s) {
        r = default_ramp();
        return SLOTH_OK;
    }
    ARG_CHECK(s->n <= SLOTH_MAX_RAMP, "shader table has more than 16 steps");
    memset(&r, 0, sizeof r);
    r.n = s->n;
    for (uint32_t i = 0; i < s->n; ++i) {
        r.thr[i] = s->thresholds[i];
        r.glyph[i] = s->glyphs[i];
    }
    r.fallback = s->fallback;
    finish_ramp(r);
    return SLOTH_OK;
}

static bool same_ramp(const Ramp& a, const Ramp& b)
{
    if (a.n != b.n || a.fallback != b.fallback) return false;
    for (uint32_t i = 0; i < a.n; ++i)
        if (memcmp(&a.thr[i], &b.thr[i], 4) != 0 || a.glyph[i] != b.glyph[i]) return false;
    return true;
}

// ---------------------------------------------------------------------------
// meshes
// ---------------------------------------------------------------------------
struct sloth_mesh {
    const float* d_xyz = nullptr;       // soup [n][3][3], or vertex rows [n_verts][3] when d_faces is set
    const uint32_t* d_faces = nullptr;  // indexed meshes: [n][3] vertex ids
    const uint8_t* d_rgb = nullptr;
    uint32_t n = 0, n_verts = 0;
    DevBuf own_xyz, own_rgb, own_faces;
    float bbmin[3] = {0, 0, 0}, bbmax[3] = {0, 0, 0};
};

// OBJ-style bounding box (geometry.rs:85-88, 129-135).
static void host_bbox(const float* xyz, uint32_t n, float* mn, float* mx)
{
    for (int c = 0; c < 3; ++c) mn[c] = mx[c] = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const float* t = xyz + (size_t)i * 9;
        for (int c = 0; c < 3; ++c) {
            float lo = fminf(t[c], fminf(t[3 + c], t[6 + c]));
            float hi = fmaxf(t[c], fmaxf(t[3 + c], t[6 + c]));
            mn[c] = fminf(lo, mn[c]);
            mx[c] = fmaxf(hi, mx[c]);
        }
    }
}

// ---------------------------------------------------------------------------
// jobs
// ---------------------------------------------------------------------------
struct DrawIn {
    uint32_t frame;
    const sloth_mesh* mesh;
    float M[16];
    uint32_t ramp;
};

struct FrameIn {
    uint32_t W, H;
    int image;
};

enum PathKind { PATH_GLOBAL = 0, PATH_TILED = 1 };

struct Job {
    std::vector<FrameIn> fin;
    std::vector<DrawIn> din;
    std::vector<Ramp> ramps;

    // laid-out tables
    std::vector<Frame> frames;
    std::vector<Draw> draws;
    uint64_t n_cells = 0, n_z = 0, n_rows = 0, n_items = 0;

    DevBuf d_frames, d_draws, d_ramps, d_cells, d_z, d_keys, d_nl, d_win;
    TiledPlan tiled;
    int path = PATH_GLOBAL;
    bool laid_out = false;

    uint32_t add_ramp(const Ramp& r)
    {
        for (size_t i = 0; i < ramps.size(); ++i)
            if (same_ramp(ramps[i], r)) return (uint32_t)i;
        ramps.push_back(r);
        return (uint32_t)(ramps.size() - 1);
    }

    // Lays out frames and draws (draws grouped by frame, order within a frame
    // kept) and allocates device tables and outputs.
    int layout()
    {
        const uint32_t F = (uint32_t)fin.size();
        std::vector<uint32_t> order(din.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (uint32_t)i;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return din[a].frame < din[b].frame; });
        frames.assign(F, Frame{});
        draws.assign(din.size(), Draw{});
        uint64_t cells = 0, z = 0, rows = 0, items = 0;
        size_t k = 0;
        bool small_frames = true;
        for (uint32_t f = 0; f < F; ++f) {
            Frame& fr = frames[f];
            fr.W = fin[f].W;
            fr.H = fin[f].H;
            fr.image = fin[f].image ? 1u : 0u;
            fr.cell_off = cells;
            fr.cell_begin = cells;
            fr.z_off = z;
            fr.nl_off = rows;
            fr.draw_begin = (uint32_t)k;
            uint32_t tri = 0;
            while (k < order.size() && din[order[k]].frame == f) {
                const DrawIn& di = din[order[k]];
                Draw& d = draws[k];
                d.xyz = di.mesh->d_xyz;
                d.faces = di.mesh->d_faces;
                d.rgb = di.mesh->d_rgb;
                d.n = di.mesh->n;
                d.tri_begin = tri;
                d.frame = f;
                d.ramp = di.ramp;
                d.item_begin = items;
                memcpy(d.M, di.M, sizeof d.M);
                if ((uint64_t)tri + d.n >= (1ull << 31) - 1)
                    return fail(SLOTH_EINVAL, "frame %u has 2^31 - 1 or more triangles", f);
                tri += d.n;
                items += d.n;
                ++k;
            }
            fr.draw_end = (uint32_t)k;
            fr.n_tris = tri;
            if (tri >= TILED_MAX_TRIS) small_frames = false;
            uint64_t WH = (uint64_t)fr.W * fr.H;
            cells += WH + (fr.image ? fr.H : 0);
            // frames start on 256-byte boundaries: a frame's rows of tile
            // stores then cover whole cache lines (an odd frame base costs
            // the streaming stores a third more lines per row segment), and
            // slots store cell pairs as one 8-byte word
            cells = (cells + 63) & ~63ull;
            z += WH;
            rows += fr.H;
        }
        n_cells = cells;
        n_z = z;
        n_rows = rows;
        n_items = items;
        // the tiled path keeps a draw's shader index in the low byte of an
        // LDS word, where 0xFF marks a triangle whose glyph is constant;
        // frames of either width parity (or both) take it
        path = (small_frames && ramps.size() < 256 && !force_global()) ? PATH_TILED : PATH_GLOBAL;
        int rc;
        if ((rc = d_frames.reserve(sizeof(Frame) * std::max<size_t>(1, frames.size())))) return rc;
        if ((rc = d_draws.reserve(sizeof(Draw) * std::max<size_t>(1, draws.size())))) return rc;
        if ((rc = d_ramps.reserve(sizeof(Ramp) * std::max<size_t>(1, ramps.size())))) return rc;
        if ((rc = d_cells.reserve(sizeof(uint32_t) * n_cells))) return rc;
        if ((rc = d_z.reserve(sizeof(float) * n_z))) return rc;
        if (path == PATH_GLOBAL) {
            if ((rc = d_keys.reserve(sizeof(uint64_t) * n_z))) return rc;
            if ((rc = d_nl.reserve(sizeof(int32_t) * std::max<uint64_t>(1, n_rows)))) return rc;
        } else {
            if ((rc = tiled.plan(frames, n_rows))) return rc;
        }
        laid_out = true;
        return SLOTH_OK;
    }

    static bool force_global()
    {
        const char* e = getenv("SLOTH_FORCE_GLOBAL");
        return e && e[0] == '1';
    }

    // SLOTH_PIPELINE=1 captures batch graphs as a two-stream pipeline of
    // frame groups (bins of one group beside the raster of the previous one).
    // Off by default: measured slower than one stream on the 1080p sweep.
    static bool pipeline()
    {
        const char* e = getenv("SLOTH_PIPELINE");
        return e && e[0] == '1';
    }

    int upload(hipStream_t s)
    {
        if (!frames.empty())
            HIP_TRY(hipMemcpyAsync(d_frames.p, frames.data(), sizeof(Frame) * frames.size(), hipMemcpyHostToDevice, s));
        if (!draws.empty())
            HIP_TRY(hipMemcpyAsync(d_draws.p, draws.data(), sizeof(Draw) * draws.size(), hipMemcpyHostToDevice, s));
        if (!ramps.empty())
            HIP_TRY(hipMemcpyAsync(d_ramps.p, ramps.data(), sizeof(Ramp) * ramps.size(), hipMemcpyHostToDevice, s));
        if (path == PATH_TILED) {
            int rc = tiled.upload(s);
            if (rc) return rc;
        }
        // the host vectors may change after return: wait for the copies
        HIP_TRY(hipStreamSynchronize(s));
        return SLOTH_OK;
    }

    static unsigned blocks_for(uint64_t n, unsigned per = 256)
    {
        return (unsigned)((n + per - 1) / per);
    }

    // Global path over frames [f_lo, f_hi).
    int launch_global(uint32_t f_lo, uint32_t f_hi, hipStream_t s, int32_t* win, bool win_raw,
                      std::vector<hipEvent_t>* ev)
    {
        const Frame& a = frames[f_lo];
        const Frame& b = frames[f_hi - 1];
        uint64_t z_lo = a.z_off, z_hi = b.z_off + (uint64_t)b.W * b.H;
        uint64_t r_lo = a.nl_off, r_hi = b.nl_off + b.H;
        uint32_t d_lo = a.draw_begin, d_hi = b.draw_end;
        if (ev) HIP_TRY(hipEventRecord((*ev)[0], s));
        {
            uint64_t n = std::max(z_hi - z_lo, r_hi - r_lo);
            unsigned g = std::min<unsigned>(blocks_for(n), 8192u);
            hipLaunchKernelGGL(k_clear_global, dim3(std::max(1u, g)), dim3(256), 0, s, d_keys.as<uint64_t>() + z_lo,
                               z_hi - z_lo, d_nl.as<int32_t>() + r_lo, r_hi - r_lo);
            HIP_TRY(hipGetLastError());
        }
        if (ev) HIP_TRY(hipEventRecord((*ev)[1], s));
        // grid (blocks of the largest draw, draws), at most 65535 draws per launch
        uint32_t max_n = 0;
        for (uint32_t d = d_lo; d < d_hi; ++d) max_n = std::max(max_n, draws[d].n);
        for (uint32_t d = d_lo; d < d_hi && max_n; d += 65535u) {
            const uint32_t nd = std::min<uint32_t>(65535u, d_hi - d);
            hipLaunchKernelGGL(k_raster_global, dim3(blocks_for(max_n), nd), dim3(256), 0, s, d_draws.as<Draw>(), d,
                               d_frames.as<Frame>(), d_keys.as<uint64_t>(), d_nl.as<int32_t>());
            HIP_TRY(hipGetLastError());
        }
        if (ev) HIP_TRY(hipEventRecord((*ev)[2], s));
        // grids (cells of the largest frame / 256, frames), at most 65535
        // frames per launch
        uint64_t max_wh = 0, max_cells = 0;
        for (uint32_t f = f_lo; f < f_hi; ++f) {
            max_wh = std::max<uint64_t>(max_wh, (uint64_t)frames[f].W * frames[f].H);
            max_cells = std::max<uint64_t>(max_cells, frame_cells(frames[f]));
        }
        for (uint32_t f = f_lo; f < f_hi && max_wh; f += 65535u) {
            const uint32_t nf = std::min<uint32_t>(65535u, f_hi - f);
            hipLaunchKernelGGL(k_winner_global, dim3(blocks_for(max_wh), nf), dim3(256), 0, s, d_draws.as<Draw>(),
                               d_ramps.as<Ramp>(), d_frames.as<Frame>(), f, d_keys.as<uint64_t>(), d_z.as<float>(), win,
                               win_raw ? 1u : 0u);
            HIP_TRY(hipGetLastError());
        }
        if (ev) HIP_TRY(hipEventRecord((*ev)[3], s));
        for (uint32_t f = f_lo; f < f_hi && max_cells; f += 65535u) {
            const uint32_t nf = std::min<uint32_t>(65535u, f_hi - f);
            hipLaunchKernelGGL(k_resolve_global, dim3(blocks_for(max_cells), nf), dim3(256), 0, s,
                               d_frames.as<Frame>(), f, d_keys.as<uint64_t>(), d_nl.as<int32_t>(),
                               d_cells.as<uint32_t>());
            HIP_TRY(hipGetLastError());
        }
        if (ev) HIP_TRY(hipEventRecord((*ev)[4], s));
        return SLOTH_OK;
    }

    // Work that must precede a launch and cannot be captured in a graph: the
    // counting run that sizes the tile lists.
    int prepare(hipStream_t s)
    {
        if (path == PATH_TILED && !tiled.sized)
            return tiled.size(frames, draws, d_frames.as<Frame>(), d_draws.as<Draw>(), s);
        return SLOTH_OK;
    }

    // After a render of frames [f_lo, f_hi) completed on `s`: if a tile list
    // overflowed, re-size the lists and render those frames again.
    // *replanned tells the caller that device pointers changed.
    int finish(uint32_t f_lo, uint32_t f_hi, hipStream_t s, int32_t* win, bool* replanned = nullptr,
               bool win_raw = false)
    {
        if (replanned) *replanned = false;
        if (path != PATH_TILED) return SLOTH_OK;
        for (int attempt = 0; attempt < 4; ++attempt) {
            bool ovf = false;
            int rc = tiled.overflowed((uint32_t)frames.size(), s, ovf);
            if (rc) return rc;
            if (!ovf) return SLOTH_OK;
            tiled.sized = false;
            if (replanned) *replanned = true;
            if ((rc = prepare(s))) return rc;
            if ((rc = launch(f_lo, f_hi, s, win, nullptr, win_raw))) return rc;
        }
        return fail(SLOTH_EHIP, "tile lists kept overflowing");
    }

    // s2 / join (optional): second stream and events for the pipelined tiled launch
    int launch(uint32_t f_lo, uint32_t f_hi, hipStream_t s, int32_t* win, std::vector<hipEvent_t>* ev = nullptr,
               bool win_raw = false, hipStream_t s2 = nullptr, const std::vector<hipEvent_t>* join = nullptr)
    {
        if (f_lo >= f_hi) return SLOTH_OK;
        if (path == PATH_TILED)
            return tiled.launch(frames, draws, d_frames.as<Frame>(), d_draws.as<Draw>(), d_ramps.as<Ramp>(), f_lo,
                                f_hi, d_cells.as<uint32_t>(), d_z.as<float>(), win, win_raw, (uint32_t)ramps.size(), s,
                                ev, s2, join);
        return launch_global(f_lo, f_hi, s, win, win_raw, ev);
    }

    int ensure_win()
    {
        return d_win.reserve(sizeof(int32_t) * std::max<uint64_t>(1, n_z));
    }
};

static hipStream_t make_stream()
{
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return s;
}

static hipEvent_t make_event()
{
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

// Makes `to` wait for everything enqueued on `from` so far.  A handle's
// uploads, overflow checks and read-backs run on its own stream; a render
// launched on a caller's stream is joined to it in both directions, so those
// never overlap the render.
static int join_streams(hipStream_t from, hipStream_t to, hipEvent_t ev)
{
    if (from == to) return SLOTH_OK;
    HIP_TRY(hipEventRecord(ev, from));
    HIP_TRY(hipStreamWaitEvent(to, ev, 0));
    return SLOTH_OK;
}

// Device scratch of the flush encoders.
struct Encoder {
    DevBuf d_off, d_block, d_total, d_out;

    // n cells of one frame; a terminal-mode frame (image false) starts with
    // crossterm's MoveTo(0, 0), CSI "1;1H" (context.rs:53-55)
    int run(const uint32_t* d_cells, uint64_t n, bool image, int mode, char* out, uint64_t cap, uint64_t* len,
            hipStream_t s)
    {
        if (mode != SLOTH_ENCODE_TEXT && mode != SLOTH_ENCODE_ANSI && mode != SLOTH_ENCODE_HTML)
            return fail(SLOTH_EINVAL, "encode mode must be SLOTH_ENCODE_TEXT, SLOTH_ENCODE_ANSI or SLOTH_ENCODE_HTML");
        static const char MOVE_TO_ORIGIN[] = "\x1b[1;1H";
        const uint64_t pre = image ? 0 : sizeof MOVE_TO_ORIGIN - 1;
        int rc;
        const uint64_t nb = encode_blocks(n);
        if ((rc = d_off.reserve(std::max<uint64_t>(8, n * 8))) || (rc = d_block.reserve(std::max<uint64_t>(8, nb * 8))) ||
            (rc = d_total.reserve(8)))
            return rc;
        uint64_t total = 0;
        if (n) {
            if ((rc = launch_encode_size(d_cells, n, mode, d_off.as<uint64_t>(), d_block.as<uint64_t>(),
                                         d_total.as<uint64_t>(), s)))
                return rc;
            HIP_TRY(hipMemcpyAsync(&total, d_total.p, 8, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        const uint64_t full = pre + total + (mode == SLOTH_ENCODE_TEXT ? 1 : 0);   // println!'s newline
        if (len) *len = full;
        if (!out) return SLOTH_OK;
        if (cap < full) return fail(SLOTH_EINVAL, "encode buffer holds %llu bytes, %llu needed",
                                    (unsigned long long)cap, (unsigned long long)full);
        memcpy(out, MOVE_TO_ORIGIN, pre);
        if (total) {
            if ((rc = d_out.reserve(total))) return rc;
            if ((rc = launch_encode_write(d_cells, n, mode, d_off.as<uint64_t>(), d_block.as<uint64_t>(),
                                          d_out.as<uint8_t>(), s)))
                return rc;
            HIP_TRY(hipMemcpyAsync(out + pre, d_out.p, total, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        if (mode == SLOTH_ENCODE_TEXT) out[pre + total] = '\n';
        return SLOTH_OK;
    }
};

// ---------------------------------------------------------------------------
// contexts
// ---------------------------------------------------------------------------
struct sloth_context {
    Job job;
    float U[16];
    hipStream_t stream = nullptr;
    hipEvent_t join_ev = nullptr;   // joins a caller's stream to `stream`
    bool rendered = false;      // device buffers match the draw list
    bool win_valid = false;     // d_win matches the draw list
    bool checked = false;       // the last render was checked for tile-list overflow
    Encoder enc;
    ~sloth_context()
    {
        if (join_ev) (void)hipEventDestroy(join_ev);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
struct sloth_batch {
    Job job;
    std::vector<float> U;       // per frame utransform
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;       // second branch of the captured pipeline
    std::vector<hipEvent_t> join;        // its fork / join events
    hipEvent_t join_ev = nullptr;        // joins a caller's stream to `stream`
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t exec_generation = 0;   // tile plan generation the graph was captured with
    bool ran = false;
    std::vector<bool> win_valid;
    bool checked = false;           // the last render was checked for tile-list overflow
    Encoder enc;
    ~sloth_batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : join)
            if (e) (void)hipEventDestroy(e);
        if (join_ev) (void)hipEventDestroy(join_ev);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* sloth_last_error(void) { return g_err.c_str(); }

int sloth_init(int device)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(SLOTH_ENODEV, "no HIP device (%s)", hipGetErrorString(e));
    }
    ARG_CHECK(device >= 0 && device < n, "device index out of range");
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, device));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return fail(SLOTH_ENODEV, "device %d is %s, this library is built for gfx950", device, p.gcnArchName);
    HIP_TRY(hipSetDevice(device));
    g_device = device;
    return noalloc_init(device);
}

static int ensure_device()
{
    static thread_local bool ok = false;
    if (ok) return SLOTH_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(SLOTH_ENODEV, "no HIP device (%s)", hipGetErrorString(e));
    }
    HIP_TRY(hipSetDevice(g_device));
    ok = true;
    return SLOTH_OK;
}

int sloth_default_ramp(sloth_ramp* out)
{
    ARG_CHECK(out, "null output");
    Ramp r = default_ramp();
    memset(out, 0, sizeof *out);
    out->n = r.n;
    for (uint32_t i = 0; i < r.n; ++i) {
        out->thresholds[i] = r.thr[i];
        out->glyphs[i] = r.glyph[i];
    }
    out->fallback = r.fallback;
    return SLOTH_OK;
}

int sloth_screen_transform(uint32_t width, uint32_t height, const float* bbox_max, uint32_t n_meshes, float* out16)
{
    ARG_CHECK(out16 && (bbox_max || n_meshes == 0), "null argument");
    screen_transform(width, height, bbox_max, n_meshes, out16);
    return SLOTH_OK;
}

int sloth_rotation_euler(float roll, float pitch, float yaw, float* out16)
{
    ARG_CHECK(out16, "null output");
    // Rotation3::from_euler_angles (main.rs:76-77) with the C library's
    // binary32 sine and cosine, as Rust's f32::sin_cos; products in the
    // order nalgebra writes them, then to_homogeneous
    const float sr = sinf(roll), cr = cosf(roll);
    const float sp = sinf(pitch), cp = cosf(pitch);
    const float sy = sinf(yaw), cy = cosf(yaw);
    const float m[16] = {cy * cp, sy * cp, -sp, 0.0f,
                         cy * sp * sr - sy * cr, sy * sp * sr + cy * cr, cp * sr, 0.0f,
                         cy * sp * cr + sy * sr, sy * sp * cr - cy * sr, cp * cr, 0.0f,
                         0.0f, 0.0f, 0.0f, 1.0f};
    memcpy(out16, m, sizeof m);
    return SLOTH_OK;
}

int sloth_mat4_mul(const float* a16, const float* b16, float* out16)
{
    ARG_CHECK(a16 && b16 && out16, "null argument");
    mat4_mul(a16, b16, out16);
    return SLOTH_OK;
}

// ---- meshes ---------------------------------------------------------------

int sloth_mesh_create(const float* positions, const uint8_t* colors, uint32_t n, sloth_mesh** out)
{
    ARG_CHECK(out, "null output");
    *out = nullptr;
    ARG_CHECK(n == 0 || (positions && colors), "null positions or colours");
    int rc = ensure_device();
    if (rc) return rc;
    sloth_mesh* m = new (std::nothrow) sloth_mesh();
    if (!m) return fail(SLOTH_ENOMEM, "host allocation failed");
    m->n = n;
    host_bbox(positions, n, m->bbmin, m->bbmax);
    if (n) {
        if ((rc = m->own_xyz.reserve((size_t)n * 36)) || (rc = m->own_rgb.reserve((size_t)n * 3))) {
            delete m;
            return rc;
        }
        hipError_t e1 = hipMemcpy(m->own_xyz.p, positions, (size_t)n * 36, hipMemcpyHostToDevice);
        hipError_t e2 = hipMemcpy(m->own_rgb.p, colors, (size_t)n * 3, hipMemcpyHostToDevice);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            delete m;
            return fail(SLOTH_EHIP, "mesh upload: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        }
        m->d_xyz = m->own_xyz.as<float>();
        m->d_rgb = m->own_rgb.as<uint8_t>();
    }
    *out = m;
    return SLOTH_OK;
}

int sloth_mesh_create_device(const float* d_positions, const uint8_t* d_colors, uint32_t n, sloth_mesh** out)
{
    ARG_CHECK(out, "null output");
    *out = nullptr;
    ARG_CHECK(n == 0 || (d_positions && d_colors), "null positions or colours");
    int rc = ensure_device();
    if (rc) return rc;
    sloth_mesh* m = new (std::nothrow) sloth_mesh();
    if (!m) return fail(SLOTH_ENOMEM, "host allocation failed");
    m->n = n;
    m->d_xyz = d_positions;
    m->d_rgb = d_colors;
    if (n) {
        std::vector<float> h;
        try {
            h.resize((size_t)n * 9);
        } catch (...) {
            delete m;
            return fail(SLOTH_ENOMEM, "host allocation failed");
        }
        hipError_t e = hipMemcpy(h.data(), d_positions, (size_t)n * 36, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            delete m;
            return fail(SLOTH_EHIP, "mesh bbox readback: %s", hipGetErrorString(e));
        }
        host_bbox(h.data(), n, m->bbmin, m->bbmax);
    }
    *out = m;
    return SLOTH_OK;
}

int sloth_mesh_create_indexed(const float* vertices, uint32_t n_vertices, const uint32_t* faces,
                              const uint8_t* colors, uint32_t n, sloth_mesh** out)
{
    ARG_CHECK(out, "null output");
    *out = nullptr;
    ARG_CHECK(n == 0 || (vertices && faces && colors), "null vertices, faces or colours");
    for (uint64_t i = 0; i < (uint64_t)n * 3; ++i)
        if (faces[i] >= n_vertices) return fail(SLOTH_EINVAL, "face %llu names vertex %u of %u", (unsigned long long)(i / 3),
                                                faces[i], n_vertices);
    int rc = ensure_device();
    if (rc) return rc;
    sloth_mesh* m = new (std::nothrow) sloth_mesh();
    if (!m) return fail(SLOTH_ENOMEM, "host allocation failed");
    m->n = n;
    m->n_verts = n_vertices;
    // the OBJ bounding box (geometry.rs:85-88, 129-135) over the triangles
    for (int c = 0; c < 3; ++c) m->bbmin[c] = m->bbmax[c] = 0.0f;
    for (uint32_t t = 0; t < n; ++t)
        for (int c = 0; c < 3; ++c) {
            const float a = vertices[3 * (size_t)faces[3 * (size_t)t] + c];
            const float b = vertices[3 * (size_t)faces[3 * (size_t)t + 1] + c];
            const float d = vertices[3 * (size_t)faces[3 * (size_t)t + 2] + c];
            m->bbmin[c] = fminf(fminf(a, fminf(b, d)), m->bbmin[c]);
            m->bbmax[c] = fmaxf(fmaxf(a, fmaxf(b, d)), m->bbmax[c]);
        }
    if (n) {
        const size_t vb = (size_t)n_vertices * 12, fb = (size_t)n * 12, cb = (size_t)n * 3;
        if ((rc = m->own_xyz.reserve(std::max<size_t>(vb, 4))) || (rc = m->own_faces.reserve(fb)) ||
            (rc = m->own_rgb.reserve(cb))) {
            delete m;
            return rc;
        }
        hipError_t e1 = vb ? hipMemcpy(m->own_xyz.p, vertices, vb, hipMemcpyHostToDevice) : hipSuccess;
        hipError_t e2 = hipMemcpy(m->own_faces.p, faces, fb, hipMemcpyHostToDevice);
        hipError_t e3 = hipMemcpy(m->own_rgb.p, colors, cb, hipMemcpyHostToDevice);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            delete m;
            return fail(SLOTH_EHIP, "indexed mesh upload failed");
        }
        m->d_xyz = m->own_xyz.as<float>();
        m->d_faces = m->own_faces.as<uint32_t>();
        m->d_rgb = m->own_rgb.as<uint8_t>();
    }
    *out = m;
    return SLOTH_OK;
}

int sloth_mesh_is_indexed(const sloth_mesh* mesh, uint32_t* n_vertices)
{
    ARG_CHECK(mesh, "null mesh");
    if (n_vertices) *n_vertices = mesh->d_faces ? mesh->n_verts : 0;
    return mesh->d_faces ? 1 : 0;
}

int sloth_mesh_bbox(const sloth_mesh* mesh, float* min3, float* max3)
{
    ARG_CHECK(mesh, "null mesh");
    if (min3) memcpy(min3, mesh->bbmin, 12);
    if (max3) memcpy(max3, mesh->bbmax, 12);
    return SLOTH_OK;
}

uint32_t sloth_mesh_triangles(const sloth_mesh* mesh) { return mesh ? mesh->n : 0; }

// ---- model files (inputs.rs:93-129, geometry.rs:83-181) ----------------------

}  // extern "C"

struct sloth_model {
    std::vector<SoupMesh> meshes;
};

// Hash of a vertex's three binary32 bit patterns.
struct VertexHash {
    size_t operator()(const std::array<uint32_t, 3>& v) const
    {
        uint64_t h = 1469598103934665603ull;
        for (uint32_t x : v) h = (h ^ x) * 1099511628211ull;
        return (size_t)h;
    }
};

extern "C" {

int sloth_model_load(const char* path, sloth_model** out)
{
    ARG_CHECK(path && out, "null argument");
    *out = nullptr;
    sloth_model* m = new (std::nothrow) sloth_model();
    if (!m) return fail(SLOTH_ENOMEM, "host allocation failed");
    std::string err;
    try {
        err = load_model_file(path, m->meshes);
    } catch (const std::bad_alloc&) {
        delete m;
        return fail(SLOTH_ENOMEM, "host allocation failed while loading %s", path);
    }
    if (!err.empty()) {
        delete m;
        return fail(SLOTH_EINVAL, "%s: %s", path, err.c_str());
    }
    *out = m;
    return SLOTH_OK;
}

int sloth_model_mesh_count(const sloth_model* model, uint32_t* n_meshes)
{
    ARG_CHECK(model && n_meshes, "null argument");
    *n_meshes = (uint32_t)model->meshes.size();
    return SLOTH_OK;
}

int sloth_model_mesh(const sloth_model* model, uint32_t i, const float** positions, const uint8_t** colors,
                     uint32_t* n_tris, float* bbox_min3, float* bbox_max3)
{
    ARG_CHECK(model, "null model");
    ARG_CHECK(i < model->meshes.size(), "mesh index out of range");
    const SoupMesh& m = model->meshes[i];
    if (positions) *positions = m.xyz.data();
    if (colors) *colors = m.rgb.data();
    if (n_tris) *n_tris = m.n();
    if (bbox_min3) memcpy(bbox_min3, m.bbmin, 12);
    if (bbox_max3) memcpy(bbox_max3, m.bbmax, 12);
    return SLOTH_OK;
}

int sloth_model_upload(const sloth_model* model, uint32_t i, sloth_mesh** out)
{
    ARG_CHECK(model && out, "null argument");
    ARG_CHECK(i < model->meshes.size(), "mesh index out of range");
    const SoupMesh& m = model->meshes[i];
    // Vertices that the soup repeats bit for bit are stored once: an indexed
    // device mesh of the same floats (renders identically to the soup)
    std::vector<float> verts;
    std::vector<uint32_t> faces;
    try {
        std::unordered_map<std::array<uint32_t, 3>, uint32_t, VertexHash> ids;
        faces.resize((size_t)m.n() * 3);
        for (size_t k = 0; k < (size_t)m.n() * 3; ++k) {
            std::array<uint32_t, 3> key;
            memcpy(key.data(), &m.xyz[k * 3], 12);
            auto it = ids.emplace(key, (uint32_t)(verts.size() / 3));
            if (it.second) verts.insert(verts.end(), &m.xyz[k * 3], &m.xyz[k * 3] + 3);
            faces[k] = it.first->second;
        }
    } catch (const std::bad_alloc&) {
        return fail(SLOTH_ENOMEM, "host allocation failed");
    }
    int rc = sloth_mesh_create_indexed(verts.data(), (uint32_t)(verts.size() / 3), faces.data(), m.rgb.data(), m.n(), out);
    if (rc) return rc;
    // the loader's bounding-box rule (an STL box does not start at the origin)
    memcpy((*out)->bbmin, m.bbmin, 12);
    memcpy((*out)->bbmax, m.bbmax, 12);
    return SLOTH_OK;
}

int sloth_model_destroy(sloth_model* model)
{
    delete model;
    return SLOTH_OK;
}

int sloth_mesh_destroy(sloth_mesh* mesh)
{
    delete mesh;
    return SLOTH_OK;
}

// ---- contexts -------------------------------------------------------------

int sloth_context_create(uint32_t width, uint32_t height, int image, sloth_context** out)
{
    ARG_CHECK(out, "null output");
    *out = nullptr;
    ARG_CHECK(width >= 1 && width <= 65535 && height >= 1 && height <= 65535,
              "width and height must be in 1..65535");
    int rc = ensure_device();
    if (rc) return rc;
    sloth_context* c = new (std::nothrow) sloth_context();
    if (!c) return fail(SLOTH_ENOMEM, "host allocation failed");
    memcpy(c->U, IDENTITY, sizeof c->U);
    c->job.fin.push_back(FrameIn{width, height, image ? 1 : 0});
    c->stream = make_stream();
    c->join_ev = make_event();
    if (!c->stream || !c->join_ev) {
        delete c;
        return fail(SLOTH_EHIP, "hipStreamCreate / hipEventCreate failed");
    }
    *out = c;
    return SLOTH_OK;
}

int sloth_context_destroy(sloth_context* ctx)
{
    if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return SLOTH_OK;
}

int sloth_context_update(sloth_context* ctx, const float* bbox_max, uint32_t n_meshes)
{
    ARG_CHECK(ctx, "null context");
    ARG_CHECK(bbox_max || n_meshes == 0, "null bbox array");
    screen_transform(ctx->job.fin[0].W, ctx->job.fin[0].H, bbox_max, n_meshes, ctx->U);
    return SLOTH_OK;
}

int sloth_context_set_utransform(sloth_context* ctx, const float* m16)
{
    ARG_CHECK(ctx && m16, "null argument");
    memcpy(ctx->U, m16, sizeof ctx->U);
    return SLOTH_OK;
}

int sloth_context_get_utransform(const sloth_context* ctx, float* m16)
{
    ARG_CHECK(ctx && m16, "null argument");
    memcpy(m16, ctx->U, sizeof ctx->U);
    return SLOTH_OK;
}

int sloth_context_clear(sloth_context* ctx)
{
    ARG_CHECK(ctx, "null context");
    ctx->job.din.clear();
    ctx->job.ramps.clear();
    ctx->rendered = false;
    ctx->win_valid = false;
    return SLOTH_OK;
}

int sloth_draw_mesh(sloth_context* ctx, const sloth_mesh* mesh, const float* transform16, const sloth_ramp* shader)
{
    ARG_CHECK(ctx && mesh && transform16, "null argument");
    Ramp r;
    int rc = to_ramp(shader, r);
    if (rc) return rc;
    DrawIn d;
    d.frame = 0;
    d.mesh = mesh;
    mat4_mul(ctx->U, transform16, d.M);
    d.ramp = ctx->job.add_ramp(r);
    ctx->job.din.push_back(d);
    ctx->rendered = false;
    ctx->win_valid = false;
    return SLOTH_OK;
}

static int context_render(sloth_context* ctx, hipStream_t s, bool want_win)
{
    int rc;
    if (!ctx->rendered) {
        if ((rc = ctx->job.layout())) return rc;
        if ((rc = ctx->job.upload(ctx->stream))) return rc;
        if ((rc = ctx->job.prepare(ctx->stream))) return rc;
        if (want_win && (rc = ctx->job.ensure_win())) return rc;
        if ((rc = join_streams(ctx->stream, s, ctx->join_ev))) return rc;
        if ((rc = ctx->job.launch(0, 1, s, want_win ? ctx->job.d_win.as<int32_t>() : nullptr))) return rc;
        if ((rc = join_streams(s, ctx->stream, ctx->join_ev))) return rc;
        ctx->rendered = true;
        ctx->checked = false;
        ctx->win_valid = want_win;
    } else if (want_win && !ctx->win_valid) {
        if ((rc = ctx->job.ensure_win())) return rc;
        if ((rc = join_streams(ctx->stream, s, ctx->join_ev))) return rc;
        if ((rc = ctx->job.launch(0, 1, s, ctx->job.d_win.as<int32_t>()))) return rc;
        if ((rc = join_streams(s, ctx->stream, ctx->join_ev))) return rc;
        ctx->checked = false;
        ctx->win_valid = true;
    }
    return SLOTH_OK;
}

int sloth_context_render(sloth_context* ctx, void* stream)
{
    ARG_CHECK(ctx, "null context");
    int rc = ensure_device();
    if (rc) return rc;
    return context_render(ctx, stream ? (hipStream_t)stream : ctx->stream, false);
}

int sloth_context_read(sloth_context* ctx, uint32_t* cells, float* z, int32_t* winners)
{
    ARG_CHECK(ctx, "null context");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    if ((rc = context_render(ctx, s, winners != nullptr))) return rc;
    if (!ctx->checked) {
        if ((rc = ctx->job.finish(0, 1, s, ctx->win_valid ? ctx->job.d_win.as<int32_t>() : nullptr))) return rc;
        ctx->checked = true;
    }
    const Job& j = ctx->job;
    const uint64_t nc = frame_cells(j.frames[0]);   // the buffer is padded past the frame's cells
    if (cells && nc) HIP_TRY(hipMemcpyAsync(cells, j.d_cells.p, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, s));
    if (z && j.n_z) HIP_TRY(hipMemcpyAsync(z, j.d_z.p, sizeof(float) * j.n_z, hipMemcpyDeviceToHost, s));
    if (winners && j.n_z)
        HIP_TRY(hipMemcpyAsync(winners, j.d_win.p, sizeof(int32_t) * j.n_z, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SLOTH_OK;
}

int sloth_context_encode(sloth_context* ctx, int mode, char* out, uint64_t cap, uint64_t* len)
{
    ARG_CHECK(ctx, "null context");
    int rc = sloth_context_read(ctx, nullptr, nullptr, nullptr);   // render + overflow check
    if (rc) return rc;
    // the frame's own cells (the layout pads the buffer past them)
    const Frame& fr = ctx->job.frames[0];
    return ctx->enc.run(ctx->job.d_cells.as<uint32_t>(), frame_cells(fr), fr.image != 0, mode, out, cap, len,
                        ctx->stream);
}

int sloth_context_device_buffers(sloth_context* ctx, uint32_t** d_cells, float** d_z, uint64_t* n_cells,
                                 uint64_t* n_z)
{
    ARG_CHECK(ctx, "null context");
    if (!ctx->job.laid_out) {
        int rc = ctx->job.layout();
        if (rc) return rc;
    }
    if (d_cells) *d_cells = ctx->job.d_cells.as<uint32_t>();
    if (d_z) *d_z = ctx->job.d_z.as<float>();
    if (n_cells) *n_cells = frame_cells(ctx->job.frames[0]);
    if (n_z) *n_z = ctx->job.n_z;
    return SLOTH_OK;
}

// ---- batches --------------------------------------------------------------

int sloth_batch_create(const sloth_frame_desc* frames, uint32_t n_frames, const sloth_draw_desc* draws,
                       uint32_t n_draws, const sloth_ramp* shader, sloth_batch** out)
{
    ARG_CHECK(out, "null output");
    *out = nullptr;
    ARG_CHECK(frames && n_frames > 0, "a batch needs at least one frame");
    ARG_CHECK(draws || n_draws == 0, "null draw array");
    for (uint32_t f = 0; f < n_frames; ++f)
        ARG_CHECK(frames[f].width >= 1 && frames[f].width <= 65535 && frames[f].height >= 1 &&
                      frames[f].height <= 65535,
                  "frame width and height must be in 1..65535");
    for (uint32_t d = 0; d < n_draws; ++d)
        ARG_CHECK(draws[d].frame < n_frames && draws[d].mesh, "draw with bad frame index or null mesh");
    int rc = ensure_device();
    if (rc) return rc;
    Ramp r;
    if ((rc = to_ramp(shader, r))) return rc;
    sloth_batch* b = new (std::nothrow) sloth_batch();
    if (!b) return fail(SLOTH_ENOMEM, "host allocation failed");
    b->stream = make_stream();
    b->stream2 = make_stream();
    b->join_ev = make_event();
    if (!b->stream || !b->stream2 || !b->join_ev) {
        delete b;
        return fail(SLOTH_EHIP, "hipStreamCreate / hipEventCreate failed");
    }
    b->join.assign(MAX_GROUPS + 1, nullptr);
    for (hipEvent_t& e : b->join)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            e = nullptr;
            delete b;
            return fail(SLOTH_EHIP, "hipEventCreate failed");
        }
    uint32_t ri = b->job.add_ramp(r);
    b->U.resize((size_t)n_frames * 16);
    for (uint32_t f = 0; f < n_frames; ++f) {
        b->job.fin.push_back(FrameIn{frames[f].width, frames[f].height, frames[f].image ? 1 : 0});
        memcpy(&b->U[(size_t)f * 16], frames[f].utransform, 64);
    }
    for (uint32_t d = 0; d < n_draws; ++d) {
        DrawIn di;
        di.frame = draws[d].frame;
        di.mesh = draws[d].mesh;
        mat4_mul(&b->U[(size_t)di.frame * 16], draws[d].transform, di.M);
        di.ramp = ri;
        b->job.din.push_back(di);
    }
    if ((rc = b->job.layout()) || (rc = b->job.upload(b->stream))) {
        delete b;
        return rc;
    }
    b->win_valid.assign(n_frames, false);
    *out = b;
    return SLOTH_OK;
}

int sloth_batch_set_transforms(sloth_batch* b, const float* transforms16)
{
    ARG_CHECK(b && transforms16, "null argument");
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (size_t d = 0; d < b->job.din.size(); ++d) {
        DrawIn& di = b->job.din[d];
        mat4_mul(&b->U[(size_t)di.frame * 16], transforms16 + 16 * d, di.M);
    }
    // the draw table is grouped by frame; recompute it in the same order
    int rc;
    if ((rc = b->job.layout()) || (rc = b->job.upload(b->stream))) return rc;
    // the frames on the device show the old transforms: a read must follow a
    // run, whose overflow check covers the new ones
    b->ran = false;
    b->checked = false;
    std::fill(b->win_valid.begin(), b->win_valid.end(), false);
    return SLOTH_OK;
}

static void drop_graph(sloth_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Checks the finished renders for tile-list overflow and re-renders if needed.
static int batch_finish(sloth_batch* b)
{
    bool replanned = false;
    int rc = b->job.finish(0, (uint32_t)b->job.frames.size(), b->stream, nullptr, &replanned);
    if (rc) return rc;
    if (replanned) drop_graph(b);
    return SLOTH_OK;
}

int sloth_batch_run(sloth_batch* b, void* stream)
{
    ARG_CHECK(b, "null batch");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : b->stream;
    if ((rc = b->job.prepare(b->stream))) return rc;
    if (b->exec && b->exec_generation != b->job.tiled.generation) drop_graph(b);
    if (!b->exec) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const bool pipe = Job::pipeline();
        rc = b->job.launch(0, (uint32_t)b->job.frames.size(), b->stream, nullptr, nullptr, false,
                           pipe ? b->stream2 : nullptr, pipe ? &b->join : nullptr);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(SLOTH_EHIP, "graph capture: %s", hipGetErrorString(e));
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->exec_generation = b->job.tiled.generation;
    }
    if ((rc = join_streams(b->stream, s, b->join_ev))) return rc;
    HIP_TRY(hipGraphLaunch(b->exec, s));
    if ((rc = join_streams(s, b->stream, b->join_ev))) return rc;
    b->ran = true;
    b->checked = false;
    std::fill(b->win_valid.begin(), b->win_valid.end(), false);
    return SLOTH_OK;
}

int sloth_batch_run_per_frame(sloth_batch* b, void* stream)
{
    ARG_CHECK(b, "null batch");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = stream ? (hipStream_t)stream : b->stream;
    if ((rc = b->job.prepare(b->stream))) return rc;
    if ((rc = join_streams(b->stream, s, b->join_ev))) return rc;
    for (uint32_t f = 0; f < (uint32_t)b->job.frames.size(); ++f) {
        if ((rc = b->job.launch(f, f + 1, s, nullptr))) return rc;
        HIP_TRY(hipStreamSynchronize(s));
    }
    b->ran = true;
    b->checked = false;
    std::fill(b->win_valid.begin(), b->win_valid.end(), false);
    return SLOTH_OK;
}

int sloth_batch_sync(sloth_batch* b)
{
    ARG_CHECK(b, "null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->ran && !b->checked) {
        int rc = batch_finish(b);
        if (rc) return rc;
        b->checked = true;
    }
    return SLOTH_OK;
}

int sloth_batch_read_frame(sloth_batch* b, uint32_t frame, uint32_t* cells, float* z, int32_t* winners)
{
    ARG_CHECK(b, "null batch");
    ARG_CHECK(frame < b->job.frames.size(), "frame index out of range");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = b->stream;
    if (!b->ran) return fail(SLOTH_ESTATE, "batch has not been run");
    if ((rc = sloth_batch_sync(b))) return rc;
    const Frame& fr = b->job.frames[frame];
    const uint64_t WH = (uint64_t)fr.W * fr.H, nc = frame_cells(fr);
    if (winners && !b->win_valid[frame]) {
        if ((rc = b->job.ensure_win())) return rc;
        if ((rc = b->job.launch(frame, frame + 1, s, b->job.d_win.as<int32_t>()))) return rc;
        // the lists were sized for these transforms by the checked run, but a
        // re-plan here must still re-render the frame and drop the graph
        bool replanned = false;
        if ((rc = b->job.finish(frame, frame + 1, s, b->job.d_win.as<int32_t>(), &replanned))) return rc;
        if (replanned) drop_graph(b);
        b->win_valid[frame] = true;
    }
    if (cells)
        HIP_TRY(hipMemcpyAsync(cells, b->job.d_cells.as<uint32_t>() + fr.cell_off, 4 * nc, hipMemcpyDeviceToHost, s));
    if (z) HIP_TRY(hipMemcpyAsync(z, b->job.d_z.as<float>() + fr.z_off, 4 * WH, hipMemcpyDeviceToHost, s));
    if (winners)
        HIP_TRY(hipMemcpyAsync(winners, b->job.d_win.as<int32_t>() + fr.z_off, 4 * WH, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SLOTH_OK;
}

int sloth_batch_encode_frame(sloth_batch* b, uint32_t frame, int mode, char* out, uint64_t cap, uint64_t* len)
{
    ARG_CHECK(b, "null batch");
    ARG_CHECK(frame < b->job.frames.size(), "frame index out of range");
    if (!b->ran) return fail(SLOTH_ESTATE, "batch has not been run");
    int rc = sloth_batch_sync(b);
    if (rc) return rc;
    const Frame& fr = b->job.frames[frame];
    const uint64_t nc = frame_cells(fr);
    return b->enc.run(b->job.d_cells.as<uint32_t>() + fr.cell_off, nc, fr.image != 0, mode, out, cap, len, b->stream);
}

int sloth_batch_encode_export(sloth_batch* b, uint32_t first, uint32_t count, int mode, char* out, uint64_t cap,
                              uint64_t* len)
{
    ARG_CHECK(b && len, "null argument");
    ARG_CHECK(count >= 1 && first < b->job.frames.size() && count <= b->job.frames.size() - first,
              "frame range out of the batch");
    ARG_CHECK(mode == SLOTH_ENCODE_HTML || mode == SLOTH_ENCODE_TEXT,
              "the export encodes frames as SLOTH_ENCODE_HTML or SLOTH_ENCODE_TEXT");
    if (!b->ran) return fail(SLOTH_ESTATE, "batch has not been run");
    int rc = sloth_batch_sync(b);
    if (rc) return rc;
    // main.rs:55-58, 85-106: the header, then per frame a backtick line, the
    // frame's flush and "`," -- "`];" after the last one
    static const char HEAD[] = "let frames = [\n", OPEN[] = "`\n", NEXT[] = "`,\n", LAST[] = "`];\n";
    uint64_t pos = 0;
    auto put = [&](const char* txt, uint64_t n) -> int {
        if (out) {
            if (pos + n > cap)
                return fail(SLOTH_EINVAL, "export buffer holds %llu bytes, more are needed", (unsigned long long)cap);
            memcpy(out + pos, txt, n);
        }
        pos += n;
        return SLOTH_OK;
    };
    if ((rc = put(HEAD, sizeof HEAD - 1))) return rc;
    for (uint32_t f = first; f < first + count; ++f) {
        if ((rc = put(OPEN, sizeof OPEN - 1))) return rc;
        const Frame& fr = b->job.frames[f];
        uint64_t n = 0;
        if ((rc = b->enc.run(b->job.d_cells.as<uint32_t>() + fr.cell_off, frame_cells(fr), fr.image != 0, mode,
                             nullptr, 0, &n, b->stream)))
            return rc;
        if (out) {
            if (pos + n > cap)
                return fail(SLOTH_EINVAL, "export buffer holds %llu bytes, more are needed", (unsigned long long)cap);
            if ((rc = b->enc.run(b->job.d_cells.as<uint32_t>() + fr.cell_off, frame_cells(fr), fr.image != 0, mode,
                                 out + pos, cap - pos, &n, b->stream)))
                return rc;
        }
        pos += n;
        if ((rc = f + 1 < first + count ? put(NEXT, sizeof NEXT - 1) : put(LAST, sizeof LAST - 1))) return rc;
    }
    *len = pos;
    return SLOTH_OK;
}

int sloth_batch_device_frame(sloth_batch* b, uint32_t frame, uint32_t** d_cells, float** d_z)
{
    ARG_CHECK(b, "null batch");
    ARG_CHECK(frame < b->job.frames.size(), "frame index out of range");
    const Frame& fr = b->job.frames[frame];
    if (d_cells) *d_cells = b->job.d_cells.as<uint32_t>() + fr.cell_off;
    if (d_z) *d_z = b->job.d_z.as<float>() + fr.z_off;
    return SLOTH_OK;
}

int sloth_batch_profile(sloth_batch* b, float* out_ms, const char** names, uint32_t* n_stages)
{
    ARG_CHECK(b && n_stages, "null argument");
    int rc = ensure_device();
    if (rc) return rc;
    const uint32_t NEV = 8;
    std::vector<hipEvent_t> ev(NEV, nullptr);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    rc = b->job.prepare(b->stream);
    if (!rc) rc = b->job.launch(0, (uint32_t)b->job.frames.size(), b->stream, nullptr, &ev);
    hipError_t se = hipStreamSynchronize(b->stream);
    std::vector<const char*> nm;
    if (b->job.path == PATH_TILED)
        nm = TiledPlan::stage_names();
    else
        nm = {"k_clear_global", "k_raster_global", "k_winner_global", "k_resolve_global"};
    uint32_t cap = *n_stages;
    *n_stages = (uint32_t)nm.size();
    if (!rc && se == hipSuccess) {
        for (uint32_t i = 0; i < nm.size() && i < cap; ++i) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
            if (out_ms) out_ms[i] = ms;
            if (names) names[i] = nm[i];
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    if (se != hipSuccess) return fail(SLOTH_EHIP, "profile run: %s", hipGetErrorString(se));
    b->ran = true;
    b->checked = false;
    std::fill(b->win_valid.begin(), b->win_valid.end(), false);
    bool replanned = false;
    if ((rc = b->job.finish(0, (uint32_t)b->job.frames.size(), b->stream, nullptr, &replanned))) return rc;
    b->checked = true;
    if (replanned) {
        drop_graph(b);
        return fail(SLOTH_ESTATE, "tile lists were re-sized during profiling; profile again");
    }
    return SLOTH_OK;
}

int sloth_batch_time(sloth_batch* b, uint32_t steps, uint32_t warmup, int mode, float* ms)
{
    ARG_CHECK(b && ms, "null argument");
    ARG_CHECK(mode == 0 || mode == 1, "mode must be 0 (graph) or 1 (per frame)");
    int rc = ensure_device();
    if (rc) return rc;
    for (int attempt = 0; attempt < 3; ++attempt) {
        for (uint32_t i = 0; i < warmup; ++i)
            if ((rc = mode == 0 ? sloth_batch_run(b, nullptr) : sloth_batch_run_per_frame(b, nullptr))) return rc;
        if ((rc = sloth_batch_sync(b))) return rc;
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, b->stream));
        for (uint32_t i = 0; i < steps && rc == SLOTH_OK; ++i)
            rc = mode == 0 ? sloth_batch_run(b, nullptr) : sloth_batch_run_per_frame(b, nullptr);
        hipError_t e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        if (e != hipSuccess) return fail(SLOTH_EHIP, "batch timing: %s", hipGetErrorString(e));
        bool replanned = false;
        if ((rc = b->job.finish(0, (uint32_t)b->job.frames.size(), b->stream, nullptr, &replanned))) return rc;
        b->checked = true;
        if (!replanned) {
            *ms = t;
            return SLOTH_OK;
        }
        drop_graph(b);   // tile lists were re-sized: the timed renders were redone; time again
    }
    return fail(SLOTH_ESTATE, "tile lists kept re-sizing while timing");
}

int sloth_batch_set_tile(sloth_batch* b, uint32_t tile_w, uint32_t tile_h)
{
    ARG_CHECK(b, "null batch");
    int rc = b->job.tiled.set_tile(tile_w, tile_h);
    if (rc) return rc;
    if (!b->job.tiled.sized) {
        if ((rc = b->job.layout()) || (rc = b->job.upload(b->stream))) return rc;
        drop_graph(b);
        b->ran = false;
    }
    return SLOTH_OK;
}

int sloth_batch_set_walk(sloth_batch* b, int mode)
{
    ARG_CHECK(b, "null batch");
    ARG_CHECK(mode >= -1 && mode <= 1, "mode must be -1 (the plan's choice), 0 or 1");
    b->job.tiled.set_walk(mode);
    return SLOTH_OK;
}

int sloth_batch_set_bin(sloth_batch* b, uint32_t tris_per_thread)
{
    ARG_CHECK(b, "null batch");
    ARG_CHECK(tris_per_thread == 1 || tris_per_thread == 2 || tris_per_thread == 4,
              "tris_per_thread must be 1, 2 or 4");
    b->job.tiled.set_bin_tpt((int)tris_per_thread);
    return SLOTH_OK;
}

}  // extern "C"

// Autotuner cache: class of a batch -> best tile size.
#include <map>
#include <mutex>
#include <tuple>

struct TuneChoice {
    uint32_t tw, th;
    int walk;
    int bin_tpt;
};
static std::mutex g_tune_mu;
static std::map<std::tuple<uint64_t, uint64_t, uint32_t>, TuneChoice> g_tune;

static std::tuple<uint64_t, uint64_t, uint32_t> tune_class(const Job& j)
{
    // frame sizes (order-independent hash), total cells, log2 of the triangle count
    uint64_t h = 1469598103934665603ull, tris = 0;
    std::vector<uint64_t> sz;
    for (const Frame& f : j.frames) sz.push_back(((uint64_t)f.W << 32) | f.H);
    std::sort(sz.begin(), sz.end());
    for (uint64_t v : sz) h = (h ^ v) * 1099511628211ull;
    for (const Draw& d : j.draws) tris += d.n;
    uint32_t lg = 0;
    while ((1ull << lg) < tris + 1) ++lg;
    return std::make_tuple(h, j.n_cells, lg);
}

extern "C" {

int sloth_batch_autotune(sloth_batch* b, uint32_t* tile_w, uint32_t* tile_h)
{
    ARG_CHECK(b, "null batch");
    int rc = ensure_device();
    if (rc) return rc;
    Job& j = b->job;
    if (j.path != PATH_TILED) {
        if (tile_w) *tile_w = 0;
        if (tile_h) *tile_h = 0;
        return SLOTH_OK;   // the general path has no tiles
    }
    auto apply = [&](const TuneChoice& c) -> int {
        int r = sloth_batch_set_tile(b, c.tw, c.th);
        if (r) return r;
        j.tiled.set_walk(c.walk);
        j.tiled.set_bin_tpt(c.bin_tpt);
        if (tile_w) *tile_w = c.tw;
        if (tile_h) *tile_h = c.th;
        return SLOTH_OK;
    };
    const auto key = tune_class(j);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) return apply(it->second);
    }
    // Tiles of 8..256 logical pixels wide (a logical pixel is two cells
    // wide) by 8..64 rows: 16..512 cells by 8..64 rows, those the raster's
    // workgroup can take.  Each candidate is timed as the
    // least of three runs of 3 graph renders, after one untimed render (which
    // also sizes the lists): single runs differ by a few per cent, as much
    // as neighbouring candidates do.
    auto time_it = [&](float& ms) -> int {
        int r = sloth_batch_time(b, 3, 1, 0, &ms);
        for (int k = 0; k < 2 && r == SLOTH_OK; ++k) {
            float m2 = 0.0f;
            r = sloth_batch_time(b, 3, 0, 0, &m2);
            ms = std::min(ms, m2);
        }
        return r;
    };
    TuneChoice best{j.tiled.TW, j.tiled.TH, -1, 4};
    float best_ms = 1e30f;
    static const uint32_t widths[] = {16, 32, 64, 128, 256, 512}, heights[] = {8, 16, 32, 64};
    j.tiled.set_walk(-1);
    j.tiled.set_bin_tpt(4);
    for (uint32_t tw : widths)
        for (uint32_t th : heights) {
            if ((rc = sloth_batch_set_tile(b, tw, th))) return rc;
            if (!j.tiled.tile_fits()) continue;   // too many slots per tile or per row
            float ms = 0.0f;
            if ((rc = time_it(ms))) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best = TuneChoice{tw, th, -1, 4};
            }
        }
    // then the raster variant: the sizing run's choice against the other
    if ((rc = sloth_batch_set_tile(b, best.tw, best.th))) return rc;
    {
        float ms = 0.0f;
        if ((rc = time_it(ms))) return rc;   // sizes the lists; the plan's own choice
        best_ms = std::min(ms, best_ms);
        const int other = j.tiled.big_walk ? 0 : 1;
        j.tiled.set_walk(other);
        float ms2 = 0.0f;
        if ((rc = time_it(ms2))) return rc;
        if (ms2 < best_ms) {
            best.walk = other;
            best_ms = ms2;
        }
        j.tiled.set_walk(best.walk);
    }
    // then the bin's triangles per thread
    for (int tpt : {2, 1}) {
        j.tiled.set_bin_tpt(tpt);
        float ms = 0.0f;
        if ((rc = time_it(ms))) return rc;
        if (ms < best_ms) {
            best.bin_tpt = tpt;
            best_ms = ms;
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = best;
    }
    return apply(best);
}

int sloth_batch_plan(const sloth_batch* b, uint32_t* tile_w, uint32_t* tile_h, int* big_walk,
                     uint32_t* bin_tris_per_thread)
{
    ARG_CHECK(b, "null batch");
    const Job& j = b->job;
    const bool tiled = j.path == PATH_TILED;
    if (tile_w) *tile_w = tiled ? j.tiled.TW : 0;
    if (tile_h) *tile_h = tiled ? j.tiled.TH : 0;
    if (big_walk) *big_walk = tiled ? (j.tiled.walk() ? 1 : 0) : -1;
    if (bin_tris_per_thread) *bin_tris_per_thread = tiled ? (uint32_t)j.tiled.bin_tpt : 0;
    return SLOTH_OK;
}

int sloth_batch_destroy(sloth_batch* b)
{
    if (b && b->stream) (void)hipStreamSynchronize(b->stream);
    delete b;
    return SLOTH_OK;
}


// ---- differentiable shading -------------------------------------------------

}  // extern "C"

struct sloth_shading {
    sloth_batch* b = nullptr;
    uint32_t n_tris = 0, n_verts = 0;
    uint64_t n_slots = 0;
    bool has_target = false;
    bool has_params = false;
    DevBuf d_faces, d_vis, d_G, d_vcolor, d_vnormal, d_light, d_target, d_out, d_partial, d_vert_acc;
    DevBuf d_bwd_tile;            // backward tiles per frame (prefix, n_frames + 1)
    uint64_t n_bwd_tiles = 0;
    DevBuf d_dcolor, d_dnormal;
    std::vector<double> partial;
    uint32_t n_partials = 0;   // per-block partial sums the last pass wrote
};

static ShadeArgs shade_args(sloth_shading* sh, bool with_target, float* out)
{
    ShadeArgs a;
    Job& j = sh->b->job;
    a.frames = j.d_frames.as<Frame>();
    a.n_frames = (uint32_t)j.frames.size();
    a.draws = j.d_draws.as<Draw>();
    a.vis = sh->d_vis.as<int32_t>();
    a.faces = sh->d_faces.as<uint32_t>();
    a.vcolor = sh->d_vcolor.as<float>();
    a.vnormal = sh->d_vnormal.as<float>();
    a.light = sh->d_light.as<float>();
    a.G = sh->d_G.as<float>();
    a.target = with_target ? sh->d_target.as<float>() : nullptr;
    a.out = out;
    a.partial = sh->d_partial.as<double>();
    a.vert_acc = sh->d_vert_acc.as<double>();
    a.n_slots = sh->n_slots;
    a.n_tris = sh->n_tris;
    a.n_verts = sh->n_verts;
    a.bwd_tile_begin = sh->d_bwd_tile.as<uint64_t>();
    a.n_bwd_tiles = sh->n_bwd_tiles;
    return a;
}

// Normal map of a draw: cof(M3) / |det M3|^(2/3) (M3 = upper-left 3x3 of
// utransform * transform), row-major, computed in double.
static void normal_map(const float* M, float* G)
{
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
    double k = std::pow(std::fabs(det), 2.0 / 3.0);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) G[r * 3 + c] = (float)(cof[r][c] / k);
}

// Renders the batch keeping (triangle << 1 | sub) per z slot.
static int shading_visibility(sloth_shading* sh)
{
    sloth_batch* b = sh->b;
    Job& j = b->job;
    int rc;
    if ((rc = j.prepare(b->stream))) return rc;
    if ((rc = j.launch(0, (uint32_t)j.frames.size(), b->stream, sh->d_vis.as<int32_t>(), nullptr, true))) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    bool replanned = false;
    if ((rc = j.finish(0, (uint32_t)j.frames.size(), b->stream, sh->d_vis.as<int32_t>(), &replanned, true)))
        return rc;
    if (replanned) drop_graph(b);
    b->ran = true;
    b->checked = true;
    return SLOTH_OK;
}

extern "C" {

int sloth_shading_create(sloth_batch* b, const uint32_t* faces, uint32_t n_tris, uint32_t n_verts,
                         sloth_shading** out)
{
    ARG_CHECK(out, "null output");
    *out = nullptr;
    ARG_CHECK(b && (faces || n_tris == 0), "null argument");
    Job& j = b->job;
    for (const Draw& d : j.draws) ARG_CHECK(d.n == n_tris, "every draw of the batch must use a mesh of n_tris triangles");
    for (uint64_t i = 0; i < (uint64_t)n_tris * 3; ++i) ARG_CHECK(faces[i] < n_verts, "face vertex index out of range");
    int rc = ensure_device();
    if (rc) return rc;
    sloth_shading* sh = new (std::nothrow) sloth_shading();
    if (!sh) return fail(SLOTH_ENOMEM, "host allocation failed");
    sh->b = b;
    sh->n_tris = n_tris;
    sh->n_verts = n_verts;
    sh->n_slots = j.n_z;
    std::vector<float> G(j.draws.size() * 9);
    for (size_t d = 0; d < j.draws.size(); ++d) normal_map(j.draws[d].M, &G[d * 9]);
    const uint32_t nb = shade_blocks(sh->n_slots);
    sh->partial.resize((size_t)nb * 4);
    std::vector<uint64_t> tile_begin(j.frames.size() + 1);
    bwd_tiles(j.frames.data(), (uint32_t)j.frames.size(), tile_begin.data());
    sh->n_bwd_tiles = tile_begin.back();
    if ((rc = sh->d_bwd_tile.reserve(tile_begin.size() * 8)) ||
        (rc = sh->d_faces.reserve(std::max<size_t>(4, (size_t)n_tris * 12))) ||
        (rc = sh->d_vis.reserve(std::max<size_t>(4, sh->n_slots * 4))) ||
        (rc = sh->d_G.reserve(std::max<size_t>(4, G.size() * 4))) ||
        (rc = sh->d_vcolor.reserve(std::max<size_t>(4, (size_t)n_verts * 12))) ||
        (rc = sh->d_vnormal.reserve(std::max<size_t>(4, (size_t)n_verts * 12))) ||
        (rc = sh->d_light.reserve(16)) ||
        (rc = sh->d_partial.reserve(std::max<size_t>(8, (size_t)nb * 32))) ||
        (rc = sh->d_vert_acc.reserve(std::max<size_t>(8, (size_t)n_verts * 6 * 8))) ||
        (rc = sh->d_dcolor.reserve(std::max<size_t>(4, (size_t)n_verts * 12))) ||
        (rc = sh->d_dnormal.reserve(std::max<size_t>(4, (size_t)n_verts * 12)))) {
        delete sh;
        return rc;
    }
    hipError_t e = hipSuccess;
    if (n_tris) e = hipMemcpy(sh->d_faces.p, faces, (size_t)n_tris * 12, hipMemcpyHostToDevice);
    if (e == hipSuccess && !G.empty()) e = hipMemcpy(sh->d_G.p, G.data(), G.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(sh->d_bwd_tile.p, tile_begin.data(), tile_begin.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        delete sh;
        return fail(SLOTH_EHIP, "shading upload: %s", hipGetErrorString(e));
    }
    if ((rc = shading_visibility(sh))) {
        delete sh;
        return rc;
    }
    *out = sh;
    return SLOTH_OK;
}

int sloth_shading_visibility(sloth_shading* sh)
{
    ARG_CHECK(sh, "null shading");
    Job& j = sh->b->job;
    std::vector<float> G(j.draws.size() * 9);
    for (size_t d = 0; d < j.draws.size(); ++d) normal_map(j.draws[d].M, &G[d * 9]);
    HIP_TRY(hipStreamSynchronize(sh->b->stream));   // no shading pass still reads the old maps
    if (!G.empty()) HIP_TRY(hipMemcpy(sh->d_G.p, G.data(), G.size() * 4, hipMemcpyHostToDevice));
    return shading_visibility(sh);
}

int sloth_shading_slots(const sloth_shading* sh, uint64_t* n_slots)
{
    ARG_CHECK(sh && n_slots, "null argument");
    *n_slots = sh->n_slots;
    return SLOTH_OK;
}

int sloth_shading_set_target(sloth_shading* sh, const float* target)
{
    ARG_CHECK(sh, "null shading");
    if (!target) {
        sh->has_target = false;
        return SLOTH_OK;
    }
    int rc = sh->d_target.reserve(std::max<size_t>(4, sh->n_slots * 12));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(sh->d_target.p, target, sh->n_slots * 12, hipMemcpyHostToDevice));
    sh->has_target = true;
    return SLOTH_OK;
}

static int shading_reduce(sloth_shading* sh, double* loss, double* dlight)
{
    const uint32_t nb = sh->n_partials;
    if (nb) HIP_TRY(hipMemcpy(sh->partial.data(), sh->d_partial.p, (size_t)nb * 32, hipMemcpyDeviceToHost));
    double acc[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < nb; ++i)
        for (int k = 0; k < 4; ++k) acc[k] += sh->partial[(size_t)i * 4 + k];
    if (loss) *loss = acc[0];
    if (dlight)
        for (int k = 0; k < 3; ++k) dlight[k] = acc[k + 1];
    return SLOTH_OK;
}

static int shading_params(sloth_shading* sh, const float* vcolor, const float* vnormal, const float* light)
{
    hipStream_t s = sh->b->stream;
    if (sh->n_verts) {
        HIP_TRY(hipMemcpyAsync(sh->d_vcolor.p, vcolor, (size_t)sh->n_verts * 12, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(sh->d_vnormal.p, vnormal, (size_t)sh->n_verts * 12, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipMemcpyAsync(sh->d_light.p, light, 12, hipMemcpyHostToDevice, s));
    sh->has_params = true;
    return SLOTH_OK;
}

// Backward launches (accumulator clear, per-tile backward with the loss,
// conversion of the vertex sums) on the shading's stream.
static int shading_backward_launch(sloth_shading* sh)
{
    hipStream_t s = sh->b->stream;
    const ShadeArgs a = shade_args(sh, true, nullptr);
    int rc = launch_shade(a, true, s);
    if (rc) return rc;
    sh->n_partials = bwd_blocks(sh->n_slots, sh->n_bwd_tiles);
    return launch_shade_finish(a, sh->d_dcolor.as<float>(), sh->d_dnormal.as<float>(), s);
}

// Copies the gradients out and sums the per-block loss and light partials.
static int shading_gradients(sloth_shading* sh, double* loss, float* d_vcolor, float* d_vnormal, float* d_light)
{
    hipStream_t s = sh->b->stream;
    if (d_vcolor && sh->n_verts)
        HIP_TRY(hipMemcpyAsync(d_vcolor, sh->d_dcolor.p, (size_t)sh->n_verts * 12, hipMemcpyDeviceToHost, s));
    if (d_vnormal && sh->n_verts)
        HIP_TRY(hipMemcpyAsync(d_vnormal, sh->d_dnormal.p, (size_t)sh->n_verts * 12, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double dl[3];
    int rc = shading_reduce(sh, loss, dl);
    if (rc) return rc;
    if (d_light)
        for (int k = 0; k < 3; ++k) d_light[k] = (float)dl[k];
    return SLOTH_OK;
}

int sloth_shading_forward(sloth_shading* sh, const float* vcolor, const float* vnormal, const float* light,
                          float* out_rgb, double* loss)
{
    ARG_CHECK(sh && vcolor && vnormal && light, "null argument");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = sh->b->stream;
    if ((rc = shading_params(sh, vcolor, vnormal, light))) return rc;
    float* d_out = nullptr;
    if (out_rgb) {
        if ((rc = sh->d_out.reserve(std::max<size_t>(4, sh->n_slots * 12)))) return rc;
        d_out = sh->d_out.as<float>();
        HIP_TRY(hipMemsetAsync(sh->d_out.p, 0, sh->n_slots * 12, s));
    }
    if ((rc = launch_shade(shade_args(sh, sh->has_target, d_out), false, s))) return rc;
    sh->n_partials = shade_blocks(sh->n_slots);
    if (out_rgb) HIP_TRY(hipMemcpyAsync(out_rgb, sh->d_out.p, sh->n_slots * 12, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return shading_reduce(sh, loss, nullptr);
}

int sloth_shading_backward(sloth_shading* sh, float* d_vcolor, float* d_vnormal, float* d_light)
{
    ARG_CHECK(sh, "null shading");
    if (!sh->has_params) return fail(SLOTH_ESTATE, "backward needs a forward call first");
    if (!sh->has_target) return fail(SLOTH_ESTATE, "backward needs a target (sloth_shading_set_target)");
    int rc = ensure_device();
    if (rc) return rc;
    if ((rc = shading_backward_launch(sh))) return rc;
    return shading_gradients(sh, nullptr, d_vcolor, d_vnormal, d_light);
}

int sloth_shading_step(sloth_shading* sh, const float* vcolor, const float* vnormal, const float* light,
                       double* loss, float* d_vcolor, float* d_vnormal, float* d_light)
{
    ARG_CHECK(sh && vcolor && vnormal && light, "null argument");
    if (!sh->has_target) return fail(SLOTH_ESTATE, "a step needs a target (sloth_shading_set_target)");
    int rc = ensure_device();
    if (rc) return rc;
    if ((rc = shading_params(sh, vcolor, vnormal, light))) return rc;
    if ((rc = shading_backward_launch(sh))) return rc;
    return shading_gradients(sh, loss, d_vcolor, d_vnormal, d_light);
}

int sloth_shading_time(sloth_shading* sh, uint32_t steps, uint32_t warmup, float* ms)
{
    ARG_CHECK(sh && ms, "null argument");
    if (!sh->has_params || !sh->has_target) return fail(SLOTH_ESTATE, "timing needs parameters and a target");
    int rc = ensure_device();
    if (rc) return rc;
    hipStream_t s = sh->b->stream;
    auto step = [&]() -> int { return shading_backward_launch(sh); };   // loss and gradients in one pass
    for (uint32_t i = 0; i < warmup; ++i)
        if ((rc = step())) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s));
    for (uint32_t i = 0; i < steps && rc == SLOTH_OK; ++i) rc = step();
    hipError_t e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SLOTH_EHIP, "shading timing: %s", hipGetErrorString(e));
    *ms = t;
    return SLOTH_OK;
}

int sloth_shading_destroy(sloth_shading* sh)
{
    if (sh && sh->b && sh->b->stream) (void)hipStreamSynchronize(sh->b->stream);
    delete sh;
    return SLOTH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// rendering without allocation
// ---------------------------------------------------------------------------
// sloth_render_frame_noalloc renders one single-draw frame on the general
// (global-key) path, which needs device memory beyond its inputs and
// outputs: the frame, draw and shader tables the kernels read, one 64-bit
// key per z slot and one newline entry per row.  Here that memory is a
// file-scope device array, one copy per device, held until the process
// ends, so the entry point calls no allocator.  Host code never passes the
// array's name to a kernel or a copy: it uses the device address that
// hipGetSymbolAddress gives for the current device.
namespace {
constexpr uint64_t NOALLOC_TABLE_BYTES = 1024;                   // Frame at 0, Draw at 256, Ramp at 512
constexpr uint64_t NOALLOC_MAX_SLOTS = 1ull << 22;               // W*H <= 4,194,304 (2560 x 1440 fits)
constexpr uint64_t NOALLOC_KEY_BYTES = 8 * NOALLOC_MAX_SLOTS;    // 32 MiB
constexpr uint64_t NOALLOC_ROW_BYTES = 4 * 65536;                // H <= 65535
constexpr uint64_t NOALLOC_BYTES = NOALLOC_TABLE_BYTES + NOALLOC_KEY_BYTES + NOALLOC_ROW_BYTES;
constexpr int NOALLOC_MAX_DEVICES = 64;
static_assert(sizeof(Frame) <= 256 && sizeof(Draw) <= 256 && sizeof(Ramp) <= 256, "table slots");
}  // namespace

static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];

// Device address of `scratch` on the current device `dev`, looked up once.
static int noalloc_scratch(int dev, uint8_t** out)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(SLOTH_ENODEV, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(SLOTH_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = p;
    }
    *out = (uint8_t*)g_scratch_ptr[dev];
    return SLOTH_OK;
}

// Sets up the scratch address of the device sloth_init selects.
int sloth::noalloc_init(int dev)
{
    uint8_t* p = nullptr;
    return noalloc_scratch(dev, &p);
}

// Writes the launch's tables into scratch from kernel arguments (no host
// buffer has to outlive the call).
__global__ void k_noalloc_tables(Frame f, Draw d, Ramp r, uint8_t* base)
{
    if (threadIdx.x == 0) *(Frame*)base = f;
    if (threadIdx.x == 1) *(Draw*)(base + 256) = d;
    if (threadIdx.x == 2) *(Ramp*)(base + 512) = r;
}

extern "C" {

int sloth_render_frame_noalloc(const float* d_positions, const uint8_t* d_colors, uint32_t n, uint32_t width,
                               uint32_t height, int image, const float* utransform16, const float* transform16,
                               const sloth_ramp* shader, uint32_t* d_cells, float* d_z, void* stream)
{
    ARG_CHECK(n == 0 || (d_positions && d_colors), "null positions or colours");
    ARG_CHECK(utransform16 && transform16 && d_cells && d_z, "null argument");
    ARG_CHECK(width >= 1 && width <= 65535 && height >= 1 && height <= 65535,
              "width and height must be in 1..65535");
    ARG_CHECK(n < (1u << 31) - 1, "a frame holds fewer than 2^31 - 1 triangles");
    const uint64_t WH = (uint64_t)width * height;
    if (WH > NOALLOC_MAX_SLOTS)
        return fail(SLOTH_EINVAL,
                    "sloth_render_frame_noalloc: %ux%u has %llu z slots; its fixed scratch holds at most %llu "
                    "(use a context or a batch for larger frames)",
                    width, height, (unsigned long long)WH, (unsigned long long)NOALLOC_MAX_SLOTS);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    uint8_t* base = nullptr;
    int rc = noalloc_scratch(dev, &base);
    if (rc) return rc;
    Ramp r;
    if ((rc = to_ramp(shader, r))) return rc;

    Frame f;
    memset(&f, 0, sizeof f);
    f.W = width;
    f.H = height;
    f.image = image ? 1u : 0u;
    f.draw_begin = 0;
    f.draw_end = 1;
    f.n_tris = n;
    Draw d;
    memset(&d, 0, sizeof d);
    d.xyz = d_positions;
    d.rgb = d_colors;
    d.n = n;
    d.ramp = 0;
    mat4_mul(utransform16, transform16, d.M);

    hipStream_t s = (hipStream_t)stream;
    const Frame* d_frame = (const Frame*)base;
    const Draw* d_draw = (const Draw*)(base + 256);
    const Ramp* d_ramp = (const Ramp*)(base + 512);
    uint64_t* keys = (uint64_t*)(base + NOALLOC_TABLE_BYTES);
    int32_t* nl = (int32_t*)(base + NOALLOC_TABLE_BYTES + NOALLOC_KEY_BYTES);
    const uint64_t cells = frame_cells(f);

    hipLaunchKernelGGL(k_noalloc_tables, dim3(1), dim3(64), 0, s, f, d, r, base);
    HIP_TRY(hipGetLastError());
    {
        // every key and newline entry of this frame starts empty: the scratch
        // holds whatever the previous call on this device left
        const unsigned g = std::max(1u, std::min<unsigned>(Job::blocks_for(std::max<uint64_t>(WH, height)), 8192u));
        hipLaunchKernelGGL(k_clear_global, dim3(g), dim3(256), 0, s, keys, WH, nl, (uint64_t)height);
        HIP_TRY(hipGetLastError());
    }
    if (n) {
        hipLaunchKernelGGL(k_raster_global, dim3(Job::blocks_for(n), 1), dim3(256), 0, s, d_draw, 0u, d_frame, keys,
                           nl);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_winner_global, dim3(Job::blocks_for(WH), 1), dim3(256), 0, s, d_draw, d_ramp, d_frame, 0u,
                       keys, d_z, (int32_t*)nullptr, 0u);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_resolve_global, dim3(Job::blocks_for(cells), 1), dim3(256), 0, s, d_frame, 0u,
                       (const uint64_t*)keys, (const int32_t*)nl, d_cells);
    HIP_TRY(hipGetLastError());
    return SLOTH_OK;
}

}  // extern "C"
