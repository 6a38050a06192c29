// Drives the batch API through its host paths (flat k_bin and raster tables,
// long-tiles order, re-plan after new transforms, per-frame launches,
// winners, encode, autotune) for a host AddressSanitizer build.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include "sloth_hip.h"

#define CK(x) do { int rc_ = (x); if (rc_) { printf("FAIL %s -> %d (%s)\n", #x, rc_, sloth_last_error()); return 1; } } while (0)

static void rot(float a, float* m) {   // column-major rotation about y
    std::memset(m, 0, 64);
    m[0] = std::cos(a); m[2] = -std::sin(a); m[5] = 1; m[8] = std::sin(a); m[10] = std::cos(a); m[15] = 1;
}
static void screen(uint32_t W, uint32_t H, float s, float* u) {
    std::memset(u, 0, 64);
    u[0] = s * W / 4.0f; u[5] = -s * H / 2.0f; u[10] = 1; u[12] = W / 4.0f; u[13] = H / 2.0f; u[15] = 1;
}

int main() {
    CK(sloth_init(0));
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(-1, 1);
    std::vector<sloth_mesh*> meshes;
    const uint32_t sizes[4] = {64, 4000, 90000, 0};
    for (uint32_t n : sizes) {
        std::vector<float> p((size_t)n * 9);
        std::vector<uint8_t> c((size_t)n * 3, 200);
        for (uint32_t t = 0; t < n; ++t) {
            float cx = U(rng), cy = U(rng), cz = U(rng), s = 0.02f + 0.05f * (U(rng) + 1);
            for (int k = 0; k < 3; ++k) { p[t * 9 + k * 3] = cx + s * U(rng); p[t * 9 + k * 3 + 1] = cy + s * U(rng); p[t * 9 + k * 3 + 2] = cz + s * U(rng); }
        }
        sloth_mesh* m = nullptr;
        CK(sloth_mesh_create(n ? p.data() : nullptr, n ? c.data() : nullptr, n, &m));
        meshes.push_back(m);
    }
    const uint32_t FW[6] = {80, 641, 33, 1280, 2, 200}, FH[6] = {24, 480, 7, 720, 3, 64};
    std::vector<sloth_frame_desc> fr(6);
    for (int f = 0; f < 6; ++f) { fr[f].width = FW[f]; fr[f].height = FH[f]; fr[f].image = f % 2 == 0; screen(FW[f], FH[f], 1.0f, fr[f].utransform); }
    std::vector<sloth_draw_desc> dr;
    for (int f = 0; f < 6; ++f)
        for (int k = 0; k < 1 + f % 3; ++k) { sloth_draw_desc d; d.frame = f; d.mesh = meshes[(f + k) % 4]; rot(0.3f * f + k, d.transform); dr.push_back(d); }
    sloth_batch* b = nullptr;
    CK(sloth_batch_create(fr.data(), 6, dr.data(), (uint32_t)dr.size(), nullptr, &b));
    for (uint32_t tpt : {1u, 2u, 4u}) {
        CK(sloth_batch_set_bin(b, tpt));
        for (int walk : {0, 1}) {
            CK(sloth_batch_set_walk(b, walk));
            CK(sloth_batch_run(b, nullptr));
            CK(sloth_batch_sync(b));
        }
    }
    std::vector<uint32_t> cells; std::vector<float> z; std::vector<int32_t> win;
    for (int f = 0; f < 6; ++f) {
        const size_t nc = (size_t)FW[f] * FH[f] + FH[f], nz = (size_t)FW[f] * FH[f];
        cells.assign(nc, 0); z.assign(nz, 0); win.assign(nz, 0);
        CK(sloth_batch_read_frame(b, f, cells.data(), z.data(), f == 5 ? win.data() : nullptr));
        uint64_t len = 0; std::vector<char> out(nc * 48 + 64);
        CK(sloth_batch_encode_frame(b, f, 0, out.data(), out.size(), &len));
    }
    // zoomed transforms: longer tile lists, re-plan on the next sync
    std::vector<float> T(dr.size() * 16);
    for (size_t d = 0; d < dr.size(); ++d) { rot(1.1f + d, &T[d * 16]); for (int k = 0; k < 12; ++k) T[d * 16 + k] *= 2.5f; }
    CK(sloth_batch_set_transforms(b, T.data()));
    CK(sloth_batch_run(b, nullptr));
    CK(sloth_batch_sync(b));
    CK(sloth_batch_run_per_frame(b, nullptr));
    CK(sloth_batch_run(b, nullptr));
    CK(sloth_batch_sync(b));
    for (int f = 0; f < 6; ++f) {
        const size_t nz = (size_t)FW[f] * FH[f];
        cells.assign(nz + FH[f], 0); z.assign(nz, 0); win.assign(nz, 0);
        CK(sloth_batch_read_frame(b, f, cells.data(), z.data(), win.data()));
    }
    uint32_t tw = 0, th = 0;
    CK(sloth_batch_autotune(b, &tw, &th));
    CK(sloth_batch_run(b, nullptr));
    CK(sloth_batch_sync(b));
    float ms[16]; const char* names[16]; uint32_t ns = 16;
    CK(sloth_batch_profile(b, ms, names, &ns));
    CK(sloth_batch_destroy(b));
    for (auto* m : meshes) CK(sloth_mesh_destroy(m));
    printf("asan batch driver ok: tile %ux%u, %u stages\n", tw, th, ns);
    return 0;
}
