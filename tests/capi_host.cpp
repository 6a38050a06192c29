// A C++ host of libsloth_hip.so that uses nothing but include/sloth_hip.h: the
// calls a Rust binding of the reference would make (main.rs:76-90 for one
// frame: update from the meshes' bounding boxes, clear, draw_mesh per mesh,
// then the frame buffer and the flush bytes).  The GPU tests feed it a spec
// file and compare what it writes with the oracle.
//
//   capi_host <spec> <out>
//
// spec (little-endian): u32 kind (0 = triangle soup, 1 = model file,
//   2 = differentiable shading of a batch of views),
//   u32 W, u32 H, u32 image, u32 encode_mode, f32 transform[16] (column-major),
//   kind 0: f32 utransform[16], u32 n, f32 positions[n][3][3], u8 colors[n][3]
//   kind 1: u32 path_len, char path[path_len]   (utransform from
//           sloth_context_update over the file's meshes)
//   kind 2 (shading): f32 utransform[16], u32 n_views, f32 view transforms
//           [n_views][16], u32 n, f32 positions[n][3][3], u8 colors[n][3],
//           u32 n_verts, u32 faces[n][3], f32 vcolor[n_verts][3],
//           f32 vnormal[n_verts][3], f32 light[3], f32 target_vcolor[n_verts][3],
//           f32 target_light[3]  (the transform in the header is unused)
// out (kinds 0, 1): u32 status, u64 n_cells, u32 cells[n_cells], f32 z[W*H],
//   i32 winners[W*H], f32 utransform[16], u64 n_bytes, u8 encoded[n_bytes]
// out (kind 2): u32 status, u64 n_slots, i32 winners[n_slots] (per view),
//   f32 target[n_slots][3], f64 loss, f32 d_vcolor[n_verts][3],
//   f32 d_vnormal[n_verts][3], f32 d_light[3] (sloth_shading_forward +
//   sloth_shading_backward), then f64 loss and the three gradients again
//   from sloth_shading_step
// On an error the status is the library's code and sloth_last_error() goes
// to stderr; nothing else is written.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sloth_hip.h"

namespace {

struct Reader {
    FILE* f;
    template <class T>
    bool get(T* p, size_t n = 1) { return fread(p, sizeof(T), n, f) == n; }
};

int fail(FILE* out, int rc, const char* what)
{
    fprintf(stderr, "capi_host: %s: %s (status %d)\n", what, sloth_last_error(), rc);
    const uint32_t st = (uint32_t)rc;
    fwrite(&st, 4, 1, out);
    return 1;
}

// The forward / backward pair and the one-pass step of the differentiable
// shading over a batch of views of one mesh, as a binding would drive them.
int shading(Reader& in, FILE* out, uint32_t W, uint32_t H, uint32_t image)
{
    float U[16];
    uint32_t V, n, nv;
    if (!in.get(U, 16) || !in.get(&V)) return fail(out, SLOTH_EINVAL, "spec");
    std::vector<float> T((size_t)V * 16);
    if (!in.get(T.data(), T.size()) || !in.get(&n)) return fail(out, SLOTH_EINVAL, "spec");
    std::vector<float> pos((size_t)n * 9);
    std::vector<uint8_t> rgb((size_t)n * 3);
    std::vector<uint32_t> faces((size_t)n * 3);
    if (!in.get(pos.data(), pos.size()) || !in.get(rgb.data(), rgb.size()) || !in.get(&nv) ||
        !in.get(faces.data(), faces.size()))
        return fail(out, SLOTH_EINVAL, "spec");
    std::vector<float> vc((size_t)nv * 3), vn((size_t)nv * 3), vct((size_t)nv * 3);
    float L[3], Lt[3];
    if (!in.get(vc.data(), vc.size()) || !in.get(vn.data(), vn.size()) || !in.get(L, 3) ||
        !in.get(vct.data(), vct.size()) || !in.get(Lt, 3))
        return fail(out, SLOTH_EINVAL, "spec");

    int rc;
    sloth_mesh* mesh = nullptr;
    if ((rc = sloth_mesh_create(pos.data(), rgb.data(), n, &mesh))) return fail(out, rc, "sloth_mesh_create");
    std::vector<sloth_frame_desc> frames(V);
    std::vector<sloth_draw_desc> draws(V);
    for (uint32_t v = 0; v < V; ++v) {
        frames[v].width = W;
        frames[v].height = H;
        frames[v].image = (int32_t)image;
        memcpy(frames[v].utransform, U, sizeof U);
        draws[v].frame = v;
        draws[v].mesh = mesh;
        memcpy(draws[v].transform, &T[(size_t)v * 16], 64);
    }
    sloth_batch* b = nullptr;
    if ((rc = sloth_batch_create(frames.data(), V, draws.data(), V, nullptr, &b))) return fail(out, rc, "sloth_batch_create");
    sloth_shading* sh = nullptr;
    if ((rc = sloth_shading_create(b, faces.data(), n, nv, &sh))) return fail(out, rc, "sloth_shading_create");
    uint64_t slots = 0;
    if ((rc = sloth_shading_slots(sh, &slots))) return fail(out, rc, "sloth_shading_slots");
    std::vector<int32_t> win(slots);
    for (uint32_t v = 0; v < V; ++v)
        if ((rc = sloth_batch_read_frame(b, v, nullptr, nullptr, win.data() + (size_t)v * W * H)))
            return fail(out, rc, "sloth_batch_read_frame");
    // the target: the forward image at the target's parameters
    std::vector<float> target(slots * 3);
    if ((rc = sloth_shading_forward(sh, vct.data(), vn.data(), Lt, target.data(), nullptr)))
        return fail(out, rc, "sloth_shading_forward (target)");
    if ((rc = sloth_shading_set_target(sh, target.data()))) return fail(out, rc, "sloth_shading_set_target");
    double loss = 0.0, loss_s = 0.0;
    std::vector<float> dc(vc.size()), dn(vn.size()), dc_s(vc.size()), dn_s(vn.size());
    float dl[3], dl_s[3];
    if ((rc = sloth_shading_forward(sh, vc.data(), vn.data(), L, nullptr, &loss))) return fail(out, rc, "sloth_shading_forward");
    if ((rc = sloth_shading_backward(sh, dc.data(), dn.data(), dl))) return fail(out, rc, "sloth_shading_backward");
    if ((rc = sloth_shading_step(sh, vc.data(), vn.data(), L, &loss_s, dc_s.data(), dn_s.data(), dl_s)))
        return fail(out, rc, "sloth_shading_step");

    const uint32_t ok = 0;
    fwrite(&ok, 4, 1, out);
    fwrite(&slots, 8, 1, out);
    fwrite(win.data(), 4, slots, out);
    fwrite(target.data(), 4, target.size(), out);
    fwrite(&loss, 8, 1, out);
    fwrite(dc.data(), 4, dc.size(), out);
    fwrite(dn.data(), 4, dn.size(), out);
    fwrite(dl, 4, 3, out);
    fwrite(&loss_s, 8, 1, out);
    fwrite(dc_s.data(), 4, dc_s.size(), out);
    fwrite(dn_s.data(), 4, dn_s.size(), out);
    fwrite(dl_s, 4, 3, out);
    sloth_shading_destroy(sh);
    sloth_batch_destroy(b);
    sloth_mesh_destroy(mesh);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: capi_host <spec> <out>\n");
        return 2;
    }
    FILE* fs = fopen(argv[1], "rb");
    FILE* out = fopen(argv[2], "wb");
    if (!fs || !out) {
        fprintf(stderr, "capi_host: cannot open the spec or the output file\n");
        return 2;
    }
    Reader in{fs};
    uint32_t kind, W, H, image, mode;
    float T[16];
    if (!in.get(&kind) || !in.get(&W) || !in.get(&H) || !in.get(&image) || !in.get(&mode) || !in.get(T, 16)) {
        fprintf(stderr, "capi_host: short spec\n");
        return 2;
    }

    int rc = sloth_init(0);
    if (rc) return fail(out, rc, "sloth_init");
    if (kind == 2) {
        rc = shading(in, out, W, H, image);
        fclose(fs);
        fclose(out);
        return rc;
    }
    sloth_context* ctx = nullptr;
    if ((rc = sloth_context_create(W, H, (int)image, &ctx))) return fail(out, rc, "sloth_context_create");

    std::vector<sloth_mesh*> meshes;
    if (kind == 0) {
        float U[16];
        uint32_t n;
        if (!in.get(U, 16) || !in.get(&n)) return fail(out, SLOTH_EINVAL, "spec");
        std::vector<float> pos((size_t)n * 9);
        std::vector<uint8_t> rgb((size_t)n * 3);
        if (n && (!in.get(pos.data(), pos.size()) || !in.get(rgb.data(), rgb.size())))
            return fail(out, SLOTH_EINVAL, "spec");
        sloth_mesh* m = nullptr;
        if ((rc = sloth_mesh_create(pos.data(), rgb.data(), n, &m))) return fail(out, rc, "sloth_mesh_create");
        meshes.push_back(m);
        if ((rc = sloth_context_set_utransform(ctx, U))) return fail(out, rc, "sloth_context_set_utransform");
    } else {
        uint32_t len;
        if (!in.get(&len)) return fail(out, SLOTH_EINVAL, "spec");
        std::string path(len, '\0');
        if (len && !in.get(&path[0], len)) return fail(out, SLOTH_EINVAL, "spec");
        sloth_model* model = nullptr;
        if ((rc = sloth_model_load(path.c_str(), &model))) return fail(out, rc, "sloth_model_load");
        uint32_t n_meshes = 0;
        if ((rc = sloth_model_mesh_count(model, &n_meshes))) return fail(out, rc, "sloth_model_mesh_count");
        std::vector<float> bbox_max;
        for (uint32_t i = 0; i < n_meshes; ++i) {
            sloth_mesh* m = nullptr;
            if ((rc = sloth_model_upload(model, i, &m))) return fail(out, rc, "sloth_model_upload");
            float lo[3], hi[3];
            if ((rc = sloth_mesh_bbox(m, lo, hi))) return fail(out, rc, "sloth_mesh_bbox");
            bbox_max.insert(bbox_max.end(), hi, hi + 3);
            meshes.push_back(m);
        }
        sloth_model_destroy(model);
        // Context::update (context.rs:93-141) from the meshes' bounding boxes
        if ((rc = sloth_context_update(ctx, bbox_max.data(), n_meshes))) return fail(out, rc, "sloth_context_update");
    }
    fclose(fs);

    if ((rc = sloth_context_clear(ctx))) return fail(out, rc, "sloth_context_clear");
    for (sloth_mesh* m : meshes)
        if ((rc = sloth_draw_mesh(ctx, m, T, nullptr))) return fail(out, rc, "sloth_draw_mesh");

    const uint64_t n_cells = (uint64_t)W * H + (image ? H : 0), n_z = (uint64_t)W * H;
    std::vector<uint32_t> cells(n_cells);
    std::vector<float> z(n_z);
    std::vector<int32_t> win(n_z);
    if ((rc = sloth_context_read(ctx, cells.data(), z.data(), win.data()))) return fail(out, rc, "sloth_context_read");
    float U[16];
    if ((rc = sloth_context_get_utransform(ctx, U))) return fail(out, rc, "sloth_context_get_utransform");
    uint64_t n_bytes = 0;
    if ((rc = sloth_context_encode(ctx, (int)mode, nullptr, 0, &n_bytes))) return fail(out, rc, "sloth_context_encode");
    std::vector<char> bytes(n_bytes);
    if ((rc = sloth_context_encode(ctx, (int)mode, bytes.data(), n_bytes, &n_bytes)))
        return fail(out, rc, "sloth_context_encode");

    const uint32_t ok = 0;
    fwrite(&ok, 4, 1, out);
    fwrite(&n_cells, 8, 1, out);
    fwrite(cells.data(), 4, n_cells, out);
    fwrite(z.data(), 4, n_z, out);
    fwrite(win.data(), 4, n_z, out);
    fwrite(U, 4, 16, out);
    fwrite(&n_bytes, 8, 1, out);
    fwrite(bytes.data(), 1, n_bytes, out);
    fclose(out);

    for (sloth_mesh* m : meshes) sloth_mesh_destroy(m);
    sloth_context_destroy(ctx);
    return 0;
}
