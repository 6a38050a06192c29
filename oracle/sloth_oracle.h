/* CPU oracle for sloth's rasterization path -- test infrastructure only.
 * Restates rasterizer.rs / geometry.rs / context.rs of ecumene/rust-sloth
 * in binary32 with the reference's operation order.  Not part of the
 * product: only tests, the smoke check and the CPU-baseline leg load it. */
#ifndef SLOTH_ORACLE_H
#define SLOTH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t n;               /* number of steps, <= 16 */
    float thresholds[16];     /* shade <= thresholds[i] -> glyphs[i] */
    uint8_t glyphs[16];
    uint8_t fallback;         /* glyph when no step matches (NaN included) */
    uint8_t pad[3];
} oracle_ramp;

#define ORACLE_NORM_SEQUENTIAL 1
#define ORACLE_MATVEC_PAIRWISE 2
void oracle_set_arith_variant(int v);
void oracle_mat4_mul(const float *A, const float *B, float *out);
void oracle_rotation_euler(float roll, float pitch, float yaw, float *out);
void oracle_tri_mul(const float *M, float *tri);
void oracle_tri_aabb(const float *tri, float *out);
void oracle_tri_normal(const float *tri, float *out);
uint8_t oracle_default_shader(float shade);
uint8_t oracle_ramp_shade(const oracle_ramp *r, float shade);
void oracle_default_ramp(oracle_ramp *r);
void oracle_draw_mesh(uint32_t W, uint32_t H, int image, const float *U, uint32_t *cells,
                      float *zb, int32_t *win, int32_t *winxy, const float *xyz, const uint8_t *rgb,
                      uint32_t n, const float *R, const oracle_ramp *ramp, uint32_t tri_base,
                      uint64_t *stats);
double oracle_shade_f64(uint32_t W, uint32_t H, const float *M, const float *xyz, const int32_t *faces,
                        uint32_t nv, const double *vcolor, const double *vnormal, const double *light,
                        const int32_t *win, const int32_t *winxy, const float *target, double *out,
                        double *dcolor, double *dnormal, double *dlight);
void oracle_update(uint32_t W, uint32_t H, const float *bbox_max, uint32_t n_meshes, float *U);
void oracle_mesh_bbox(const float *xyz, uint32_t n, float *out);
void oracle_clear(uint32_t W, uint32_t H, int image, uint32_t *cells, float *zb);
int oracle_render_frames(uint32_t W, uint32_t H, int image, const float *U, const float *R,
                         uint32_t F, const float *xyz, const uint8_t *rgb, uint32_t n,
                         uint32_t *cells_out, float *z_out, const oracle_ramp *ramp, int nthreads);
uint32_t oracle_shader_exhaustive_mismatch(uint32_t lo_hi16, uint32_t hi_hi16);
#include <stddef.h>
size_t oracle_encode(const uint32_t *cells, uint64_t n, int mode, int image, char *out, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
