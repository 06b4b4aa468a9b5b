/*
 * ct_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * C API of the CPU oracle: a plain-C++ restatement of cube-tracing's
 * per-frame ray-tracing path (the GLSL shaders under src/shaders/glsl and
 * the host logic of src/main.cpp / src/accel_struct_mngr.cpp).  Only the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and only as the checker; the product library never links or calls it.
 *
 * Boundary structs (ct_instance, ct_aabb, ct_reservoir, ...) come from
 * include/ct_abi.h so both sides agree on the reference's scalar layouts.
 */
#ifndef CT_ORACLE_H
#define CT_ORACLE_H

#include <stdint.h>

#include "../include/ct_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* random.glsl:26-77, prng.glsl:198-215 */
uint32_t oc_tea(uint32_t v0, uint32_t v1);
uint32_t oc_lcg(uint32_t *state);
float oc_rnd(uint32_t *state);
uint32_t oc_urnd_interval(uint32_t *state, uint32_t mn, uint32_t mx);

/* primitives.glsl:233-257 */
void oc_compute_ray_origin(const float pos[3], const float nrm[3], float out[3]);
/* primitives.glsl:273-355 with a unit-rotation Box; returns hit (0/1). */
int oc_intersect_box(const float center[3], float radius, const float origin[3],
                     const float direction[3], int can_start_in_box,
                     int oriented, float *t, float normal[3]);
/* primitives.glsl:499-507 */
void oc_cube_like_normal(const float n[3], float out[3]);
/* mat.glsl:137-202 */
void oc_sky_color(float time, int is_afternoon, const float dir[3], float out[3]);

/* 30-bit Morton code of three 10-bit cells. */
uint32_t oc_morton3(uint32_t qx, uint32_t qy, uint32_t qz);

/* LBVH over n items: keys from the item centres, bounds from lo/hi.
 * Outputs (each may be NULL): keys[n], ids[n], child[2(n-1)],
 * boxes[12(n-1)], parent[2n-1]; returns depth (internal nodes on the
 * longest root-to-leaf path) or -1 on bad input. */
int oc_lbvh_build(uint32_t n, const float *centers, const float *lo,
                  const float *hi, uint32_t *keys, uint32_t *ids,
                  uint32_t *child, float *boxes, uint32_t *parent);

typedef struct oc_scene oc_scene;
/* world_to_object: [n_inst][16]; prims/aabbs indexed by
 * first_primitive_index + primitive id (capacity prim_capacity).  An
 * instance with primitive_count 0 is an unused id (not in the TLAS). */
oc_scene *oc_scene_create(uint32_t n_inst, const ct_instance *instances,
                          const float *world_to_object, uint32_t prim_capacity,
                          const ct_aabb *aabbs, const ct_primitive *prims,
                          uint32_t n_mat, const ct_material *materials,
                          uint32_t n_cube, const ct_light *cube_lights,
                          uint32_t n_point, const ct_light *point_lights,
                          uint32_t n_env, const ct_light *env_lights,
                          const ct_light_config *config);
void oc_scene_destroy(oc_scene *scene);
/* The previous frame's scene for the frame after an update: instance
 * transforms [n_inst][16] and first primitive indices [n_inst] before the
 * update (NULL: the current ones), and the remapping tables by slot and by
 * cube light (0 = unchanged, CT_INVALID_INDEX = deleted), read when the
 * frame's is_active has CT_REMAP_BIT. */
int oc_scene_set_previous(oc_scene *scene, const float *prev_transforms, const uint32_t *prev_first,
                          const uint32_t *remap_prims, uint32_t n_remap_prims,
                          const uint32_t *remap_lights, uint32_t n_remap_lights);
/* Arrays of one BLAS (instance) or of the TLAS (CT_INVALID_INDEX). */
int oc_scene_read_bvh(oc_scene *scene, uint32_t instance, ct_bvh_arrays *arrays);

/* brute != 0: test every primitive (no BVH).  stats may be NULL. */
void oc_trace_closest(oc_scene *scene, uint32_t n, const float *origins,
                      const float *directions, float tmax, int brute,
                      float *t_out, uint32_t *ids_out, ct_trace_stats *stats);
void oc_trace_any(oc_scene *scene, uint32_t n, const float *origins,
                  const float *directions, const float *distances,
                  const uint32_t *targets, int brute, uint8_t *occluded_out,
                  ct_trace_stats *stats);

/* intersect() (bounce.glsl:9-130), same outputs as ct_trace_bounce;
 * brute != 0 tests every voxel instead of traversing. */
void oc_trace_bounce(oc_scene *scene, uint32_t n, const float *origins,
                     const float *directions, int brute, uint32_t *ids_out,
                     float *pos_out, float *nrm_out, float *dist_out,
                     float *wo_out, ct_trace_stats *stats);
/* The build's sin/cos (Cephes polynomials) and random_cosine_direction. */
void oc_sincos(float x, float *s, float *c);
void oc_random_cosine_direction(uint32_t *seed, const float n[3], float out[3]);

/* The differentiable path tracer of ct_adjoint_render / _backward on host
 * arrays (gradients accumulated in double, then added). */
int oc_path_trace(oc_scene *scene, const ct_adjoint_desc *desc, const ct_adjoint_view *views,
                  const ct_material *materials, uint32_t material_count, float *image);
void oc_scatter_material(const ct_material *m, const float dir[3], const float nrm[3], uint32_t *seed,
                         float sd[3], float att[3], int *ok);
void oc_refraction(const float uv[3], const float n[3], float eta, float out[3]);
float oc_reflectance(float cosine, float ref_idx);
int oc_adjoint_render(oc_scene *scene, const ct_adjoint_desc *desc,
                      const ct_adjoint_view *views, const float *albedo,
                      const float *emission, float *image);
int oc_adjoint_backward(oc_scene *scene, const ct_adjoint_desc *desc,
                        const ct_adjoint_view *views, const float *albedo,
                        const float *emission, const float *grad_image,
                        float *grad_albedo, float *grad_emission);

typedef struct oc_frame oc_frame;
oc_frame *oc_frame_create(oc_scene *scene, uint32_t width, uint32_t height);
/* Later frames render `scene` (the scene after an update) while keeping
 * the frame history. */
int oc_frame_set_scene(oc_frame *frame, oc_scene *scene);
void oc_frame_destroy(oc_frame *frame);
/* One frame (prepass, shading, TAA) with up to `threads` OpenMP threads
 * (0 = default).  Returns CT_OK or a negative CT_E_* code.
 * row_begin/row_end restrict the prepass and shading passes to a band of
 * rows (used only for the bounded CPU-baseline sample); pass 0 and height
 * for a full frame. */
int oc_frame_render(oc_frame *frame, const ct_frame_params *params,
                    int threads, uint32_t row_begin, uint32_t row_end);
/* Same buffer ids and layouts as ct_frame_read. */
int oc_frame_read(oc_frame *frame, uint32_t which, void *dst, uint64_t bytes);
int oc_frame_get_stats(oc_frame *frame, ct_frame_stats *out);
/* The scene before the last update, traced by temporal reuse's
 * target-at-previous query when CT_REMAP_BIT is set (NULL: the current). */
int oc_frame_set_previous_scene(oc_frame *frame, oc_scene *scene);
/* The last frame's shading output (its image, or with TAA the TAA input)
 * with its final reservoirs held fixed, for the materials' diffuse albedo
 * and emission [n_mats][3] (the scene's count); image [H][W][4].  Equals
 * the frame's shading output at the scene's own values. */
int oc_frame_reshade(oc_frame *frame, const ct_frame_params *params, const float *diffuse, const float *emission,
                     uint32_t n_mats, float *image);
/* Gradient of sum(grad_image * image) of that function at the scene's own
 * values, added into grad_diffuse / grad_emission [materials][3]. */
int oc_frame_adjoint(oc_frame *frame, const ct_frame_params *params, const float *grad_image, float *grad_diffuse,
                     float *grad_emission);

/* Probes (tests only): single passes of the frame on given inputs. */
int oc_probe_ris(oc_scene *scene, const ct_frame_params *p, const ct_di_info *di, float confidence, uint32_t n,
                 const uint32_t *seeds, ct_reservoir *out);
int oc_probe_temporal(oc_scene *scene, const ct_frame_params *p, const ct_di_info *di, const ct_reservoir *r,
                      const ct_di_info *prev_di, const ct_reservoir *prev, uint32_t seed, ct_reservoir *out);
int oc_probe_pairwise(oc_scene *scene, const ct_frame_params *p, const ct_di_info *center, const ct_reservoir *r,
                      uint32_t num, const ct_di_info *ndi, const ct_reservoir *nres, uint32_t seed,
                      ct_reservoir *out);
int oc_probe_retrace(oc_scene *scene, const ct_frame_params *p, const ct_di_info *dst, const ct_path_reservoir *res,
                     int *ok, float pos[3], float nrm[3], float wo[3], uint32_t ids[2], float tp[3]);
int oc_probe_shift(oc_scene *scene, const ct_frame_params *p, const ct_di_info *dst, const ct_di_info *src,
                   int normalize_dst_wo, const ct_path_reservoir *res, float *jacobian, float integrand[3]);

#ifdef __cplusplus
}
#endif
#endif
