/*
 * ct_abi.h -- C ABI of the MI355X-native cube-tracing per-frame ray-tracing path.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no HIP or torch
 * types.  Streams are passed as `void*` holding a hipStream_t (NULL selects
 * the library's own stream).  Every entry point returns an int status
 * (CT_OK == 0, negative on error) and never aborts; the message of the last
 * error on the calling thread is returned by ct_last_error_message().
 *
 * The boundary structs below keep the reference's scalar layouts byte for
 * byte (src/shaders/shared.inl, `layout(scalar)` in
 * src/shaders/include/defines.glsl:126-158) so that a host that fills the
 * reference's buffers can fill these unchanged.  Inside the library all
 * per-pixel and per-primitive data are stored structure-of-arrays.
 */
#ifndef CT_ABI_H
#define CT_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define CT_OK 0
#define CT_E_INVALID (-1)     /* bad argument or handle */
#define CT_E_CAPACITY (-2)    /* a max_* capacity would be exceeded */
#define CT_E_HIP (-3)         /* a HIP runtime call failed */
#define CT_E_UNSUPPORTED (-4) /* feature not implemented by this library */
#define CT_E_STATE (-5)       /* call not valid in the current state */
#define CT_E_NODEVICE (-6)    /* no usable GPU */

/* ------------------------------------------------ constants (shared.inl) */
#define CT_MAX_INSTANCES 983040u        /* shared.inl:18 */
#define CT_MAX_PRIMITIVES 134217728u    /* shared.inl:19 */
#define CT_VOXEL_EXTENT 0.03125f        /* shared.inl:45 */
#define CT_HALF_VOXEL_EXTENT 0.015625f  /* shared.inl:46 */
#define CT_PRIMITIVE_ALIGNMENT 32u      /* accel_struct_mngr.hpp:22 */
#define CT_INVALID_INDEX 0xFFFFFFFFu    /* (u32)-1 used as "none" */

/* feature bits of Status.is_active (shared.inl:73-79) */
#define CT_TAA_BIT (1u << 0)
#define CT_RIS_POINT_LIGHT_BIT (1u << 1)
#define CT_RIS_ENV_LIGHT_BIT (1u << 2)
#define CT_RIS_CUBE_LIGHT_BIT (1u << 3)
/* The BRDF strategy of RIS (restir_resampling.glsl:190-244): brdf_count
 * cosine-distributed candidates per pixel, each traced to its closest voxel;
 * a candidate selected last skips the visibility pass and keeps W_y = 0.
 * Its rays count into ct_frame_stats.bounce_rays. */
#define CT_RIS_BRDF_BIT (1u << 4)
#define CT_REMAP_BIT (1u << 30)

/* LIGHT.type (shared.inl:336-342) */
#define CT_LIGHT_NONE 0u
#define CT_LIGHT_POINT 1u
#define CT_LIGHT_CUBE 2u
#define CT_LIGHT_ENV_MAP 5u

/* MATERIAL.type (shared.inl:295-301) */
#define CT_MATERIAL_LAMBERTIAN 0u
#define CT_MATERIAL_METAL 1u           /* shared.inl:295-298 */
#define CT_MATERIAL_DIELECTRIC 2u
#define CT_MATERIAL_CONSTANT_MEDIUM 3u
#define CT_MATERIAL_TYPE_MASK 0xFFu

/* ------------------------------------------- boundary structs (AoS) */
typedef struct ct_aabb { /* shared.inl:89-93, 24 B */
  float minimum[3];
  float maximum[3];
} ct_aabb;

typedef struct ct_object_info { /* shared.inl:95-99 */
  uint32_t instance_id;
  uint32_t primitive_id; /* index inside the instance's BLAS slice */
} ct_object_info;

typedef struct ct_instance { /* shared.inl:265-270, 72 B */
  float transform[16];       /* object-to-world, column-major (glm/daxa) */
  uint32_t first_primitive_index;
  uint32_t primitive_count;
} ct_instance;

typedef struct ct_primitive { /* shared.inl:289-293 */
  uint32_t material_index;
  uint32_t light_index; /* CT_INVALID_INDEX when the voxel is not a light */
} ct_primitive;

typedef struct ct_material { /* shared.inl:305-321, 84 B */
  uint32_t type;
  float ambient[3];
  float diffuse[3];
  float specular[3];
  float transmittance[3];
  float emission[3];
  float shininess;
  float roughness;
  float ior;
  float dissolve;
  int32_t illum;
} ct_material;

typedef struct ct_light { /* shared.inl:344-351, 40 B */
  float position[3];
  float emissive[3];
  ct_object_info instance_info;
  float size;
  uint32_t type;
} ct_light;

typedef struct ct_light_config { /* shared.inl:203-218 */
  uint32_t point_light_count;
  float point_light_pdf;
  uint32_t cube_light_count;
  float cube_light_pdf;
  uint32_t sphere_light_count;
  float sphere_light_pdf;
  uint32_t analytic_light_count;
  float analytic_light_pdf;
  uint32_t env_map_count;
  float env_map_pdf;
  uint32_t brdf_count;
  float brdf_pdf;
  uint32_t light_count;
} ct_light_config;

typedef struct ct_reservoir { /* shared.inl:368-377, 36 B */
  uint32_t Y;
  uint32_t l_type;
  uint32_t seed;
  float W_y;
  float W_sum;
  float M;
  float F[3];
} ct_reservoir;

typedef struct ct_di_info { /* shared.inl:430-440, 60 B */
  float position[3];
  float distance;
  float normal[3];
  float ray_origin[3];
  uint32_t seed;
  ct_object_info instance_hit;
  uint32_t mat_index;
  float confidence;
} ct_di_info;

/* ------------------------------------------------------------- errors */
const char *ct_last_error_message(void);
/* 1 when a gfx950 device is visible to the HIP runtime, else 0. */
int ct_device_available(void);

/* =================================================================
 * Acceleration-structure manager.
 * Replaces ACCEL_STRUCT_MNGR (src/accel_struct_mngr.hpp:18-664).
 * ================================================================= */
typedef struct ct_accel ct_accel;

typedef struct ct_accel_desc {
  uint32_t max_instance_count;   /* create() argument, accel_struct_mngr.hpp:196 */
  uint32_t max_primitive_count;
  uint32_t max_cube_light_count;
  int32_t device;                /* HIP device ordinal */
} ct_accel_desc;

/* ACCEL_STRUCT_MNGR::create, accel_struct_mngr.cpp:83-248 */
int ct_accel_create(const ct_accel_desc *desc, ct_accel **out);
/* ACCEL_STRUCT_MNGR::destroy, accel_struct_mngr.cpp:250 */
int ct_accel_destroy(ct_accel *accel);

/* Host staging the loader writes before queueing a build.  Mirrors
 * get_instances / get_primitives / get_aabb_host_address / get_cube_lights /
 * get_host_instance_count / get_host_primitive_count
 * (accel_struct_mngr.hpp:265-311).  Pointers stay valid until destroy. */
typedef struct ct_staging {
  ct_instance *instances;          /* [max_instance_count] */
  ct_primitive *primitives;        /* [max_primitive_count] */
  ct_aabb *aabbs;                  /* [max_primitive_count] */
  ct_light *cube_lights;           /* [max_cube_light_count] */
  uint32_t host_instance_count;    /* temp_instance_count */
  uint32_t host_primitive_count;   /* temp_primitive_count */
  uint32_t cube_light_count;       /* the value create() gets by pointer */
  uint32_t max_instance_count;
  uint32_t max_primitive_count;
  uint32_t max_cube_light_count;
} ct_staging;
int ct_accel_get_staging(ct_accel *accel, ct_staging *out);
/* light_config->cube_light_count, kept by the manager for the frame. */
int ct_accel_set_cube_light_count(ct_accel *accel, uint32_t count);
/* request_aabb_host_buffer_count, accel_struct_mngr.hpp:279-285 */
int ct_accel_request_aabb_host_buffer_count(ct_accel *accel, uint32_t count,
                                            ct_aabb **base, uint32_t *offset);
/* request_primitive_index_host_buffer_count, accel_struct_mngr.hpp:293-299:
 * staging for the primitive indices (within the instance) whose AABBs an
 * UPDATE_BLAS_FROM_CPU task replaces. */
int ct_accel_request_primitive_index_host_buffer_count(ct_accel *accel, uint32_t count,
                                                       uint32_t **base, uint32_t *offset);

/* TASK, accel_struct_mngr.hpp:40-107; processed by update_scene in queue
 * order (process_task_queue, accel_struct_mngr.cpp:2115-2443):
 *
 * BUILD_BLAS_FROM_CPU: instance_count staged instances, each in the lowest
 *   free instance id, its primitives in a 32-aligned slot range.
 * DELETE_BLAS_FROM_CPU: frees the instance id and its slot range (its
 *   cube lights stay in the light list, as in delete_blas_process).
 * DELETE_PRIMITIVE_BLAS_FROM_CPU: removes primitive del_primitive_index of
 *   the instance by moving the instance's last primitive into its place;
 *   if it was a cube light, the last light of the list moves into the
 *   light's place.  The back-references (PRIMITIVE.light_index and
 *   LIGHT.instance_info) of everything moved are rewritten.  The BLAS is
 *   rebuilt.  (delete_aabb_device_buffer :607-732, delete_light_device_buffer
 *   :1155-1196; the build applies the swap the same way whether or not the
 *   deleted primitive is the instance's last, and resets the light state per
 *   task, where the reference carries it over between tasks.)
 * DELETE_PRIMITIVE_BLAS_FROM_GPU: the caller's own pass has already
 *   compacted the instance's primitives on the device (and written the
 *   remapping table); the instance loses its last del_prim_count primitives
 *   and is rebuilt, or is deleted when that leaves none (:2258-2290).  The
 *   cube lights of the primitives dropped this way are removed and those of
 *   moved primitives relinked, over all such tasks of the update at once.
 * UPDATE_BLAS_FROM_CPU: transform = transform * task transform (column-major
 *   product, :2303-2304); then, for primitive_count > 0, the AABBs staged at
 *   aabb_buf_offset replace those of the primitive indices staged at
 *   primitive_index_buf_offset (indices >= the instance's count are
 *   skipped), and the BLAS is refitted in place: same topology, bounds
 *   recomputed bottom up (update_blases :1672-1852).
 * UNDO_OP_CPU: reverts the most recent task completed by an earlier
 *   update_scene call.  A primitive deletion is restored exactly (primitive,
 *   AABB, light and every moved record back in place); other kinds are
 *   popped without effect, as process_undo_task_queue does (:853-913).
 *
 * Remapping tables (REMAPPED_PRIMITIVE_BUFFER / REMAPPED_CUBE_LIGHT_BUFFER,
 * shared.inl:280-281), one u32 per primitive slot and per cube light, map a
 * previous frame's primitive (by slot) or light (by index) to the current
 * primitive index within its instance, or light index: 0 = unchanged,
 * CT_INVALID_INDEX = deleted.  They are filled by deletions and undos and
 * cleared by the next update_scene call (process_settling_task_queue
 * :2586-2680). */
#define CT_TASK_BUILD_BLAS_FROM_CPU 0u
#define CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_CPU 1u
#define CT_TASK_DELETE_BLAS_FROM_CPU 2u
#define CT_TASK_UPDATE_BLAS_FROM_CPU 3u
#define CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_GPU 4u
#define CT_TASK_UNDO_OP_CPU 5u

typedef struct ct_task {
  uint32_t type;
  union {
    struct { /* BLAS_BUILD_FROM_CPU */
      uint32_t instance_count;
      uint32_t primitive_count;
      float transform[16];
    } blas_build_from_cpu;
    struct { /* BLAS_PRIMITIVE_DELETE_FROM_CPU */
      uint32_t instance_index;
      uint32_t del_primitive_index;
      uint32_t remap_primitive_index;
      uint32_t del_light_index;
      uint32_t remap_light_index;
      uint32_t remap_primitive_light_index;
    } blas_delete_primitive_from_cpu;
    struct { /* BLAS_DELETE_FROM_CPU */
      uint32_t instance_index;
      uint32_t first_primitive_index;
      uint32_t deleted_primitive_count;
    } blas_delete_from_cpu;
    struct { /* BLAS_UPDATE */
      uint32_t instance_index;
      float transform[16];
      uint32_t primitive_count;
      uint32_t primitive_index_buf_offset;
      uint32_t aabb_buf_offset;
    } blas_update;
    struct { /* BLAS_DEL_PRIM_FROM_GPU */
      uint32_t instance_index;
      uint32_t del_prim_count;
    } blas_del_prim_gpu;
  } u;
} ct_task;

/* task_queue_add, accel_struct_mngr.hpp:321-330 */
int ct_accel_task_queue_add(ct_accel *accel, const ct_task *task);

/* update_scene, accel_struct_mngr.hpp:332-436, called once per frame before
 * the frame as App::update does (main.cpp:1420).  The reference spreads an
 * update over three calls (UPDATING, SWITCHING, SETTLING); here one call
 * settles the previous update (clears the remapping tables, makes the
 * previous instance transforms equal the current ones: status IDLE), then
 * processes the whole queue and rebuilds what changed (status SETTLING
 * until the next call).  A frame rendered while SETTLING sees the instance
 * transforms from before the update as the previous frame's
 * (get_geometry_previous_transform_from_instance_id, motion_vectors.glsl:72)
 * and, with CT_REMAP_BIT set, applies the remapping tables in temporal
 * reuse (restir_resampling.glsl:338-387) and traces its target-at-previous
 * visibility against the previous set (restir_resampling.glsl:369 ->
 * light.glsl:102).  An update that runs tasks first copies the committed
 * scene on the device (about 116 B per primitive slot): while SETTLING,
 * `previous` != 0 in ct_accel_get_tlas / ct_accel_get_world names that
 * copy; otherwise previous and current are the same set.  The work is ordered on the
 * accel's stream after `stream` has drained; with synchronize != 0 the call
 * returns after the build.  *did_work (optional) receives 1 when a task or
 * a settle ran, 0 when idle. */
int ct_accel_update_scene(ct_accel *accel, int synchronize, void *stream,
                          int *did_work);

/* AS_MANAGER_STATUS values, accel_struct_mngr.hpp:24-33 */
#define CT_AS_IDLE 0u
#define CT_AS_UPDATING 1u
#define CT_AS_SWITCHING 2u
#define CT_AS_SETTLING 4u
int ct_accel_get_status(ct_accel *accel, uint32_t *status);
/* is_remapping_primitive_active, accel_struct_mngr.hpp:252-254 */
int ct_accel_is_remapping_active(ct_accel *accel, int *active);

/* Top-level structure handle; get_current_tlas / get_previous_tlas
 * (accel_struct_mngr.hpp:216-223): previous != 0 while SETTLING gives the
 * TLAS (and through it the BLASes) from before the last update. Device
 * pointers. */
typedef struct ct_tlas {
  const void *nodes;     /* BVH2 nodes, 64 B each */
  const void *leaves;    /* instance ids, Morton order */
  uint32_t root_ref;     /* bit 31 set: leaf */
  uint32_t instance_count;
  uint32_t depth;
  uint32_t generation;   /* increments on every TLAS build */
} ct_tlas;
int ct_accel_get_tlas(ct_accel *accel, int previous, ct_tlas *out);

/* Device buffers of the WORLD table (shared.inl:272-286) for the current
 * or previous buffer set; get_{current,previous}_*_buffer,
 * accel_struct_mngr.hpp:225-262.  The previous set differs while SETTLING:
 * its instances, primitives and AABBs are those from before the last update
 * (a caller's compaction ahead of a GPU deletion task has already written
 * the AABB and primitive slots by then); cube lights and the remapping
 * tables have one copy. */
typedef struct ct_world {
  const void *instances;   /* ct_instance AoS */
  void *primitives;        /* ct_primitive AoS by slot */
  void *aabbs;             /* ct_aabb AoS by slot */
  const void *cube_lights; /* ct_light AoS */
  void *remapped_primitives;  /* u32 by slot (remapped_primitive_address) */
  void *remapped_cube_lights; /* u32 by light (remapped_cube_light_address) */
  uint32_t instance_count;
  uint32_t primitive_capacity;
  uint32_t cube_light_count;
  uint32_t cube_light_capacity;
} ct_world;
/* Writable members are for a caller's own compaction pass ahead of a
 * DELETE_PRIMITIVE_BLAS_FROM_GPU task (rearrangement.slang writes the same
 * buffers). */
int ct_accel_get_world(ct_accel *accel, int previous, ct_world *out);
/* Host copies of the committed WORLD buffers as the device holds them:
 * aabbs and primitives by slot [prim_capacity], cube lights
 * [light_capacity], and the two remapping tables.  Any pointer may be NULL. */
int ct_accel_read_world(ct_accel *accel, ct_aabb *aabbs, ct_primitive *prims,
                        uint32_t prim_capacity, ct_light *lights, uint32_t light_capacity,
                        uint32_t *remapped_primitives, uint32_t *remapped_cube_lights);

/* Read back one BLAS (instance_index) or the TLAS (instance_index ==
 * CT_INVALID_INDEX) as built on the GPU, for parity tests.  Arrays are
 * host buffers the caller sizes with n = primitive (or instance) count:
 * keys[n], ids[n], child[2*(n-1)], boxes[12*(n-1)] (child0 lo,hi, child1
 * lo,hi), parent[2n-1] (internal nodes then leaves).  Any may be NULL. */
typedef struct ct_bvh_arrays {
  uint32_t *keys;
  uint32_t *ids;
  uint32_t *child;
  float *boxes;
  uint32_t *parent;
  uint32_t n;        /* out: leaf count */
  uint32_t depth;    /* out: longest root-to-leaf path in nodes */
  uint32_t root_ref; /* out */
  uint32_t pad;
} ct_bvh_arrays;
int ct_accel_read_bvh(ct_accel *accel, uint32_t instance_index,
                      ct_bvh_arrays *arrays);
/* Host copy of an instance's world-to-object matrix (inverse of the
 * transform), as the kernels use it. */
/* Host copy of the committed instance table, the contents of
 * get_current_instance_buffer (accel_struct_mngr.hpp:225-230), by instance
 * id: count = number of slots; an unused slot has primitive_count 0.  At
 * most `capacity` records are written. */
int ct_accel_read_instances(ct_accel *accel, ct_instance *out, uint32_t capacity,
                            uint32_t *count);
int ct_accel_get_world_to_object(ct_accel *accel, uint32_t instance_index,
                                 float out[16]);

/* =================================================================
 * Per-stage traversal on caller-owned host arrays (tests, tools).
 * Closest hit follows traceRayEXT + rint.glsl:10-43 + rchit semantics;
 * any hit follows is_vertex_visible, light.glsl:82-142.
 * ================================================================= */
typedef struct ct_trace_stats {
  uint64_t rays;
  uint64_t nodes;     /* BVH2 internal nodes fetched (TLAS + BLAS) */
  uint64_t leaves;    /* primitive leaves tested */
  uint64_t instances; /* instance records entered */
} ct_trace_stats;

/* origins, directions: [n][3]; t_out: closest t or -1; ids_out: [n][2]
 * (instance_id, primitive_id) or CT_INVALID_INDEX pairs on a miss. */
int ct_trace_closest(ct_accel *accel, uint32_t n, const float *origins,
                     const float *directions, float tmax, float *t_out,
                     uint32_t *ids_out, ct_trace_stats *stats);
/* distances: [n]; targets: [n][2] voxel excluded from the test, or
 * CT_INVALID_INDEX pairs; occluded_out: [n] bytes, 1 = blocked. */
int ct_trace_any(ct_accel *accel, uint32_t n, const float *origins,
                 const float *directions, const float *distances,
                 const uint32_t *targets, uint8_t *occluded_out,
                 ct_trace_stats *stats);

/* ct_trace_any without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same rays, same result bytes as
 * ct_trace_any, but every array is in DEVICE memory and the call is only
 * enqueued on `stream` (null: the accel's stream); nothing is copied back
 * and nothing is waited for.  The staged rays and the work counters live in
 * a fixed 36 MiB device array per GPU, so n is at most 2^20 (CT_E_CAPACITY
 * above); the array is shared by every call on that GPU, so calls that
 * overlap on one GPU race. */
int ct_trace_any_noalloc(ct_accel *accel, uint32_t n, const float *origins,
                         const float *directions, const float *distances,
                         const uint32_t *targets, uint8_t *occluded_out,
                         void *stream);

/* Path-segment closest hit, intersect() of bounce.glsl:9-130: the box
 * tests start at compute_ray_origin(origin, direction) and the closest
 * voxel is returned (the shader commits the first one its query
 * generates).  Outputs per ray: ids [n][2] (CT_INVALID_INDEX on a miss),
 * world hit position and normal [n][3], distance to the un-pushed origin
 * [n] and wo [n][3], as the INTERSECT record holds them. */
int ct_trace_bounce(ct_accel *accel, uint32_t n, const float *origins,
                    const float *directions, uint32_t *ids_out, float *pos_out,
                    float *nrm_out, float *dist_out, float *wo_out,
                    ct_trace_stats *stats);

/* =================================================================
 * Frame context: the buffers and passes App::draw records
 * (src/main.cpp:1436-1703) -- prepass, shading pass, TAA.
 * ================================================================= */
typedef struct ct_ctx ct_ctx;

typedef struct ct_ctx_desc {
  uint32_t max_width;   /* the reference sizes for 3840x2160 (shared.inl:68-71) */
  uint32_t max_height;
  int32_t device;
  uint32_t pad;
} ct_ctx_desc;

int ct_ctx_create(const ct_ctx_desc *desc, ct_ctx **out);
int ct_ctx_destroy(ct_ctx *ctx);
/* load_materials, main.cpp:512-558 (Lambertian materials only). */
int ct_ctx_set_materials(ct_ctx *ctx, const ct_material *materials,
                         uint32_t count);
/* create_point_lights / create_environment_light, main.cpp:348-400 */
int ct_ctx_set_point_lights(ct_ctx *ctx, const ct_light *lights, uint32_t count);
int ct_ctx_set_env_lights(ct_ctx *ctx, const ct_light *lights, uint32_t count);
/* load_materials with an offset: appends after the materials already
 * loaded and returns the index of the first one in *base. */
int ct_ctx_append_materials(ct_ctx *ctx, const ct_material *materials,
                            uint32_t count, uint32_t *base);
/* load_lights, main.cpp:418-438 */
int ct_ctx_set_light_config(ct_ctx *ctx, const ct_light_config *config);
int ct_ctx_get_light_config(ct_ctx *ctx, ct_light_config *out);
/* load_reservoirs, main.cpp:234-318: zero every reservoir, DI, velocity
 * and image buffer. */
int ct_ctx_reset(ct_ctx *ctx, void *stream);

/* ReSTIR tunables: defines.glsl:34-62 and taa.glsl:13.  The reference
 * values are returned by ct_restir_params_reference(). */
#define CT_MAX_NEIGHBORS 8u
typedef struct ct_restir_params {
  uint32_t min_ris_point_samples;  /* 1 */
  uint32_t max_ris_point_samples;  /* 4 */
  uint32_t min_ris_env_samples;    /* 1 */
  uint32_t max_ris_env_samples;    /* 4 */
  uint32_t min_ris_cube_samples;   /* 4 */
  uint32_t max_ris_cube_samples;   /* 32 */
  float max_influence_from_the_past; /* 20: confidence scale and temporal M cap */
  uint32_t min_neighbors;          /* 3 */
  uint32_t max_neighbors;          /* 4, at most CT_MAX_NEIGHBORS */
  float min_neighbor_radius;       /* 2 px */
  float max_neighbor_radius;       /* 50 px */
  float max_distance_to_hit;       /* 1000 */
  float taa_alpha;                 /* 0.8: weight of the current frame */
  uint32_t taa_truncate_dispatch;  /* 1: TAA covers (W/8*8)x(H/8*8) like main.cpp:1642-1643 */
  uint32_t direct_emittance;       /* 1: DIRECT_EMITTANCE_ON */
  /* Indirect illumination by ReSTIR PT (INDIRECT_ILLUMINATION_ON with
   * RESTIR_PT_ON, rgen.glsl:271-315, 427-452; SCENE_PARAMS rgen.glsl:280-287). */
  uint32_t indirect;               /* 0: INDIRECT_ILLUMINATION_ON (defines.glsl:20) */
  uint32_t max_depth;              /* 3: Status.max_depth = MAX_DEPTH (shared.inl:61) */
  uint32_t pt_temporal;            /* 1: RESTIR_PT_TEMPORAL_ON (defines.glsl:15) */
  uint32_t pt_spatial;             /* 1: RESTIR_PT_SPATIAL_ON */
  uint32_t pt_shift_mapping;       /* CT_SHIFT_RECONNECTION (rgen.glsl:282) */
  uint32_t pt_neighbor_count;      /* 3: NEIGHBOR_COUNT, at most CT_MAX_PT_NEIGHBORS */
  int32_t pt_neighbor_radius;      /* 1: NEIGHBOR_RADIUS (pixels) */
  float pt_m_cap;                  /* 20: MAX_INFLUENCE_FROM_THE_PAST_THRESHOLD_PT */
  float near_field_distance;       /* 0.0078125: NEAR_FIELD_DISTANCE */
  float roughness_threshold;       /* 0.2: SPECULAR_ROUGHNESS_THRESHOLD */
  uint32_t pt_russian_roulette;    /* 0: USE_RUSSIAN_ROULETTE */
  uint32_t pad;
} ct_restir_params;
#define CT_SHIFT_RECONNECTION 0u  /* SHIFT_MAPPING_RECONNECTION, defines.glsl:85 */
#define CT_SHIFT_HYBRID 2u        /* SHIFT_MAPPING_HYBRID, defines.glsl:87.  The path
                                     builder marks the first connectable vertex (far field,
                                     rough after rough); a shift replays the source's seed
                                     from the destination's primary hit to the vertex before
                                     the reconnection vertex (no vertex before it connectable,
                                     that one rough, else rejected), then reconnects with the
                                     cached Jacobian, its replayed prefix counting 1;
                                     near-field connections are rejected. */
#define CT_MAX_PT_NEIGHBORS 8u
#define CT_MAX_PT_DEPTH 14u       /* path lengths are stored in 4 bits */

/* PATH_RESERVOIR (shared.inl:405-423), scalar layout, 92 bytes. */
typedef struct ct_object_hit { /* OBJECT_HIT, shared.inl:101-105 */
  ct_object_info object;
  float hit[3];
} ct_object_hit;
typedef struct ct_path_reservoir {
  float M;
  float weight;
  uint32_t path_flags;
  uint32_t rc_random_seed;
  float F[3];
  float light_pdf;
  float cached_jacobian[3];
  uint32_t init_random_seed;
  ct_object_hit rc_vertex_hit;
  float rc_vertex_wi[3];
  float rc_vertex_irradiance[3];
} ct_path_reservoir;
void ct_restir_params_reference(ct_restir_params *out);

typedef struct ct_frame_params {
  uint32_t width;
  uint32_t height;
  uint32_t frame_number;   /* Status.frame_number */
  uint32_t is_active;      /* Status.is_active feature bits */
  float time;              /* Status.time */
  uint32_t is_afternoon;   /* Status.is_afternoon */
  uint32_t has_previous;   /* a previous frame exists (TAA precondition, main.cpp:1483) */
  uint32_t pad;
  float inv_view[16];      /* camera_view.inv_view (column-major) */
  float inv_proj[16];      /* camera_view.inv_proj with y.y negated (main.cpp:1460) */
  float prev_view_proj[16];/* inverse(prev_inv_proj) * inverse(prev_inv_view), motion_vectors.glsl:143 */
  ct_restir_params restir;
} ct_frame_params;

/* Record the frame's passes on `stream`. A frame of at least 2^20 pixels
 * whose width is a multiple of 16 runs as two halves of the image, the
 * even and the odd rows of 16-pixel tiles: the passes up to temporal path
 * reuse, then the spatial passes, run for both halves at once, the second
 * half on a stream of the context that forks from and joins back into
 * `stream`; TAA follows on `stream`. Per-pass events or traversal counters
 * (ct_ctx_set_instrumentation 1 or 2) run the frame whole on `stream`.
 * Results are those of the whole-frame sequence either way. */
int ct_frame_render(ct_ctx *ctx, ct_accel *accel, const ct_frame_params *params,
                    void *stream);

/* Buffers readable after a frame (host copies, AoS reference layouts). */
#define CT_BUF_IMAGE 0u          /* float4 per pixel: final image */
#define CT_BUF_PREV_IMAGE 1u     /* float4: previous final image */
#define CT_BUF_TAA_FRAME 2u      /* float4: shading output TAA reads */
#define CT_BUF_DI 3u             /* ct_di_info: current frame DI */
#define CT_BUF_PREV_DI 4u        /* ct_di_info: previous-frame DI */
#define CT_BUF_INT_RESERVOIR 5u  /* ct_reservoir: intermediate */
#define CT_BUF_PREV_RESERVOIR 6u /* ct_reservoir: previous frame */
#define CT_BUF_VELOCITY 7u       /* float2 */
#define CT_BUF_PATH_RESERVOIR 8u          /* ct_path_reservoir: output (current frame) */
#define CT_BUF_TEMPORAL_PATH_RESERVOIR 9u /* ct_path_reservoir: temporal (next frame's history) */
#define CT_BUF_INDIRECT 10u               /* float3: indirect colour of the first pass */
int ct_frame_read(ct_ctx *ctx, uint32_t which, void *host_dst, size_t bytes);

typedef struct ct_frame_stats {
  uint64_t primary_rays;
  uint64_t shadow_rays;     /* visibility queries of every pass */
  uint64_t shadow_rays_ris;
  uint64_t shadow_rays_temporal;
  uint64_t shadow_rays_spatial;
  uint64_t nodes;           /* any-hit traversal: BVH2 nodes fetched (counters on) */
  uint64_t leaves;          /* any-hit traversal: leaves tested */
  uint64_t instances;       /* any-hit traversal: instance records entered */
  uint64_t closest_nodes;   /* primary closest-hit traversal, same counters */
  uint64_t closest_leaves;
  uint64_t closest_instances;
  float ms_total;           /* HIP-event time of the frame's kernels */
  float ms_primary;
  float ms_ris;
  float ms_temporal;
  float ms_spatial;
  float ms_taa;
  float ms_trace_any;       /* sum over the any-hit traversal launches */
  float ms_trace_closest;
  uint64_t bounce_rays;     /* closest-hit rays of intersect(): ReSTIR PT path
                               segments and the RIS BRDF strategy's rays */
  uint64_t shadow_rays_pt;  /* ReSTIR PT reconnection visibility queries */
  uint64_t pt_nodes;        /* path-segment traversal (counters on) */
  uint64_t pt_leaves;
  uint64_t pt_instances;
  float ms_pt;              /* initial paths + temporal path reuse */
  float ms_pt_spatial;      /* spatial path reuse + final composition */
  float ms_pt_segments;     /* sum over the path-segment traversal launches */
  float pad;
} ct_frame_stats;
/* Enable per-pass HIP events (1, frames then run whole on one stream),
 * traversal node/leaf counters (2, likewise), or only the frame's start
 * and end events for ms_total (4). */
int ct_ctx_set_instrumentation(ct_ctx *ctx, uint32_t flags);
int ct_frame_get_stats(ct_ctx *ctx, ct_frame_stats *out);
/* Adjoint of the last rendered frame with its final reservoirs held fixed
 * (the checkpoint the frame leaves in its buffers).  No counterpart in the
 * reference (its renderer has no derivatives).  The function
 * differentiated is the frame's shading output (the image, or with TAA the
 * TAA input) as a function of every material's diffuse albedo d and
 * emission e: a shaded pixel shows F_y W_y + I_ind + e[m_px] (e with
 * direct emittance), where F_y = brdf(d[m_px]) * Le * G replays the final
 * reservoir's light sample from its seed with the visibility the frame
 * found, Le is the emission of the light voxel's material for a cube light
 * (a point or environment light keeps its own), and W_y and the indirect
 * term I_ind are held fixed (a biased gradient; the resampling weights are
 * not differentiated).  Channels the frame's NaN scrub zeroes have no
 * gradient.  grad_image: device [H][W][4] (dLoss/dimage, alpha ignored);
 * the gradients are added into grad_diffuse and grad_emission, device
 * [materials][3] (zero them first).  Call after ct_frame_render (or a batch
 * render of this context) with the same params, before the context renders
 * again and before the accel or the materials change (CT_E_STATE once the
 * accel has a newer build).  Ordered on `stream` (null: the context's). */
int ct_frame_adjoint(ct_ctx *ctx, ct_accel *accel, const ct_frame_params *params, const float *grad_image,
                     float *grad_diffuse, float *grad_emission, void *stream);
/* Random replay (trace_random_replay_path_hybrid_simple,
 * path_tracer.glsl:889-966), whose only caller in the reference is
 * commented out (indirect_illumination.glsl:301-302): after a frame with
 * paths, pixel (x, y) replays the path reservoir of pixel (x+dx, y+dy) --
 * its init seed, up to its reconnection length -- from its own primary
 * hit, with the frame's path sampling steps.  `params` are the frame's
 * (same size).  out: device pointer, 12 u32 per pixel: ok (the vertex at
 * the reconnection length exists), length reached, the reservoir's path
 * and reconnection lengths, that vertex's instance, primitive and hit
 * position (f32 bits), the vertex before it (instance, primitive), 0.
 * With dx = dy = 0 and no path reuse the replay reproduces each pixel's own
 * reconnection vertex. */
int ct_pt_random_replay(ct_ctx *ctx, ct_accel *accel, const ct_frame_params *params, int32_t dx, int32_t dy,
                        uint32_t *out, void *stream);

/* =================================================================
 * Batched frames.  No counterpart in the reference, which renders one
 * view per App::draw (main.cpp:1436); for callers with many views.
 *
 * A batch binds n (context, accel) pairs: views of one or several scenes,
 * each view with its own context (its frame history) and resolution.
 * ct_batch_render runs frame params[i] of every view with one launch per
 * pass over the pixels of each group of views (the passes and results of
 * ct_frame_render, view by view).  With use_graph the launch sequence is captured once into
 * a hipGraph and replayed; it is captured again when resolutions, path
 * depth, neighbour count or the tuned grid change.  Views must differ in
 * context; contexts and accels must share one device; a context of a
 * batch holds at most 2^22 pixels.
 * ================================================================= */
typedef struct ct_batch ct_batch;
int ct_batch_create(uint32_t n_views, ct_ctx *const *ctxs, ct_accel *const *accels,
                    ct_batch **out);
int ct_batch_destroy(ct_batch *batch);
int ct_batch_render(ct_batch *batch, const ct_frame_params *params, int use_graph,
                    void *stream);
/* Totals over the views of the last batch render: rays and ray counters,
 * ms_total = HIP-event time of the whole batch (other times are 0). */
int ct_batch_get_stats(ct_batch *batch, ct_frame_stats *out);
/* Launch-shape autotuner: renders the given frames `repeats` times for
 * each candidate shape and keeps the fastest; reports the choice. A shape
 * is the number of view groups (1 to 4: contiguous ranges of views of
 * about equal pixel count, each with its own queues, run on concurrent
 * streams so one group's traversal tail overlaps another's passes), the
 * grid of the persistent traversal kernels (3, 4 or 5 blocks per CU; a
 * grid never exceeds the blocks that can be resident at once) and the ray
 * queue's order (shadow rays at least far_ray long are queued and traced
 * first: 1, 2, 4 world units or none): 48 shapes.  No shape changes a
 * result.  A new batch starts with 2 groups, 5 blocks per CU and 2 units.
 * Switched off in this build: returns CT_E_UNSUPPORTED unless the library
 * is compiled with CT_BATCH_AUTOTUNE=1 (one run of it ended in a host
 * segmentation fault whose cause is not found); batches keep the default
 * shape. */
typedef struct ct_batch_tuning {
  uint32_t blocks_per_cu;
  float ms;            /* mean batch time of the choice */
  uint32_t candidates; /* number of shapes tried */
  uint32_t groups;     /* view groups of the choice */
  float far_ray;       /* the queue's far/near split of the choice */
} ct_batch_tuning;
int ct_batch_autotune(ct_batch *batch, const ct_frame_params *params, uint32_t repeats,
                      ct_batch_tuning *out);

/* =================================================================
 * Camera (src/camera.h:75-93, main.cpp:1452-1460) and matrices.
 * ================================================================= */
int ct_camera_matrices(const float position[3], const float forward[3],
                       const float up[3], float fov_radians, float aspect,
                       float z_near, float z_far, float inv_view[16],
                       float inv_proj[16]);
int ct_camera_prev_view_proj(const float prev_inv_view[16],
                             const float prev_inv_proj[16], float out[16]);
int ct_mat4_inverse(const float m[16], float out[16]);

/* =================================================================
 * Synthetic voxel scenes (the map loader's role, map_loader.cpp:224-300).
 * ================================================================= */
typedef struct ct_scene ct_scene;

typedef struct ct_scene_desc {
  uint32_t grid[3];          /* voxels per axis */
  uint32_t chunk[3];         /* voxels per BLAS chunk per axis */
  uint32_t occupied_count;   /* exact number of occupied voxels */
  uint32_t emissive_count;   /* voxels that become cube lights */
  float le_min, le_max;      /* emitted radiance range (white) */
  float albedo_min, albedo_max;
  uint32_t seed;
  uint32_t palette_size;     /* 0: one material per voxel; else a shared palette
                                of this many entries (at most 256), one uint8
                                entry per voxel */
  float voxel_extent;        /* 0.03125 */
  uint32_t palette_emissive; /* palette mode: entries that emit, with a
                                radiance in [le_min, le_max] per channel; every
                                voxel of such an entry is a cube light and
                                emissive_count is ignored */
} ct_scene_desc;

typedef struct ct_scene_info {
  uint32_t instance_count;
  uint32_t primitive_count;     /* occupied voxels */
  uint32_t material_count;
  uint32_t cube_light_count;
  const ct_instance *instances; /* first_primitive_index: host order, contiguous */
  const ct_aabb *aabbs;         /* [primitive_count] */
  const ct_primitive *primitives;
  const ct_material *materials;
  const ct_light *cube_lights;  /* instance_info = (instance, primitive in BLAS) */
  float grid_min[3];
  float grid_max[3];
} ct_scene_info;

int ct_scene_generate(const ct_scene_desc *desc, ct_scene **out);
int ct_scene_get_info(const ct_scene *scene, ct_scene_info *out);
int ct_scene_destroy(ct_scene *scene);
/* load_scene (main.cpp:1045-1126): write the staging buffers, queue one
 * BUILD_BLAS_FROM_CPU task, run update_scene(true), upload materials and
 * the light configuration (cube lights only, brdf_count = 1 as at
 * main.cpp:1317-1320). */
int ct_scene_load(const ct_scene *scene, ct_accel *accel, ct_ctx *ctx,
                  void *stream);
/* load_model(name, transform) (main.cpp:1006-1043): the same, with
 * `transform` (column-major 4x4, as INSTANCE.transform) in the
 * BUILD_BLAS_FROM_CPU task for every new instance; null = identity. */
int ct_scene_load_transformed(const ct_scene *scene, ct_accel *accel, ct_ctx *ctx,
                              const float transform[16], void *stream);
/* The context half of ct_scene_load, for more contexts rendering a scene
 * already loaded into an accel with a fresh context: the same materials
 * and light configuration.  The context must hold no materials yet. */
int ct_scene_bind(const ct_scene *scene, ct_ctx *ctx);

/* MagicaVoxel .vox files as scenes: the voxel -> AABB / PRIMITIVE /
 * MATERIAL / LIGHT rules of the map loader (map_loader.cpp:97-325, called
 * by load_model / load_scene, main.cpp:1006-1126) over the library's own
 * parser of the VOX chunk format (SIZE, XYZI, RGBA, MATL, MATT, nTRN,
 * nGRP, nSHP).  The reference parses the file with gvox
 * (map_loader.cpp:479-493); what gvox decides is fixed here: a model (or
 * each chunk-aligned box of it) is one region and one instance; with a
 * scene graph a model of size s sits at v - floor(s/2) + translation;
 * emissive color indices are MATL `_type` `_emit` with `_emit` > 0 or MATT
 * type 3 with weight > 0, and emit their palette color.  The result is a
 * ct_scene for ct_scene_load / ct_scene_bind like a generated one. */
#define CT_AXIS_X_BOTTOM_TOP 0u /* AXIS_DIRECTION, defines.h:118-125 */
#define CT_AXIS_Y_BOTTOM_TOP 1u
#define CT_AXIS_Z_BOTTOM_TOP 2u
#define CT_AXIS_X_TOP_BOTTOM 3u
#define CT_AXIS_Y_TOP_BOTTOM 4u
#define CT_AXIS_Z_TOP_BOTTOM 5u
typedef struct ct_vox_desc {
  uint32_t axis;      /* CT_AXIS_*; main.cpp:1009 loads with X_BOTTOM_TOP */
  uint32_t chunk[3];  /* region size in file voxels per axis; 0: a whole model */
  int32_t model;      /* -1: every model (placed by the scene graph if any); k: model k alone at the origin */
  float voxel_extent; /* 0: CT_VOXEL_EXTENT */
  float flux;         /* emission scale, 0: 1 (the loader's flux, map_loader.cpp:129) */
  uint32_t pad;
} ct_vox_desc;
/* Parses `bytes` bytes at `data` (a whole .vox file).  CT_E_INVALID for a
 * malformed or empty file, CT_E_UNSUPPORTED for a rotated scene-graph
 * node, CT_E_CAPACITY beyond CT_MAX_PRIMITIVES voxels. */
int ct_scene_from_vox(const void *data, size_t bytes, const ct_vox_desc *desc, ct_scene **out);

/* =================================================================
 * Voxel editing on the GPU: the squared deletion brush
 * (squared_brush.slang:37-133), the rearrangement pass
 * (rearrangement.slang:8-204) and check/process_voxel_modifications
 * (accel_struct_mngr.cpp:2717-2871).
 *
 * Marks accumulate in the accel (an instance bitmask, a primitive-slot
 * bitmask and the BRUSH_COUNTER pair) until processed.  Processing
 * compacts every marked instance on the device -- the k-th marked local
 * index below n - d receives the primitive and AABB of the k-th unmarked
 * local index at or above n - d (n primitives, d marked) -- writes the
 * remapping table (moved: new local index; deleted: ~0u), queues one
 * DELETE_PRIMITIVE_BLAS_FROM_GPU task per instance in id order and clears
 * the marks.  The next ct_accel_update_scene applies the tasks (BLAS
 * rebuild, or instance deletion when nothing is left); it takes the moves
 * into its host copy of the scene from this call's record, so the slots of
 * a compacted instance are left alone until then (a caller's own pass that
 * compacts the instance again queues a task of its own, and the update then
 * reads the instance's slots back from the device).  Unlike the
 * reference, which leaves them in place, that update also removes the
 * cube lights of deleted voxels (the last light moves into each hole and
 * its owner is relinked, highest index first), points the lights of moved
 * voxels at their new primitive ids and fills the light remapping table.
 * All three calls synchronize the device first; `stream` is accepted for
 * symmetry and unused.
 * ================================================================= */
/* entry_brush: every pixel of the square [x-window, x+window] x
 * [y-window, y+window] (clipped to the last frame, in u32 arithmetic as
 * the shader: a cursor closer than `window` to the left or top edge
 * selects nothing) whose stored DI record (the previous DI,
 * squared_brush.slang:95-96) hits the instance under (x, y) marks its
 * voxel.  Nothing happens when (x, y) itself shows the sky. */
int ct_brush_apply(ct_ctx *ctx, ct_accel *accel, uint32_t x, uint32_t y, uint32_t window, void *stream);
/* delete_primitive_from_instance for a caller's list of voxels (instance
 * id, primitive id within the instance); ids out of range are ignored. */
int ct_brush_mark(ct_accel *accel, const ct_object_info *voxels, uint32_t n, void *stream);
/* BRUSH_COUNTER: newly marked instances and primitives since the last processing. */
int ct_brush_get_counters(ct_accel *accel, uint32_t *instance_count, uint32_t *primitive_count);
/* check_voxel_modifications: does nothing (marks kept) unless the accel
 * is idle (:2855).  *instances / *primitives (may be NULL) receive the
 * number of tasks queued and of voxels they delete. */
int ct_accel_process_voxel_modifications(ct_accel *accel, void *stream, uint32_t *instances,
                                         uint32_t *primitives);

/* =================================================================
 * Differentiable rendering.  No counterpart in the reference (its
 * renderer has no derivatives); the build's own adjoint of radiance with
 * respect to per-voxel albedo and emission.
 *
 * The forward model is a unidirectional path tracer over the accel's
 * voxels: every segment is intersect() of bounce.glsl:9-130 (closest
 * voxel, boxes tested from the origin pushed along the ray), scattering
 * is cosine-weighted around the face normal (random_cosine_direction,
 * prng.glsl:84-108, whose Lambertian weight f*cos/pdf is the albedo),
 * emission is collected at every vertex and a ray that leaves the scene
 * carries nothing.  A sample of pixel (x, y) is
 *   L = sum_k beta_k * e(v_k),  beta_0 = 1,  beta_{k+1} = beta_k * a(v_k)
 * over vertices v_0..v_K, K <= max_depth.  The camera ray passes through
 * (x + u, y + v), u, v the first two numbers of the sample's stream
 * tea(tea(global pixel, seed), sample).
 *
 * The adjoint replays each path from its seed (nothing is stored between
 * the passes; the <= max_depth+1 vertices of the path being replayed stay
 * in registers) and adds dLoss/da and dLoss/de into the gradient buffers
 * with atomics.
 * ================================================================= */
#define CT_ADJOINT_MAX_DEPTH 7u
typedef struct ct_adjoint_view {
  uint32_t width;
  uint32_t height;
  float inv_view[16]; /* as ct_frame_params */
  float inv_proj[16];
} ct_adjoint_view;
typedef struct ct_adjoint_desc {
  uint32_t n_views;
  uint32_t spp;       /* samples per pixel; the image is their mean */
  uint32_t max_depth; /* scattering events after the primary hit, <= CT_ADJOINT_MAX_DEPTH */
  uint32_t seed;
} ct_adjoint_desc;
/* Device pointers.  albedo [slots][3] and emission [slots] are indexed by
 * the voxel's primitive slot (first_primitive_index + primitive id);
 * image holds the views one after another, [H][W][3] each.  Calls are
 * ordered on `stream` (null: the accel's stream). */
int ct_adjoint_render(ct_accel *accel, const ct_adjoint_desc *desc,
                      const ct_adjoint_view *views, const float *albedo,
                      const float *emission, float *image, void *stream);
/* grad_image has image's layout; the gradients are added into
 * grad_albedo [slots][3] and grad_emission [slots] (zero them first). */
int ct_adjoint_backward(ct_accel *accel, const ct_adjoint_desc *desc,
                        const ct_adjoint_view *views, const float *albedo,
                        const float *emission, const float *grad_image,
                        float *grad_albedo, float *grad_emission, void *stream);

/* Plain (non-ReSTIR) path tracing with the material callables
 * (indirect_illumination_path_tracing, indirect_illumination.glsl:454-539;
 * METAL / DIELECTRIC / CONSTANT_MEDIUM / LAMBERTIAN scatter,
 * rcall_scatter.glsl:11-88).  The camera rays, samples and segments are
 * those of ct_adjoint_render; at every vertex the voxel's material
 * (materials[PRIMITIVE.material_index]) adds its emission times the
 * throughput, then its callable gives the next direction and the
 * throughput is multiplied by the attenuation: LAMBERTIAN cosine-weighted
 * around the face normal, times diffuse; METAL the mirror direction blended
 * with a cosine lobe by roughness, times specular, absorbed when it points
 * below the face; DIELECTRIC Schlick-weighted reflection or Snell
 * refraction with ior (1/ior leaving the voxel), times 1; CONSTANT_MEDIUM a
 * uniform direction, times diffuse.  A voxel whose material index is out of
 * range absorbs.  image: device pointer, views one after another,
 * [H][W][3].  The materials are packed and kept in the accel: a later call
 * may pass materials = NULL (and material_count 0) to reuse them.  Calls
 * are ordered on `stream` (null: the accel's stream). */
int ct_path_trace(ct_accel *accel, const ct_adjoint_desc *desc, const ct_adjoint_view *views,
                  const ct_material *materials, uint32_t material_count, float *image, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CT_ABI_H */
