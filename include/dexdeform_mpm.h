/* dexdeform_mpm.h — C ABI of the MI355X (gfx950) MLS-MPM substep library
 * libdexdeform_mpm.so.
 *
 * Two groups of entry points:
 *
 *  1. Per-stage launchers with the names, argument order and meaning of the
 *     reference's extern "C" launchers in mpm/csrc/integrator.cu (bound from Python by
 *     ctypes in mpm/types.py:103-290). Pointers are device pointers to FP32 AoS arrays
 *     (vec3 = 3 floats, mat3 = 9 floats row-major, quat = 4 floats (w,x,y,z)); an
 *     `ivec3 const&` becomes `const int *` to 3 host ints (the same bits a ctypes
 *     POINTER(ivec3) passes); `cudaStream_t` becomes hipStream_t. Unlike the reference
 *     (which returns void and never checks launches, common.h:20-25) every function
 *     returns 0 on success or a nonzero hipError_t / DDM_E* code.
 *
 *  2. Fused, multi-scene entry points (ddm_*) used by the Python simulator: one launch
 *     per stage for all scenes of a batch, block-sparse grids in a 4x4x4 blocked
 *     layout, LDS-tiled scatters. They have no single counterpart in the reference:
 *     ddm_substep_forward replaces the sequence MPMSimulator.substep
 *     (mpm/simulator.py:561-571) and ddm_substep_backward replaces
 *     MPMSimulator.substep_grad (mpm/simulator.py:573-585), for every scene at once.
 *
 * Groups 1 to 4 enqueue all work on the given stream and neither synchronise the device
 * nor allocate device memory, so their sequences can be captured in a HIP graph (group 0
 * allocates, copies and synchronises, as the reference's helpers do).
 */
#ifndef DEXDEFORM_MPM_H
#define DEXDEFORM_MPM_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes beyond hipError_t */
#define DDM_EINVAL 10001   /* bad argument (null pointer, size, dims not multiple of 4) */
#define DDM_ETOOMANYBODIES 10002
#define DDM_ESCRATCH 10003 /* launch addresses more grid blocks than the fixed workspace of
                              ddm_substep_forward_scratch holds (DDM_SCRATCH_BLOCKS) */

/* ---------------------------------------------------------------- group 0
 * Device memory, copy and stream helpers under the names the reference's ctypes layer
 * binds (mpm/types.py:103-128), so that its `array` class and MPMSimulator's streams run
 * on this library unchanged. Each replaces the integrator.cu function of the same name
 * (integrator.cu:1987-2072). Where the reference returns void these return 0 or the
 * failing hipError_t; cuda_alloc and cuda_stream_create return NULL on failure. */

/* replaces cuda_alloc, integrator.cu:1990-1995 */
void *cuda_alloc(size_t size);
/* replaces cuda_free, integrator.cu:2011 */
int cuda_free(void *ptr);
/* replaces print_memory_info, integrator.cu:1997-2009 (called by mpm/simulator.py:226) */
void print_memory_info(void);
/* replaces cuda_stream_create, integrator.cu:2013-2018 (default priority) */
hipStream_t cuda_stream_create(void);
/* replaces cuda_stream_destroy, integrator.cu:2020-2023 */
int cuda_stream_destroy(hipStream_t stream);
/* replaces cuda_stream_sync, integrator.cu:2069-2072 */
int cuda_stream_sync(hipStream_t stream);
/* replaces cuda_upload, integrator.cu:2025-2028 (synchronous host -> device) */
int cuda_upload(void *device_ptr, void *host_ptr, size_t size);
/* replaces cuda_download, integrator.cu:2030-2033 (synchronous device -> host) */
int cuda_download(void *host_ptr, void *device_ptr, size_t size);
/* replaces cuda_copy, integrator.cu:2035-2038 (device -> device) */
int cuda_copy(void *dst, void *src, size_t size);
/* replaces cuda_copy2d, integrator.cu:2040-2045 (device -> device, pitched) */
int cuda_copy2d(void *dst, size_t dpitch, void *src, size_t spitch, size_t width, size_t height);
/* replaces cuda_zero, integrator.cu:2047 */
int cuda_zero(void *ptr, size_t size);
/* replaces cuda_zero_async, integrator.cu:2048 */
int cuda_zero_async(void *ptr, size_t size, hipStream_t stream);
/* replaces cuda_upload_async, integrator.cu:2050-2055 */
int cuda_upload_async(void *device_ptr, void *host_ptr, size_t size, hipStream_t stream);
/* replaces cuda_download_async, integrator.cu:2057-2062 */
int cuda_download_async(void *host_ptr, void *device_ptr, size_t size, hipStream_t stream);
/* replaces cuda_copy_async, integrator.cu:2064-2067 */
int cuda_copy_async(void *dst, void *src, size_t size, hipStream_t stream);

/* ---------------------------------------------------------------- group 1
 * Reference-shaped launchers. n / dim = number of particles. */

/* replaces compute_grid_lower, integrator.cu:70-82, 1619-1625: min-reduces
 * floor((x - 10 dx) / dx) over the particles into the 3 ints at grid_lower (device) */
int compute_grid_lower(float *particle_x, float dx, float inv_dx, int *grid_lower, int dim, hipStream_t stream);

/* replaces compute_svd, integrator.cu:1627-1638 */
int compute_svd(float *F, float *C, float *newF, float *U, float *V, float *sig, float dt, int dim,
                hipStream_t stream);

/* replaces compute_svd_grad, integrator.cu:1640-1657 */
int compute_svd_grad(float *F, float *C, float *U, float *V, float *sig, float *newF_grad, float *U_grad,
                     float *V_grad, float *sig_grad, float *F_grad, float *C_grad, float dt, int dim,
                     hipStream_t stream);

/* replaces p2g, integrator.cu:1659-1690 */
int p2g(float *particle_x, float *particle_v, float *particle_m, float *particle_vol, float *particle_F,
        float *particle_U, float *particle_sig, float *particle_V, float *particle_C, float *particle_mu_lam_yield,
        int *grid_lower, const int *grid_dim, float dx, float inv_dx, float dt, float *out_particle_F,
        float *out_grid_mv, float *out_grid_m, int dim, hipStream_t stream);

/* replaces p2g_grad, integrator.cu:1692-1742 */
int p2g_grad(float *particle_x, float *particle_v, float *particle_m, float *particle_vol, float *particle_F,
             float *particle_U, float *particle_sig, float *particle_V, float *particle_C,
             float *particle_mu_lam_yield, int *grid_lower, const int *grid_dim, float dx, float inv_dx, float dt,
             float *out_particle_F, float *out_grid_mv, float *out_grid_m, float *particle_x_grad,
             float *particle_v_grad, float *particle_F_grad, float *particle_C_grad, float *particle_U_grad,
             float *particle_sig_grad, float *particle_V_grad, float *out_particle_F_grad, float *out_grid_v_grad,
             float *out_grid_m_grad, int dim, hipStream_t stream);

/* replaces grid_op_v2, integrator.cu:1754-1779 (grid_body_v_in may be NULL: no tape) */
int grid_op_v2(float *grid_m, float *grid_v_in, float *grid_body_v_in, int *grid_lower, float *gravity,
               float *body_pos, float *body_rot, float *next_body_pos, float *next_body_rot,
               float *body_type_friction_softness_round, float *body_args, float dx, float inv_dx, float dt,
               float ground_friction, float *out_grid_v, const int *grid_dim, int n_bodies, hipStream_t stream);

/* replaces grid_op_v2_grad, integrator.cu:1781-1819 (reads the tape written by grid_op_v2) */
int grid_op_v2_grad(float *grid_m, float *grid_v_in, float *grid_body_v_in, int *grid_lower, float *gravity,
                    float *body_pos, float *body_rot, float *next_body_pos, float *next_body_rot,
                    float *body_type_friction_softness_round, float *body_args, float *grid_m_grad,
                    float *grid_v_in_grad, float *body_pos_grad, float *body_rot_grad, float *next_body_pos_grad,
                    float *next_body_rot_grad, float dx, float inv_dx, float dt, float ground_friction,
                    float *out_grid_v, float *out_grid_v_grad, const int *grid_dim, int n_bodies,
                    hipStream_t stream);

/* replaces g2p, integrator.cu:1821-1835 */
int g2p(float *particle_x, float *grid_v, int *grid_lower, float dx, float inv_dx, float dt, const int *grid_dim,
        float *out_particle_v, float ground_height, float *out_particle_C, float *out_particle_x, int dim,
        hipStream_t stream);

/* replaces g2p_grad, integrator.cu:1837-1861 */
int g2p_grad(float *particle_x, float *grid_v, int *grid_lower, float dx, float inv_dx, float dt,
             const int *grid_dim, float *out_particle_v, float ground_height, float *out_particle_C,
             float *out_particle_x, int dim, float *particle_x_grad, float *grid_v_grad, float *out_particle_v_grad,
             float *out_particle_C_grad, float *out_particle_x_grad, hipStream_t stream);

/* replaces compute_dist, integrator.cu:1935-1960 */
int compute_dist(float *particle_x, float *body_pos, float *body_rot, float *body_type_friction_softness_round,
                 float *body_args, float *dist, int n_bodies, float *particle_x_grad, float *body_pos_grad,
                 float *body_rot_grad, float *dist_grad, int compute_grad, int dim, hipStream_t stream);

/* replaces particle2mass, integrator.cu:1962-1984 */
int particle2mass(float *particle_x, float *particle_m, int *grid_lower, const int *grid_dim, float dx,
                  float inv_dx, float *out_grid_m, float *out_grid_m_grad, float *particle_x_grad,
                  int *particle_ids, int id, int compute_grad, int dim, hipStream_t stream);

/* ---------------------------------------------------------------- group 2
 * Multi-scene fused substep. Scenes are independent: each has its own particle
 * arrays, material, primitives and poses, addressed through per-scene records in
 * device memory, so a launch can mix scenes of different sizes, grids and rollout
 * phases. Grid work arrays are shared arenas of 4x4x4 node blocks; each scene owns
 * n_blocks consecutive blocks starting at blk_begin. */

#define DDM_MAX_BODIES 64
#define DDM_GRAD_OVERWRITE 1  /* ddm_batch.options */
#define DDM_DETERMINISTIC 2   /* ddm_batch.options; needs ddm_batch.det */
#define DDM_TAPE_COMPLETE 4   /* ddm_batch.options, ddm_substep_backward_taped only: the caller
                                 guarantees that every launched scene keeps a snapshot
                                 (tape_mmv set) whose tape_capacity is at least the number of
                                 blocks of its grid, so no snapshot can overflow and the
                                 recompute kernels of overflowing scenes are not launched
                                 (ignored with DDM_DETERMINISTIC, whose adjoint reads the
                                 particle order its recompute pass builds) */

typedef struct ddm_scene {
  int32_t p_count;             /* particles of the scene (chunks define the work) */
  int32_t blk_begin, n_blocks; /* node blocks of the scene in the grid arenas */
  int32_t dim[3];              /* grid nodes per axis, multiples of 4 */
  int32_t n_bodies;            /* <= DDM_MAX_BODIES */
  int32_t adjoint_mode;        /* 0 = reference adjoint, 1 = exact (finite-difference consistent) */
  float dx, inv_dx, dt, ground_friction;
  float ground_height;         /* in cells (3 in the reference, mpm/simulator.py:493-494) */
  float gravity[3];            /* as seen by the kernel: configured gravity * 30 (mpm/simulator.py:385) */
  const float *mass, *vol;     /* [p_count] */
  const float *mu_lam_yield;   /* [p_count][3] */
  const float *body_prim;      /* [n_bodies][4] (type, friction, softness, round) */
  const float *body_args;      /* [n_bodies][4] */
  float *tmp_qu, *tmp_qv, *tmp_sig; /* SVD of F_tmp = (I + dt C) F for the backward: U and V as
                                      unit quaternions (w, x, y, z) [p_count][4] each, sigma
                                      [p_count][3] (F_tmp is recomputed from F and C) */
  /* Grid snapshot of this scene for one substep, used only by the taped entry points
   * (no counterpart in the reference, whose backward always recomputes P2G and the grid
   * update, simulator.py:573-585). NULL tape_mmv = no snapshot. */
  float *tape_mmv;             /* [tape_capacity][64][4] (m, mv) of the scene's listed blocks */
  float *tape_v;               /* [tape_capacity][64][4] node velocity after the grid update */
  int32_t *tape_list;          /* [tape_capacity] block indices */
  int32_t *tape_status;        /* [2]: [0] blocks the scene listed (taped forward); [1] 1 if the
                                  taped backward had to recompute the scene, else 0 */
  int32_t tape_capacity;       /* blocks the snapshot holds */
  int32_t reserved;
} ddm_scene;                   /* 168 bytes */

/* Workspace of the deterministic sorted-reduction mode (ddm_batch.options bit
 * DDM_DETERMINISTIC; no counterpart in the reference, whose atomics make every run's
 * rounding differ). Each substep bins the launch's particles by stencil-base node
 * (counting sort, particles of a node in ascending particle index), and every grid node
 * sums its particles' contributions in that fixed order (27 stencil offsets, then the
 * node's particles) instead of through atomics; body-pose gradients are reduced per node
 * block and then over blocks in block order. Results are then bitwise reproducible, and
 * identical for a scene alone and in any batch. "Launch slot" = chunk * 256 + position
 * in the chunk. */
typedef struct ddm_det {
  int32_t *cell_count;         /* [64*n_blocks] particles per stencil-base node; zero between
                                  substeps (ddm_batch_reset clears it) */
  int32_t *cell_start;         /* [64*n_blocks] first sorted position of the node's particles */
  int32_t *order;              /* [256*n_chunks] launch slots in arrival order within each node */
  int32_t *order_p;            /* [256*n_chunks] particle index of each of those entries */
  int32_t *key_slot;           /* [256*n_chunks][2] per launch slot: base node, then the arrival
                                  rank and finally the sorted position */
  float *records;              /* [256*n_chunks][16] per-particle stencil record per launch slot */
  float *sorted;               /* [256*n_chunks][16] the records in sorted order (gathered) */
  float *body_partial;         /* [n_blocks][max_bodies][16] per-block body-pose gradients */
  int32_t *range_count;        /* [(n_blocks + 1023) / 1024] flagged blocks per range */
  int32_t *counter;            /* [1] sorted-segment allocator; zero between substeps */
} ddm_det;                     /* 80 bytes */

typedef struct ddm_batch {
  int32_t n_scenes;
  int32_t n_chunks;            /* particle chunks of <= 256 particles, none straddling scenes */
  int32_t n_blocks;            /* node blocks addressable by this launch (every scene's
                                  blk_begin + n_blocks <= n_blocks <= arena size); sizes
                                  the grid kernels' launch */
  int32_t max_bodies;          /* max n_bodies over the scenes */
  const ddm_scene *scenes;     /* device [n_scenes] */
  const int32_t *chunks;       /* device [n_chunks][4] = (scene, first, end, 0), local particle indices */
  float *grid_mmv;             /* [64*n_blocks][4] (m, mv); zero between substeps */
  float *grid_v;               /* [64*n_blocks][4] node velocity after the grid update */
  float *grid_dv;              /* [64*n_blocks][4] adjoint of grid_v; zero between substeps */
  float *grid_dmmv;            /* [64*n_blocks][4] adjoint of (m, mv) */
  uint32_t *block_flags;       /* [n_blocks]; zero between substeps */
  int32_t *block_list;         /* [n_blocks][2] (block, scene index) of the blocks P2G flagged, in
                                  flagging order; the grid kernels visit exactly these */
  int32_t *block_count;        /* [1] entries in block_list; zero between substeps */
  const int32_t *live;         /* optional device [2] = (chunks, scenes) actually used, <= n_chunks
                                  and n_scenes: lets one captured launch of fixed size serve
                                  batches of any size (NULL = n_chunks, n_scenes) */
  const ddm_det *det;          /* workspace (host pointer to a record whose arrays are device
                                  memory) for DDM_DETERMINISTIC; NULL otherwise */
  int32_t options;             /* DDM_GRAD_OVERWRITE: the adjoint substep writes the current
                                  state's x/v/F/C gradients instead of adding to them;
                                  DDM_DETERMINISTIC: sorted reductions instead of atomics */
  int32_t reserved;
} ddm_batch;

/* Per-scene state record (device memory, one per scene, in scene order). */
typedef struct ddm_state {
  float *x, *v, *F, *C;        /* [p_count][3], [..][3], [..][9], [..][9] */
  float *body_pos, *body_rot;  /* [n_bodies][3], [n_bodies][4] */
} ddm_state;

/* One forward substep of every scene: cur[s] -> next[s] (next[s].body_* must hold the
 * poses of substep f+1). Writes next x, v, F, C. cur and next are DEVICE arrays of
 * n_scenes records. Replaces MPMSimulator.substep. */
int ddm_substep_forward(const ddm_batch *b, const ddm_state *cur, const ddm_state *next, hipStream_t stream);

/* Adjoint of one substep of every scene. Reads cur, next (state and poses) and the
 * next_grad x, v, F, C; ADDS into the cur_grad x, v, F, C, body_pos, body_rot and into
 * next_grad body_pos, body_rot (with b->options & DDM_GRAD_OVERWRITE it WRITES the cur_grad
 * x, v, F, C instead; pose gradients are still added). All four are DEVICE arrays of
 * n_scenes records. Replaces MPMSimulator.substep_grad. */
int ddm_substep_backward(const ddm_batch *b, const ddm_state *cur, const ddm_state *next, const ddm_state *cur_grad,
                         const ddm_state *next_grad, hipStream_t stream);

/* Instrumented variants: identical work, plus hipEventRecord on `stream` around each
 * stage kernel (events[0..3] for the forward: before P2G, after P2G, after the grid
 * update, after G2P; events[0..5] for the backward: before the P2G recompute, after
 * it, after the grid update, after the G2P adjoint, after the grid adjoint, after the
 * P2G adjoint). Used by the benchmark to time each kernel on the stream it runs on. */
int ddm_substep_forward_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next, hipStream_t stream,
                              hipEvent_t *events);
int ddm_substep_backward_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                               const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream,
                               hipEvent_t *events);

/* Taped substeps: the forward keeps what its adjoint needs, per scene, in the scene
 * records (give each taped substep its own scene table): the SVD temporaries in
 * tmp_qu/tmp_qv/tmp_sig and a snapshot of the scene's listed grid blocks in tape_*.
 * No counterpart in the reference; for callers with the memory to keep them (120 B per
 * particle plus 2 KiB per active block per substep). Scenes without a snapshot (NULL
 * tape_mmv) run as in ddm_substep_forward. */
int ddm_substep_forward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                              hipStream_t stream);

/* Adjoint substep from the snapshots recorded by ddm_substep_forward_taped with the same
 * scene table, states and poses: the grid is restored instead of recomputed. A scene
 * that listed more blocks than its snapshot holds (or has none) is recomputed instead,
 * decided on the device (no host synchronisation); its tape_status[1] is then 1. Same
 * results as ddm_substep_backward up to FP32 accumulation order. */
int ddm_substep_backward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                               const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream);

/* ddm_substep_forward (taped = 0) or ddm_substep_forward_taped (taped = 1, b's scene
 * table that of this substep) in a sequence of substeps (t-1, then t, ...): finish = 0
 * leaves this substep's G2P (cur -> next) to the next call, which passes prev = this
 * call's cur and runs that G2P first (prev -> its cur), inside its P2G launch where the
 * launch is small enough to scan block flags, else as its own launch. One launch fewer
 * per substep for small scenes. prev NULL and finish 1 is the plain substep; the last
 * substep of a sequence passes finish = 1. No counterpart in the reference; for callers
 * that run forward substeps in sequence, such as rollouts over a checkpoint segment. */
int ddm_substep_forward_ahead(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                              const ddm_state *prev, int finish, int taped, hipStream_t stream);

/* ddm_substep_backward_taped in a sequence of substeps (t, then t-1, ...): with
 * prev_scenes (the device scene table of substep t-1, recorded by its taped forward with
 * the same batch) the P2G adjoint launch of substep t also puts back substep t-1's
 * snapshots, in extra workgroups, and the call for substep t-1 then passes restored = 1
 * and launches no restore of its own: one launch fewer per substep. prev_scenes NULL and
 * restored 0 is ddm_substep_backward_taped. Not with DDM_DETERMINISTIC (DDM_EINVAL). No
 * counterpart in the reference; for callers that run the taped adjoint substep after
 * substep, such as rollouts over a checkpoint segment. */
int ddm_substep_backward_taped_ahead(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                     const ddm_state *cur_grad, const ddm_state *next_grad, int restored,
                                     const ddm_scene *prev_scenes, hipStream_t stream);

/* The taped substeps with hipEvents recorded between their stages (as the _timed
 * variants above): forward events[0..3] around P2G, grid update, G2P; backward events[0..4]
 * around restore (+ recompute where needed), G2P adjoint, grid adjoint, P2G adjoint. */
int ddm_substep_forward_taped_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                    hipStream_t stream, hipEvent_t *events);
int ddm_substep_backward_taped_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                     const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream,
                                     hipEvent_t *events);

/* Allocation-free forward substep. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same work and results as ddm_substep_forward, but the
 * grid workspace (the batch's grid_mmv, grid_v, block_flags, block_list, block_count,
 * which are ignored and may be NULL) lives in a fixed array inside the library, one per
 * GPU, held until the process ends. The caller still provides the scene, chunk and state
 * tables and the particle arrays. Limits: b->n_blocks <= DDM_SCRATCH_BLOCKS (8192 node
 * blocks of 4x4x4, e.g. two 64^3 grids; 16.1 MiB per GPU), else DDM_ESCRATCH; the
 * deterministic mode is rejected (DDM_EINVAL). Each call clears the part of the workspace
 * it uses with fill kernels on `stream` and allocates nothing. Every call on one GPU shares
 * that GPU's workspace: calls that overlap on one GPU (different streams, or host threads)
 * race and give wrong results; serialise them on one stream. */
#define DDM_SCRATCH_BLOCKS 8192
int ddm_substep_forward_scratch(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                hipStream_t stream);

/* ---------------------------------------------------------------- group 3
 * Point-cloud scores between simulated particles and a target shape, computed without
 * forming the n x m cost matrix. No counterpart in the reference's library: the reference
 * scores shapes with the geomloss package (SamplesLoss("sinkhorn", p=1, blur=1e-4),
 * tutorials/4_computing_score.ipynb cells 1-7), whose Sinkhorn iterations reduce over all
 * pairs with these two operations; the Chamfer distance uses the nearest-neighbour one.
 * Points are float [n][3]; C(x, y) = |x - y| (p = 1) or |x - y|^2 / 2 (p = 2). The j range
 * is split into `segments` pieces (more workgroups for small n); `workspace` holds
 * ddm_pairwise_workspace_bytes(n, segments) bytes of per-segment partial results, combined
 * in a fixed order (results are deterministic). */
size_t ddm_pairwise_workspace_bytes(int n, int segments);
/* out[i] = -eps * log(sum_j exp(h[j] - C(x_i, y_j) / eps)) */
int ddm_softmin(const float *x, int n, const float *y, const float *h, int m, float eps, int p, int segments,
                void *workspace, size_t workspace_bytes, float *out, hipStream_t stream);
/* grad[i] = sum_j exp(a[i] + b[j] - C(x_i, y_j) / eps) * dC/dx(x_i, y_j)  (float [n][3]) */
int ddm_plan_grad(const float *x, const float *a, int n, const float *y, const float *b, int m, float eps, int p,
                  int segments, void *workspace, size_t workspace_bytes, float *grad, hipStream_t stream);
/* dist2[i] = min_j |x_i - y_j|^2, index[i] = its j (the smallest on ties) */
int ddm_nearest(const float *x, int n, const float *y, int m, int segments, void *workspace, size_t workspace_bytes,
                float *dist2, int32_t *index, hipStream_t stream);

/* ---------------------------------------------------------------- group 4
 * Renderer: the reference's launchers particle_sdf (integrator.cu:1908-1933) and render
 * (integrator.cu:1863-1906), bound by mpm/types.py:284-286 and called by
 * Renderer.bake_particles / Renderer.render (mpm/renderer.py:89-169). Same names,
 * argument order and meaning; pointers are device pointers (vec3 = 3 floats, mat3 = 9
 * floats row-major, quat = 4 floats (w, x, y, z)), `ivec3 const&` becomes `const int *`
 * to 3 host ints. Unlike the reference's render, which cudaMallocs and frees one curand
 * state per pixel on every call, render allocates nothing: its random numbers are a
 * counter-based Philox4x32-10 stream per (pixel, sample) keyed by `seed`, so its images
 * are not the reference's bit for bit (different random numbers; same estimator).
 * n_bodies <= DDM_MAX_BODIES. */
/* volume: int64 per voxel, filled with (1 << 33) - 1 by the caller; bbox_max unused (as in
 * the reference); sdf, sdf_tmp: float per voxel; color: vec3 per voxel. */
int particle_sdf(long long *volume, float *particle_x, int *particle_color, float *bbox_min, float *bbox_max,
                 int bake_size, const int *grid_dim, float inv_dx, float *sdf, float *color, float *sdf_tmp,
                 int n_particles, hipStream_t stream);
/* color_buffer: vec3 per pixel, depth_buffer: float per pixel, pixel p = row * W + column
 * with image_dim = (H, W, 3); both are overwritten. */
int render(float *sdf_volume, float *box_min, float *box_max, float *color_volume, float *body_pos, float *body_rot,
           float *body_type_friction_softness_round, float *body_args, float *camera_rot, float *camera_pos,
           float *camera_intrinsic, float *color_buffer, float *depth_buffer, float sdf_threshold,
           const int *grid_dim, int n_bodies, bool visualize_shape, const int *image_dim, int max_ray_depth, int spp,
           float ground_height, int seed, float *light_direction, hipStream_t stream);

/* create_volume / destroy_volume (integrator.cu:2074-2124, bound by mpm/types.py:121-124):
 * the reference uploads a float4 host volume into a 3-D CUDA array with a linearly
 * filtered texture object; nothing in the reference samples it. gfx950 has no texture
 * objects (hipErrorNotSupported), so `volume` is a plain device copy (float4 per voxel,
 * x fastest, then y, then z) and `texture` is 0. Both allocate/free, as the reference's
 * do; on failure create_volume returns a null volume; destroy_volume returns 0 or the
 * hipError_t of the free. */
typedef struct {
  void *volume;
  unsigned long long texture;
} ddm_texture_resources; /* 16 bytes: the layout of types.py's texture_resources */
ddm_texture_resources create_volume(float *data, int x, int y, int z);
int destroy_volume(ddm_texture_resources tex);

/* Zeroes the batch's grid arenas and block flags (after an error or on setup). */
int ddm_batch_reset(const ddm_batch *b, hipStream_t stream);

/* Library version (major*10000 + minor*100 + patch). */
int ddm_version(void);

#ifdef __cplusplus
}
#endif
#endif
