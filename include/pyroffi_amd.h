/*
 * pyroffi_amd — C ABI of the MI355X (gfx950) batched-kinematics kernels.
 *
 * Every entry point takes a HIP stream (as `void*`, i.e. a `hipStream_t`),
 * device pointers that the caller owns, and plain sizes.  Nothing here
 * allocates, frees or synchronises: outputs are preallocated by the caller,
 * work is enqueued on `stream`, and the call returns as soon as the launch is
 * queued.  This is the ownership model of an XLA FFI custom call (XLA owns all
 * buffers and the stream), so each function below can sit behind a ~20-line
 * FFI handler that forwards `Buffer::typed_data()` pointers and dimensions.
 *
 * Conventions
 *   - poses are float32 [w, x, y, z, tx, ty, tz] ("wxyz_xyz", Hamilton
 *     quaternion then translation);
 *   - twists are float32 [vx, vy, vz, wx, wy, wz] (linear part first);
 *   - index tables are int32; -1 means "none" (root parent, fixed joint ...);
 *   - signed distances are positive when separated, negative in penetration.
 *
 * Return value: PK_OK, or PK_ERR_INVALID / PK_ERR_HIP with a message that
 * pk_last_error() returns (thread-local).  An empty batch is a successful
 * no-op, as in the reference handlers.
 *
 * Robot tables ("the 8 robot buffers") are the reference's JointInfo arrays,
 * cast to float32/int32 (src/pyroffi/_robot_urdf_parser.py:15-68):
 *   twists (J,6), parent_tf (J,7), parent_idx (J), act_idx (J),
 *   mimic_mul (J), mimic_off (J), mimic_act_idx (J), topo_inv (J).
 */
#ifndef PYROFFI_AMD_H_
#define PYROFFI_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PK_OK 0
#define PK_ERR_INVALID 1
#define PK_ERR_HIP 2

#define PK_MAX_JOINTS 64
#define PK_MAX_ACT 16
#define PK_MAX_EE 4
#define PK_MAX_GRIDS 2

/*
 * A world grid built by pk_world_grid_f32 (device arrays + its geometry),
 * passed by host pointer to the LM entry points.
 */
typedef struct pk_world_grid {
  const int32_t* cells;  /* (dims[0]*dims[1]*dims[2], 2): start, count in items */
  const int32_t* items;  /* world sphere indices, increasing within a cell      */
  float lo[3];           /* corner of cell (0, 0, 0)                            */
  float cell;            /* cell side                                           */
  int32_t dims[3];       /* nx, ny, nz                                          */
  float pad;             /* reach the grid was built with                       */
} pk_world_grid;

/* Message of the last failing call on this thread ("" when none). */
const char* pk_last_error(void);

/* ABI version; bumped whenever a signature below changes. */
int32_t pk_abi_version(void);

/*
 * Forward kinematics: world pose of every joint.
 *
 * Replaces the XLA FFI target FkCudaFfi
 * (src/pyroffi/cuda_kernels/_fk_cuda_kernel.cu:456-471, bound in
 * src/pyroffi/cuda_kernels/_fk_cuda.py:109-124) with the same buffers in the
 * same order.  fk_level_starts/fk_level_joints are accepted for signature
 * parity and not read: one lane evaluates one configuration's whole chain in
 * topological order (topo_inv), so the depth schedule is not needed.
 *
 *   cfg        (batch, n_act)        float32
 *   out        (batch, n_joints, 7)  float32, joint order of the tables
 * Errors: n_joints > PK_MAX_JOINTS or n_act > PK_MAX_ACT -> PK_ERR_INVALID.
 */
int pk_fk_f32(void* stream, const float* cfg, int64_t batch, int32_t n_act,
              const float* twists, const float* parent_tf,
              const int32_t* parent_idx, const int32_t* act_idx,
              const float* mimic_mul, const float* mimic_off,
              const int32_t* mimic_act_idx, const int32_t* topo_inv,
              const int32_t* fk_level_starts, const int32_t* fk_level_joints,
              int32_t n_joints, int32_t n_levels, float* out);

/*
 * Forward kinematics straight to link poses: Robot.forward_kinematics
 * (src/pyroffi/_robot.py:65-109) = FkCudaFfi followed by the link gather of
 * _link_poses_from_joint_poses (src/pyroffi/_robot.py:180-196), fused so the
 * joint poses never reach HBM.  link_parent_joint (n_links) int32 is
 * LinkInfo.parent_joint_indices; links whose entry is -1 get the identity.
 *
 *   out_links  (batch, n_links, 7)   float32
 */
int pk_fk_links_f32(void* stream, const float* cfg, int64_t batch, int32_t n_act,
                    const float* twists, const float* parent_tf,
                    const int32_t* parent_idx, const int32_t* act_idx,
                    const float* mimic_mul, const float* mimic_off,
                    const int32_t* mimic_act_idx, const int32_t* topo_inv,
                    int32_t n_joints, const int32_t* link_parent_joint,
                    int32_t n_links, float* out_links);

/*
 * Adjoint of pk_fk_links_f32: the vector-Jacobian product of the link poses.
 * Given a cotangent on every wxyz_xyz link pose, returns dq = J^T cot, the
 * gradient reverse-mode AD of Robot.forward_kinematics
 * (src/pyroffi/_robot.py:198-263, through _link_poses_from_joint_poses
 * :180-196) returns for the actuated configuration; mimic joints add their
 * multiplier times their own term to their source.  The reference has no
 * native kernel for it (JAX differentiates its FK); this is the backward of a
 * jax.custom_vjp over FkCudaFfi.  FK is recomputed from cfg, so the forward
 * needs to keep no residual but the configuration.
 *
 *   cfg        (batch, n_act)        float32
 *   cot_links  (batch, n_links, 7)   float32, cotangent of out_links
 *   out_dq     (batch, n_act)        float32
 */
int pk_fk_vjp_f32(void* stream, const float* cfg, int64_t batch, int32_t n_act,
                  const float* twists, const float* parent_tf,
                  const int32_t* parent_idx, const int32_t* act_idx,
                  const float* mimic_mul, const float* mimic_off,
                  const int32_t* mimic_act_idx, const int32_t* topo_inv,
                  int32_t n_joints, const int32_t* link_parent_joint,
                  int32_t n_links, const float* cot_links, float* out_dq);

/*
 * World collision per robot sphere: out[b, k, m] = signed distance of sphere k
 * (centre c[:, b, k], radius r[b, k]) to obstacle m of the packed world, 1e9
 * for a padding slot (r < 0).  Obstacles are indexed spheres first, then
 * capsules, boxes and half-spaces (M = Ms + Mc + Mb + Mh), as in the
 * reference.
 *
 * Replaces CollisionWorldSphereFfi
 * (src/pyroffi/cuda_kernels/_collision_cuda_kernel.cu:659-669, kernels
 * :64-161, bound in src/pyroffi/cuda_kernels/_collision_cuda_ffi.py:70-107).
 *
 *   sphere_centers   (3, batch, k) float32 SoA
 *   sphere_radii     (batch, k)    float32
 *   world_spheres    (Ms, 4)   [x, y, z, r]
 *   world_capsules   (Mc, 7)   [a_xyz, b_xyz, r]
 *   world_boxes      (Mb, 15)  [centre, axis1, axis2, axis3, half lengths]
 *   world_halfspaces (Mh, 6)   [unit normal, point]
 *   out              (batch, k, M) float32
 */
int pk_collision_world_sphere_f32(void* stream, const float* sphere_centers,
                                  const float* sphere_radii, int64_t batch, int32_t k,
                                  const float* world_spheres, int32_t n_world_spheres,
                                  const float* world_capsules, int32_t n_world_capsules,
                                  const float* world_boxes, int32_t n_world_boxes,
                                  const float* world_halfspaces, int32_t n_world_halfspaces,
                                  float* out);

/*
 * World collision, sphere robot, reduced over each link's S spheres:
 * out[b, n, m] = min_s dist(sphere (b, s*N+n), obstacle m).  Sphere slots with
 * r < 0 are padding and skipped; a link with no real sphere reports 1e9.
 * Obstacle kinds and their order as for pk_collision_world_sphere_f32; a
 * sphere-only world runs the fused sweep kernel, any other world the
 * mixed-kind kernel.
 *
 * Replaces CollisionWorldSphereReducedFfi
 * (src/pyroffi/cuda_kernels/_collision_cuda_kernel.cu:673-839, bound in
 * src/pyroffi/cuda_kernels/_collision_cuda_ffi.py:109-151).
 *
 *   sphere_centers (3, batch, k) float32 SoA,  k = S * n_links
 *   sphere_radii   (batch, k)    float32
 *   out            (batch, n_links, M) float32
 */
int pk_collision_world_sphere_reduced_f32(
    void* stream, const float* sphere_centers, const float* sphere_radii,
    int64_t batch, int32_t k, const float* world_spheres, int32_t n_world_spheres,
    const float* world_capsules, int32_t n_world_capsules,
    const float* world_boxes, int32_t n_world_boxes,
    const float* world_halfspaces, int32_t n_world_halfspaces,
    int32_t n_links, float* out);

/*
 * Self collision between link pairs: out[b, p] = min over all sphere pairs
 * (si of link pair_i[p], sj of link pair_j[p]) of the sphere distance, padding
 * (r < 0) skipped, 1e9 when either link has no real sphere.
 *
 * Replaces CollisionSelfSphereFfi
 * (src/pyroffi/cuda_kernels/_collision_cuda_kernel.cu:1006-1110, bound in
 * src/pyroffi/cuda_kernels/_collision_cuda_ffi.py:191-216).
 *
 *   centers (batch, S, N, 3) float32,  radii (batch, S, N) float32
 *   pair_i, pair_j (n_pairs) int32,    out (batch, n_pairs) float32
 */
int pk_collision_self_sphere_f32(void* stream, const float* centers,
                                 const float* radii, int64_t batch, int32_t s,
                                 int32_t n_links, const int32_t* pair_i,
                                 const int32_t* pair_j, int32_t n_pairs,
                                 float* out);

/*
 * World distances of a spherized robot from link poses over the four
 * obstacle kinds, with the reference CUDA kernels' semantics (padding slots
 * skipped, 1e9 for a link without spheres): out_world[b, l, m] = min over
 * link l's spheres of the distance to obstacle m, out_world_min[b] (or NULL)
 * = its minimum over (l, m).  The world path of CUDARobotCollisionChecker for
 * worlds with capsules, boxes or half-spaces (the fused sweep serves sphere
 * worlds); the reference reaches it through CollisionWorldSphereReducedFfi
 * after placing the spheres (src/pyroffi/collision/_cuda_collision.py:500-540).
 *   link_poses (batch, n_links, 7), spheres_padded (n_links, s_per_link, 4)
 *   out_world (batch, n_links, M) required, out_world_min (batch) or NULL
 */
int pk_sphere_world_links_f32(void* stream, const float* link_poses, int64_t batch, int32_t n_links,
                              const float* spheres_padded, int32_t s_per_link,
                              const float* world_spheres, int32_t n_world_spheres,
                              const float* world_capsules, int32_t n_world_capsules,
                              const float* world_boxes, int32_t n_world_boxes,
                              const float* world_halfspaces, int32_t n_world_halfspaces,
                              float* out_world, float* out_world_min);

/*
 * Capsule robot at a configuration: link-frame capsules capsules_local
 * (n_links, 7) = [a, b, r] moved by link_poses (batch, n_links, 7), written as
 * out_soa [7, batch, n_links] and/or out_aos [batch, n_links, 7] (either may be
 * NULL, not both).  Replaces RobotCollision.at_config followed by the
 * checker's packing (src/pyroffi/collision/_robot_collision.py:323-345,
 * src/pyroffi/collision/_cuda_collision.py:443-474).
 */
int pk_capsules_place_f32(void* stream, const float* link_poses, int64_t batch, int32_t n_links,
                          const float* capsules_local, float* out_soa, float* out_aos);

/*
 * out[b] = min over the row_len floats of row b of dist (batch, row_len), 1e9
 * for an empty row: the per-configuration minimum the reference's coarse
 * guard takes with jnp.all(coarse_out > 0) over a capsule model's distances
 * (src/pyroffi/collision/_cuda_collision.py:560-577).
 */
int pk_row_min_f32(void* stream, const float* dist, int64_t batch, int64_t row_len, float* out);

/*
 * World collision of a capsule robot: out[b, n, m] = distance of link
 * capsule n (segment caps[0..2], caps[3..5], radius caps[6], component-major
 * (7, batch, n_links)) to obstacle m of the packed world (kinds and order as
 * for pk_collision_world_sphere_f32).  Replaces CollisionWorldCapsuleFfi
 * (src/pyroffi/cuda_kernels/_collision_cuda_kernel.cu:993-1002, kernels
 * :203-407): the capsule-box distance is the box SDF at the segment point
 * nearest the box centre, as there.
 *   out (batch, n_links, M) float32
 */
int pk_collision_world_capsule_f32(void* stream, const float* caps, int64_t batch, int32_t n_links,
                                   const float* world_spheres, int32_t n_world_spheres,
                                   const float* world_capsules, int32_t n_world_capsules,
                                   const float* world_boxes, int32_t n_world_boxes,
                                   const float* world_halfspaces, int32_t n_world_halfspaces,
                                   float* out);

/*
 * Self collision of a capsule robot: out[b, p] = capsule-capsule distance of
 * links pair_i[p] and pair_j[p] from AoS capsules (batch, n_links, 7), 1e9 for
 * an out-of-range link index.  Replaces CollisionSelfCapsuleFfi
 * (src/pyroffi/cuda_kernels/_collision_cuda_kernel.cu:1114-1214, kernel :474-495).
 */
int pk_collision_self_capsule_f32(void* stream, const float* caps, int64_t batch, int32_t n_links,
                                  const int32_t* pair_i, const int32_t* pair_j, int32_t n_pairs,
                                  float* out);

/*
 * The JAX path's distance conventions for a spherized robot, from link poses:
 *   out_world[b, l, m] = min over slots s of |(o_m - c_ls) + 1e-6| - (r_ls + R_m)
 *     (RobotCollisionSpherized.compute_world_collision_distance,
 *     src/pyroffi/collision/_robot_collision.py:830-895)
 *   out_self[b, p]     = min over slots s of |(c_js - c_is) + 1e-6| - (r_is + r_js)
 *     for the pair (i, j) = (pair_i[p], pair_j[p]): same-slot ("diagonal")
 *     sphere pairs only (compute_self_collision_distance, :796-828)
 * with the 1e-6 added to every component inside the norm
 * (src/pyroffi/collision/_utils.py:45).  Padding slots (r = -1e9) take part
 * as in JAX and read about +1e9.  Either output may be NULL.  The reference
 * computes these with JAX only; there is no FFI target to replace.
 *
 *   link_poses      (batch, n_links, 7)        float32
 *   spheres_padded  (n_links, s_per_link, 4)   link-frame [x, y, z, r], r = -1e9 padding
 *   world_spheres   (M, 4)
 *   out_world       (batch, n_links, M),  out_self (batch, n_pairs)
 */
int pk_sphere_distances_jax_f32(void* stream, const float* link_poses, int64_t batch,
                                int32_t n_links, const float* spheres_padded,
                                int32_t s_per_link, const float* world_spheres,
                                int32_t n_world_spheres, const int32_t* pair_i,
                                const int32_t* pair_j, int32_t n_pairs, float* out_world,
                                float* out_self);

/*
 * Fused sphere sweep from link poses (no reference counterpart as one call:
 * it fuses what CUDARobotCollisionChecker does in three steps -- the vmapped
 * sphere transform, AoS->SoA transposes and the two distance kernels
 * (src/pyroffi/collision/_cuda_collision.py:358-540) -- for the spherized
 * model of src/pyroffi/collision/_robot_collision.py:445-549).
 *
 *   link_poses        (batch, n_links, 7) float32   (pk_fk_links_f32 output)
 *   spheres_local     (n_spheres, 4) float32, [x, y, z, r] in the frame of
 *                     their link, grouped by link (padding removed); adjacent
 *                     spheres of one link with equal radius share one square
 *                     root per obstacle, so order each link's spheres by radius
 *   link_sphere_start (n_links + 1) int32: spheres of link n are
 *                     [link_sphere_start[n], link_sphere_start[n+1])
 *   world_spheres     (Ms, 4) float32
 *   self_entries      (n_entries) uint32: one per sphere pair to test,
 *                     (a | b << 10 | p << 20): sphere indices a, b into
 *                     spheres_local and active pair p, entries of one
 *                     pair adjacent; n_spheres <= 1023 and n_pairs <= 2047
 *                     when n_entries > 0 (else PK_ERR_INVALID); entries out
 *                     of range are skipped
 *   outputs (any may be NULL):
 *     out_world       (batch, n_links, Ms)  per-link min over its spheres,
 *                     starting from 1e9 and skipping NaN distances as the
 *                     reference does  == compute_world_collision_distance
 *     out_world_min   (batch)               min over all links and obstacles
 *     out_self        (batch, n_pairs)      == compute_self_collision_distance
 *                     (full S_i x S_j minimum, the reference CUDA semantics)
 */
int pk_sphere_sweep_f32(void* stream, const float* link_poses, int64_t batch,
                        int32_t n_links, const float* spheres_local,
                        const int32_t* link_sphere_start, int32_t n_spheres,
                        const float* world_spheres, int32_t n_world_spheres,
                        const uint32_t* self_entries, int32_t n_entries,
                        int32_t n_pairs, float* out_world,
                        float* out_world_min, float* out_self);

/*
 * Pose + hinge collision cost and its gradient (hand-written adjoint).
 *
 *   cost[b] = sum_c (W_c r_c)^2 + sum_{k,m} max(0, margin - d_km)^2
 *   r = [p_ee - p_t; log(q_ee q_t^-1)] of link ee_link against targets[b]
 *   (ee_link = -1 drops the pose term), W = (pos_weight x3, ori_weight x3),
 *   d_km = |c_k - o_m| - r_k - R_m over robot spheres and world spheres.
 *   out_grad[b] = dcost/dq (n_act), computed by reverse accumulation of link
 *   wrenches over the joint tree -- no Jacobian is formed; NULL skips it.
 *
 * No single reference counterpart: the reference differentiates the CHOMP
 * hinge^2 cost (src/pyroffi/optimization_engines/_chomp_optimization.py:140-154)
 * with JAX AD and its CUDA trajopt kernels use finite differences.  This is
 * the backward of a custom VJP over FK + collision.
 *
 *   link_poses (batch, n_links, 7) from pk_fk_links_f32; spheres as in
 *   pk_sphere_sweep_f32; targets (batch, 7); joint tables as in pk_fk_f32 plus
 *   link_parent_joint (n_links).
 */
int pk_pose_hinge_cost_f32(void* stream, const float* link_poses, int64_t batch,
                           int32_t n_links, const float* spheres_local,
                           const int32_t* link_sphere_start, int32_t n_spheres,
                           const float* world_spheres, int32_t n_world_spheres,
                           float margin, int32_t ee_link, const float* targets,
                           float pos_weight, float ori_weight, const float* twists,
                           const int32_t* parent_idx, const int32_t* act_idx,
                           const int32_t* mimic_act_idx, const float* mimic_mul,
                           const int32_t* topo_inv, const int32_t* link_parent_joint,
                           int32_t n_joints, int32_t n_act, float* out_cost,
                           float* out_grad);

/*
 * Multi-seed Levenberg-Marquardt IK, one seed per lane.
 *
 * Replaces LsIkCudaFfi (src/pyroffi/cuda_kernels/_ls_ik_cuda_kernel.cu:400-529,
 * bound in src/pyroffi/cuda_kernels/_ls_ik_cuda.py:152-182): same buffers,
 * attributes and outputs, same algorithm (weighted geometric-Jacobian LM with
 * Jacobi column scaling, norm-clipped trust region, 5-alpha line search,
 * all-time-best tracking, merit-only collision penalty).  The damped normal
 * equations are solved in FP32 (the reference uses FP64).  The collision
 * penalty takes the four obstacle kinds of the reference kernel (spheres,
 * capsules, boxes, half-spaces; the non-sphere ones at most 8192 floats in
 * all).
 *
 *   seeds        (n_problems, n_seeds, n_act)
 *   target_jnts  (n_ee) int32, ancestor_masks (n_ee, n_joints) int32
 *   target_T     (n_problems, n_ee, 7)
 *   robot_spheres_local (n_robot_spheres, 4) in joint frames,
 *   robot_sphere_joint_idx (n_robot_spheres) int32
 *   lower, upper (n_act) float32, fixed_mask (n_act) int32
 *   out_cfgs (n_problems, n_seeds, n_act), out_err (n_problems, n_seeds)
 */
int pk_ls_ik_f32(void* stream, const float* seeds, int32_t n_problems,
                 int32_t n_seeds, int32_t n_act, const float* twists,
                 const float* parent_tf, const int32_t* parent_idx,
                 const int32_t* act_idx, const float* mimic_mul,
                 const float* mimic_off, const int32_t* mimic_act_idx,
                 const int32_t* topo_inv, int32_t n_joints,
                 const int32_t* target_jnts, const int32_t* ancestor_masks,
                 int32_t n_ee, const float* target_T,
                 const float* robot_spheres_local,
                 const int32_t* robot_sphere_joint_idx, int32_t n_robot_spheres,
                 const float* world_spheres, int32_t n_world_spheres,
                 const float* world_capsules, int32_t n_world_capsules,
                 const float* world_boxes, int32_t n_world_boxes,
                 const float* world_halfspaces, int32_t n_world_halfspaces,
                 const float* lower, const float* upper, const int32_t* fixed_mask,
                 int32_t max_iter, float pos_weight, float ori_weight,
                 float lambda_init, float eps_pos, float eps_ori,
                 int32_t enable_collision, float collision_weight,
                 float collision_margin, float* out_cfgs, float* out_err);

/*
 * Uniform grid over a sphere world for exact culling of margin-limited costs,
 * in CSR form.  Cell (ix, iy, iz) is the box lo + cell * [i, i+1) per axis and
 * lists every world sphere whose distance to the box is at most R_m + pad.
 * With pad = (robot sphere radius) + margin and the grid box = world bounding
 * box grown by pad, a robot sphere whose centre lies in a cell can be within
 * the margin only of spheres listed there, and one outside the grid of none.
 * No reference counterpart (an acceleration structure for the collision costs).
 * Two passes:
 *   pk_world_grid_count_f32 writes cells[2c + 1] = count of cell c;
 *   the caller writes cells[2c] = exclusive prefix sum of the counts and
 *   allocates n_items = total count indices;
 *   pk_world_grid_fill_f32 writes each cell's indices, in increasing order, to
 *   items[cells[2c] ...] (never past n_items).
 *   cells (nx*ny*nz, 2) int32, items (n_items) int32
 */
int pk_world_grid_count_f32(void* stream, const float* world_spheres, int32_t n_world_spheres,
                            float pad, float lo_x, float lo_y, float lo_z, float cell, int32_t nx,
                            int32_t ny, int32_t nz, int32_t* cells);
int pk_world_grid_fill_f32(void* stream, const float* world_spheres, int32_t n_world_spheres,
                           float pad, float lo_x, float lo_y, float lo_z, float cell, int32_t nx,
                           int32_t ny, int32_t nz, const int32_t* cells, int32_t* items,
                           int64_t n_items);

/*
 * Uneven batched LM-IK: many problems of one robot with different seed
 * counts in one persistent launch.  Same algorithm and buffers as
 * pk_ls_ik_f32, except
 *   seeds            (total_seeds, n_act), problem p owns the rows
 *                    [problem_seed_start[p], problem_seed_start[p+1])
 *   target_T         (n_problems, n_ee, 7)
 *   grids            0..PK_MAX_GRIDS world grids (pk_world_grid_f32) in
 *                    increasing pad order; a robot sphere of radius r uses
 *                    the first with pad - collision_margin >= r, and tests
 *                    every world sphere when none qualifies, so the penalty
 *                    is the same for any set of grids
 *   iter_begin/end   run iterations [iter_begin, iter_end) of max_iter; a
 *                    partial range reads/writes `state` (total_seeds,
 *                    2 n_act + 3), so a solve can be split into launches
 *                    (e.g. nodes of a captured graph); outputs are written
 *                    by the launch that reaches max_iter
 *   grid_waves       number of persistent workgroups (0: one per 64-seed tile)
 *   out_cfgs (total_seeds, n_act), out_err (total_seeds)
 *   work_counter     NULL, or a device uint64 the launch adds the number of
 *                    sphere-pair tests of its collision penalty to (the
 *                    work measure of the solver's roofline; the caller
 *                    clears it)
 * Replaces one ls_ik_solve_cuda_batch call per robot
 * (src/pyroffi/optimization_engines/_ls_ik.py:1002-1172), which needs the same
 * seed count for every problem, and the per-call path for uneven ones.
 */
int pk_ls_ik_batched_f32(
    void* stream, const float* seeds, const int32_t* problem_seed_start, int32_t n_problems,
    int64_t total_seeds, int32_t n_act, const float* twists, const float* parent_tf,
    const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
    const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv,
    int32_t n_joints, const int32_t* target_jnts, const int32_t* ancestor_masks, int32_t n_ee,
    const float* target_T, const float* robot_spheres_local,
    const int32_t* robot_sphere_joint_idx, int32_t n_robot_spheres, const float* world_spheres,
    int32_t n_world_spheres, const pk_world_grid* grids, int32_t n_grids, const float* lower,
    const float* upper, const int32_t* fixed_mask, int32_t max_iter, int32_t iter_begin,
    int32_t iter_end, float* state, float pos_weight, float ori_weight, float lambda_init,
    float eps_pos, float eps_ori, int32_t enable_collision, float collision_weight,
    float collision_margin, int32_t grid_waves, float* out_cfgs, float* out_err,
    uint64_t* work_counter);

/*
 * Winner selection after pk_ls_ik_batched_f32: per problem, over its CSR seed
 * range, the seed minimising err + continuity_weight * |cfg - previous|^2.
 */
int pk_ls_ik_select_segments_f32(void* stream, const float* cfgs, const float* errs,
                                 const float* previous, const int32_t* problem_seed_start,
                                 int32_t n_problems, int32_t n_act, float continuity_weight,
                                 float* out_best, int32_t* out_idx);

/*
 * Winner selection after pk_ls_ik_f32: per problem, the seed minimising
 * err + continuity_weight * |cfg - previous|^2 (first index on ties; a NaN
 * score ranks as +inf, where jnp.argmin would return the first NaN -- the
 * solver only produces NaN errors for a robot it rejects, all seeds alike).
 * Replaces the argmin of _ls_ik_solve_cuda_batch_jit
 * (src/pyroffi/optimization_engines/_ls_ik.py:974-999).
 *
 *   cfgs (n_problems, n_seeds, n_act), errs (n_problems, n_seeds),
 *   previous (n_problems, n_act)
 *   out_best (n_problems, n_act), out_idx (n_problems) int32 or NULL
 */
int pk_ls_ik_select_f32(void* stream, const float* cfgs, const float* errs,
                        const float* previous, int32_t n_problems,
                        int32_t n_seeds, int32_t n_act, float continuity_weight,
                        float* out_best, int32_t* out_idx);

/*
 * The reference's pure-JAX LM solvers, per seed, with its residual
 * f = log(T_ee^-1 T_tgt) (src/pyroffi/optimization_engines/_ik_primitives.py:93-103,
 * jaxlie SE3.log) and the exact Jacobian of it:
 *   adaptive_weights = 0: _ls_ik_single (src/pyroffi/optimization_engines/_ls_ik.py:192-350),
 *     weights W = (pos_weight x3, ori_weight x3); limit_prior_weight and
 *     kick_scale must be 0 to match it;
 *   adaptive_weights = 1: _lm_refine_single (src/pyroffi/optimization_engines/_hjcd_ik.py:140-365),
 *     per-EE adaptive weights, the soft joint-limit prior of weight
 *     limit_prior_weight, and after six non-improving steps a kick of
 *     kick_scale * N(0, 1) per free joint (counter-based normal keyed by
 *     kick_key and the seed's flat index p * n_seeds + s) with lambda reset.
 * The reference evaluates both inside jax.vmap over seeds; there is no FFI
 * target to replace.  Seeds, targets and the robot buffers are those of
 * pk_ls_ik_f32 (n_ee in [1, PK_MAX_EE]).
 *   out_cfgs (n_problems, n_seeds, n_act) the all-time best configuration per seed
 *   out_err  (n_problems, n_seeds) the winner metric at it: the sum over EEs of
 *            |W f|^2 (adaptive_weights = 0, ls_ik_solve :503-516) or |f|^2
 *            (adaptive_weights = 1, hjcd_solve _hjcd_ik.py:616-630)
 * A robot whose moving joints exceed the padded width gets NaN errors.
 */
int pk_ik_jax_lm_f32(void* stream, const float* seeds, int32_t n_problems, int32_t n_seeds,
                     int32_t n_act, const float* twists, const float* parent_tf,
                     const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
                     const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv,
                     int32_t n_joints, const int32_t* target_jnts, const int32_t* ancestor_masks,
                     int32_t n_ee, const float* target_T, const float* lower, const float* upper,
                     const int32_t* fixed_mask, int32_t max_iter, float pos_weight, float ori_weight,
                     float lambda_init, int32_t adaptive_weights, float limit_prior_weight,
                     float kick_scale, uint64_t kick_key, float* out_cfgs, float* out_err);

/*
 * HJCD-IK coarse phase: _coarse_search_single
 * (src/pyroffi/optimization_engines/_hjcd_ik.py:55-137) for every seed, k_max
 * greedy Hamiltonian coordinate-descent steps on end effector ee_of_seed[s]
 * (hjcd_solve splits the seeds equally over the EEs, :570-589), skipped once
 * |f_pos| < epsilon and |f_rot| < nu.
 *   out_cfgs (n_problems, n_seeds, n_act)
 *   out_err  (n_problems, n_seeds) the sum over all EEs of |f|^2 (:591-597)
 */
int pk_hjcd_coarse_f32(void* stream, const float* seeds, int32_t n_problems, int32_t n_seeds,
                       int32_t n_act, const float* twists, const float* parent_tf,
                       const int32_t* parent_idx, const int32_t* act_idx, const float* mimic_mul,
                       const float* mimic_off, const int32_t* mimic_act_idx, const int32_t* topo_inv,
                       int32_t n_joints, const int32_t* target_jnts, const int32_t* ancestor_masks,
                       int32_t n_ee, const float* target_T, const int32_t* ee_of_seed,
                       const float* lower, const float* upper, const int32_t* fixed_mask,
                       int32_t k_max, float epsilon, float nu, float* out_cfgs, float* out_err);

/*
 * HJCD-IK refinement seeds (src/pyroffi/optimization_engines/_hjcd_ik.py:599-611):
 * per problem the top_k coarse results by error (stable: ties by index, NaN
 * last), tiled repeats times; copy r >= top_k adds
 * noise[p, r] * 0.15 * (upper - lower) (zero on fixed joints) and is clipped.
 *   coarse_cfgs (n_problems, n_seeds, n_act), coarse_err (n_problems, n_seeds)
 *   noise       (n_problems, top_k * repeats, n_act) standard normal draws
 *   out_seeds   (n_problems, top_k * repeats, n_act)
 */
int pk_hjcd_refine_seeds_f32(void* stream, const float* coarse_cfgs, const float* coarse_err,
                             int32_t n_problems, int32_t n_seeds, int32_t n_act, int32_t top_k,
                             int32_t repeats, const float* noise, const float* lower,
                             const float* upper, const int32_t* fixed_mask, float* out_seeds);

/*
 * Trajectory model for the trajectory cost and optimiser (device pointers).
 * The robot is given with its fixed joints folded into the moving joints
 * ("nodes", topological order), as pyroffi_amd.trajopt builds it:
 *   nodes          (n_nodes, 20) float: parent-node -> joint transform before
 *                  motion (wxyz_xyz, 7), twist v (3), twist w (3), mimic
 *                  multiplier, mimic offset, actuated source index (-1: none),
 *                  parent node (-1: world), exp kind (1 rotation, 2
 *                  translation, 3 general), 2 unused
 *   link_node      (n_links) node carrying each link (-1: world-fixed); links
 *                  may be in any order (grouping them by node is faster)
 *   link_sph_start (n_links + 1) CSR of spheres per link, in the link's slot
 *                  order of the reference's padded sphere table
 *   spheres        (n_spheres, 4) centre in the node frame, radius
 *   link_bound     (n_links, 4) bounding sphere of each link's spheres (node frame)
 *   pair_links     (n_pairs, 2) active self pairs (link i, link j)
 *   pair_start     (n_pairs + 1) CSR into pair_items
 *   pair_items     (n_items, 2) sphere pairs of a link pair (same slot)
 *   world          (n_world, 4) world spheres
 *   grids          0..PK_MAX_GRIDS grids of the world (increasing pad; pad -
 *                  margin >= a robot sphere's radius selects a grid for it)
 *   lower, upper   (n_act) joint limits
 *   margin, mode   activation margin; mode 0 = max(0, margin - d)^2 (CHOMP),
 *                  1 = -min(colldist_from_sdf(d, margin), 0) (L-BFGS trajopt)
 */
typedef struct pk_traj_model {
  const float* nodes;
  int32_t n_nodes, n_act;
  const int32_t* link_node;
  const int32_t* link_sph_start;
  const float* spheres;
  const float* link_bound;
  int32_t n_links, n_spheres;
  const int32_t* pair_links;
  const int32_t* pair_start;
  const int32_t* pair_items;
  int32_t n_pairs, n_items;
  const float* world;
  int32_t n_world;
  const pk_world_grid* grids;
  int32_t n_grids;
  const float* lower;
  const float* upper;
  float margin;
  int32_t mode;
} pk_traj_model;

/* LbfgsTrajOptConfig (src/pyroffi/optimization_engines/_lbfgs_trajopt_optimization.py:36-70). */
typedef struct pk_lbfgs_config {
  int32_t n_iters, m_hist, escalation_interval;
  float w_smooth, w_vel, w_acc, w_jerk, w_limits;
  float w_collision, w_collision_max, penalty_scale;
} pk_lbfgs_config;

/*
 * Trajectory cost components and gradient, many trajectories of different
 * lengths in one launch: trajectory b is rows [traj_start[b], traj_start[b+1])
 * of x (n_waypoints, n_act), at most max_len rows.  One workgroup per
 * trajectory, of block_threads threads (128, 256, 512; 0 = chosen from
 * max_len); with traj_subset (device int32, n_subset entries) only those
 * trajectories are processed, so length classes can run as separate launches.
 *   weights   {w_smooth, w_vel, w_acc, w_jerk, w_limits, w_coll}
 *   out_comps (n_traj, 3) {smooth, limits, coll} without w_smooth/w_limits/w_coll
 *   out_grad  (n_waypoints, n_act) gradient of w_smooth smooth + w_limits
 *             limits + w_coll coll, or NULL
 * Replaces the cost_fn + jax.value_and_grad of
 * src/pyroffi/optimization_engines/_lbfgs_trajopt_optimization.py:98-118 (the
 * gradient is a hand-written adjoint, no reference counterpart).
 */
int pk_traj_cost_grad_f32(void* stream, const pk_traj_model* model, const float* x,
                          const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints,
                          int32_t max_len, int32_t block_threads, const int32_t* traj_subset,
                          int32_t n_subset, const float* weights, float* out_comps,
                          float* out_grad);

/* Floats of the state buffer pk_traj_lbfgs_f32 needs. */
int64_t pk_traj_lbfgs_state_floats(int64_t n_waypoints, int32_t n_act, int32_t n_traj,
                                   int32_t m_hist);

/*
 * Direct L-BFGS trajectory optimisation (_lbfgs_direct_solve,
 * src/pyroffi/optimization_engines/_lbfgs_trajopt_optimization.py:76-205),
 * one workgroup per trajectory (block_threads / traj_subset as for
 * pk_traj_cost_grad_f32).  Endpoints stay where x_init has them (the
 * caller pins them, _lbfgs_trajopt_jax :229-230).  Iterations
 * [iter_begin, iter_end) of cfg->n_iters run per call and resume from `state`
 * (pk_traj_lbfgs_state_floats floats), so a solve can be split into launches.
 *   out_best      (n_waypoints, n_act) best iterate (at the maximum collision weight)
 *   out_best_cost (n_traj) its cost, written by the call that reaches n_iters
 */
int pk_traj_lbfgs_f32(void* stream, const pk_traj_model* model, const float* x_init,
                      const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints,
                      int32_t max_len, int32_t block_threads, const int32_t* traj_subset,
                      int32_t n_subset, const pk_lbfgs_config* cfg, int32_t iter_begin,
                      int32_t iter_end, float* state, float* out_best, float* out_best_cost);

/* ChompTrajOptConfig (src/pyroffi/optimization_engines/_chomp_optimization.py:31-90). */
typedef struct pk_chomp_config {
  int32_t n_iters, use_covariant_update, n_cg_iters, early_stop_patience;
  float step_size, w_smooth, w_acc, w_jerk, w_collision, w_collision_max, collision_penalty_scale;
  float w_limits, smoothness_reg, cg_tol, grad_clip_norm, max_delta_per_step, min_cost_improve;
} pk_chomp_config;

/*
 * CHOMP trajectory optimisation of many trajectories of different lengths
 * (x_init / traj_start as for pk_traj_lbfgs_f32; endpoints stay where x_init
 * has them), one workgroup per trajectory, all iterations in one launch:
 * _optimize_single_traj of src/pyroffi/optimization_engines/_chomp_optimization.py:232-330
 * (clipped gradient, covariant direction by truncated CG per joint, six
 * clamped trials, early stop, collision-weight continuation).  The model's
 * mode must be 0 (hinge^2) for the reference's cost.  out_traj (W, A) = final
 * iterates, out_cost (B) = their cost at w_collision_max (_eval_cost,
 * :333-348); out_trace (B, n_iters, 3) or NULL = per iteration the current
 * cost, the picked trial (0-5) and whether it was accepted (-1 once stopped).
 */
int pk_traj_chomp_f32(void* stream, const pk_traj_model* model, const float* x_init,
                      const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints, int32_t max_len,
                      int32_t block_threads, const pk_chomp_config* cfg, float* out_traj,
                      float* out_cost, float* out_trace);

/* ScoTrajOptConfig (src/pyroffi/optimization_engines/_sco_optimization.py:49-110). */
typedef struct pk_sco_config {
  int32_t n_outer_iters, n_inner_iters, m_lbfgs;
  float w_smooth, w_vel, w_acc, w_jerk, w_collision, w_collision_max, penalty_scale;
  float collision_margin, w_trust, w_limits, smooth_min_temperature;
} pk_sco_config;

/*
 * SCO's linearisation (_collision_dists_reduced + _compute_coll_dists_and_jacs,
 * src/pyroffi/optimization_engines/_sco_optimization.py:224-288) of every
 * waypoint of x (layout as pk_traj_lbfgs_f32): per collision group the smooth
 * minimum -T logsumexp(-d / T) of its JAX distances and its gradient.  Group
 * 0 = the model's active self pairs, group 1 + g = every (link, world sphere)
 * pair of world spheres [world_group_start[g], world_group_start[g + 1]) (the
 * reference's one group per world geometry; at most 7).  An empty group reads
 * +inf with a zero row.  out_d (W, 1 + n_world_groups), out_J (W, 1 +
 * n_world_groups, n_act).  The model's margin, mode and grids are not used.
 */
int pk_traj_sco_linearize_f32(void* stream, const pk_traj_model* model, const float* x,
                              const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints, int32_t max_len,
                              int32_t block_threads, const int32_t* world_group_start, int32_t n_world_groups,
                              float temperature, float* out_d, float* out_J);

/*
 * SCO's inner problem (_lbfgs_inner_solve, _sco_optimization.py:291-418) for
 * every trajectory from its linearisation point x: n_inner_iters L-BFGS steps
 * on smoothness + limits + w_collision * linearised hinge^2 (d, J from
 * pk_traj_sco_linearize_f32, n_groups = 1 + n_world_groups) + w_trust |x -
 * x_lin|^2, endpoints fixed; out_x (W, A) = each trajectory's best inner
 * iterate.  One workgroup per trajectory with every vector in LDS:
 * 4 (9 + 2 m) max_len n_act bytes plus the linearisation must fit 150 KiB.
 */
int pk_traj_sco_inner_f32(void* stream, const float* x, const int64_t* traj_start, int32_t n_traj,
                          int64_t n_waypoints, int32_t max_len, int32_t n_act, int32_t n_groups, const float* d,
                          const float* J, const float* lower, const float* upper, const pk_sco_config* cfg,
                          float w_collision, float* out_x);

/*
 * pk_traj_lbfgs_f32 without a workspace, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The solver state
 * lives in a static device array of 4 Mi floats (16 MiB) that each GPU the
 * library runs on holds until the process ends: a solve needs
 * pk_traj_lbfgs_state_floats(n_waypoints, n_act, n_traj, m_hist) <= 4,194,304
 * floats (about 40,000 waypoints of a 7-joint robot at m_hist = 6), larger
 * ones are rejected with PK_ERR_INVALID before anything runs.  The call
 * allocates nothing.  Every call on one GPU shares that GPU's array and leaves
 * its contents there: a solve split over calls (iter_begin > 0) resumes from
 * it, so no other solve may run on that GPU in between, and calls that
 * overlap on one GPU (two streams or host threads) race.
 */
int pk_traj_lbfgs_noalloc_f32(void* stream, const pk_traj_model* model, const float* x_init,
                              const int64_t* traj_start, int32_t n_traj, int64_t n_waypoints,
                              int32_t max_len, int32_t block_threads, const int32_t* traj_subset,
                              int32_t n_subset, const pk_lbfgs_config* cfg, int32_t iter_begin,
                              int32_t iter_end, float* out_best, float* out_best_cost);

#ifdef __cplusplus
}
#endif

#endif /* PYROFFI_AMD_H_ */
