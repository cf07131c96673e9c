/*
 * zllbm_hip.h -- C ABI of the MI355X (gfx950) lattice Boltzmann time step on ZLLBM's
 * Morton-keyed AMR grid.
 *
 * One context owns all device memory of one solver instance on the current HIP device.  Host
 * arrays passed in are borrowed for the duration of the call; output arrays are caller owned.
 * Calls are not thread safe on one context; every device operation is enqueued on the
 * context's stream and only the download / conservation / timing calls synchronise.
 * No function aborts or throws: errors are returned as a negative status and described by
 * zlh_last_error().
 *
 * Each entry point names the reference interface it replaces (paths relative to the ZLLBM
 * source tree).
 */
#ifndef ZLLBM_HIP_H_
#define ZLLBM_HIP_H_
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct zlh_ctx zlh_ctx;
typedef int32_t zlh_status;

enum {
    ZLH_OK = 0,
    ZLH_EINVAL = -1,    /* bad argument */
    ZLH_EDEVICE = -2,   /* HIP runtime error */
    ZLH_ENOMEM = -3,    /* device allocation failed */
    ZLH_ESTATE = -4     /* call out of order (e.g. step before build) */
};

/* storage / arithmetic form of the populations */
enum {
    ZLH_FP64 = 0,           /* f stored and computed in double */
    ZLH_FP32_SHIFTED = 1,   /* f - w_i stored and computed in float */
    ZLH_BF16_SHIFTED = 2    /* f - w_i stored in bfloat16 (round to nearest even), computed in
                               float; not with the learned closure (ZLH_EINVAL) */
};

/* ELbmBoundaryConditionScheme (boundary_conditions/lbm_boundary_conditions.h) */
enum { ZLH_BC_NONE = 0, ZLH_BC_BOUNCE_BACK = 1, ZLH_BC_PERIODIC = 2, ZLH_BC_NEE = 3 };
/* EDomainBoundaryDirection (grid/grid_enumerates.h:37-45) */
enum { ZLH_XMIN = 1, ZLH_XMAX = 2, ZLH_YMIN = 3, ZLH_YMAX = 4, ZLH_ZMIN = 5, ZLH_ZMAX = 6 };
/* InterfaceLayerInfo maps (grid/grid_info_interface.h:117-127) */
enum { ZLH_OUTER_F2C = 0, ZLH_OUTER_C2F = 1, ZLH_INNER_F2C = 2, ZLH_INNER_C2F = 3 };
/* collision operators (ELbmCollisionOperatorType, collision_model/lbm_collision_models.h) */
enum { ZLH_SRT = 0, ZLH_MRT = 1 };
/* sub-grid closures */
enum { ZLH_LES_NONE = 0, ZLH_LES_SMAGORINSKY = 1, ZLH_LES_MLP = 2 };
/* Smagorinsky relaxation time: reference = tau_eff is the SGS increment alone (the shipped
 * behaviour of lbm_solver_interface.cpp:276-295), physical = tau0 + increment */
enum { ZLH_LES_TAU_REFERENCE = 0, ZLH_LES_TAU_PHYSICAL = 1 };
/* operands of the FP32 solver's matrix-core closure network (FP32 accumulation either way):
 * BF16 (8-bit significand) or FP16 (11-bit significand, same rate) */
enum { ZLH_MLP_OPERANDS_BF16 = 0, ZLH_MLP_OPERANDS_FP16 = 1 };

/* Solver settings: SolverLbmInterface members set by the drivers (lbm_interface.h:288-474)
 * and the per-direction boundary settings (SetDomainBoundaryCondition, SetValues). */
typedef struct zlh_solver_desc {
    int32_t dims;               /* 2 (D2Q9) or 3 (D3Q19) */
    int32_t max_level;          /* finest level, 0 = uniform */
    int32_t precision;          /* ZLH_FP64 / ZLH_FP32_SHIFTED */
    int32_t compressible;       /* k0BoolCompressible_ */
    double nu0;                 /* k0LbmViscosity_: viscosity at level 0 */
    double rho0;                /* k0Rho_ */
    double u0[3];               /* k0Velocity_ */
    int32_t les;                /* ZLH_LES_* */
    int32_t les_tau_mode;       /* ZLH_LES_TAU_* */
    double cs;                  /* Smagorinsky constant */
    int32_t has_force;          /* bool_forces_ */
    int32_t mlp_operands;       /* ZLH_MLP_OPERANDS_*, FP32 solver with ZLH_LES_MLP */
    double const_force[3];      /* k0ConstForce_ */
    int32_t collision[16];      /* per level ZLH_SRT / ZLH_MRT */
    int32_t bc_scheme[7];       /* per direction 1..6 */
    int32_t pad1;
    double bc_value[7][3];      /* wall / boundary velocity per direction */
    uint32_t dom_min[3];        /* background node index bounds */
    uint32_t dom_max[3];
} zlh_solver_desc;

/* SolverCreatorLbmD2Q9/D3Q19::CreateSolver + SolverLbmInterface::SolverInitial
 * (lbm_solver_interface.cpp:19-44) */
zlh_status zlh_create(const zlh_solver_desc *desc, zlh_ctx **out);
/* nodes of one level: map_grid_node_ keys and flag_status_ (grid/grid_info_interface.h:258);
 * keys sorted ascending and unique, no axis index 0 or at the top of the 64-bit code */
zlh_status zlh_set_level_nodes(zlh_ctx *ctx, int level, const uint64_t *keys, const int32_t *flags,
                               int64_t n);
/* one interface layer map (InterfaceLayerInfo::vec_{outer,inner}_{coarse2fine,fine2coarse}_) */
zlh_status zlh_set_interface_layer(zlh_ctx *ctx, int level, int kind, int layer,
                                   const uint64_t *keys, int64_t n);
/* immersed-boundary markers of one geometry at one level (GeometryVertexImmersedBoundary
 * coordinate_/velocity_/area_, immersed_boundary.h:23-28); xyz and vel are [n][3] */
zlh_status zlh_ib_set_markers(zlh_ctx *ctx, int level, const double *xyz, const double *area,
                              const double *vel, int64_t n, int stencil_dis, double stiffness);
/* velocity profile of a non-equilibrium-extrapolation face, before zlh_build: 0 uniform (the
 * reference's one velocity per face, BoundaryNonEqExtrapolation::SetValues), 1 parabolic: the face
 * velocity times the product over the transverse axes of 4 t (1 - t), t the node's position across
 * the domain (no reference counterpart; a channel inflow) */
#define ZLH_PROFILE_UNIFORM 0
#define ZLH_PROFILE_PARABOLIC 1
zlh_status zlh_set_boundary_profile(zlh_ctx *ctx, int dir, int profile);
/* per-node velocities of a non-equilibrium-extrapolation face on one level, before zlh_build (no
 * reference counterpart; the reference's BoundaryNonEqExtrapolation::SetValues takes one velocity
 * per face, lbm_noneq_extrapolation.cpp:14-20): u is [n][3] for the level's nodes on that face in
 * key order, n their count (zlh_build checks it); replaces the face velocity and profile there.
 * A seeded perturbed inflow, for example. */
zlh_status zlh_set_boundary_velocities(zlh_ctx *ctx, int level, int dir, const double *u, int64_t n);
/* interpolation of the coarse-to-fine ghost nodes of fine level `level` (>= 1), before zlh_build
 * (GridInfoInterface::interp_method_, grid_info_interface.h:317, and ChooseInterpolationMethod,
 * grid/grid_interpolation.cpp:20-82): ZLH_INTERP_LAGRANGIAN (the default, 4-point Lagrange
 * weights, InterpolationLagrangian{2D,3D}) or ZLH_INTERP_LINEAR (InterpolationLinear{2D,3D}) */
#define ZLH_INTERP_LAGRANGIAN 0
#define ZLH_INTERP_LINEAR 1
zlh_status zlh_set_interpolation_method(zlh_ctx *ctx, int level, int method);
/* time interpolation at coarse-fine interfaces (no reference counterpart; 0, the default, is the
 * reference's schedule of transfers after even fine substeps only, lbm_solver_interface.cpp:
 * 224-238).  1: the ghost layers of every refined level are also refilled after its odd
 * substeps, each coarse contributor giving the mean of its transferred values before and after
 * the coarse level's last step (a linear interpolation in time), each fine fallback contributor
 * its current values; the even-substep transfers are unchanged.  zlh_transfer_levels and
 * zlh_transfer_from_coarse after an odd substep then do that refill (zlh_transfer_to_coarse there
 * stays ZLH_ESTATE).  Any time, taking effect from the next substep; set it between background
 * steps, since the earlier coarse state is taken from the state at the call (and at each upload). */
zlh_status zlh_set_time_interpolation(zlh_ctx *ctx, int enable);
/* where the FP32 learned-closure kernels run the MRT moment transform m = M f and its inverse
 * (no reference counterpart; the reference's MRT collision multiplies by dense matrices,
 * lbm_collision_model_mrt.cpp:114-143): ZLH_TRANSFORM_VALU (the default, pair sums and
 * differences of opposite populations on the vector units) or ZLH_TRANSFORM_MFMA (19 x 19
 * contractions on v_mfma_f32_32x32x2f32, FP32 operands and sums).  Other solver forms always use
 * the vector units. */
#define ZLH_TRANSFORM_VALU 0
#define ZLH_TRANSFORM_MFMA 1
zlh_status zlh_set_moment_transform(zlh_ctx *ctx, int where);
/* largest number of 64-node bricks one level may hold with these settings (no reference
 * counterpart; no device call): slot indices are 32-bit, and the FP32 learned closure's fused
 * kernel addresses its pulls with 32-bit byte offsets (bricks x Q x 256 bytes <= 2^32, about
 * 883k bricks = 56.5M nodes in D3Q19).  zlh_build refuses a level above it (ZLH_EINVAL). */
zlh_status zlh_max_level_bricks(const zlh_solver_desc *desc, int64_t *max_bricks);
/* device slot and neighbour tables, boundary / interface / IB tables
 * (replaces the per-access unordered_map lookups of lbm_d3q19_solver.cpp:20-135) */
zlh_status zlh_build(zlh_ctx *ctx);
/* every node at feq(rho0, u0) (SetInitialDisFuncBasedOnReferenceMacros, lbm_solver_interface.cpp:344-367) */
zlh_status zlh_init_feq(zlh_ctx *ctx);
/* per node initial macros in key order, f = f_collide = feq(rho, u); u is [n][dims] */
zlh_status zlh_init_macros(zlh_ctx *ctx, int level, const double *rho, const double *u);
/* post-stream populations in key order [n][q] (GridNodeLbm::f_); also sets f_collide = f */
zlh_status zlh_upload_populations(zlh_ctx *ctx, int level, const double *f);
/* one level substep: immersed-boundary forcing, collision, streaming and domain boundaries
 * (SolverLbmInterface::RunSolverOnGivenGrid, lbm_solver_interface.cpp:246-262, with the IB hook
 * UpdateCriterion and ComputeDomainBoundaryCondition of GridInfoLbmInteface::AdvancingAtCurrentTime,
 * lbm_grid_interface.cpp:61-153); time_step_level counts this level's substeps from 1 */
zlh_status zlh_step_level(zlh_ctx *ctx, int level, int time_step_level);
/* coarse-to-fine interpolation into the fine ghost layers, then fine-to-coarse copy into the
 * coarse interface layer (SolverLbmInterface::InformationFromGridOfDifferentLevel,
 * lbm_solver_interface.cpp:224-238; the caller applies it after even substeps of fine_level,
 * ZLH_ESTATE after an odd one unless time interpolation is on, zlh_set_time_interpolation) */
zlh_status zlh_transfer_levels(zlh_ctx *ctx, int fine_level);
/* its two halves, in this order: the coarse-to-fine interpolation
 * (GridInfoLbmInteface::TransferInfoFromCoarseGrid, lbm_grid_interface.cpp:408-515) and the
 * fine-to-coarse copy (TransferInfoToCoarseGrid, lbm_grid_interface.cpp:524-564) */
zlh_status zlh_transfer_from_coarse(zlh_ctx *ctx, int fine_level);
zlh_status zlh_transfer_to_coarse(zlh_ctx *ctx, int fine_level);
/* background steps in the coarse-to-fine multi time stepping order (AmrManager::TimeMarching,
 * amr_manager.cpp:279-300; MultiTimeSteppingC2F, grid/grid_manager.cpp:997-1024) */
zlh_status zlh_time_marching(zlh_ctx *ctx, int nsteps);
/* zlh_time_marching replays pairs of background steps as captured HIP graphs (default on);
 * off launches every kernel directly (no reference counterpart) */
zlh_status zlh_graphs_enable(zlh_ctx *ctx, int enable);
/* run each level's substeps of zlh_time_marching on a stream of its own, ordered by events
 * around the transfers, so that coarse levels overlap finer ones (default 1; no reference
 * counterpart).  Results are identical either way; timed runs (zlh_timing_enable) use one
 * stream. */
zlh_status zlh_level_streams_enable(zlh_ctx *ctx, int enable);
/* post-stream populations GridNodeLbm::f_ in key order [n][q] (synchronising) */
zlh_status zlh_download_populations(zlh_ctx *ctx, int level, double *f);
/* density and velocity of the post-stream populations, key order; u is [n][dims] (synchronising) */
zlh_status zlh_download_macros(zlh_ctx *ctx, int level, double *rho, double *u);
/* pull-source table [n][q]: key-order index of the node x - c_q, -1 when absent; the device
 * replacement of the FindX/Y/Z{Pos,Neg} + map lookups (synchronising) */
zlh_status zlh_download_source_table(zlh_ctx *ctx, int level, int32_t *src);
/* node forces of a level in key order, [n][3] (GridNodeLbm::force_, lbm_interface.h:34-79, as
 * the checkpoint writes them): the immersed-boundary force of the last forcing step on its
 * supports, zero elsewhere (synchronising) */
zlh_status zlh_download_node_forces(zlh_ctx *ctx, int level, double *force);
/* Lagrangian marker forces of the last step [n][3] (synchronising) */
zlh_status zlh_ib_download_forces(zlh_ctx *ctx, int level, double *force);
/* FP64 mass and momentum sums of the post-stream populations of a level: out = {m, jx, jy, jz}
 * (no reference counterpart; synchronising) */
zlh_status zlh_conservation(zlh_ctx *ctx, int level, double *out4);
/* node counts: total nodes of a level, and how many take the deferred (aux) path */
zlh_status zlh_level_info(zlh_ctx *ctx, int level, int64_t *n_nodes, int64_t *n_aux, int64_t *n_bricks);
/* interpolation tables of a refined level (no reference counterpart): out[ZLH_XFER_INFO_N] =
 * {ghost nodes, coarse source rows, fine-to-coarse nodes, gather workgroups, rows staged by all
 * workgroups, contributors, contributors taken from a fine fallback node, LDS bytes of the largest
 * workgroup, 1 if the transfer into the next finer level writes a node of this level that this
 * level's own transfer touches, or reads one it writes (time marching then orders the two)};
 * zeros for level 0 */
#define ZLH_XFER_INFO_N 9
zlh_status zlh_transfer_info(zlh_ctx *ctx, int level, int64_t *out);
zlh_status zlh_synchronize(zlh_ctx *ctx);
/* device time of the fused stream-collide kernel: enable recording, then read the total in
 * milliseconds and the number of launches since enabling, of one level or of all (level -1)
 * (synchronising; no reference counterpart) */
zlh_status zlh_timing_enable(zlh_ctx *ctx, int enable);
zlh_status zlh_timing_read(zlh_ctx *ctx, int level, double *total_ms, int64_t *launches);

/* ---- learned sub-grid closure (no counterpart in the reference; replaces the Smagorinsky
 * relaxation time of les_models/lbm_les_smagorinsky.cpp:22-94 when desc.les = ZLH_LES_MLP) ----
 * 9 -> 32 -> 32 -> 1 MLP, SiLU hidden layers, softplus output times nu_scale = eddy viscosity in
 * lattice units; tau_eff = tau0 + 3 nu_t.  params: W1[32][9], b1[32], W2[32][32], b2[32],
 * w3[32], b3 (1409 floats).  Inputs are normalised as (x - mean) / std; where any raw input
 * lies outside [fmin, fmax] the node uses Smagorinsky (physical tau mode, desc.cs). */
zlh_status zlh_mlp_set(zlh_ctx *ctx, const float *params, const float *mean, const float *std_dev,
                       const float *fmin, const float *fmax, double nu_scale);
/* full-batch mean-squared-error gradient of softplus(net(x)) against t over n samples
 * (normalised inputs [n][9]); grad is [1409], loss may be null */
zlh_status zlh_mlp_gradient(zlh_ctx *ctx, const float *params, const float *x, const float *t, int64_t n,
                            float *grad, float *loss);
/* zlh_mlp_gradient without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight training
 * loops, harnesses that forbid hipMalloc).  The parameters, samples, targets and the gradient
 * and loss accumulators are staged in one static device array of the library, 2.6 MiB per GPU,
 * held until the process ends; it holds at most ZLH_MLP_NOALLOC_MAX_SAMPLES samples, and a
 * larger n returns ZLH_EINVAL.  Same arguments, kernel (zlh_mlp_training_form) and results as
 * zlh_mlp_gradient.  Every call on one GPU uses that GPU's one array, which keeps the last
 * call's contents (each call clears its accumulators first): calls that overlap on one GPU,
 * from several threads or contexts, race. */
#define ZLH_MLP_NOALLOC_MAX_SAMPLES 65536
zlh_status zlh_mlp_gradient_noalloc(zlh_ctx *ctx, const float *params, const float *x, const float *t,
                                    int64_t n, float *grad, float *loss);
/* Adam training on the device: `steps` minibatches of `batch` samples drawn with a seeded hash
 * (batch >= n: every step uses all n samples in order); params is updated in place;
 * loss_history (may be null) receives each minibatch loss */
zlh_status zlh_mlp_train(zlh_ctx *ctx, const float *x, const float *t, int64_t n, int steps, int batch,
                         double lr, uint64_t seed, float *params, float *loss_history);
/* network forms: FP32 on the vector units (the FP64 solver's collision) and BF16 operands with
 * FP32 accumulation on the matrix cores, 64 nodes per wavefront (the FP32 solver's collision) */
enum { ZLH_MLP_FP32 = 0, ZLH_MLP_BF16_MFMA = 1, ZLH_MLP_FP16_MFMA = 2 };
/* kernel of zlh_mlp_gradient and zlh_mlp_train: ZLH_MLP_FP32 (default; FP32 on the vector units,
 * one thread per sample) or ZLH_MLP_BF16_MFMA (forward and backward of 64 samples per wavefront
 * on the matrix cores, BF16 operands and FP32 accumulation); ZLH_EINVAL for any other form */
zlh_status zlh_mlp_training_form(zlh_ctx *ctx, int form);
/* network output (nu_t / nu_scale) of normalised inputs [n][9] with the parameters set by
 * zlh_mlp_set, in the given form */
zlh_status zlh_mlp_eval(zlh_ctx *ctx, const float *x, int64_t n, int form, float *y);
/* per node of a level (key order): raw closure inputs [n][9], the closure's nu_t and whether the
 * node falls back to Smagorinsky, from the current post-stream populations (synchronising) */
zlh_status zlh_closure_features(zlh_ctx *ctx, int level, float *features, float *nu_t, int32_t *fallback);

const char *zlh_last_error(const zlh_ctx *ctx);
void zlh_destroy(zlh_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
