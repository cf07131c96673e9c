/*
 * zlbm_oracle.h -- FP64 CPU restatement of ZLLBM's lattice Boltzmann time step on the
 * Morton-keyed AMR grid.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP implementation in
 * zllbm_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product path never links or calls it.
 *
 * Every function states the reference file:line whose semantics it restates (paths relative
 * to the ZLLBM source tree).  The restatement mirrors the reference's operation order so that
 * FP64 results follow the same rounding sequence.
 */
#ifndef ZLBM_ORACLE_H_
#define ZLBM_ORACLE_H_
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct zo_sim zo_sim;

/* boundary directions, numbered as EDomainBoundaryDirection (grid/grid_enumerates.h:37-45) */
enum { ZO_XMIN = 1, ZO_XMAX = 2, ZO_YMIN = 3, ZO_YMAX = 4, ZO_ZMIN = 5, ZO_ZMAX = 6 };
/* boundary schemes, numbered as ELbmBoundaryConditionScheme (boundary_conditions/lbm_boundary_conditions.h) */
enum { ZO_BC_NONE = 0, ZO_BC_BOUNCE_BACK = 1, ZO_BC_PERIODIC = 2, ZO_BC_NEE = 3 };
/* interface layer kinds (InterfaceLayerInfo, grid/grid_info_interface.h:117-127) */
enum { ZO_OUTER_F2C = 0, ZO_OUTER_C2F = 1, ZO_INNER_F2C = 2, ZO_INNER_C2F = 3 };
/* node flags (NodeBitStatus, grid/grid_manager.h:64-73) */
enum { ZO_FLAG_C2F0 = 1 << 1, ZO_FLAG_C2F_GHOST = 1 << 2, ZO_FLAG_F2C0 = 1 << 3,
       ZO_FLAG_F2C_GHOST = 1 << 4 };

typedef struct zo_params {
    int32_t dims;               /* 2 or 3 */
    int32_t max_level;          /* 0 = uniform */
    double nu0;                 /* LBM viscosity at level 0 (k0LbmViscosity_) */
    double rho0;                /* k0Rho_ */
    double u0[3];               /* k0Velocity_ */
    int32_t compressible;       /* k0BoolCompressible_ */
    int32_t les;                /* 0 none, 1 Smagorinsky */
    int32_t les_tau_mode;       /* 0 reference (tau_eff = tau_t), 1 physical (tau_eff = tau0 + tau_t) */
    double cs;                  /* Smagorinsky constant (k0Smagorinsky = 18 cs^2) */
    int32_t has_force;          /* bool_forces_ */
    double const_force[3];      /* k0ConstForce_ */
    int32_t collision[16];      /* per level: 0 SRT, 1 MRT */
    int32_t bc_scheme[7];       /* index by direction 1..6 */
    double bc_value[7][3];      /* wall / boundary velocity per direction */
    uint32_t dom_min[3];        /* background index bounds (k0MinIndexOfBackgroundNode_) */
    uint32_t dom_max[3];        /* (k0MaxIndexOfBackgroundNode_) */
} zo_params;

/* lattice constants; q = 9 (dims 2) or 19 (dims 3) */
int zo_lattice(int dims, double *cx, double *cy, double *cz, double *w, int32_t *opp);
/* moment matrix M and its inverse as used by the MRT operator (row-major q x q) */
int zo_mrt_matrices(int dims, double *m, double *minv, double *s_base);

/* Morton helpers (grid/morton_aux.cpp, grid/morton_aux.hpp) */
uint64_t zo_encode(int dims, uint32_t x, uint32_t y, uint32_t z);
void zo_decode(int dims, uint64_t key, uint32_t *xyz);
uint64_t zo_find_neighbor(int dims, uint64_t key, int axis, int dir);

zo_sim *zo_create(const zo_params *p);
void zo_destroy(zo_sim *s);
const char *zo_last_error(const zo_sim *s);
/* keys must be sorted ascending and unique */
int zo_set_level(zo_sim *s, int level, int64_t n, const uint64_t *keys, const int32_t *flags);
int zo_set_layer(zo_sim *s, int level, int kind, int layer, int64_t n, const uint64_t *keys);
int zo_finalize(zo_sim *s);
/* initial state: every node at feq(rho0, u0) (lbm_solver_interface.cpp:344-367) */
int zo_init_feq(zo_sim *s);
/* per node (rho, u) initial condition, f = f_collide = feq(rho, u) */
int zo_init_macro(zo_sim *s, int level, const double *rho, const double *u);
int zo_set_f(zo_sim *s, int level, const double *f);      /* [n][q], sets f and f_collide */
int zo_get_f(const zo_sim *s, int level, double *f);      /* post-stream f_ [n][q] */
int zo_get_fc(const zo_sim *s, int level, double *fc);    /* f_collide_ [n][q] */
int zo_get_node_macro(const zo_sim *s, int level, double *rho, double *u);  /* rho_, velocity_ */
int zo_set_node_force(zo_sim *s, int level, const double *force);          /* force_ [n][dims] */
int zo_get_node_force(const zo_sim *s, int level, double *force);
int64_t zo_level_size(const zo_sim *s, int level);
/* pull-source index table [n][q] (-1 = absent) as built from the Morton neighbour rules */
int zo_get_source_table(const zo_sim *s, int level, int32_t *src);
/* per-level boundary node lists, direction 1..6; returns count, fills idx when non-null */
int64_t zo_get_boundary(const zo_sim *s, int level, int dir, int32_t *idx);

/* one level substep (GridInfoLbmInteface::AdvancingAtCurrentTime, lbm_grid_interface.cpp:104-153) */
int zo_advance_level(zo_sim *s, int level, int time_step_level);
/* one background step with the coarse-to-fine multi time stepping order (amr_manager.cpp:279-300) */
int zo_time_marching(zo_sim *s, int nsteps);
/* level order of one background step (grid/grid_manager.cpp:997-1024); returns its length */
int zo_time_stepping_order(int max_level, int32_t *order, int cap);

/* velocity profile of a non-equilibrium-extrapolation face: 0 uniform (the reference), 1 parabolic
 * across the transverse axes with the face velocity on the centre line */
int zo_set_bc_profile(zo_sim *s, int dir, int profile);
/* interpolation of the ghost nodes of fine level `level` (>= 1): 0 Lagrangian (default), 1 linear */
int zo_set_interp_method(zo_sim *s, int level, int method);
/* 1: also refill fine ghost layers after odd substeps, each coarse contributor giving the mean of
 * its values before and after the coarse level's last step (no reference counterpart; 0, the
 * default, is the reference's behaviour) */
int zo_set_time_interp(zo_sim *s, int on);
/* per-node velocities [n][3] of a non-equilibrium-extrapolation face on one level, in key order of
 * the level's nodes on that face (no reference counterpart); checked against the count at finalize */
int zo_set_bc_velocities(zo_sim *s, int level, int dir, const double *u, int64_t n);
/* immersed boundary markers for one level (immersed_boundary.cpp:35-384) */
int zo_set_ib(zo_sim *s, int level, int64_t n, const double *xyz, const double *area,
              const double *vel, int stencil_dis, double stiffness, double dx_background);
int zo_get_ib_force(const zo_sim *s, int level, double *force);   /* Lagrangian force [n][dims] */

/* learned closure for les == 2 (params: W1[32][9], b1, W2[32][32], b2, w3, b3) */
int zo_set_mlp(zo_sim *s, const double *params, const double *mean, const double *std_dev, const double *fmin,
               const double *fmax, double nu_scale);

/* mass and momentum sums over a level (FP64) */
int zo_conservation(const zo_sim *s, int level, double *out4);

/* test hook: run the coarse/fine transfer of one refined level on the current populations */
int zo_debug_transfer(zo_sim *s, int fine_level);
/* test hook: replace the non-viscous MRT relaxation rates */
int zo_debug_set_s_base(zo_sim *s, const double *s_base);

/* set OpenMP thread count used by the loops (0 = runtime default) */
void zo_set_threads(int n);
#ifdef __cplusplus
}
#endif
#endif
