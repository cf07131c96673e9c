/*
 * zlbm_oracle.c -- FP64 CPU restatement of the ZLLBM lattice Boltzmann time step.
 *
 * TEST INFRASTRUCTURE ONLY (see zlbm_oracle.h).  Written from the reference's documented
 * behaviour; every routine cites the reference function it restates.  Node storage is a sorted
 * key array per level (the reference uses an unordered_map keyed by the Morton code,
 * defs_libs.h:37-45); lookups are binary searches, which give the same node for the same key.
 *
 * Iteration order: the reference iterates unordered_maps (unspecified order).  Every loop here
 * whose result could depend on order is noted; the others are order independent.
 */
#include "zlbm_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ZO_MAXQ 19
#define ZO_MAXL 16
#define ZO_KEPS 1.e-6   /* kEps, defs_libs.h:103 */

typedef struct zo_layers {
    int64_t n[4][3];
    int32_t *idx[4][3];   /* node indices in this level, ascending key order */
} zo_layers;

typedef struct zo_ib {
    int64_t n;
    double *xyz, *area, *vel, *force;
    int stencil_dis;
    double stiffness;
} zo_ib;

typedef struct zo_level {
    int64_t n;
    uint64_t *key;
    int32_t *flag;
    int32_t *src;        /* [n][q] pull source index (x - c_q), -1 absent */
    double *f, *fc;      /* [n][q] */
    double *rho, *u;     /* [n], [n][3] */
    double *force;       /* [n][3] node (immersed boundary) force */
    double *f_prev;      /* [n][q] f at the start of the level's last substep (time interpolation) */
    int64_t nb[7];
    int32_t *bnd[7];     /* boundary node indices per direction */
    double *bnd_u[7];    /* per-node velocities [nb][3] of a face (zo_set_bc_velocities), or NULL */
    int64_t nbu[7];
    zo_layers layers;
    double dt, tau0;
    double tau_stream_c2f, tau_stream_f2c;
    double imsm[ZO_MAXQ * ZO_MAXQ], imdm[ZO_MAXQ * ZO_MAXQ];
    zo_ib ib;
    int have;
} zo_level;

struct zo_sim {
    zo_params p;
    int q;
    double cx[ZO_MAXQ], cy[ZO_MAXQ], cz[ZO_MAXQ], w[ZO_MAXQ];
    int32_t opp[ZO_MAXQ];
    double m[ZO_MAXQ * ZO_MAXQ], minv[ZO_MAXQ * ZO_MAXQ], s_base[ZO_MAXQ];
    zo_level lv[ZO_MAXL];
    char err[512];
    int finalized;
    /* learned closure (les == 2): W1[32][9], b1, W2[32][32], b2, w3, b3; normalisation, range */
    double mlp_p[1409], mlp_mean[9], mlp_istd[9], mlp_min[9], mlp_max[9], mlp_scale;
    int mlp_set;
    int64_t mlp_fallbacks;
    int bc_profile[7];   /* per face: 0 uniform boundary velocity, 1 parabolic (zo_set_bc_profile) */
    int interp_method[ZO_MAXL];   /* per fine level: 0 Lagrangian (the reference default), 1 linear */
    int time_interp;     /* 1: refill fine ghost layers after odd substeps too (zo_set_time_interp) */
};

static void zo_seterr(zo_sim *s, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(s->err, sizeof(s->err), fmt, ap);
    va_end(ap);
}

void zo_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* Lattice constants.  D2Q9 ordering lbm_d2q9.h:53-61, D3Q19 ordering lbm_d3q19.h:18-39,73-85. */
/* ------------------------------------------------------------------------------------------ */
static const int kD2C[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1},
                               {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
static const int kD3C[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                                {0, 0, 1}, {0, 0, -1}, {1, 1, 0}, {-1, 1, 0}, {1, -1, 0},
                                {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
                                {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};

int zo_lattice(int dims, double *cx, double *cy, double *cz, double *w, int32_t *opp) {
    int q = dims == 2 ? 9 : 19;
    for (int i = 0; i < q; ++i) {
        int c0 = dims == 2 ? kD2C[i][0] : kD3C[i][0];
        int c1 = dims == 2 ? kD2C[i][1] : kD3C[i][1];
        int c2 = dims == 2 ? 0 : kD3C[i][2];
        cx[i] = c0; cy[i] = c1; cz[i] = c2;
        int c2sum = c0 * c0 + c1 * c1 + c2 * c2;
        if (dims == 2) w[i] = c2sum == 0 ? 4. / 9. : (c2sum == 1 ? 1. / 9. : 1. / 36.);
        else w[i] = c2sum == 0 ? 1. / 3. : (c2sum == 1 ? 1. / 18. : 1. / 36.);
        for (int j = 0; j < q; ++j) {
            int d0 = dims == 2 ? kD2C[j][0] : kD3C[j][0];
            int d1 = dims == 2 ? kD2C[j][1] : kD3C[j][1];
            int d2 = dims == 2 ? 0 : kD3C[j][2];
            if (d0 == -c0 && d1 == -c1 && d2 == -c2) opp[i] = j;
        }
    }
    return q;
}

/*
 * Moment basis built from the polynomial of each row (d'Humieres basis, the rows of
 * kMatrixMMrt_ in lbm_d2q9.h:70-80 and lbm_d3q19.h:98-118).  The rows are orthogonal, so the
 * inverse is M^T diag(1 / |row|^2) (the rationals of kMatrixImMrt_).  Base relaxation rates
 * k0VectorSMrt_ (lbm_d2q9.h:92, lbm_d3q19.h:159-160); the viscous entries are set per level.
 */
int zo_mrt_matrices(int dims, double *m, double *minv, double *s_base) {
    int q = dims == 2 ? 9 : 19;
    for (int i = 0; i < q; ++i) {
        double x, y, z;
        if (dims == 2) { x = kD2C[i][0]; y = kD2C[i][1]; z = 0; }
        else { x = kD3C[i][0]; y = kD3C[i][1]; z = kD3C[i][2]; }
        double c2 = x * x + y * y + z * z;
        double row[19];
        if (dims == 2) {
            row[0] = 1; row[1] = 3 * c2 - 4; row[2] = (9 * c2 * c2 - 21 * c2) / 2 + 4;
            row[3] = x; row[4] = (3 * c2 - 5) * x; row[5] = y; row[6] = (3 * c2 - 5) * y;
            row[7] = x * x - y * y; row[8] = x * y;
        } else {
            row[0] = 1; row[1] = 19 * c2 - 30; row[2] = (21 * c2 * c2 - 53 * c2 + 24) / 2;
            row[3] = x; row[4] = (5 * c2 - 9) * x; row[5] = y; row[6] = (5 * c2 - 9) * y;
            row[7] = z; row[8] = (5 * c2 - 9) * z;
            row[9] = 3 * x * x - c2; row[10] = (3 * c2 - 5) * (3 * x * x - c2);
            row[11] = y * y - z * z; row[12] = (3 * c2 - 5) * (y * y - z * z);
            row[13] = x * y; row[14] = y * z; row[15] = x * z;
            row[16] = (y * y - z * z) * x; row[17] = (z * z - x * x) * y; row[18] = (x * x - y * y) * z;
        }
        for (int k = 0; k < q; ++k) m[k * q + i] = row[k];
    }
    for (int k = 0; k < q; ++k) {
        double nrm = 0;
        for (int i = 0; i < q; ++i) nrm += m[k * q + i] * m[k * q + i];
        for (int i = 0; i < q; ++i) minv[i * q + k] = m[k * q + i] / nrm;
    }
    if (dims == 2) {
        const double sb[9] = {0., 1.4, 1.4, 0., 1.2, 0., 1.2, 0., 0.};
        memcpy(s_base, sb, sizeof(sb));
    } else {
        const double sb[19] = {0., 1.5, 1.4, 0., 1.2, 0., 1.2, 0., 1.2, 0., 1.4, 0., 1.4,
                               0., 0., 0., 1.98, 1.98, 1.98};
        memcpy(s_base, sb, sizeof(sb));
    }
    return q;
}

/* viscous rows (InitialMrtSMatrix / UpdateMrtSMatrix, lbm_d2q9_mandatory.cpp:189-247,
 * lbm_d3q19_mandatory.cpp:304-364) */
static void zo_s_vector(int dims, const double *s_base, double relax, double *s) {
    int q = dims == 2 ? 9 : 19;
    for (int i = 0; i < q; ++i) s[i] = s_base[i];
    if (dims == 2) { s[7] = relax; s[8] = relax; }
    else { s[9] = relax; s[11] = relax; s[13] = relax; s[14] = relax; s[15] = relax; }
}

/* M^{-1} diag(d) M with the loop order of SetImSMMatrix (lbm_collision_model_mrt.cpp:63-82) */
static void zo_im_diag_m(int q, const double *m, const double *minv, const double *d, double *out) {
    double dm[ZO_MAXQ * ZO_MAXQ];
    for (int j = 0; j < q; ++j)
        for (int i = 0; i < q; ++i) {
            double acc = 0.;
            for (int k = 0; k < q; ++k) acc = acc + (i == k ? d[i] : 0.) * m[k * q + j];
            dm[i * q + j] = acc;
        }
    for (int j = 0; j < q; ++j)
        for (int i = 0; i < q; ++i) {
            double acc = 0.;
            for (int k = 0; k < q; ++k) acc = acc + minv[i * q + k] * dm[k * q + j];
            out[i * q + j] = acc;
        }
}

/* ------------------------------------------------------------------------------------------ */
/* Morton codes: x at bit 3i, y at 3i+1, z at 3i+2 (3D); x at 2i, y at 2i+1 (2D)               */
/* (SFBitsetAux3D::SFBitsetInitial / SFBitsetEncoding, grid/morton_aux.cpp:238-286).            */
/* ------------------------------------------------------------------------------------------ */
static uint64_t zo_axis_mask(int dims, int axis) {
    uint64_t mask = 0;
    int bits = 64 / dims;
    for (int i = 0; i < bits; ++i) mask |= (uint64_t)1 << (i * dims + axis);
    return mask;
}

uint64_t zo_encode(int dims, uint32_t x, uint32_t y, uint32_t z) {
    uint64_t code = 0;
    int bits = 64 / dims;
    uint32_t v[3] = {x, y, z};
    for (int i = 0; i < bits && i < 32; ++i)
        for (int a = 0; a < dims; ++a)
            if ((v[a] >> i) & 1u) code |= (uint64_t)1 << (i * dims + a);
    return code;
}

void zo_decode(int dims, uint64_t key, uint32_t *xyz) {
    int bits = 64 / dims;
    xyz[0] = xyz[1] = xyz[2] = 0;
    for (int i = 0; i < bits && i < 32; ++i)
        for (int a = 0; a < dims; ++a)
            if ((key >> (i * dims + a)) & 1u) xyz[a] |= 1u << i;
}

/* FindXNeg/FindXPos and the y/z versions (grid/morton_aux.hpp:65-129, 184-297): Neg returns the
 * key unchanged when the axis component is zero; Pos adds one inside the axis lane. */
uint64_t zo_find_neighbor(int dims, uint64_t key, int axis, int dir) {
    uint64_t cur = zo_axis_mask(dims, axis);
    uint64_t others = key & ~cur;
    if (dir < 0) {
        uint64_t a = key & cur;
        if (a == 0) return key;
        return ((a - 1) & cur) | others;
    }
    uint64_t t = (key | ~cur) + 1;
    return (t & cur) | others;
}

static uint64_t zo_step_key(int dims, uint64_t key, const int *delta) {
    for (int a = 0; a < dims; ++a) {
        if (delta[a] > 0) key = zo_find_neighbor(dims, key, a, +1);
        else if (delta[a] < 0) key = zo_find_neighbor(dims, key, a, -1);
    }
    return key;
}

static int64_t zo_find(const zo_level *L, uint64_t key) {
    int64_t lo = 0, hi = L->n - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) >> 1;
        if (L->key[mid] == key) return mid;
        if (L->key[mid] < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* ------------------------------------------------------------------------------------------ */
zo_sim *zo_create(const zo_params *p) {
    if (!p || (p->dims != 2 && p->dims != 3) || p->max_level < 0 || p->max_level >= ZO_MAXL)
        return NULL;
    zo_sim *s = (zo_sim *)calloc(1, sizeof(zo_sim));
    s->p = *p;
    s->q = zo_lattice(p->dims, s->cx, s->cy, s->cz, s->w, s->opp);
    zo_mrt_matrices(p->dims, s->m, s->minv, s->s_base);
    return s;
}

static void zo_free_level(zo_level *L) {
    free(L->key); free(L->flag); free(L->src); free(L->f); free(L->fc);
    free(L->rho); free(L->u); free(L->force); free(L->f_prev);
    for (int d = 0; d < 7; ++d) { free(L->bnd[d]); free(L->bnd_u[d]); }
    for (int k = 0; k < 4; ++k) for (int l = 0; l < 3; ++l) free(L->layers.idx[k][l]);
    free(L->ib.xyz); free(L->ib.area); free(L->ib.vel); free(L->ib.force);
    memset(L, 0, sizeof(*L));
}

void zo_destroy(zo_sim *s) {
    if (!s) return;
    for (int l = 0; l < ZO_MAXL; ++l) zo_free_level(&s->lv[l]);
    free(s);
}

const char *zo_last_error(const zo_sim *s) { return s ? s->err : "null"; }

int zo_set_level(zo_sim *s, int level, int64_t n, const uint64_t *keys, const int32_t *flags) {
    if (level < 0 || level > s->p.max_level) { zo_seterr(s, "bad level"); return -1; }
    zo_level *L = &s->lv[level];
    zo_free_level(L);
    for (int64_t i = 1; i < n; ++i)
        if (keys[i] <= keys[i - 1]) { zo_seterr(s, "keys not sorted/unique at %lld", (long long)i); return -1; }
    int q = s->q;
    L->n = n;
    L->key = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    L->flag = (int32_t *)calloc(n ? n : 1, sizeof(int32_t));
    memcpy(L->key, keys, sizeof(uint64_t) * n);
    if (flags) memcpy(L->flag, flags, sizeof(int32_t) * n);
    L->src = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1) * q);
    L->f = (double *)calloc((n ? n : 1) * q, sizeof(double));
    L->fc = (double *)calloc((n ? n : 1) * q, sizeof(double));
    L->rho = (double *)calloc(n ? n : 1, sizeof(double));
    L->u = (double *)calloc((n ? n : 1) * 3, sizeof(double));
    L->force = (double *)calloc((n ? n : 1) * 3, sizeof(double));
    L->have = 1;
    s->finalized = 0;
    return 0;
}

int zo_set_layer(zo_sim *s, int level, int kind, int layer, int64_t n, const uint64_t *keys) {
    if (level < 0 || level > s->p.max_level || kind < 0 || kind > 3 || layer < 0 || layer > 2) {
        zo_seterr(s, "bad layer args"); return -1;
    }
    zo_level *L = &s->lv[level];
    if (!L->have) { zo_seterr(s, "level not set"); return -1; }
    free(L->layers.idx[kind][layer]);
    L->layers.idx[kind][layer] = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    L->layers.n[kind][layer] = n;
    for (int64_t i = 0; i < n; ++i) {
        int64_t j = zo_find(L, keys[i]);
        if (j < 0) { zo_seterr(s, "layer node not in level %d", level); return -1; }
        L->layers.idx[kind][layer][i] = (int32_t)j;
    }
    return 0;
}

/* per level constants: LbmCollisionOptInterface ctor and CalRelaxationTime{,Ratio}
 * (collision_model/lbm_collision_models.cpp:105-151) */
static void zo_level_constants(zo_sim *s, int level) {
    zo_level *L = &s->lv[level];
    double nu = s->p.nu0;
    L->dt = 1. / (double)((uint32_t)1 << level);
    L->tau0 = nu * 3. / L->dt + 0.5;
    double relax_tau_f = nu * 3. / L->dt * 2. + 0.5;
    double relax_tau_c = nu * 3. / L->dt + 0.5;
    L->tau_stream_c2f = 0.5 * relax_tau_f / relax_tau_c;
    relax_tau_f = nu * 3. / L->dt + 0.5;
    relax_tau_c = nu * 3. / L->dt / 2 + 0.5;
    L->tau_stream_f2c = 2. * relax_tau_c / relax_tau_f;
    double sv[ZO_MAXQ], dv[ZO_MAXQ];
    double relax = 1. / L->tau0;
    zo_s_vector(s->p.dims, s->s_base, relax, sv);
    for (int i = 0; i < s->q; ++i) dv[i] = 1. - 0.5 * sv[i];
    zo_im_diag_m(s->q, s->m, s->minv, sv, L->imsm);
    zo_im_diag_m(s->q, s->m, s->minv, dv, L->imdm);
}

/* pull sources, as in SolverLbmD3Q19::Stream (lbm_d3q19_solver.cpp:20-135): the value of
 * direction q at x comes from x - c_q */
static void zo_build_sources(zo_sim *s, zo_level *L) {
    int q = s->q, dims = s->p.dims;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L->n; ++i) {
        for (int iq = 0; iq < q; ++iq) {
            int delta[3] = {-(int)s->cx[iq], -(int)s->cy[iq], -(int)s->cz[iq]};
            if (iq == 0) { L->src[i * q] = (int32_t)i; continue; }
            uint64_t k = zo_step_key(dims, L->key[i], delta);
            L->src[i * q + iq] = (int32_t)zo_find(L, k);
        }
    }
}

/* boundary lists: CheckIfNodeOutsideCubicDomain (grid/grid_info.cpp:84-142) with the level's
 * domain bounds (GridManager3D::CalDomainBoundsAtGivenLevel, grid_manager3d.cpp:165-205) */
static void zo_build_boundaries(zo_sim *s, int level) {
    zo_level *L = &s->lv[level];
    int dims = s->p.dims;
    for (int d = 0; d < 7; ++d) { free(L->bnd[d]); L->bnd[d] = NULL; L->nb[d] = 0; }
    for (int d = 1; d <= 2 * dims; ++d) L->bnd[d] = (int32_t *)malloc(sizeof(int32_t) * (L->n ? L->n : 1));
    for (int64_t i = 0; i < L->n; ++i) {
        uint32_t xyz[3];
        zo_decode(dims, L->key[i], xyz);
        for (int a = 0; a < dims; ++a) {
            uint64_t lo = (uint64_t)s->p.dom_min[a] << level, hi = (uint64_t)s->p.dom_max[a] << level;
            if (xyz[a] == lo) { int d = 1 + 2 * a; L->bnd[d][L->nb[d]++] = (int32_t)i; }
            if (xyz[a] == hi) { int d = 2 + 2 * a; L->bnd[d][L->nb[d]++] = (int32_t)i; }
        }
    }
}

int zo_finalize(zo_sim *s) {
    for (int l = 0; l <= s->p.max_level; ++l) {
        if (!s->lv[l].have) { zo_seterr(s, "level %d missing", l); return -1; }
        zo_level_constants(s, l);
        zo_build_sources(s, &s->lv[l]);
        zo_build_boundaries(s, l);
        for (int d = 1; d <= 2 * s->p.dims; ++d)
            if (s->lv[l].bnd_u[d] && s->lv[l].nbu[d] != s->lv[l].nb[d]) {
                zo_seterr(s, "level %d face %d: %lld boundary velocities for %lld boundary nodes", l, d,
                          (long long)s->lv[l].nbu[d], (long long)s->lv[l].nb[d]);
                return -1;
            }
    }
    for (int d = 1; d <= 2 * s->p.dims; ++d)
        if (s->p.bc_scheme[d] == ZO_BC_NONE) {
            zo_seterr(s, "boundary condition for direction %d not set", d);
            return -1;
        }
    s->finalized = 1;
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Per node physics (lbm_solver_interface.cpp)                                                 */
/* ------------------------------------------------------------------------------------------ */
/* CalMacro{2D,3D}{Compressible,Incompressible} (:477-532) */
static void zo_macro(const zo_sim *s, const double *f, double *rho, double *u) {
    int q = s->q;
    double r = f[0], ux = 0, uy = 0, uz = 0;
    for (int iq = 1; iq < q; ++iq) {
        r += f[iq];
        ux += s->cx[iq] * f[iq];
        uy += s->cy[iq] * f[iq];
        if (s->p.dims == 3) uz += s->cz[iq] * f[iq];
    }
    *rho = r; u[0] = ux; u[1] = uy; u[2] = uz;
    if (s->p.compressible) {
        u[0] /= r; u[1] /= r;
        if (s->p.dims == 3) u[2] /= r;
    }
}

/* CalMacroForce{2D,3D}{Compressible,Incompressible} (:541-592) */
static void zo_macro_force(const zo_sim *s, double dt, const double *f, const double *force,
                           double *rho, double *u) {
    int q = s->q;
    double r = f[0], ux = 0, uy = 0, uz = 0;
    for (int iq = 1; iq < q; ++iq) {
        r += f[iq];
        ux += s->cx[iq] * f[iq];
        uy += s->cy[iq] * f[iq];
        if (s->p.dims == 3) uz += s->cz[iq] * f[iq];
    }
    ux += 0.5 * force[0] * dt;
    uy += 0.5 * force[1] * dt;
    if (s->p.dims == 3) uz += 0.5 * force[2] * dt;
    *rho = r; u[0] = ux; u[1] = uy; u[2] = uz;
    if (s->p.compressible) {
        u[0] /= r; u[1] /= r;
        if (s->p.dims == 3) u[2] /= r;
    }
}

/* CalFeq{2D,3D}{Compressible,Incompressible} (:374-435) */
static void zo_feq(const zo_sim *s, double rho, const double *u, double *feq) {
    int q = s->q;
    double usq = u[0] * u[0] + u[1] * u[1];
    if (s->p.dims == 3) usq += u[2] * u[2];
    for (int iq = 0; iq < q; ++iq) {
        double cu = u[0] * s->cx[iq] + u[1] * s->cy[iq];
        if (s->p.dims == 3) cu += u[2] * s->cz[iq];
        if (s->p.compressible)
            feq[iq] = s->w[iq] * rho * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * usq);
        else
            feq[iq] = s->w[iq] * (rho + 3. * cu + 4.5 * cu * cu - 1.5 * usq);
    }
}

/* CalForcingTerming{2D,3D} (:443-470), Guo forcing */
static void zo_forcing(const zo_sim *s, const double *u, const double *F, double *fi) {
    for (int iq = 0; iq < s->q; ++iq) {
        if (s->p.dims == 2) {
            fi[iq] = s->w[iq] * (3. * ((s->cx[iq] - u[0]) * F[0] + (s->cy[iq] - u[1]) * F[1])
                + 9. * (u[0] * s->cx[iq] + u[1] * s->cy[iq]) * (s->cx[iq] * F[0] + s->cy[iq] * F[1]));
        } else {
            fi[iq] = s->w[iq] * (3. * ((s->cx[iq] - u[0]) * F[0] + (s->cy[iq] - u[1]) * F[1]
                + (s->cz[iq] - u[2]) * F[2])
                + 9. * (u[0] * s->cx[iq] + u[1] * s->cy[iq] + u[2] * s->cz[iq])
                * (s->cx[iq] * F[0] + s->cy[iq] * F[1] + s->cz[iq] * F[2]));
        }
    }
}

/* GetAllForcesForANode (:132-144): constant force plus the node's force slot */
static void zo_all_forces(const zo_sim *s, const double *node_force, double *F) {
    F[0] = s->p.const_force[0] + node_force[0];
    F[1] = s->p.const_force[1] + node_force[1];
    F[2] = s->p.dims == 3 ? s->p.const_force[2] + node_force[2] : 0.;
}

/* LesModelSmagorinsky::CalSgsRelaxationTime{Without,With}Force (les_models/lbm_les_smagorinsky.cpp:22-94) */
static double zo_les_tau_t(const zo_sim *s, double dt, double tau0, const double *f, const double *feq,
                           const double *F, const double *u) {
    double pxx = 0, pxy = 0, pyy = 0, pzz = 0, pxz = 0, pyz = 0;
    for (int iq = 1; iq < s->q; ++iq) {
        double d = f[iq] - feq[iq];
        pxx += s->cx[iq] * s->cx[iq] * d;
        pxy += s->cx[iq] * s->cy[iq] * d;
        pyy += s->cy[iq] * s->cy[iq] * d;
        if (s->p.dims == 3) {
            pzz += s->cz[iq] * s->cz[iq] * d;
            pxz += s->cx[iq] * s->cz[iq] * d;
            pyz += s->cy[iq] * s->cz[iq] * d;
        }
    }
    if (F) {
        pxx += dt * F[0] * u[0];
        pyy += dt * F[1] * u[1];
        pxy += 0.5 * dt * (F[0] * u[0] + F[1] * u[1]);
        if (s->p.dims == 3) {
            pzz += dt * F[2] * u[2];
            pxz += 0.5 * dt * (F[0] * u[0] + F[2] * u[2]);
            pyz += 0.5 * dt * (F[1] * u[1] + F[2] * u[2]);
        }
    }
    double pi_neq2 = pxx * pxx + pyy * pyy + pzz * pzz + 2. * (pxy * pxy + pxz * pxz + pyz * pyz);
    double k = 18. * s->p.cs * s->p.cs;
    return 0.5 * (sqrt(tau0 * tau0 + k * sqrt(2. * pi_neq2)) - tau0);
}

/* learned closure: features of P = Pi/rho and the FP64 network (mirrors the device closure) */
static double zo_mlp_tau_t(zo_sim *s, const double *pi, double rho, const double *u, double tau0, int *fb) {
    double ir = 1. / rho;
    double pxx = pi[0] * ir, pyy = pi[1] * ir, pzz = pi[2] * ir, pxy = pi[3] * ir, pxz = pi[4] * ir, pyz = pi[5] * ir;
    double pp = pxx * pxx + pyy * pyy + pzz * pzz + 2. * (pxy * pxy + pxz * pxz + pyz * pyz);
    double a = sqrt(2. * pp);
    double i1 = pxx + pyy + pzz;
    double i2 = 0.5 * (i1 * i1 - pp);
    double i3 = pxx * (pyy * pzz - pyz * pyz) - pxy * (pxy * pzz - pyz * pxz) + pxz * (pxy * pyz - pyy * pxz);
    double x[9] = {a, i1, i2, i3, u[0] * u[0] + u[1] * u[1] + u[2] * u[2], rho - 1., tau0, a * tau0, a * a};
    for (int k = 0; k < 9; ++k) {
        /* the device compares the float value of each input */
        float xf = (float)x[k];
        if (!(xf >= (float)s->mlp_min[k] && xf <= (float)s->mlp_max[k])) {
            *fb = 1;
            double q = sqrt(2. * (pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                                  2. * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5])));
            double k18 = 18. * s->p.cs * s->p.cs;
            return 0.5 * (sqrt(tau0 * tau0 + k18 * q) - tau0);
        }
    }
    *fb = 0;
    double xn[9], h1[32];
    for (int k = 0; k < 9; ++k) xn[k] = (x[k] - s->mlp_mean[k]) * s->mlp_istd[k];
    const double *W1 = s->mlp_p, *b1 = W1 + 288, *W2 = b1 + 32, *b2 = W2 + 1024, *w3 = b2 + 32, b3 = w3[32];
    for (int j = 0; j < 32; ++j) {
        double z = b1[j];
        for (int i = 0; i < 9; ++i) z += W1[j * 9 + i] * xn[i];
        h1[j] = z / (1. + exp(-z));
    }
    double out = b3;
    for (int j = 0; j < 32; ++j) {
        double z = b2[j];
        for (int i = 0; i < 32; ++i) z += W2[j * 32 + i] * h1[i];
        out += w3[j] * (z / (1. + exp(-z)));
    }
    double sp = out > 30. ? out : log1p(exp(out));
    return 3. * s->mlp_scale * sp;
}

int zo_set_mlp(zo_sim *s, const double *params, const double *mean, const double *std_dev, const double *fmin,
               const double *fmax, double nu_scale) {
    memcpy(s->mlp_p, params, sizeof(s->mlp_p));
    for (int k = 0; k < 9; ++k) {
        s->mlp_mean[k] = mean[k];
        s->mlp_istd[k] = 1. / std_dev[k];
        s->mlp_min[k] = fmin[k];
        s->mlp_max[k] = fmax[k];
    }
    s->mlp_scale = nu_scale;
    s->mlp_set = 1;
    return 0;
}

/* func_collision_node_ (lbm_solver_interface.cpp:266-300) with the SRT/MRT operators
 * (lbm_collision_model_srt.cpp:20-39, lbm_collision_model_mrt.cpp:114-143) */
static void zo_collide_node(zo_sim *s, const zo_level *L, int collision, int64_t i) {
    int q = s->q;
    const double *f = L->f + i * q;
    double *fc = L->fc + i * q;
    double feq[ZO_MAXQ], fi[ZO_MAXQ], F[3] = {0, 0, 0};
    double rho, u[3];
    int forces = s->p.has_force;
    if (forces) {
        zo_all_forces(s, L->force + i * 3, F);
        zo_macro_force(s, L->dt, f, F, &rho, u);
    } else {
        zo_macro(s, f, &rho, u);
    }
    L->rho[i] = rho;
    L->u[i * 3] = u[0]; L->u[i * 3 + 1] = u[1]; L->u[i * 3 + 2] = u[2];
    zo_feq(s, rho, u, feq);
    if (forces) zo_forcing(s, u, F, fi);
    double tau_eff = L->tau0;
    int les = s->p.les;
    if (les == 1) {
        double tau_t = zo_les_tau_t(s, L->dt, L->tau0, f, feq, forces ? F : NULL, u);
        tau_eff = s->p.les_tau_mode == 0 ? tau_t : L->tau0 + tau_t;
    } else if (les == 2) {
        double pi[6] = {0, 0, 0, 0, 0, 0};
        for (int iq = 1; iq < q; ++iq) {
            double d = f[iq] - feq[iq];
            pi[0] += s->cx[iq] * s->cx[iq] * d; pi[1] += s->cy[iq] * s->cy[iq] * d; pi[2] += s->cz[iq] * s->cz[iq] * d;
            pi[3] += s->cx[iq] * s->cy[iq] * d; pi[4] += s->cx[iq] * s->cz[iq] * d; pi[5] += s->cy[iq] * s->cz[iq] * d;
        }
        if (forces) {
            pi[0] += L->dt * F[0] * u[0];
            pi[1] += L->dt * F[1] * u[1];
            pi[3] += 0.5 * L->dt * (F[0] * u[0] + F[1] * u[1]);
            if (s->p.dims == 3) {
                pi[2] += L->dt * F[2] * u[2];
                pi[4] += 0.5 * L->dt * (F[0] * u[0] + F[2] * u[2]);
                pi[5] += 0.5 * L->dt * (F[1] * u[1] + F[2] * u[2]);
            }
        }
        int fb = 0;
        tau_eff = L->tau0 + zo_mlp_tau_t((zo_sim *)s, pi, rho, u, L->tau0, &fb);
    }
    if (collision == 0) {
        for (int iq = 0; iq < q; ++iq) {
            if (forces)
                fc[iq] = f[iq] - (f[iq] - feq[iq]) / tau_eff + (1. - 0.5 / tau_eff) * fi[iq] * L->dt;
            else
                fc[iq] = f[iq] - (f[iq] - feq[iq]) / tau_eff;
        }
    } else {
        const double *imsm = L->imsm, *imdm = L->imdm;
        double imsm_n[ZO_MAXQ * ZO_MAXQ], imdm_n[ZO_MAXQ * ZO_MAXQ];
        if (les) {
            double sv[ZO_MAXQ], dv[ZO_MAXQ];
            zo_s_vector(s->p.dims, s->s_base, 1. / tau_eff, sv);
            zo_im_diag_m(q, s->m, s->minv, sv, imsm_n);
            imsm = imsm_n;
            if (forces) {
                for (int k = 0; k < q; ++k) dv[k] = 1. - 0.5 * sv[k];
                zo_im_diag_m(q, s->m, s->minv, dv, imdm_n);
                imdm = imdm_n;
            }
        }
        for (int iq = 0; iq < q; ++iq) {
            double f_mrt = 0.;
            for (int is = 0; is < q; ++is) {
                if (forces)
                    f_mrt += imsm[iq * q + is] * (f[is] - feq[is]) - imdm[iq * q + is] * fi[is] * L->dt;
                else
                    f_mrt += imsm[iq * q + is] * (f[is] - feq[is]);
            }
            fc[iq] = f[iq] - f_mrt;
        }
    }
}

/* SolverLbmInterface::Collision (:306-323) */
static void zo_collision(zo_sim *s, int level) {
    zo_level *L = &s->lv[level];
    int coll = s->p.collision[level];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L->n; ++i) zo_collide_node(s, L, coll, i);
}

/* Stream: f[q](x) = f_collide[q](x - c_q) when the source exists, otherwise unchanged */
static void zo_stream(zo_sim *s, int level) {
    zo_level *L = &s->lv[level];
    int q = s->q;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L->n; ++i) {
        L->f[i * q] = L->fc[i * q];
        for (int iq = 1; iq < q; ++iq) {
            int32_t j = L->src[i * q + iq];
            if (j >= 0) L->f[i * q + iq] = L->fc[(int64_t)j * q + iq];
        }
    }
}

/* boundary index sets k0QIndicesNeg_/Pos_ (lbm_d3q19.h:86-95, lbm_d2q9.h:62-67) and
 * GetBoundaryNInverseIndices (lbm_boundary_conditions.cpp:102-156): at a min boundary the
 * incoming directions are those with c_axis = +1, the inverse ones have c_axis = -1 */
static int zo_bc_indices(const zo_sim *s, int dir, int32_t *ind, int32_t *inv) {
    int axis = (dir - 1) / 2;
    int sign_in = (dir % 2 == 1) ? +1 : -1;
    int n = 0;
    /* ascending index order; the writes of one boundary touch distinct entries, so the order of
     * the reference's lists does not change the result */
    for (int iq = 1; iq < s->q; ++iq) {
        double c = axis == 0 ? s->cx[iq] : (axis == 1 ? s->cy[iq] : s->cz[iq]);
        if ((int)c == sign_in) { ind[n] = iq; inv[n] = s->opp[iq]; ++n; }
    }
    return n;
}

static void zo_bc_bounce_back(zo_sim *s, zo_level *L, int dir) {
    int q = s->q;
    int32_t ind[8], inv[8];
    int nin = zo_bc_indices(s, dir, ind, inv);
    const double *v = s->p.bc_value[dir];
    double rho0 = s->p.rho0;
    /* BoundaryBounceBack{2D,3D}::CalBoundaryCondition (lbm_bounce_back.cpp:29-91) */
    for (int64_t b = 0; b < L->nb[dir]; ++b) {
        int64_t i = L->bnd[dir][b];
        for (int k = 0; k < nin; ++k) {
            int o = inv[k];
            double cu = s->cx[o] * v[0] + s->cy[o] * v[1];
            if (s->p.dims == 3) cu += s->cz[o] * v[2];
            L->f[i * q + ind[k]] = L->fc[i * q + o] - 6 * rho0 * s->w[o] * cu;
        }
    }
}

static void zo_bc_periodic(zo_sim *s, int level, int dir) {
    zo_level *L = &s->lv[level];
    int q = s->q, dims = s->p.dims;
    int axis = (dir - 1) / 2;
    int32_t ind[8], inv[8];
    int nin = zo_bc_indices(s, dir, ind, inv);
    uint64_t cur = zo_axis_mask(dims, axis);
    uint32_t tw = (dir % 2 == 1) ? s->p.dom_max[axis] : s->p.dom_min[axis];
    uint32_t tv[3] = {0, 0, 0};
    tv[axis] = tw << level;
    uint64_t counterpart_bits = zo_encode(dims, tv[0], tv[1], tv[2]);
    /* BoundaryPeriodic{2D,3D}::CalBoundaryCondition (lbm_periodic.cpp:22-133) */
    for (int64_t b = 0; b < L->nb[dir]; ++b) {
        int64_t i = L->bnd[dir][b];
        uint64_t ck = (L->key[i] & ~cur) | counterpart_bits;
        int64_t j = zo_find(L, ck);
        if (j < 0) continue;
        for (int k = 0; k < nin; ++k) L->f[i * q + ind[k]] = L->f[j * q + ind[k]];
    }
}

/* Parabolic inflow (no reference counterpart; the reference's faces take one velocity): the face
 * velocity times the product over the transverse axes of 4 t (1 - t), t the node's position across
 * the domain, 0 and 1 on the walls, so the profile peaks at the face velocity on the centre line
 * and vanishes on the walls. */
static double zo_profile_factor(const zo_sim *s, int level, int axis, uint64_t key) {
    uint32_t c[3] = {0, 0, 0};
    zo_decode(s->p.dims, key, c);
    double fac = 1.0;
    for (int a = 0; a < s->p.dims; ++a) {
        if (a == axis) continue;
        double lo = (double)((int64_t)s->p.dom_min[a] << level), hi = (double)((int64_t)s->p.dom_max[a] << level);
        double t = ((double)c[a] - lo) / (hi - lo);
        fac *= 4.0 * t * (1.0 - t);
    }
    return fac;
}

int zo_set_bc_profile(zo_sim *s, int dir, int profile) {
    if (dir < 1 || dir > 2 * s->p.dims || profile < 0 || profile > 1) {
        zo_seterr(s, "bad boundary profile %d on face %d", profile, dir);
        return -1;
    }
    s->bc_profile[dir] = profile;
    return 0;
}

/* per-node face velocities (no reference counterpart: the reference's faces take one velocity,
 * lbm_noneq_extrapolation.cpp:14-20), [n][3] in key order of the level's nodes on the face */
int zo_set_bc_velocities(zo_sim *s, int level, int dir, const double *u, int64_t n) {
    if (level < 0 || level > s->p.max_level || dir < 1 || dir > 2 * s->p.dims || n < 0 || (n > 0 && !u) ||
        s->p.bc_scheme[dir] != ZO_BC_NEE) {
        zo_seterr(s, "bad boundary velocities (level %d, face %d)", level, dir);
        return -1;
    }
    zo_level *L = &s->lv[level];
    free(L->bnd_u[dir]);
    L->bnd_u[dir] = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
    if (n) memcpy(L->bnd_u[dir], u, sizeof(double) * 3 * n);
    L->nbu[dir] = n;
    return 0;
}

static void zo_bc_nee(zo_sim *s, int level, int dir) {
    zo_level *L = &s->lv[level];
    int q = s->q, dims = s->p.dims;
    int axis = (dir - 1) / 2;
    int step = (dir % 2 == 1) ? +1 : -1;
    double feq_b[ZO_MAXQ], feq_in[ZO_MAXQ], rho_in, u_in[3];
    const double *bv = s->p.bc_value[dir];
    /* BoundaryNonEqExtrapolation{2D,3D}::CalBoundaryCondition (lbm_noneq_extrapolation.cpp:29-144) */
    for (int64_t b = 0; b < L->nb[dir]; ++b) {
        int64_t i = L->bnd[dir][b];
        uint64_t ik = zo_find_neighbor(dims, L->key[i], axis, step);
        int64_t j = zo_find(L, ik);
        if (j < 0) continue;
        double fac = s->bc_profile[dir] ? zo_profile_factor(s, level, axis, L->key[i]) : 1.0;
        double ub[3] = {bv[0] * fac, bv[1] * fac, bv[2] * fac};
        if (L->bnd_u[dir])
            for (int a = 0; a < 3; ++a) ub[a] = L->bnd_u[dir][b * 3 + a];
        if (s->p.has_force) {
            double F[3];
            zo_all_forces(s, L->force + j * 3, F);
            zo_macro_force(s, L->dt, L->f + j * q, F, &rho_in, u_in);
        } else {
            zo_macro(s, L->f + j * q, &rho_in, u_in);
        }
        zo_feq(s, rho_in, u_in, feq_in);
        zo_feq(s, s->p.rho0, ub, feq_b);
        for (int iq = 0; iq < q; ++iq) L->f[i * q + iq] = feq_b[iq] + L->f[j * q + iq] - feq_in[iq];
    }
}

/* ComputeDomainBoundaryCondition (lbm_grid_interface.cpp:225-314): all minimum faces in axis
 * order, then all maximum faces */
static void zo_boundaries(zo_sim *s, int level) {
    int dims = s->p.dims;
    int order[6];
    for (int a = 0; a < dims; ++a) { order[a] = 1 + 2 * a; order[dims + a] = 2 + 2 * a; }
    for (int k = 0; k < 2 * dims; ++k) {
        int dir = order[k];
        switch (s->p.bc_scheme[dir]) {
            case ZO_BC_BOUNCE_BACK: zo_bc_bounce_back(s, &s->lv[level], dir); break;
            case ZO_BC_PERIODIC: zo_bc_periodic(s, level, dir); break;
            case ZO_BC_NEE: zo_bc_nee(s, level, dir); break;
            default: break;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Coarse <-> fine transfer                                                                    */
/* ------------------------------------------------------------------------------------------ */
static int zo_periodic_axis(const zo_sim *s, int axis, int max_side) {
    int dir = 1 + 2 * axis + (max_side ? 1 : 0);
    return s->p.bc_scheme[dir] == ZO_BC_PERIODIC;
}

/*
 * One axis of FindNodesInPeriodicRegionCorner (grid/sfbitset_aux.cpp:380-557 (2D),
 * 2493-2878 (3D)).  Fills the 2*L axis coordinates of the region whose local index L-1 is the
 * start value; entries past a non-periodic domain edge stay invalid (-1).  At a periodic edge
 * the walk jumps to the opposite domain plane (which duplicates the edge plane).  Returns the
 * axis' limit on the valid half length.
 */
static int zo_region_axis(int64_t start, int L, int64_t lo, int64_t hi, int per_lo, int per_hi,
                          int64_t *seq) {
    int index_min = L;
    for (int k = 0; k < 2 * L; ++k) seq[k] = -1;
    int64_t v = start;
    for (int ix = 0; ix < L; ++ix) {            /* negative direction */
        seq[L - ix - 1] = v;
        if (v == lo) {
            if (per_lo) { v = hi + 1; }
            else { if (index_min > ix + 1) index_min = ix + 1; break; }
        }
        v = v > 0 ? v - 1 : v;
    }
    v = start;
    int go = 1;
    if (v == hi) {
        if (per_hi) v = lo > 0 ? lo - 1 : lo;
        else { index_min = 0; go = 0; }
    }
    if (go) {
        for (int ix = 0; ix < L; ++ix) {        /* positive direction */
            v = v + 1;
            seq[L + ix] = v;
            if (v == hi) {
                if (per_hi) v = lo > 0 ? lo - 1 : lo;
                else { if (index_min > ix + 1) index_min = ix + 1; break; }
            }
        }
    }
    return index_min;
}

/*
 * One axis of FindNodesInPeriodicRegionCornerOverlap (grid/sfbitset_aux.cpp:581-850 (2D),
 * 2893-3825 (3D)), the region walk of the immersed boundary in a periodic domain
 * (immersed_boundary.cpp:250-256).  Unlike zo_region_axis, a walk reaching a periodic domain
 * plane continues past its duplicate on the opposite side (lo -> hi - 1, hi -> lo + 1), and the
 * duplicate is returned in twin[] at the entry of the plane node (-1 elsewhere): it receives the
 * entry's spread force as well (immersed_boundary.cpp:341-363).  At a start node on a periodic
 * maximum plane the reference gives the duplicate the weight of the next entry, and the
 * duplicate rows of a region's first and last rows their neighbours' or none; here every
 * duplicate takes its own entry's weight (the reference's weights at those places are not
 * reproduced).  Returns the axis' bound on the valid half length.
 */
static int zo_region_axis_twin(int64_t start, int L, int64_t lo, int64_t hi, int per_lo, int per_hi,
                               int64_t *seq, int64_t *twin) {
    int index_min = L;
    for (int k = 0; k < 2 * L; ++k) seq[k] = twin[k] = -1;
    int64_t v = start;
    for (int ix = 0; ix < L; ++ix) {            /* negative direction */
        seq[L - ix - 1] = v;
        if (v == lo) {
            if (per_lo) { twin[L - ix - 1] = hi; v = hi; }
            else { if (index_min > ix + 1) index_min = ix + 1; break; }
        }
        v = v - 1;
    }
    v = start;
    if (v == hi) {
        if (per_hi) { twin[L - 1] = lo; v = lo; }
        else return 0;
    }
    for (int ix = 0; ix < L; ++ix) {            /* positive direction */
        v = v + 1;
        seq[L + ix] = v;
        if (v == hi) {
            if (per_hi) { twin[L + ix] = lo; v = lo; }
            else { if (index_min > ix + 1) index_min = ix + 1; break; }
        }
    }
    return index_min;
}

/* CalculateLagrangianInterpCoeff (grid/grid_interpolation.cpp:87-112) */
static void zo_lagrange_coeff(int half, double *c0, double *c1) {
    int num = 2 * half, offset = half - 1;
    for (int iy = 0; iy < num; ++iy) {
        c1[iy] = 1.;
        c0[iy] = 0.;
        double y = iy;
        for (int ix = 0; ix < num; ++ix) {
            if (iy != ix) {
                double x = ix;
                c1[iy] *= (0.5 + offset - x) / (y - x);
            }
        }
    }
    if (num > 0) c0[offset] = 1.;
}

/* NodeInfoCoarse2fine (lbm_grid_interface.cpp:570-588) with PostStreamCoarse2Fine
 * (lbm_collision_models.cpp:244-251) using the coarse level's tau ratio */
static void zo_node_c2f(const zo_sim *s, const zo_level *C, const double *fj, double *rho, double *u, double *f) {
    double feq[ZO_MAXQ];
    zo_macro(s, fj, rho, u);
    zo_feq(s, *rho, u, feq);
    for (int iq = 0; iq < s->q; ++iq)
        f[iq] = feq[iq] + C->tau_stream_c2f * (fj[iq] - feq[iq]);
}

/* Time interpolation (no reference counterpart: the reference refills fine ghost layers only
 * after even substeps, when both levels have reached the same time, lbm_solver_interface.cpp:
 * 224-238).  When enabled, the ghost layers are also refilled after odd substeps, at the fine
 * level's half-way time between two coarse steps: every coarse contributor then gives the mean of
 * its transferred values (zo_node_c2f) before and after the coarse level's last step, i.e. a
 * linear interpolation in time; fine fallback contributors give their current values.  The
 * even-substep transfers are unchanged. */
int zo_set_time_interp(zo_sim *s, int on) {
    s->time_interp = on != 0;
    return 0;
}

/* interpolation method of a fine level's ghost nodes (GridInfoInterface::interp_method_,
 * grid_info_interface.h:317; ChooseInterpolationMethod, grid/grid_interpolation.cpp:20-82) */
int zo_set_interp_method(zo_sim *s, int level, int method) {
    if (level < 1 || level > s->p.max_level || (method != 0 && method != 1)) {
        zo_seterr(s, "bad interpolation method %d at level %d", method, level);
        return -1;
    }
    s->interp_method[level] = method;
    return 0;
}

/* InterpolationLagrangian{2D,3D} (grid/grid_interpolation.cpp:224-304, 433-525) and
 * InterpolationLinear{2D,3D} (:126-208, 319-417) for one fine ghost node; returns 0 or -1 when a
 * contributor is missing in both levels.  Linear is the Lagrangian form with max_interp_length_
 * 1 and the half-length-1 weights ((1, 0) at even, (1/2, 1/2) at odd fine coordinates) whatever
 * the region's valid length. */
static int zo_interp_node(zo_sim *s, int fine_level, int64_t i, int tmean) {
    zo_level *F = &s->lv[fine_level], *C = &s->lv[fine_level - 1];
    int dims = s->p.dims, q = s->q, clevel = fine_level - 1;
    const int linear = s->interp_method[fine_level] == 1;
    /* max_interp_length_ (grid_info_interface.h:317-318; 1 for kLinear, grid_interpolation.cpp:23) */
    const int Lmax = linear ? 1 : 2;
    uint64_t parent = F->key[i] >> dims;
    uint32_t pxyz[3];
    zo_decode(dims, parent, pxyz);
    int64_t seq[3][4];
    int valid = Lmax;
    for (int a = 0; a < dims; ++a) {
        int64_t lo = (int64_t)s->p.dom_min[a] << clevel, hi = (int64_t)s->p.dom_max[a] << clevel;
        int v = zo_region_axis(pxyz[a], Lmax, lo, hi, zo_periodic_axis(s, a, 0), zo_periodic_axis(s, a, 1),
                               seq[a]);
        if (v < valid) valid = v;
    }
    /* SetNodeVariablesAsZeros (lbm_grid_interface.cpp:45-55) */
    F->rho[i] = 0.;
    F->u[i * 3] = F->u[i * 3 + 1] = F->u[i * 3 + 2] = 0.;
    for (int iq = 0; iq < q; ++iq) F->f[i * q + iq] = 0.;
    const int half = linear ? 1 : valid;
    int num = 2 * half;
    double c0[4], c1[4], cf[3][4];
    zo_lagrange_coeff(half, c0, c1);
    uint32_t fxyz[3];
    zo_decode(dims, F->key[i], fxyz);
    for (int a = 0; a < dims; ++a)
        for (int k = 0; k < num; ++k) cf[a][k] = (fxyz[a] & 1u) ? c1[k] : c0[k];
    int nz = dims == 3 ? num : 1;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < num; ++iy)
            for (int ix = 0; ix < num; ++ix) {
                double coeff = dims == 3 ? cf[0][ix] * cf[1][iy] * cf[2][iz] : cf[0][ix] * cf[1][iy];
                if (!(fabs(coeff) > ZO_KEPS)) continue;
                int off = Lmax - half;
                if (seq[0][ix + off] < 0 || seq[1][iy + off] < 0 || (dims == 3 && seq[2][iz + off] < 0)) return -1;
                uint32_t c[3] = {(uint32_t)seq[0][ix + off], (uint32_t)seq[1][iy + off],
                                 dims == 3 ? (uint32_t)seq[2][iz + off] : 0};
                uint64_t ck = zo_encode(dims, c[0], c[1], c[2]);
                int64_t j = zo_find(C, ck);
                double r, uu[3], ff[ZO_MAXQ];
                const double *fr;
                if (j >= 0 && !(C->flag[j] & ZO_FLAG_C2F_GHOST)) {
                    zo_node_c2f(s, C, C->f + j * q, &r, uu, ff);
                    if (tmean) {
                        /* mean of the contributor's values before and after the coarse step */
                        double r0, u0[3] = {0., 0., 0.}, f0[ZO_MAXQ];
                        zo_node_c2f(s, C, C->f_prev + j * q, &r0, u0, f0);
                        r = 0.5 * (r0 + r);
                        for (int a = 0; a < dims; ++a) uu[a] = 0.5 * (u0[a] + uu[a]);
                        for (int iq = 0; iq < q; ++iq) ff[iq] = 0.5 * (f0[iq] + ff[iq]);
                    }
                    fr = ff;
                } else {
                    int64_t k = zo_find(F, ck << dims);
                    if (k < 0) return -1;
                    r = F->rho[k];
                    uu[0] = F->u[k * 3]; uu[1] = F->u[k * 3 + 1]; uu[2] = F->u[k * 3 + 2];
                    fr = F->f + k * q;
                }
                /* GridNodeLbm::InterpolationAdditionAssignCoefficient (lbm_grid_interface.cpp:26-40) */
                F->rho[i] += r * coeff;
                for (int a = 0; a < dims; ++a) F->u[i * 3 + a] = F->u[i * 3 + a] + coeff * uu[a];
                for (int iq = 0; iq < q; ++iq) F->f[i * q + iq] = F->f[i * q + iq] + coeff * fr[iq];
            }
    return 0;
}

/* TransferInfoFromCoarseGrid (lbm_grid_interface.cpp:408-515): fine layers [1, 3) of every
 * interface, inner then outer per layer.  List order is ascending key (the reference's is the
 * unordered_map order; the results do not depend on it for grids whose ghost nodes never
 * serve as fallback contributors of each other). */
static int zo_transfer_from_coarse(zo_sim *s, int fine_level, int tmean) {
    zo_level *F = &s->lv[fine_level];
    int errs = 0;
    if (tmean && !s->lv[fine_level - 1].f_prev) {
        zo_seterr(s, "time interpolation into level %d before a step of level %d", fine_level, fine_level - 1);
        return -1;
    }
    for (int layer = 1; layer < 3; ++layer) {
        for (int kind = ZO_INNER_F2C; kind >= ZO_OUTER_F2C; kind -= 2) {
            for (int64_t k = 0; k < F->layers.n[kind][layer]; ++k)
                if (zo_interp_node(s, fine_level, F->layers.idx[kind][layer][k], tmean)) ++errs;
        }
    }
    if (errs) zo_seterr(s, "%d interpolation stencils incomplete at level %d", errs, fine_level);
    return errs ? -1 : 0;
}

/* TransferInfoToCoarseGrid (lbm_grid_interface.cpp:524-564) with NodeInfoFine2Coarse
 * (:594-610) and PostStreamFine2Coarse (lbm_collision_models.cpp:259-267) */
static void zo_transfer_to_coarse(zo_sim *s, int fine_level) {
    zo_level *F = &s->lv[fine_level], *C = &s->lv[fine_level - 1];
    int q = s->q, dims = s->p.dims;
    for (int kind = ZO_INNER_C2F; kind >= ZO_OUTER_C2F; kind -= 2) {
        for (int64_t k = 0; k < C->layers.n[kind][1]; ++k) {
            int64_t j = C->layers.idx[kind][1][k];
            int64_t i = zo_find(F, C->key[j] << dims);
            if (i < 0) continue;
            double feq[ZO_MAXQ];
            C->rho[j] = F->rho[i];
            C->u[j * 3] = F->u[i * 3]; C->u[j * 3 + 1] = F->u[i * 3 + 1]; C->u[j * 3 + 2] = F->u[i * 3 + 2];
            zo_feq(s, F->rho[i], F->u + i * 3, feq);
            for (int iq = 0; iq < q; ++iq)
                C->f[j * q + iq] = feq[iq] + F->tau_stream_f2c * (F->f[i * q + iq] - feq[iq]);
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Immersed boundary direct forcing (immersed_boundary.cpp:35-384)                            */
/* ------------------------------------------------------------------------------------------ */
static double zo_stencil_one(double d) { return fabs(d) < 1. ? 1. - fabs(d) : 0.; }
static double zo_stencil_two(double d) {
    if (d < 2.) {
        double a = fabs(d), sq = d * d;
        if (a < 1.) return 0.125 * (3. - 2. * a + sqrt(1. + 4. * a - 4. * sq));
        return 0.125 * (5. - 2. * a - sqrt(-7. + 12. * a - 4. * sq));
    }
    return 0.;
}

static int zo_ib_level(zo_sim *s, int level) {
    zo_level *L = &s->lv[level];
    zo_ib *ib = &L->ib;
    if (ib->n == 0) return 0;
    int dims = s->p.dims, q = s->q;
    double dx = 1. / (double)((uint32_t)1 << level);   /* background spacing 1 */
    unsigned char *touched = (unsigned char *)calloc(L->n ? L->n : 1, 1);
    int errs = 0;
    /* DomainInfo::bool_periodic_domain_: any periodic face selects the overlap walk */
    int periodic = 0;
    for (int d = 1; d <= 2 * dims; ++d) periodic |= s->p.bc_scheme[d] == ZO_BC_PERIODIC;
    for (int64_t m = 0; m < ib->n; ++m) {      /* marker order: index order */
        double iref[3], dis[3];
        uint32_t idx[3] = {0, 0, 0};
        for (int a = 0; a < dims; ++a) {
            iref[a] = ib->xyz[m * 3 + a] / dx;
            idx[a] = (uint32_t)(iref[a] + ZO_KEPS);
            dis[a] = iref[a] - idx[a];
        }
        int Ls = ib->stencil_dis;
        int64_t seq[3][8], twin[3][8];
        int valid = Ls;
        for (int a = 0; a < dims; ++a) {
            int64_t lo = (int64_t)s->p.dom_min[a] << level, hi = (int64_t)s->p.dom_max[a] << level;
            int v = periodic ? zo_region_axis_twin(idx[a], Ls, lo, hi, zo_periodic_axis(s, a, 0),
                                                   zo_periodic_axis(s, a, 1), seq[a], twin[a])
                             : zo_region_axis(idx[a], Ls, lo, hi, 0, 0, seq[a]);
            if (!periodic)
                for (int k = 0; k < 2 * Ls; ++k) twin[a][k] = -1;
            if (v < valid) valid = v;
        }
        for (int a = dims; a < 3; ++a)
            for (int k = 0; k < 8; ++k) { seq[a][k] = 0; twin[a][k] = -1; }
        double (*stencil)(double) = valid == 2 ? zo_stencil_two : zo_stencil_one;
        int n1 = 2 * valid;
        int64_t nodes[512];
        double delta[512];
        int entry_off[512][3];   /* the entry's offsets (ox, oy, oz) */
        int cnt = 0;
        double uref[3] = {0, 0, 0};
        int nz = dims == 3 ? n1 : 1;
        for (int iz = 0; iz < nz; ++iz) {
            int oz = iz - valid;
            double dz = dims == 3 ? stencil(fabs(dis[2] - oz - 1)) : 1.;
            for (int iy = 0; iy < n1; ++iy) {
                int oy = iy - valid;
                double dy = stencil(fabs(dis[1] - oy - 1));
                for (int ix = 0; ix < n1; ++ix) {
                    int ox = ix - valid;
                    double dxw = stencil(fabs(dis[0] - ox - 1));
                    int64_t c[3] = {seq[0][Ls + ox], seq[1][Ls + oy], dims == 3 ? seq[2][Ls + oz] : 0};
                    entry_off[cnt][0] = ox; entry_off[cnt][1] = oy; entry_off[cnt][2] = dims == 3 ? oz : 0;
                    if (c[0] < 0 || c[1] < 0 || c[2] < 0) { ++errs; nodes[cnt] = -1; delta[cnt++] = 0; continue; }
                    uint64_t key = zo_encode(dims, (uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2]);
                    int64_t j = zo_find(L, key);
                    if (j < 0) { ++errs; nodes[cnt] = -1; delta[cnt++] = 0; continue; }
                    if (!touched[j]) {
                        /* lazy reset: zero the node force and recompute u with the remaining forces */
                        double F[3];
                        L->force[j * 3] = L->force[j * 3 + 1] = L->force[j * 3 + 2] = 0.;
                        zo_all_forces(s, L->force + j * 3, F);
                        zo_macro_force(s, L->dt, L->f + j * q, F, &L->rho[j], L->u + j * 3);
                        touched[j] = 1;
                    }
                    double d = dims == 3 ? dxw * dy * dz : dxw * dy;
                    nodes[cnt] = j; delta[cnt++] = d;
                    uref[0] += L->u[j * 3] * d;
                    uref[1] += L->u[j * 3 + 1] * d;
                    if (dims == 3) uref[2] += L->u[j * 3 + 2] * d;
                }
            }
        }
        double Fm[3];
        for (int a = 0; a < dims; ++a) Fm[a] = ib->stiffness * (ib->vel[m * 3 + a] - uref[a]);
        for (int a = 0; a < 3; ++a) ib->force[m * 3 + a] = a < dims ? Fm[a] : 0.;
        double area = dims == 3 ? ib->area[m] / L->dt / dx / dx : ib->area[m] / L->dt / dx;
        for (int k = 0; k < cnt; ++k) {
            int64_t j = nodes[k];
            if (j < 0) continue;
            double dt = delta[k] * area;
            for (int a = 0; a < dims; ++a) L->force[j * 3 + a] += Fm[a] * dt;
        }
        if (periodic) {
            /* duplicates of the stencil nodes on periodic planes: every combination of an entry's
             * coordinate or its twin per axis, other than the entry itself, in entry order */
            for (int k = 0; k < cnt; ++k) {
                if (nodes[k] < 0) continue;
                int64_t opt[3][2];
                int nopt[3];
                for (int a = 0; a < 3; ++a) {
                    int o = entry_off[k][a];
                    opt[a][0] = a < dims ? seq[a][Ls + o] : 0;
                    nopt[a] = 1;
                    if (a < dims && twin[a][Ls + o] >= 0) opt[a][nopt[a]++] = twin[a][Ls + o];
                }
                for (int cz = 0; cz < nopt[2]; ++cz)
                    for (int cy = 0; cy < nopt[1]; ++cy)
                        for (int cx = 0; cx < nopt[0]; ++cx) {
                            if (cx == 0 && cy == 0 && cz == 0) continue;
                            uint64_t key = zo_encode(dims, (uint32_t)opt[0][cx], (uint32_t)opt[1][cy],
                                                     (uint32_t)opt[2][cz]);
                            int64_t j = zo_find(L, key);
                            if (j < 0) { ++errs; continue; }
                            if (!touched[j]) {
                                double F[3];
                                L->force[j * 3] = L->force[j * 3 + 1] = L->force[j * 3 + 2] = 0.;
                                zo_all_forces(s, L->force + j * 3, F);
                                zo_macro_force(s, L->dt, L->f + j * q, F, &L->rho[j], L->u + j * 3);
                                touched[j] = 1;
                            }
                            double dt = delta[k] * area;
                            for (int a = 0; a < dims; ++a) L->force[j * 3 + a] += Fm[a] * dt;
                        }
            }
        }
    }
    free(touched);
    if (errs) { zo_seterr(s, "%d IB stencil nodes missing at level %d", errs, level); return -1; }
    return 0;
}

int zo_set_ib(zo_sim *s, int level, int64_t n, const double *xyz, const double *area,
              const double *vel, int stencil_dis, double stiffness, double dx_background) {
    if (level < 0 || level > s->p.max_level || stencil_dis < 1 || stencil_dis > 4 || dx_background != 1.) {
        zo_seterr(s, "bad IB args"); return -1;
    }
    zo_ib *ib = &s->lv[level].ib;
    free(ib->xyz); free(ib->area); free(ib->vel); free(ib->force);
    ib->n = n;
    ib->xyz = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
    ib->vel = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
    ib->force = (double *)calloc(3 * (n ? n : 1), sizeof(double));
    ib->area = (double *)malloc(sizeof(double) * (n ? n : 1));
    memcpy(ib->xyz, xyz, sizeof(double) * 3 * n);
    memcpy(ib->vel, vel, sizeof(double) * 3 * n);
    memcpy(ib->area, area, sizeof(double) * n);
    ib->stencil_dis = stencil_dis;
    ib->stiffness = stiffness;
    return 0;
}

int zo_get_ib_force(const zo_sim *s, int level, double *force) {
    const zo_ib *ib = &s->lv[level].ib;
    memcpy(force, ib->force, sizeof(double) * 3 * ib->n);
    return 0;
}

/* test hook: the coarse/fine transfer of one refined level on the current populations */
int zo_debug_transfer(zo_sim *s, int fine_level) {
    if (fine_level < 1 || fine_level > s->p.max_level) { zo_seterr(s, "bad level"); return -1; }
    int rc = zo_transfer_from_coarse(s, fine_level, 0);
    zo_transfer_to_coarse(s, fine_level);
    return rc;
}

/* test hook: replace the non-viscous MRT relaxation rates (all levels) */
int zo_debug_set_s_base(zo_sim *s, const double *s_base) {
    memcpy(s->s_base, s_base, sizeof(double) * s->q);
    for (int l = 0; l <= s->p.max_level; ++l)
        if (s->lv[l].have) zo_level_constants(s, l);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
int zo_advance_level(zo_sim *s, int level, int time_step_level) {
    if (!s->finalized) { zo_seterr(s, "not finalized"); return -1; }
    int rc = 0;
    zo_level *L = &s->lv[level];
    if (s->time_interp && level < s->p.max_level) {     /* the state the next finer level interpolates from */
        if (!L->f_prev && !(L->f_prev = (double *)malloc(sizeof(double) * (size_t)L->n * s->q + 1))) {
            zo_seterr(s, "out of memory");
            return -1;
        }
        memcpy(L->f_prev, L->f, sizeof(double) * (size_t)L->n * s->q);
    }
    if (zo_ib_level(s, level)) rc = -1;                  /* UpdateCriterion (:61-95) */
    zo_collision(s, level);                              /* RunSolverOnGivenGrid (:246-262) */
    zo_stream(s, level);
    zo_boundaries(s, level);                             /* ComputeDomainBoundaryCondition */
    if (level > 0 && time_step_level % 2 == 0) {         /* InformationFromGridOfDifferentLevel (:224-238) */
        if (zo_transfer_from_coarse(s, level, 0)) rc = -1;
        zo_transfer_to_coarse(s, level);
    } else if (level > 0 && s->time_interp) {            /* time interpolation (zo_set_time_interp) */
        if (zo_transfer_from_coarse(s, level, 1)) rc = -1;
    }
    return rc;
}

/* MultiTimeSteppingC2F restated as the recursion it produces: a level step is followed by two
 * steps of the next finer level, the finest level's two steps being consecutive. */
static void zo_order_rec(int level, int max_level, int32_t *order, int *n, int cap) {
    if (*n < cap) order[*n] = level;
    ++*n;
    if (level < max_level) {
        zo_order_rec(level + 1, max_level, order, n, cap);
        zo_order_rec(level + 1, max_level, order, n, cap);
    }
}

int zo_time_stepping_order(int max_level, int32_t *order, int cap) {
    int n = 0;
    zo_order_rec(0, max_level, order, &n, cap);
    return n;
}

int zo_time_marching(zo_sim *s, int nsteps) {
    int32_t order[1 << 12];
    int n = zo_time_stepping_order(s->p.max_level, order, 1 << 12);
    if (n > (1 << 12)) { zo_seterr(s, "too many levels"); return -1; }
    int rc = 0;
    for (int t = 0; t < nsteps; ++t) {
        int tsl[ZO_MAXL] = {0};
        for (int k = 0; k < n; ++k) {
            int l = order[k];
            ++tsl[l];
            if (zo_advance_level(s, l, tsl[l])) rc = -1;
        }
    }
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
int zo_init_feq(zo_sim *s) {
    double feq[ZO_MAXQ];
    zo_feq(s, s->p.rho0, s->p.u0, feq);
    for (int l = 0; l <= s->p.max_level; ++l) {
        zo_level *L = &s->lv[l];
        for (int64_t i = 0; i < L->n; ++i) {
            memcpy(L->f + i * s->q, feq, sizeof(double) * s->q);
            memcpy(L->fc + i * s->q, feq, sizeof(double) * s->q);
            L->rho[i] = s->p.rho0;
            for (int a = 0; a < 3; ++a) L->u[i * 3 + a] = s->p.u0[a];
        }
    }
    return 0;
}

int zo_init_macro(zo_sim *s, int level, const double *rho, const double *u) {
    zo_level *L = &s->lv[level];
    int d = s->p.dims;
    for (int64_t i = 0; i < L->n; ++i) {
        double uu[3] = {u[i * d], u[i * d + 1], d == 3 ? u[i * d + 2] : 0.};
        zo_feq(s, rho[i], uu, L->f + i * s->q);
        memcpy(L->fc + i * s->q, L->f + i * s->q, sizeof(double) * s->q);
        L->rho[i] = rho[i];
        for (int a = 0; a < 3; ++a) L->u[i * 3 + a] = uu[a];
    }
    return 0;
}

int zo_set_f(zo_sim *s, int level, const double *f) {
    zo_level *L = &s->lv[level];
    memcpy(L->f, f, sizeof(double) * L->n * s->q);
    memcpy(L->fc, f, sizeof(double) * L->n * s->q);
    return 0;
}
int zo_get_f(const zo_sim *s, int level, double *f) {
    memcpy(f, s->lv[level].f, sizeof(double) * s->lv[level].n * s->q);
    return 0;
}
int zo_get_fc(const zo_sim *s, int level, double *fc) {
    memcpy(fc, s->lv[level].fc, sizeof(double) * s->lv[level].n * s->q);
    return 0;
}
int zo_get_node_macro(const zo_sim *s, int level, double *rho, double *u) {
    const zo_level *L = &s->lv[level];
    memcpy(rho, L->rho, sizeof(double) * L->n);
    for (int64_t i = 0; i < L->n; ++i)
        for (int a = 0; a < s->p.dims; ++a) u[i * s->p.dims + a] = L->u[i * 3 + a];
    return 0;
}
int zo_set_node_force(zo_sim *s, int level, const double *force) {
    zo_level *L = &s->lv[level];
    for (int64_t i = 0; i < L->n; ++i)
        for (int a = 0; a < s->p.dims; ++a) L->force[i * 3 + a] = force[i * s->p.dims + a];
    return 0;
}
int zo_get_node_force(const zo_sim *s, int level, double *force) {
    const zo_level *L = &s->lv[level];
    for (int64_t i = 0; i < L->n; ++i)
        for (int a = 0; a < s->p.dims; ++a) force[i * s->p.dims + a] = L->force[i * 3 + a];
    return 0;
}
int64_t zo_level_size(const zo_sim *s, int level) { return s->lv[level].n; }
int zo_get_source_table(const zo_sim *s, int level, int32_t *src) {
    memcpy(src, s->lv[level].src, sizeof(int32_t) * s->lv[level].n * s->q);
    return 0;
}
int64_t zo_get_boundary(const zo_sim *s, int level, int dir, int32_t *idx) {
    const zo_level *L = &s->lv[level];
    if (dir < 1 || dir > 6) return -1;
    if (idx && L->nb[dir]) memcpy(idx, L->bnd[dir], sizeof(int32_t) * L->nb[dir]);
    return L->nb[dir];
}

int zo_conservation(const zo_sim *s, int level, double *out4) {
    const zo_level *L = &s->lv[level];
    double m = 0, jx = 0, jy = 0, jz = 0;
    for (int64_t i = 0; i < L->n; ++i)
        for (int iq = 0; iq < s->q; ++iq) {
            double f = L->f[i * s->q + iq];
            m += f; jx += s->cx[iq] * f; jy += s->cy[iq] * f; jz += s->cz[iq] * f;
        }
    out4[0] = m; out4[1] = jx; out4[2] = jy; out4[3] = jz;
    return 0;
}
