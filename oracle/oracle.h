/*
 * oracle.h -- CPU restatement of v2mesh's relaxation (relax_cpu and the
 * per-node bodies it runs), in plain C with OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY. It is the checker the HIP path is compared with,
 * bit for bit, and nothing in the product links it. It follows
 * src/v2m_particles.cpp:728-976 and src/opencl/{v2m_grid_body,v2m_seed_body,
 * v2m_particle_body}.cl operation for operation (same association, no
 * contraction: build with -ffp-contract=off), for label volumes (gm == 0)
 * and interface trapping (voxel_trap == 0).
 *
 * Parity pins (tests/test_oracle_*.py): every stage against fixtures written by
 * the compiled reference (tests/golden), and against the reference library
 * itself where it is built (oracle/_ref); closed-form and brute-force checks
 * for the hash, sort, neighbour lists and forces.
 */
#ifndef TN_ORACLE_H
#define TN_ORACLE_H

#include <stdint.h>

#include "../include/trussnet_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One record per iteration of the loop, for stage-by-stage comparison. */
typedef struct orc_iter_trace {
    int32_t it;
    int32_t rebuilt;      /* a rebuild ran at the end of this iteration */
    int32_t nhalf;        /* nodes past skin/2 since the last rebuild */
    int32_t nrefresh;     /* nodes that refreshed their own list (past a full skin) */
    float mmax, p99;
    float fdt, falpha;    /* FIRE's step and steering after this iteration's update */
    double power;         /* FIRE's power: sum of pw, serial, in double */
    int32_t hist[32];
} orc_iter_trace;

/* hash geometry (src/v2m_particles.cpp:738-761) */
void orc_hash_geometry(const tn_grid_view* g, const tn_relax_params* prm, tn_hash_info* H);

/* the first rebuild on the given nodes (src/v2m_particles.cpp:767-820) */
int orc_rebuild(const tn_grid_view* g, const tn_relax_params* prm, const tn_nodes_view* nd, int32_t* key,
                int32_t* sorted, int32_t* cstart, int32_t* nbr, int32_t* nnb, float* hn);

/* v2m_force for every node (src/opencl/v2m_particle_body.cl:161-184) */
void orc_force(const tn_relax_params* prm, const tn_nodes_view* nd, const float* hn, const int32_t* nbr,
               const int32_t* nnb, float* F);

/* v2m_move for every node (src/opencl/v2m_particle_body.cl:844-935) */
int orc_move(const tn_grid_view* g, const tn_relax_params* prm, tn_nodes_view* nd, float* hn, const float* F,
             float* V, float fdt, float falpha, float* pw, float* mv);

/* relax_cpu (src/v2m_particles.cpp:728-976); trace may be null, else it holds
 * max_iters records. nthreads <= 0: OpenMP's default. */
int orc_relax(const tn_grid_view* g, const tn_relax_params* prm, tn_nodes_view* nd, tn_relax_stats* st,
              orc_iter_trace* trace, int32_t nthreads);

/* point queries: phi_a, psi_ab and its gradient, label_of and the runner-up */
/* label l's analytic field (and gradient, grad != 0) of a shape program at m points */
void orc_sdf_points(const float* prog, int32_t l, int64_t m, const float* X, int32_t grad, float* v, float* g);
void orc_field_points(const tn_grid_view* g, int32_t a, int32_t b, int64_t m, const float* X, float* phi_a,
                      float* psi, float* grad, int32_t* label, int32_t* second);

/* the histogram bin of a move (src/v2m_particles.cpp:881) */
int32_t orc_hist_bin(float mv);

/* The differentiable relaxation of tn_adj_* in double precision (oracle/adjoint.c):
 * the loss after p->iters steps (xT, terms may be null), and the loss with its
 * gradients with respect to x0 [3n] and theta [gx gy gz]. */
double orc_adj_loss(const tn_adj_problem* p, const double* x0, const double* theta, double* xT, double* terms);
double orc_adj_grad(const tn_adj_problem* p, const double* x0, const double* theta, double* gx0, double* gtheta,
                    double* terms);

/* oracle/thin.c: thin_nodes (src/v2m_particles.cpp:1223-1357) with thin factor B, and
 * protect_features (src/v2m_particles.cpp:1002-1221) with the grid's feature curves
 * (Grid::feat, nfeat floats): keep[i] = 1 for the nodes kept (the reference compacts
 * them in index order); the count removed is returned. part3 may be null (no third
 * partners). */
int64_t orc_thin_nodes(const tn_grid_view* g, float B, int64_t n, const float* P, const uint16_t* lab,
                       const uint8_t* typ, const uint16_t* part, const uint16_t* part3, uint8_t* keep);
int64_t orc_protect_features(const tn_grid_view* g, const float* feat, int64_t nfeat, int64_t n, const float* P,
                             const uint8_t* typ, uint8_t* keep);

const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
