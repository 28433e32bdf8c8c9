/* sph_oracle.h — CPU f64 restatement of phys-rs's SPH step (test infrastructure only). */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H
#include <stdint.h>
#include "../include/sphx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Particle state in id order (index i is the entity with ECS order id[i]). */
typedef struct OrcState {
    uint32_t n;
    double *x, *y, *z, *vx, *vy, *vz, *m, *h;
    double *rho, *p;
    double *ax, *ay, *az;    /* Acceleration                */
    double *apx, *apy, *apz; /* PreviousAcceleration        */
    double *bx, *by, *bz, *psi, *q; /* optional (NULL)      */
    uint32_t* id;
    double *u, *du, *dup;    /* SPHX_EOS_ADIABATIC: specific internal energy, du/dt, previous du/dt */
} OrcState;

/* ---- pair geometry shared by the pair form (sph_oracle.c) and the gather
 *      form (sph_omp.c) -------------------------------------------------- */
#include <math.h>
#include "../include/sphx.h"

/* pair vector: core/coordinates.rs:51-65 (one sector: 0 + (b - a)),
 * plus minimum image for the new periodic box */
static inline void pair_vec(const OrcState* s, const SphxParams* p, uint32_t a, uint32_t b,
                            double* dx, double* dy, double* dz) {
    double d[3] = {s->x[b] - s->x[a], s->y[b] - s->y[a], s->z[b] - s->z[a]};
    if (p->boundary_flags & SPHX_BND_PERIODIC) {
        for (int k = 0; k < 3; ++k) {
            double L = p->box_hi[k] - p->box_lo[k];
            if (d[k] > 0.5 * L) d[k] -= L;
            else if (d[k] < -0.5 * L) d[k] += L;
        }
    }
    *dx = d[0]; *dy = d[1]; *dz = d[2];
}

/* glam DVec3::length: sqrt((x*x + y*y) + z*z) */
static inline double vlen(double x, double y, double z) { return sqrt((x * x + y * y) + z * z); }
static inline double vdot(double ax, double ay, double az, double bx, double by, double bz) {
    return (ax * bx + ay * by) + az * bz;
}

/* neighbour predicate, sph.rs:102-108 */
static inline int is_pair(const OrcState* s, const SphxParams* p, uint32_t a, uint32_t b) {
    double dx, dy, dz;
    pair_vec(s, p, a, b, &dx, &dy, &dz);
    double r = vlen(dx, dy, dz);
    double h_avg = 0.5 * (s->h[a] + s->h[b]);
    return r < 2.0 * h_avg;
}

double orc_kernel_w(double r, double h);
double orc_kernel_dw(double r, double h);
uint64_t orc_pairs_brute(const OrcState* s, const SphxParams* p, uint32_t* pairs, uint64_t cap);
uint64_t orc_pairs_cells(const OrcState* s, const SphxParams* p, uint32_t* pairs, uint64_t cap);
void orc_density(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np);
void orc_eos(OrcState* s, const SphxParams* p);
void orc_eval(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np);
void orc_eval_state(OrcState* s, const SphxParams* p);
void orc_vv_position(OrcState* s, const SphxParams* p);
void orc_vv_velocity(OrcState* s, const SphxParams* p);
void orc_xsph(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np);
void orc_mhd_induction(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np);
void orc_div_clean(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np);
void orc_step(OrcState* s, const SphxParams* p, uint32_t nsteps);
void orc_energy(const OrcState* s, const SphxParams* p, double out[8]);
/* sph_adjoint.c */
int orc_adjoint(OrcState* s, const SphxParams* p, uint32_t nsteps, const double* target, double* loss,
                double* gx, double* gv, double* galpha);
double orc_loss(OrcState* s, const SphxParams* p, uint32_t nsteps, const double* target);

#ifdef __cplusplus
}
#endif
#endif
