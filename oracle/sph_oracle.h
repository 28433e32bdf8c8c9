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
