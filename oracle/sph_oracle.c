/*
 * sph_oracle.c — CPU f64 restatement of phys-rs's SPH step (TEST INFRASTRUCTURE).
 *
 * This file is the oracle the HIP path is checked against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product (libsphx.so) never links or calls it.
 *
 * It restates the reference CPU path in the reference's own structure:
 *   - an undirected pair list (a, b) with id_a < id_b, the set CandidatePairs
 *     holds once each SPH system applies its r < 2*h_avg filter
 *     (src/physics/broadphase.rs:57-96, src/physics/sph.rs:102-108);
 *   - two-sided accumulation per pair, one system at a time, in the order of
 *     the velocity-Verlet schedule (src/main.rs:103-157);
 *   - the reference's arithmetic order term by term, compiled with
 *     -ffp-contract=off (Rust does not contract a*b+c).
 * Pair order is (id_a, id_b) ascending; the reference's order follows kiddo and
 * bevy iteration, so sums agree with it up to summation order.
 *
 * Physics the reference lacks (uniform gravity, reflective and periodic boxes,
 * polytropic EOS, Maxwell-stress force) is new; each such function says so and
 * is pinned by closed forms and invariants in tests/test_oracle.py.
 *
 * Parity pins (what checks each function): see the table in the header of
 * tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sphx.h"
#include "sph_oracle.h"

#define ORC_PI 3.141592653589793 /* std::f64::consts::PI */

/* ---- kernels: src/physics/sph.rs:29-43 and 50-64 ---------------------- */
double orc_kernel_w(double r, double h) {
    double q = r / h;
    if (q >= 2.0) return 0.0;
    double sigma = 1.0 / (ORC_PI * ((h * h) * h)); /* h.powi(3) */
    if (q < 1.0) {
        return sigma * ((1.0 - 1.5 * (q * q)) + 0.75 * ((q * q) * q));
    }
    double t = 2.0 - q;
    return (sigma * 0.25) * ((t * t) * t);
}

double orc_kernel_dw(double r, double h) {
    double q = r / h;
    if (q >= 2.0 || r < 1e-10) return 0.0;
    double sigma = 1.0 / (ORC_PI * ((h * h) * (h * h))); /* h.powi(4) */
    if (q < 1.0) {
        return sigma * ((-3.0 * q) + 2.25 * (q * q));
    }
    double t = 2.0 - q;
    return sigma * (-0.75 * (t * t));
}

static int cmp_pair(const void* A, const void* B) {
    const uint32_t* a = (const uint32_t*)A;
    const uint32_t* b = (const uint32_t*)B;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
    return 0;
}

/* Orders each pair so that id[first] < id[second] (entity_a < entity_b,
 * broadphase.rs:87) and sorts the list by (id_a, id_b). */
static void orient_and_sort(const OrcState* s, uint32_t* pairs, uint64_t np) {
    for (uint64_t k = 0; k < np; ++k) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        if (s->id[a] > s->id[b]) { pairs[2 * k] = b; pairs[2 * k + 1] = a; }
    }
    /* sort by ids: map to ids, sort, map back through an inverse table */
    uint32_t maxid = 0;
    for (uint32_t i = 0; i < s->n; ++i) if (s->id[i] > maxid) maxid = s->id[i];
    uint32_t* inv = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)maxid + 1));
    for (uint32_t i = 0; i < s->n; ++i) inv[s->id[i]] = i;
    for (uint64_t k = 0; k < 2 * np; ++k) pairs[k] = s->id[pairs[k]];
    qsort(pairs, np, 2 * sizeof(uint32_t), cmp_pair);
    for (uint64_t k = 0; k < 2 * np; ++k) pairs[k] = inv[pairs[k]];
    free(inv);
}

/* Brute-force O(N^2) pair builder: the predicate the reference's own tests
 * use to inject CandidatePairs (sph.rs:551-556, 617-628; tests/gpu_bench.rs:113-124),
 * filtered by r < 2*h_avg as every SPH system does. Returns the total count;
 * writes at most cap pairs (particle indices, oriented by id, sorted by id). */
uint64_t orc_pairs_brute(const OrcState* s, const SphxParams* p, uint32_t* pairs, uint64_t cap) {
    uint64_t np = 0;
    for (uint32_t a = 0; a < s->n; ++a)
        for (uint32_t b = a + 1; b < s->n; ++b)
            if (is_pair(s, p, a, b)) {
                if (np < cap) { pairs[2 * np] = a; pairs[2 * np + 1] = b; }
                ++np;
            }
    if (np <= cap) orient_and_sort(s, pairs, np);
    return np;
}

/* Cell-list pair builder with the identical predicate (used at sizes where
 * O(N^2) is too slow). Cell size >= 2*max(h); periodic boxes wrap cells. */
uint64_t orc_pairs_cells(const OrcState* s, const SphxParams* p, uint32_t* pairs, uint64_t cap) {
    uint32_t n = s->n;
    if (n == 0) return 0;
    double hmax = 0.0;
    for (uint32_t i = 0; i < n; ++i) if (s->h[i] > hmax) hmax = s->h[i];
    double cs = 2.0 * hmax * (1.0 + 1e-9);
    int periodic = (p->boundary_flags & SPHX_BND_PERIODIC) != 0;
    double lo[3], hi[3];
    int nc[3];
    if (periodic) {
        for (int k = 0; k < 3; ++k) { lo[k] = p->box_lo[k]; hi[k] = p->box_hi[k]; }
    } else {
        lo[0] = hi[0] = s->x[0]; lo[1] = hi[1] = s->y[0]; lo[2] = hi[2] = s->z[0];
        for (uint32_t i = 1; i < n; ++i) {
            double c[3] = {s->x[i], s->y[i], s->z[i]};
            for (int k = 0; k < 3; ++k) { if (c[k] < lo[k]) lo[k] = c[k]; if (c[k] > hi[k]) hi[k] = c[k]; }
        }
    }
    for (int k = 0; k < 3; ++k) {
        double ext = hi[k] - lo[k];
        if (periodic) {
            nc[k] = (int)floor(ext / cs);
            if (nc[k] < 3) nc[k] = 1; /* too small to wrap a 3-stencil: one cell, all pairs tested */
        } else {
            nc[k] = (int)(ext / cs) + 1;
        }
        if (nc[k] < 1) nc[k] = 1;
        if (nc[k] > 1024) nc[k] = 1024;
    }
    double w[3];
    for (int k = 0; k < 3; ++k) w[k] = periodic ? (hi[k] - lo[k]) / nc[k] : cs;
    /* widen cells (never narrow them) until the grid has at most 2^22 cells */
    while ((size_t)nc[0] * nc[1] * nc[2] > ((size_t)1 << 22)) {
        for (int k = 0; k < 3; ++k) {
            if (nc[k] < 6) continue;
            if (periodic) { nc[k] /= 2; w[k] = (hi[k] - lo[k]) / nc[k]; }
            else { w[k] *= 2.0; nc[k] = (int)((hi[k] - lo[k]) / w[k]) + 1; }
        }
    }
    size_t ncell = (size_t)nc[0] * nc[1] * nc[2];
    uint32_t* head = (uint32_t*)malloc(sizeof(uint32_t) * ncell);
    uint32_t* next = (uint32_t*)malloc(sizeof(uint32_t) * n);
    int* cidx = (int*)malloc(sizeof(int) * 3 * n);
    for (size_t c = 0; c < ncell; ++c) head[c] = UINT32_MAX;
    for (uint32_t i = n; i-- > 0;) {
        double c[3] = {s->x[i], s->y[i], s->z[i]};
        int ci[3];
        for (int k = 0; k < 3; ++k) {
            long v = (long)floor((c[k] - lo[k]) / w[k]);
            if (v < 0) v = 0;
            if (v >= nc[k]) v = nc[k] - 1;
            ci[k] = (int)v;
            cidx[3 * i + k] = ci[k];
        }
        size_t cell = ((size_t)ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
        next[i] = head[cell];
        head[cell] = i;
    }
    uint64_t np = 0;
    for (uint32_t a = 0; a < n; ++a) {
        int ci[3] = {cidx[3 * a], cidx[3 * a + 1], cidx[3 * a + 2]};
        int span[3][3], ns[3];
        for (int k = 0; k < 3; ++k) {
            ns[k] = 0;
            for (int o = -1; o <= 1; ++o) {
                int v = ci[k] + o;
                if (periodic) { v = (v + nc[k]) % nc[k]; }
                else if (v < 0 || v >= nc[k]) continue;
                int dup = 0;
                for (int t = 0; t < ns[k]; ++t) if (span[k][t] == v) dup = 1;
                if (!dup) span[k][ns[k]++] = v;
            }
        }
        for (int iz = 0; iz < ns[2]; ++iz)
            for (int iy = 0; iy < ns[1]; ++iy)
                for (int ix = 0; ix < ns[0]; ++ix) {
                    size_t cell = ((size_t)span[2][iz] * nc[1] + span[1][iy]) * nc[0] + span[0][ix];
                    for (uint32_t b = head[cell]; b != UINT32_MAX; b = next[b]) {
                        if (b <= a) continue;
                        if (is_pair(s, p, a, b)) {
                            if (np < cap) { pairs[2 * np] = a; pairs[2 * np + 1] = b; }
                            ++np;
                        }
                    }
                }
    }
    free(head); free(next); free(cidx);
    if (np <= cap) orient_and_sort(s, pairs, np);
    return np;
}

/* ---- density: sph_density_system, sph.rs:72-116 ----------------------- */
void orc_density(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np) {
    for (uint32_t i = 0; i < s->n; ++i) s->rho[i] = s->m[i] * orc_kernel_w(0.0, s->h[i]);
    for (uint64_t k = 0; k < np; ++k) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        double dx, dy, dz;
        pair_vec(s, p, a, b, &dx, &dy, &dz);
        double r = vlen(dx, dy, dz);
        double h_avg = 0.5 * (s->h[a] + s->h[b]);
        if (r < 2.0 * h_avg) {
            double w = orc_kernel_w(r, h_avg);
            s->rho[a] += s->m[b] * w;
            s->rho[b] += s->m[a] * w;
        }
    }
}

/* ---- EOS: sph_eos_system, sph.rs:149-169 (Tait); polytropic is new;
 *      isothermal follows ideal_gas_pressure, thermodynamics.rs:232-238;
 *      adiabatic (new) is the ideal gas at the particle's specific internal
 *      energy u = InternalEnergy / Mass (components/thermal.rs:9) ---------- */
void orc_eos(OrcState* s, const SphxParams* p) {
    for (uint32_t i = 0; i < s->n; ++i) {
        double P;
        if (p->eos == SPHX_EOS_TAIT) {
            double B = p->rest_density * (p->speed_of_sound * p->speed_of_sound) / p->gamma;
            double ratio = s->rho[i] / p->rest_density;
            P = B * (pow(ratio, p->gamma) - 1.0);
            if (P < 0.0) P = 0.0;
        } else if (p->eos == SPHX_EOS_POLYTROPIC) {
            P = p->poly_k * pow(s->rho[i], p->gamma);
        } else if (p->eos == SPHX_EOS_ADIABATIC) {
            P = (p->gamma - 1.0) * s->rho[i] * s->u[i];
        } else {
            P = s->rho[i] * p->iso_cs2;
        }
        s->p[i] = P;
    }
}

/* ---- force systems: accumulate into F (forces.rs:15-19 zeroes it) ----- */

/* Uniform gravity (new): F += m g, placed where the reference's gravity
 * systems sit in the chain, before the SPH forces (main.rs:109-110). */
static void orc_gravity(const OrcState* s, const SphxParams* p, double* F) {
    for (uint32_t i = 0; i < s->n; ++i) {
        F[3 * i + 0] += s->m[i] * p->gravity[0];
        F[3 * i + 1] += s->m[i] * p->gravity[1];
        F[3 * i + 2] += s->m[i] * p->gravity[2];
    }
}

/* lorentz_force_system, lorentz.rs:21-39: F += q (E + v x B) */
static void orc_lorentz(const OrcState* s, const SphxParams* p, double* F) {
    if (!s->q || !s->bx) return;
    for (uint32_t i = 0; i < s->n; ++i) {
        double vx = s->vx[i], vy = s->vy[i], vz = s->vz[i];
        double bx = s->bx[i], by = s->by[i], bz = s->bz[i];
        double cx = vy * bz - vz * by, cy = vz * bx - vx * bz, cz = vx * by - vy * bx;
        double q = s->q[i];
        F[3 * i + 0] += q * (p->e_field[0] + cx);
        F[3 * i + 1] += q * (p->e_field[1] + cy);
        F[3 * i + 2] += q * (p->e_field[2] + cz);
    }
}

/* The energy equation of the adiabatic EOS (new; the reference has no SPH
 * energy equation): du_a/dt = sum_b m_b (P_a/rho_a^2 + Pi_ab/2) v_ab . grad_a W_ab
 * (Monaghan 1992, Annu. Rev. Astron. Astrophys. 30:543, eq. 2.23 with the
 * viscous term), the rate that makes sum (m v^2/2 + m u) conserved under the
 * pressure and viscosity forces below.  Accumulated by those two pair loops. */
static inline int energy_on(const OrcState* s, const SphxParams* p) {
    return p->eos == SPHX_EOS_ADIABATIC && s->u && s->du;
}

/* sph_pressure_force_system, sph.rs:177-228 */
static void orc_pressure(const OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np,
                         double* F) {
    for (uint64_t k = 0; k < np; ++k) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        double dx, dy, dz;
        pair_vec(s, p, a, b, &dx, &dy, &dz);
        double r = vlen(dx, dy, dz);
        double h_avg = 0.5 * (s->h[a] + s->h[b]);
        if (r > 1e-10 && r < 2.0 * h_avg) {
            double rx = dx / r, ry = dy / r, rz = dz / r;
            double dw = orc_kernel_dw(r, h_avg);
            double term = (s->p[a] / (s->rho[a] * s->rho[a])) + (s->p[b] / (s->rho[b] * s->rho[b]));
            double mag = ((s->m[a] * s->m[b]) * term) * dw;
            double fx = rx * mag, fy = ry * mag, fz = rz * mag;
            F[3 * a + 0] += fx; F[3 * a + 1] += fy; F[3 * a + 2] += fz;
            F[3 * b + 0] -= fx; F[3 * b + 1] -= fy; F[3 * b + 2] -= fz;
            if (energy_on(s, p)) {
                /* v_ab . grad_a W = -((v_a - v_b) . (x_b - x_a)) dW / r, the same for both ends */
                double vg = -vdot(s->vx[a] - s->vx[b], s->vy[a] - s->vy[b], s->vz[a] - s->vz[b], dx, dy, dz) *
                            (dw / r);
                s->du[a] += s->m[b] * (s->p[a] / (s->rho[a] * s->rho[a])) * vg;
                s->du[b] += s->m[a] * (s->p[b] / (s->rho[b] * s->rho[b])) * vg;
            }
        }
    }
}

/* sph_viscosity_system, sph.rs:260-322.  REFERENCE convention: the test is
 * (v_a - v_b).(x_b - x_a) < 0 exactly as line 295 evaluates it; MONAGHAN
 * flips the sign of the test (approaching pairs) and leaves the rest. */
static void orc_viscosity(const OrcState* s, const SphxParams* p, const uint32_t* pairs,
                          uint64_t np, double* F) {
    double c = p->speed_of_sound;
    double eps = 1e-2;
    for (uint64_t k = 0; k < np; ++k) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        double dx, dy, dz;
        pair_vec(s, p, a, b, &dx, &dy, &dz);
        double vrx = s->vx[a] - s->vx[b], vry = s->vy[a] - s->vy[b], vrz = s->vz[a] - s->vz[b];
        double sdot = vdot(vrx, vry, vrz, dx, dy, dz);
        int active = (p->visc_convention == SPHX_VISC_MONAGHAN) ? (sdot > 0.0) : (sdot < 0.0);
        if (!active) continue;
        double r = vlen(dx, dy, dz);
        double h_avg = 0.5 * (s->h[a] + s->h[b]);
        if (r < 2.0 * h_avg) {
            double rx = dx / r, ry = dy / r, rz = dz / r;
            double dw = orc_kernel_dw(r, h_avg);
            double rho_bar = 0.5 * (s->rho[a] + s->rho[b]);
            /* MONAGHAN measures mu along x_a - x_b, i.e. with the opposite sign */
            double sgn_s = (p->visc_convention == SPHX_VISC_MONAGHAN) ? -sdot : sdot;
            double mu = (h_avg * sgn_s) / ((r * r) + eps * (h_avg * h_avg));
            double pi_ij = ((-p->visc_alpha) * c * mu + p->visc_beta * (mu * mu)) / rho_bar;
            double mag = ((s->m[a] * s->m[b]) * pi_ij) * dw;
            double fx = rx * mag, fy = ry * mag, fz = rz * mag;
            F[3 * a + 0] += fx; F[3 * a + 1] += fy; F[3 * a + 2] += fz;
            F[3 * b + 0] -= fx; F[3 * b + 1] -= fy; F[3 * b + 2] -= fz;
            if (energy_on(s, p)) {
                double vg = -sdot * (dw / r); /* v_ab . grad_a W */
                s->du[a] += s->m[b] * (0.5 * pi_ij) * vg;
                s->du[b] += s->m[a] * (0.5 * pi_ij) * vg;
            }
        }
    }
}

/* SPMHD Maxwell-stress force (new; mu0 = 1):
 * F_a += m_a m_b (M_a/rho_a^2 + M_b/rho_b^2) . grad_a W_ab, M = B B - |B|^2 I / 2,
 * grad_a W_ab = -(dW/dr) r_hat with r_hat = (x_b - x_a)/r, and the opposite on b. */
static void orc_maxwell(const OrcState* s, const SphxParams* p, const uint32_t* pairs,
                        uint64_t np, double* F) {
    if (!s->bx) return;
    for (uint64_t k = 0; k < np; ++k) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        double dx, dy, dz;
        pair_vec(s, p, a, b, &dx, &dy, &dz);
        double r = vlen(dx, dy, dz);
        double h_avg = 0.5 * (s->h[a] + s->h[b]);
        if (!(r > 1e-10 && r < 2.0 * h_avg)) continue;
        double dw = orc_kernel_dw(r, h_avg);
        double g[3] = {-(dx / r) * dw, -(dy / r) * dw, -(dz / r) * dw}; /* grad_a W */
        double Ba[3] = {s->bx[a], s->by[a], s->bz[a]};
        double Bb[3] = {s->bx[b], s->by[b], s->bz[b]};
        double ia = 1.0 / (s->rho[a] * s->rho[a]), ib = 1.0 / (s->rho[b] * s->rho[b]);
        double Ba2 = vdot(Ba[0], Ba[1], Ba[2], Ba[0], Ba[1], Ba[2]);
        double Bb2 = vdot(Bb[0], Bb[1], Bb[2], Bb[0], Bb[1], Bb[2]);
        double Bag = vdot(Ba[0], Ba[1], Ba[2], g[0], g[1], g[2]);
        double Bbg = vdot(Bb[0], Bb[1], Bb[2], g[0], g[1], g[2]);
        double mm = s->m[a] * s->m[b];
        for (int c = 0; c < 3; ++c) {
            double Mag = Ba[c] * Bag - 0.5 * Ba2 * g[c];
            double Mbg = Bb[c] * Bbg - 0.5 * Bb2 * g[c];
            double f = mm * (Mag * ia + Mbg * ib);
            F[3 * a + c] += f;
            F[3 * b + c] -= f;
        }
    }
}

/* sph_boundary_system, sph.rs:443-484 */
static void orc_boundary_penalty(const OrcState* s, const SphxParams* p, double* F) {
    double ext = p->half_extent, ad = p->activation_distance, K = p->stiffness;
    for (uint32_t i = 0; i < s->n; ++i) {
        double c[3] = {s->x[i], s->y[i], s->z[i]};
        for (int k = 0; k < 3; ++k) {
            double d_low = c[k] - (-ext);
            if (d_low < ad && d_low >= 0.0) F[3 * i + k] += K * (ad - d_low);
            double d_high = ext - c[k];
            if (d_high < ad && d_high >= 0.0) F[3 * i + k] -= K * (ad - d_high);
        }
    }
}

/* One force evaluation of the schedule (main.rs:105-119):
 * reset_forces, [gravity], lorentz, density, eos, pressure, viscosity,
 * [maxwell], boundary, compute_acceleration (integrators.rs:110-118). */
void orc_eval(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np) {
    uint32_t n = s->n;
    double* F = (double*)calloc((size_t)3 * (n ? n : 1), sizeof(double));
    if (energy_on(s, p)) memset(s->du, 0, sizeof(double) * n);
    orc_gravity(s, p, F);
    if (p->mhd_flags & SPHX_MHD_LORENTZ_CHARGE) orc_lorentz(s, p, F);
    orc_density(s, p, pairs, np);
    orc_eos(s, p);
    orc_pressure(s, p, pairs, np, F);
    orc_viscosity(s, p, pairs, np, F);
    if (p->mhd_flags & SPHX_MHD_MAXWELL_STRESS) orc_maxwell(s, p, pairs, np, F);
    if (p->boundary_flags & SPHX_BND_PENALTY_CUBE) orc_boundary_penalty(s, p, F);
    for (uint32_t i = 0; i < n; ++i) {
        s->ax[i] = F[3 * i + 0] / s->m[i];
        s->ay[i] = F[3 * i + 1] / s->m[i];
        s->az[i] = F[3 * i + 2] / s->m[i];
    }
    free(F);
}

/* velocity_verlet_position_system, integrators.rs:63-83, followed by the new
 * reflective / periodic box handling. Reflection mirrors x about the wall and
 * flips that component of v and of the stored acceleration. */
void orc_vv_position(OrcState* s, const SphxParams* p) {
    double dt = p->dt;
    double half_dt_sq = 0.5 * dt * dt;
    for (uint32_t i = 0; i < s->n; ++i) {
        double* x[3] = {&s->x[i], &s->y[i], &s->z[i]};
        double* v[3] = {&s->vx[i], &s->vy[i], &s->vz[i]};
        double* a[3] = {&s->ax[i], &s->ay[i], &s->az[i]};
        double* ap[3] = {&s->apx[i], &s->apy[i], &s->apz[i]};
        for (int k = 0; k < 3; ++k) {
            *ap[k] = *a[k];
            *x[k] += *v[k] * dt + *a[k] * half_dt_sq;
        }
        if (energy_on(s, p)) s->dup[i] = s->du[i];
        for (int k = 0; k < 3; ++k) {
            double lo = p->box_lo[k], hi = p->box_hi[k];
            if (p->boundary_flags & SPHX_BND_REFLECT) {
                if (*x[k] < lo) { *x[k] = lo + (lo - *x[k]); *v[k] = -*v[k]; *ap[k] = -*ap[k]; }
                else if (*x[k] > hi) { *x[k] = hi - (*x[k] - hi); *v[k] = -*v[k]; *ap[k] = -*ap[k]; }
            } else if (p->boundary_flags & SPHX_BND_PERIODIC) {
                double L = hi - lo;
                if (*x[k] >= hi) *x[k] -= L;
                else if (*x[k] < lo) *x[k] += L;
            }
        }
    }
}

/* velocity_verlet_velocity_system, integrators.rs:92-104 */
void orc_vv_velocity(OrcState* s, const SphxParams* p) {
    double half_dt = 0.5 * p->dt;
    for (uint32_t i = 0; i < s->n; ++i) {
        s->vx[i] += (s->apx[i] + s->ax[i]) * half_dt;
        s->vy[i] += (s->apy[i] + s->ay[i]) * half_dt;
        s->vz[i] += (s->apz[i] + s->az[i]) * half_dt;
        /* the internal energy is advanced like the velocity (new) */
        if (energy_on(s, p)) s->u[i] += (s->dup[i] + s->du[i]) * half_dt;
    }
}

/* sph_xsph_system, sph.rs:349-405 (Jacobi: all corrections from pre-pass v) */
void orc_xsph(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np) {
    uint32_t n = s->n;
    double* dv = (double*)calloc((size_t)3 * (n ? n : 1), sizeof(double));
    double eps = p->xsph_epsilon;
    for (uint64_t k = 0; k < np; ++k) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        double dx, dy, dz;
        pair_vec(s, p, a, b, &dx, &dy, &dz);
        double r = vlen(dx, dy, dz);
        double h_avg = 0.5 * (s->h[a] + s->h[b]);
        if (r < 2.0 * h_avg) {
            double w = orc_kernel_w(r, h_avg);
            double rho_bar = 0.5 * (s->rho[a] + s->rho[b]);
            double vd[3] = {s->vx[b] - s->vx[a], s->vy[b] - s->vy[a], s->vz[b] - s->vz[a]};
            double fa = eps * (s->m[b] / rho_bar), fb = eps * (s->m[a] / rho_bar);
            for (int c = 0; c < 3; ++c) {
                dv[3 * a + c] += (fa * vd[c]) * w;
                dv[3 * b + c] += (fb * (-vd[c])) * w;
            }
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        s->vx[i] += dv[3 * i]; s->vy[i] += dv[3 * i + 1]; s->vz[i] += dv[3 * i + 2];
    }
    free(dv);
}

/* mhd_induction_system, mhd.rs:68-152 */
void orc_mhd_induction(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np) {
    uint32_t n = s->n;
    double* dB = (double*)calloc((size_t)3 * (n ? n : 1), sizeof(double));
    double dt = p->dt, eta = p->resistivity;
    for (uint64_t k = 0; k < np; ++k) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        double dx, dy, dz;
        pair_vec(s, p, a, b, &dx, &dy, &dz);
        double r = vlen(dx, dy, dz);
        double h_avg = 0.5 * (s->h[a] + s->h[b]);
        if (!(r > 1e-10 && r < 2.0 * h_avg)) continue;
        double dw = orc_kernel_dw(r, h_avg);
        double g[3] = {(dx / r) * dw, (dy / r) * dw, (dz / r) * dw};
        double vij[3] = {s->vx[a] - s->vx[b], s->vy[a] - s->vy[b], s->vz[a] - s->vz[b]};
        double B1[3] = {s->bx[a], s->by[a], s->bz[a]};
        double B2[3] = {s->bx[b], s->by[b], s->bz[b]};
        double v_dot_g = vdot(vij[0], vij[1], vij[2], g[0], g[1], g[2]);
        double b1g = vdot(B1[0], B1[1], B1[2], g[0], g[1], g[2]);
        double b2g = vdot(B2[0], B2[1], B2[2], g[0], g[1], g[2]);
        double f12 = s->m[b] / (s->rho[a] * s->rho[b]);
        double f21 = s->m[a] / (s->rho[b] * s->rho[a]);
        double r_dot_g = vdot(dx, dy, dz, g[0], g[1], g[2]);
        double denom = (r * r) + 0.01 * (h_avg * h_avg);
        for (int c = 0; c < 3; ++c) {
            double db1_ideal = f12 * (B1[c] * v_dot_g - vij[c] * b1g);
            double db2_ideal = f21 * (B2[c] * (-v_dot_g) - (-vij[c]) * b2g);
            double bd = B2[c] - B1[c];
            double db1_res = 2.0 * eta * (s->m[b] / s->rho[b]) * bd * r_dot_g / denom;
            double db2_res = 2.0 * eta * (s->m[a] / s->rho[a]) * (-bd) * (-r_dot_g) / denom;
            dB[3 * a + c] += (db1_ideal + db1_res) * dt;
            dB[3 * b + c] += (db2_ideal + db2_res) * dt;
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        s->bx[i] += dB[3 * i]; s->by[i] += dB[3 * i + 1]; s->bz[i] += dB[3 * i + 2];
    }
    free(dB);
}

/* divergence_cleaning_system, mhd.rs:172-259 */
void orc_div_clean(OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np) {
    uint32_t n = s->n;
    double* dB = (double*)calloc((size_t)3 * (n ? n : 1), sizeof(double));
    double* dpsi = (double*)calloc(n ? n : 1, sizeof(double));
    double dt = p->dt, ch = p->cleaning_speed, ch2 = ch * ch, sigma = p->cleaning_damping;
    for (uint64_t k = 0; k < np; ++k) {
        uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        double dx, dy, dz;
        pair_vec(s, p, a, b, &dx, &dy, &dz);
        double r = vlen(dx, dy, dz);
        double h_avg = 0.5 * (s->h[a] + s->h[b]);
        if (!(r > 1e-10 && r < 2.0 * h_avg)) continue;
        double dw = orc_kernel_dw(r, h_avg);
        double g[3] = {(dx / r) * dw, (dy / r) * dw, (dz / r) * dw};
        double bd[3] = {s->bx[b] - s->bx[a], s->by[b] - s->by[a], s->bz[b] - s->bz[a]};
        double div1 = (s->m[b] / s->rho[b]) * vdot(bd[0], bd[1], bd[2], g[0], g[1], g[2]);
        double div2 = (s->m[a] / s->rho[a]) * vdot(-bd[0], -bd[1], -bd[2], -g[0], -g[1], -g[2]);
        double psid = s->psi[b] - s->psi[a];
        for (int c = 0; c < 3; ++c) {
            double gp1 = (s->m[b] / s->rho[b]) * psid * g[c];
            double gp2 = (s->m[a] / s->rho[a]) * (-psid) * (-g[c]);
            dB[3 * a + c] += -gp1 * dt;
            dB[3 * b + c] += -gp2 * dt;
        }
        dpsi[a] += -ch2 * div1 * dt;
        dpsi[b] += -ch2 * div2 * dt;
    }
    for (uint32_t i = 0; i < n; ++i) {
        s->bx[i] += dB[3 * i]; s->by[i] += dB[3 * i + 1]; s->bz[i] += dB[3 * i + 2];
        s->psi[i] += dpsi[i];
    }
    for (uint32_t i = 0; i < n; ++i) {
        double decay = exp(-sigma * ch * dt / s->h[i]);
        s->psi[i] *= decay;
    }
    free(dB); free(dpsi);
}

/* (Re)build the pair list for the current positions. */
static uint32_t* build_pairs(const OrcState* s, const SphxParams* p, uint64_t* np_out) {
    uint64_t cap = (uint64_t)s->n * 64 + 1024;
    for (;;) {
        uint32_t* pairs = (uint32_t*)malloc(sizeof(uint32_t) * 2 * cap);
        uint64_t np = orc_pairs_cells(s, p, pairs, cap);
        if (np <= cap) { *np_out = np; return pairs; }
        free(pairs);
        cap = np;
    }
}

/* The velocity-Verlet tick, main.rs:103-157, nsteps times.  Each force
 * evaluation rebuilds the pair list from the current positions, as the
 * schedule runs build_broadphase/broadphase_query before each chain. */
void orc_step(OrcState* s, const SphxParams* p, uint32_t nsteps) {
    for (uint32_t t = 0; t < nsteps; ++t) {
        uint64_t np;
        uint32_t* pairs = build_pairs(s, p, &np);
        orc_eval(s, p, pairs, np);
        free(pairs);
        orc_vv_position(s, p);
        pairs = build_pairs(s, p, &np);
        orc_eval(s, p, pairs, np);
        orc_vv_velocity(s, p);
        if (p->xsph_epsilon != 0.0) orc_xsph(s, p, pairs, np);
        if (s->bx && (p->mhd_flags & SPHX_MHD_INDUCTION)) orc_mhd_induction(s, p, pairs, np);
        if (s->bx && s->psi && (p->mhd_flags & SPHX_MHD_CLEANING)) orc_div_clean(s, p, pairs, np);
        free(pairs);
    }
}

/* One force evaluation at the current state with a freshly built pair list. */
void orc_eval_state(OrcState* s, const SphxParams* p) {
    uint64_t np;
    uint32_t* pairs = build_pairs(s, p, &np);
    orc_eval(s, p, pairs, np);
    free(pairs);
}

/* Energy diagnostic (new; the reference's EnergyMonitor, energy.rs:9-77, has
 * no SPH terms).  Internal energy u(rho) = integral of P/rho^2 d rho from the
 * rest state, so that sum(KE + m g z + m u) is the conserved quantity of the
 * summation-density / symmetric-pressure discretisation. */
static double specific_internal(const SphxParams* p, double rho) {
    if (p->eos == SPHX_EOS_TAIT) {
        double rho0 = p->rest_density, g = p->gamma;
        if (rho <= rho0) return 0.0;
        double B = rho0 * (p->speed_of_sound * p->speed_of_sound) / g;
        return B * ((pow(rho / rho0, g - 1.0) - 1.0) / ((g - 1.0) * rho0) + 1.0 / rho - 1.0 / rho0);
    } else if (p->eos == SPHX_EOS_POLYTROPIC) {
        return p->poly_k * pow(rho, p->gamma - 1.0) / (p->gamma - 1.0);
    }
    return p->iso_cs2 * log(rho / p->rest_density);
}

void orc_energy(const OrcState* s, const SphxParams* p, double out[8]) {
    double ke = 0, pe = 0, ie = 0, me = 0, mom[3] = {0, 0, 0};
    for (uint32_t i = 0; i < s->n; ++i) {
        double m = s->m[i];
        double v2 = vdot(s->vx[i], s->vy[i], s->vz[i], s->vx[i], s->vy[i], s->vz[i]);
        ke += 0.5 * m * v2;
        pe -= m * ((p->gravity[0] * s->x[i] + p->gravity[1] * s->y[i]) + p->gravity[2] * s->z[i]);
        ie += m * ((p->eos == SPHX_EOS_ADIABATIC && s->u) ? s->u[i] : specific_internal(p, s->rho[i]));
        if (s->bx) {
            double b2 = vdot(s->bx[i], s->by[i], s->bz[i], s->bx[i], s->by[i], s->bz[i]);
            me += 0.5 * (m / s->rho[i]) * b2;
        }
        mom[0] += m * s->vx[i]; mom[1] += m * s->vy[i]; mom[2] += m * s->vz[i];
    }
    out[0] = ke; out[1] = pe; out[2] = ie; out[3] = me;
    out[4] = mom[0]; out[5] = mom[1]; out[6] = mom[2]; out[7] = 0.0;
}
