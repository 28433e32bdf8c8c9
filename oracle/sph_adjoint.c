/*
 * sph_adjoint.c — CPU f64 reverse-mode gradient of the SPH step (TEST INFRASTRUCTURE).
 *
 * Checker for the HIP adjoint.  The forward is the oracle's velocity-Verlet
 * tick (sph_oracle.c; main.rs:103-157): a_n = A(x_n, v_n); reflect
 * x_n + v_n dt + a_n dt^2/2 into the box (signs s in {+1,-1} per component,
 * v' = s v_n, a_prev = s a_n); a* = A(x_{n+1}, v'); v_{n+1} = v' + (a_prev + a*) dt/2.
 * A = gravity + pressure (sph.rs:177-228) + viscosity (sph.rs:260-322),
 * density by summation (sph.rs:72-116) with the EOS of sph.rs:149-169.
 *
 * The loss is L = sum |x_N - x_target|^2 / (3 n).  Reverse pass per tick:
 *   vbar' = vbar_{n+1}; abar* = abar_prev = vbar_{n+1} dt/2
 *   (xbar_{n+1}, vbar') += VJP_A(x_{n+1}, v'; abar*)
 *   ybar = s xbar_{n+1}; vbar_n = s vbar' + dt ybar; abar_n = s abar_prev + dt^2/2 ybar
 *   (xbar_n, vbar_n) += VJP_A(x_n, v_n; abar_n), xbar_n += ybar
 * VJP_A, per undirected pair (a, b) with d = x_b - x_a and
 * u = abar_a / m_a - abar_b / m_b, c = u.d, phi = m_a m_b (T + Pi) G(r),
 * G = (dW/dr) / r:  L_pair = c phi.  Its derivatives in d, v, rho and alpha
 * feed xbar, vbar, rhobar, alphabar; a second pass carries rhobar through the
 * density sum.  The neighbour sets and the viscosity switch are held fixed
 * (the map is piecewise smooth; the gradient is that of the active piece).
 *
 * Unsupported (returns -1): penalty walls, MHD, charge, XSPH, periodic boxes,
 * the adiabatic EOS.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sphx.h"
#include "sph_oracle.h"

/* pair list builder (same as the oracle's cell list) */
static uint32_t* adj_pairs(const OrcState* s, const SphxParams* p, uint64_t* np_out) {
    uint64_t cap = (uint64_t)s->n * 64 + 1024;
    for (;;) {
        uint32_t* pairs = (uint32_t*)malloc(sizeof(uint32_t) * 2 * cap);
        uint64_t np = orc_pairs_cells(s, p, pairs, cap);
        if (np <= cap) { *np_out = np; return pairs; }
        free(pairs);
        cap = np;
    }
}

static double d2w(double r, double h) { /* d^2 W / dr^2 of the cubic spline */
    double q = r / h;
    if (q >= 2.0) return 0.0;
    double sigma = 1.0 / (3.141592653589793 * ((h * h) * (h * h)) * h);
    if (q < 1.0) return sigma * (-3.0 + 4.5 * q);
    return sigma * (1.5 * (2.0 - q));
}

static double eos_dp(const SphxParams* p, double rho) {
    if (p->eos == SPHX_EOS_TAIT) {
        double B = p->rest_density * (p->speed_of_sound * p->speed_of_sound) / p->gamma;
        double ratio = rho / p->rest_density;
        if (B * (pow(ratio, p->gamma) - 1.0) < 0.0) return 0.0;
        return B * p->gamma * pow(ratio, p->gamma - 1.0) / p->rest_density;
    }
    if (p->eos == SPHX_EOS_POLYTROPIC) return p->poly_k * p->gamma * pow(rho, p->gamma - 1.0);
    return p->iso_cs2;
}

/* VJP of the acceleration map at the state s (rho, p already evaluated on
 * the pair list).  abar: 3n (interleaved).  Adds into xbar, vbar (3n) and
 * *alphabar. */
static void vjp_accel(const OrcState* s, const SphxParams* p, const uint32_t* pairs, uint64_t np,
                      const double* abar, double* xbar, double* vbar, double* alphabar) {
    const uint32_t n = s->n;
    double* rhobar = (double*)calloc(n ? n : 1, sizeof(double));
    const double c_s = p->speed_of_sound, alpha = p->visc_alpha, beta = p->visc_beta;
    const double sc = (p->visc_convention == SPHX_VISC_MONAGHAN) ? -1.0 : 1.0;
    for (uint64_t k = 0; k < np; ++k) {
        const uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        double d[3] = {s->x[b] - s->x[a], s->y[b] - s->y[a], s->z[b] - s->z[a]};
        double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        double r = sqrt(r2);
        double h = 0.5 * (s->h[a] + s->h[b]);
        if (!(r > 1e-10 && r < 2.0 * h)) continue;
        double dw = orc_kernel_dw(r, h);
        double G = dw / r;
        double Gp = (d2w(r, h) - G) / r; /* dG/dr */
        double ma = s->m[a], mb = s->m[b], mm = ma * mb;
        double u[3];
        for (int c = 0; c < 3; ++c) u[c] = abar[3 * a + c] / ma - abar[3 * b + c] / mb;
        double cdot = (u[0] * d[0] + u[1] * d[1]) + u[2] * d[2];
        double ra = s->rho[a], rb = s->rho[b];
        double T = s->p[a] / (ra * ra) + s->p[b] / (rb * rb);
        double vij[3] = {s->vx[a] - s->vx[b], s->vy[a] - s->vy[b], s->vz[a] - s->vz[b]};
        double sdot = (vij[0] * d[0] + vij[1] * d[1]) + vij[2] * d[2];
        int act = (p->visc_convention == SPHX_VISC_MONAGHAN) ? (sdot > 0.0) : (sdot < 0.0);
        double Pi = 0.0, mu = 0.0, Dn = r2 + 0.01 * h * h, rhobar_ab = 0.5 * (ra + rb), Pi_mu = 0.0;
        if (act) {
            mu = h * (sc * sdot) / Dn;
            Pi = (-alpha * c_s * mu + beta * mu * mu) / rhobar_ab;
            Pi_mu = (-alpha * c_s + 2.0 * beta * mu) / rhobar_ab;
        }
        double phi = mm * (T + Pi) * G;
        /* dL/dd = phi u + c [ mm (T + Pi) G' d/r + mm G Pi_mu dmu/dd ] */
        double dLdd[3];
        for (int c = 0; c < 3; ++c) {
            double dmu = act ? (h * sc * vij[c] - 2.0 * mu * d[c]) / Dn : 0.0;
            dLdd[c] = phi * u[c] + cdot * (mm * (T + Pi) * Gp * d[c] / r + mm * G * Pi_mu * dmu);
            xbar[3 * a + c] -= dLdd[c];
            xbar[3 * b + c] += dLdd[c];
            if (act) {
                double dv = cdot * mm * G * Pi_mu * (h * sc * d[c] / Dn);
                vbar[3 * a + c] += dv;
                vbar[3 * b + c] -= dv;
            }
        }
        double dTa = eos_dp(p, ra) / (ra * ra) - 2.0 * s->p[a] / (ra * ra * ra);
        double dTb = eos_dp(p, rb) / (rb * rb) - 2.0 * s->p[b] / (rb * rb * rb);
        double dPi_drho = act ? -Pi / (2.0 * rhobar_ab) : 0.0;
        rhobar[a] += cdot * mm * G * (dTa + dPi_drho);
        rhobar[b] += cdot * mm * G * (dTb + dPi_drho);
        if (act) *alphabar += cdot * mm * G * (-c_s * mu / rhobar_ab);
    }
    /* density part: rho_a gets m_b W(r), rho_b gets m_a W(r) */
    for (uint64_t k = 0; k < np; ++k) {
        const uint32_t a = pairs[2 * k], b = pairs[2 * k + 1];
        double d[3] = {s->x[b] - s->x[a], s->y[b] - s->y[a], s->z[b] - s->z[a]};
        double r = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        double h = 0.5 * (s->h[a] + s->h[b]);
        if (!(r < 2.0 * h) || !(r > 0.0)) continue;
        double coef = (rhobar[a] * s->m[b] + rhobar[b] * s->m[a]) * orc_kernel_dw(r, h) / r;
        for (int c = 0; c < 3; ++c) {
            xbar[3 * a + c] -= coef * d[c];
            xbar[3 * b + c] += coef * d[c];
        }
    }
    free(rhobar);
}

typedef struct {
    double *x, *y, *z, *vx, *vy, *vz;
} Snap;

static void snap_save(Snap* S, const OrcState* s) {
    size_t b = sizeof(double) * s->n;
    memcpy(S->x, s->x, b); memcpy(S->y, s->y, b); memcpy(S->z, s->z, b);
    memcpy(S->vx, s->vx, b); memcpy(S->vy, s->vy, b); memcpy(S->vz, s->vz, b);
}
static void snap_load(const Snap* S, OrcState* s) {
    size_t b = sizeof(double) * s->n;
    memcpy(s->x, S->x, b); memcpy(s->y, S->y, b); memcpy(s->z, S->z, b);
    memcpy(s->vx, S->vx, b); memcpy(s->vy, S->vy, b); memcpy(s->vz, S->vz, b);
}
static void snap_alloc(Snap* S, uint32_t n) {
    double** f[] = {&S->x, &S->y, &S->z, &S->vx, &S->vy, &S->vz};
    for (int k = 0; k < 6; ++k) *f[k] = (double*)malloc(sizeof(double) * (n ? n : 1));
}
static void snap_free(Snap* S) {
    free(S->x); free(S->y); free(S->z); free(S->vx); free(S->vy); free(S->vz);
}

/* Gradient of L with respect to the initial positions (gx, may be NULL) and
 * velocities (gv; both 3n interleaved) and the viscosity alpha; s holds the
 * initial state and is left at x_N.  target: 3n interleaved.  Returns 0, or
 * -1 for unsupported options. */
int orc_adjoint(OrcState* s, const SphxParams* p, uint32_t nsteps, const double* target, double* loss,
                double* gx, double* gv, double* galpha) {
    if ((p->boundary_flags & (SPHX_BND_PENALTY_CUBE | SPHX_BND_PERIODIC)) || s->bx || s->q ||
        p->xsph_epsilon != 0.0 || p->eos == SPHX_EOS_ADIABATIC)
        return -1;
    const uint32_t n = s->n;
    const double dt = p->dt, hdt = 0.5 * p->dt, hdt2 = 0.5 * p->dt * p->dt;
    Snap* A = (Snap*)malloc(sizeof(Snap) * (nsteps + 1)); /* x_t, v_t */
    Snap* B = (Snap*)malloc(sizeof(Snap) * (nsteps ? nsteps : 1)); /* x_{t+1}, v'_t */
    uint8_t* flip = (uint8_t*)malloc((size_t)n * (nsteps ? nsteps : 1));
    for (uint32_t t = 0; t <= nsteps; ++t) snap_alloc(&A[t], n);
    for (uint32_t t = 0; t < nsteps; ++t) snap_alloc(&B[t], n);
    /* forward, recording the tape */
    for (uint32_t t = 0; t < nsteps; ++t) {
        snap_save(&A[t], s);
        uint64_t np;
        uint32_t* pairs = adj_pairs(s, p, &np);
        orc_eval(s, p, pairs, np);
        free(pairs);
        /* reflection signs from the pre-update state */
        for (uint32_t i = 0; i < n; ++i) {
            double xs[3] = {s->x[i], s->y[i], s->z[i]}, vs[3] = {s->vx[i], s->vy[i], s->vz[i]};
            double as[3] = {s->ax[i], s->ay[i], s->az[i]};
            uint8_t bits = 0;
            for (int k = 0; k < 3; ++k) {
                double y = xs[k] + (vs[k] * dt + as[k] * hdt2);
                if ((p->boundary_flags & SPHX_BND_REFLECT) && (y < p->box_lo[k] || y > p->box_hi[k]))
                    bits |= (uint8_t)(1u << k);
            }
            flip[(size_t)t * n + i] = bits;
        }
        orc_vv_position(s, p);
        snap_save(&B[t], s);
        pairs = adj_pairs(s, p, &np);
        orc_eval(s, p, pairs, np);
        free(pairs);
        orc_vv_velocity(s, p);
    }
    snap_save(&A[nsteps], s);
    /* loss and its gradient */
    double* xb = (double*)calloc((size_t)3 * (n ? n : 1), sizeof(double));
    double* vb = (double*)calloc((size_t)3 * (n ? n : 1), sizeof(double));
    double* ab = (double*)calloc((size_t)3 * (n ? n : 1), sizeof(double));
    double* apb = (double*)calloc((size_t)3 * (n ? n : 1), sizeof(double));
    double L = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double xs[3] = {s->x[i], s->y[i], s->z[i]};
        for (int c = 0; c < 3; ++c) {
            double e = xs[c] - target[3 * i + c];
            L += e * e;
            xb[3 * i + c] = 2.0 * e / (3.0 * n);
        }
    }
    *loss = L / (3.0 * n);
    double alphab = 0.0;
    /* reverse */
    for (uint32_t tt = nsteps; tt-- > 0;) {
        /* at (x_{t+1}, v'_t) */
        snap_load(&B[tt], s);
        uint64_t np;
        uint32_t* pairs = adj_pairs(s, p, &np);
        orc_density(s, p, pairs, np);
        orc_eos(s, p);
        for (uint32_t k = 0; k < 3 * n; ++k) { ab[k] = hdt * vb[k]; apb[k] = hdt * vb[k]; }
        vjp_accel(s, p, pairs, np, ab, xb, vb, &alphab); /* vb now holds vbar' */
        free(pairs);
        for (uint32_t i = 0; i < n; ++i) {
            uint8_t bits = flip[(size_t)tt * n + i];
            for (int c = 0; c < 3; ++c) {
                double sg = (bits >> c) & 1u ? -1.0 : 1.0;
                double ybar = sg * xb[3 * i + c];
                vb[3 * i + c] = sg * vb[3 * i + c] + dt * ybar;
                ab[3 * i + c] = sg * apb[3 * i + c] + hdt2 * ybar;
                xb[3 * i + c] = ybar;
            }
        }
        /* at (x_t, v_t) */
        snap_load(&A[tt], s);
        pairs = adj_pairs(s, p, &np);
        orc_density(s, p, pairs, np);
        orc_eos(s, p);
        vjp_accel(s, p, pairs, np, ab, xb, vb, &alphab);
        free(pairs);
    }
    for (uint32_t k = 0; k < 3 * n; ++k) gv[k] = vb[k];
    if (gx)
        for (uint32_t k = 0; k < 3 * n; ++k) gx[k] = xb[k];
    *galpha = alphab;
    snap_load(&A[nsteps], s);
    for (uint32_t t = 0; t <= nsteps; ++t) snap_free(&A[t]);
    for (uint32_t t = 0; t < nsteps; ++t) snap_free(&B[t]);
    free(A); free(B); free(flip); free(xb); free(vb); free(ab); free(apb);
    return 0;
}

/* Loss only (forward), for finite differences. */
double orc_loss(OrcState* s, const SphxParams* p, uint32_t nsteps, const double* target) {
    orc_step(s, p, nsteps);
    double L = 0.0;
    for (uint32_t i = 0; i < s->n; ++i) {
        double xs[3] = {s->x[i], s->y[i], s->z[i]};
        for (int c = 0; c < 3; ++c) {
            double e = xs[c] - target[3 * i + c];
            L += e * e;
        }
    }
    return L / (3.0 * s->n);
}
