/*
 * sph_omp.c — all-core gather form of the oracle's force evaluation and
 * velocity-Verlet tick (TEST INFRASTRUCTURE: the all-core CPU baseline of
 * bench.py).
 *
 * The pair arithmetic is the oracle's (sph_oracle.c, restating sph.rs:72-484
 * and thermodynamics.rs:232-238), but each particle sums over its own
 * neighbour list instead of scattering each undirected pair to both ends, so
 * the particles are split over OpenMP threads with no shared writes.  Sums
 * agree with the pair form up to summation order (tests/test_oracle.py).
 * Lists are rebuilt for every evaluation from a uniform grid of cells of
 * width >= 2 h_max, as the reference rebuilds CandidatePairs each tick
 * (main.rs:103-157).  XSPH and the MHD induction / cleaning passes are not in
 * this form.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "sph_oracle.h"

typedef struct OrcNbr {
    uint32_t n;
    uint32_t per;     /* list capacity per particle */
    uint64_t cap;     /* capacity of nb */
    uint32_t* nb;     /* neighbours of particle i at nb[i * per .. i * per + cnt[i]) */
    uint32_t* cnt;    /* n */
} OrcNbr;

void* orc_omp_new(void) { return calloc(1, sizeof(OrcNbr)); }

void orc_omp_free(void* h) {
    OrcNbr* L = (OrcNbr*)h;
    if (!L) return;
    free(L->nb); free(L->cnt);
    free(L);
}

int orc_omp_threads(void) { return omp_get_max_threads(); }

/* The neighbour lists of every particle (exact predicate r < 2 h_avg).
 * Returns the number of directed pairs. */
uint64_t orc_omp_build(void* h, const OrcState* s, const SphxParams* p) {
    OrcNbr* L = (OrcNbr*)h;
    const uint32_t n = s->n;
    if (L->n < n || !L->cnt) {
        free(L->cnt);
        L->cnt = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    }
    L->n = n;
    if (n == 0) return 0;
    if (L->per == 0) L->per = 128;
    double hmax = 0.0;
    for (uint32_t i = 0; i < n; ++i) if (s->h[i] > hmax) hmax = s->h[i];
    const double cs = 2.0 * hmax * (1.0 + 1e-9);
    const int periodic = (p->boundary_flags & SPHX_BND_PERIODIC) != 0;
    double lo[3], hi[3], w[3];
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
    /* cells never narrower than 2 h_max; widened until the grid has <= 2^24 cells */
    for (double wc = cs;; wc *= 1.25) {
        size_t prod = 1;
        for (int k = 0; k < 3; ++k) {
            double ext = hi[k] - lo[k];
            nc[k] = periodic ? (int)floor(ext / wc) : (int)(ext / wc) + 1;
            if (periodic && nc[k] < 3) nc[k] = 1; /* too few cells to wrap a 3-stencil: one cell */
            if (nc[k] < 1) nc[k] = 1;
            w[k] = periodic ? ext / nc[k] : wc;
            prod *= (size_t)nc[k];
        }
        if (prod <= ((size_t)1 << 24)) break;
    }
    const size_t ncell = (size_t)nc[0] * nc[1] * nc[2];
    uint32_t* cell = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* start = (uint32_t*)calloc(ncell + 1, sizeof(uint32_t));
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * n);
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; ++i) {
        double c[3] = {s->x[i], s->y[i], s->z[i]};
        size_t key = 0;
        for (int k = 2; k >= 0; --k) {
            long v = (long)floor((c[k] - lo[k]) / w[k]);
            if (v < 0) v = 0;
            if (v >= nc[k]) v = nc[k] - 1;
            key = key * (size_t)nc[k] + (size_t)v;
        }
        cell[i] = (uint32_t)key;
    }
    for (uint32_t i = 0; i < n; ++i) start[cell[i] + 1]++;
    for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    {
        uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * ncell);
        memcpy(fill, start, sizeof(uint32_t) * ncell);
        for (uint32_t i = 0; i < n; ++i) order[fill[cell[i]]++] = i; /* ascending index within a cell */
        free(fill);
    }
    /* one pass over the candidates into per-particle slots of L->per
     * entries; repeated with twice the slots while some list overflows */
    for (;;) {
        const uint64_t need = (uint64_t)n * L->per;
        if (need > L->cap) {
            free(L->nb);
            L->cap = need;
            L->nb = (uint32_t*)malloc(sizeof(uint32_t) * L->cap);
        }
        uint32_t most = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : most)
        for (uint32_t i = 0; i < n; ++i) {
            int ci[3];
            uint32_t key = cell[i];
            ci[0] = (int)(key % (uint32_t)nc[0]); key /= (uint32_t)nc[0];
            ci[1] = (int)(key % (uint32_t)nc[1]); ci[2] = (int)(key / (uint32_t)nc[1]);
            int span[3][3], ns[3];
            for (int k = 0; k < 3; ++k) {
                ns[k] = 0;
                for (int o = -1; o <= 1; ++o) {
                    int v = ci[k] + o;
                    if (periodic) v = (v + nc[k]) % nc[k];
                    else if (v < 0 || v >= nc[k]) continue;
                    int dup = 0;
                    for (int t = 0; t < ns[k]; ++t) if (span[k][t] == v) dup = 1;
                    if (!dup) span[k][ns[k]++] = v;
                }
            }
            uint32_t c = 0;
            uint32_t* out = L->nb + (uint64_t)i * L->per;
            for (int iz = 0; iz < ns[2]; ++iz)
                for (int iy = 0; iy < ns[1]; ++iy)
                    for (int ix = 0; ix < ns[0]; ++ix) {
                        size_t cc = ((size_t)span[2][iz] * nc[1] + span[1][iy]) * nc[0] + span[0][ix];
                        for (uint32_t q = start[cc]; q < start[cc + 1]; ++q) {
                            uint32_t j = order[q];
                            if (j == i || !is_pair(s, p, i, j)) continue;
                            if (c < L->per) out[c] = j;
                            ++c;
                        }
                    }
            L->cnt[i] = c;
            if (c > most) most = c;
        }
        if (most <= L->per) break;
        while (L->per < most) L->per *= 2;
    }
    free(cell); free(start); free(order);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) total += L->cnt[i];
    return total;
}

static double eos_p(const OrcState* s, const SphxParams* p, uint32_t i) {
    double rho = s->rho[i];
    if (p->eos == SPHX_EOS_TAIT) {
        double B = p->rest_density * (p->speed_of_sound * p->speed_of_sound) / p->gamma;
        double P = B * (pow(rho / p->rest_density, p->gamma) - 1.0);
        return P < 0.0 ? 0.0 : P;
    }
    if (p->eos == SPHX_EOS_POLYTROPIC) return p->poly_k * pow(rho, p->gamma);
    if (p->eos == SPHX_EOS_ADIABATIC) return (p->gamma - 1.0) * rho * s->u[i];
    return rho * p->iso_cs2;
}

/* One force evaluation on the lists of the last orc_omp_build: rho, P, a
 * (and du/dt for the adiabatic EOS), as orc_eval computes them. */
void orc_omp_eval(void* h, OrcState* s, const SphxParams* p) {
    const OrcNbr* L = (const OrcNbr*)h;
    const uint32_t n = s->n;
    const int energy = p->eos == SPHX_EOS_ADIABATIC && s->u && s->du;
    const int stress = (p->mhd_flags & SPHX_MHD_MAXWELL_STRESS) && s->bx;
    const int charge = (p->mhd_flags & SPHX_MHD_LORENTZ_CHARGE) && s->q && s->bx;
    const double c = p->speed_of_sound, eps = 1e-2;
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t i = 0; i < n; ++i) { /* sph.rs:72-116, 149-169 */
        double rho = s->m[i] * orc_kernel_w(0.0, s->h[i]);
        for (const uint32_t *q = L->nb + (uint64_t)i * L->per, *e = q + L->cnt[i]; q < e; ++q) {
            const uint32_t j = *q;
            double dx, dy, dz;
            pair_vec(s, p, i, j, &dx, &dy, &dz);
            rho += s->m[j] * orc_kernel_w(vlen(dx, dy, dz), 0.5 * (s->h[i] + s->h[j]));
        }
        s->rho[i] = rho;
        s->p[i] = eos_p(s, p, i);
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t i = 0; i < n; ++i) {
        double F[3] = {s->m[i] * p->gravity[0], s->m[i] * p->gravity[1], s->m[i] * p->gravity[2]};
        if (charge) { /* lorentz.rs:21-39 */
            double cx = s->vy[i] * s->bz[i] - s->vz[i] * s->by[i];
            double cy = s->vz[i] * s->bx[i] - s->vx[i] * s->bz[i];
            double cz = s->vx[i] * s->by[i] - s->vy[i] * s->bx[i];
            F[0] += s->q[i] * (p->e_field[0] + cx);
            F[1] += s->q[i] * (p->e_field[1] + cy);
            F[2] += s->q[i] * (p->e_field[2] + cz);
        }
        const double pi2 = s->p[i] / (s->rho[i] * s->rho[i]);
        double du = 0.0;
        for (const uint32_t *q = L->nb + (uint64_t)i * L->per, *e = q + L->cnt[i]; q < e; ++q) {
            const uint32_t j = *q;
            double dx, dy, dz;
            pair_vec(s, p, i, j, &dx, &dy, &dz);
            double r = vlen(dx, dy, dz);
            double h_avg = 0.5 * (s->h[i] + s->h[j]);
            if (!(r > 1e-10)) continue; /* dW/dr = 0 (sph.rs:52) */
            double dw = orc_kernel_dw(r, h_avg);
            double rx = dx / r, ry = dy / r, rz = dz / r;
            double mm = s->m[i] * s->m[j];
            /* pressure, sph.rs:208-221 */
            double mag = (mm * (pi2 + s->p[j] / (s->rho[j] * s->rho[j]))) * dw;
            F[0] += rx * mag; F[1] += ry * mag; F[2] += rz * mag;
            /* viscosity, sph.rs:292-314 */
            double sdot = vdot(s->vx[i] - s->vx[j], s->vy[i] - s->vy[j], s->vz[i] - s->vz[j], dx, dy, dz);
            int active = (p->visc_convention == SPHX_VISC_MONAGHAN) ? (sdot > 0.0) : (sdot < 0.0);
            double pi_ij = 0.0;
            if (active) {
                double rho_bar = 0.5 * (s->rho[i] + s->rho[j]);
                double sgn_s = (p->visc_convention == SPHX_VISC_MONAGHAN) ? -sdot : sdot;
                double mu = (h_avg * sgn_s) / ((r * r) + eps * (h_avg * h_avg));
                pi_ij = ((-p->visc_alpha) * c * mu + p->visc_beta * (mu * mu)) / rho_bar;
                double vm = (mm * pi_ij) * dw;
                F[0] += rx * vm; F[1] += ry * vm; F[2] += rz * vm;
            }
            if (energy) du += s->m[j] * (pi2 + 0.5 * pi_ij) * (-sdot * (dw / r));
            if (stress) { /* SPMHD Maxwell stress, M = B B - |B|^2 I / 2 */
                double g[3] = {-rx * dw, -ry * dw, -rz * dw};
                double Ba[3] = {s->bx[i], s->by[i], s->bz[i]}, Bb[3] = {s->bx[j], s->by[j], s->bz[j]};
                double ia = 1.0 / (s->rho[i] * s->rho[i]), ib = 1.0 / (s->rho[j] * s->rho[j]);
                double Ba2 = vdot(Ba[0], Ba[1], Ba[2], Ba[0], Ba[1], Ba[2]);
                double Bb2 = vdot(Bb[0], Bb[1], Bb[2], Bb[0], Bb[1], Bb[2]);
                double Bag = vdot(Ba[0], Ba[1], Ba[2], g[0], g[1], g[2]);
                double Bbg = vdot(Bb[0], Bb[1], Bb[2], g[0], g[1], g[2]);
                for (int a = 0; a < 3; ++a)
                    F[a] += mm * ((Ba[a] * Bag - 0.5 * Ba2 * g[a]) * ia + (Bb[a] * Bbg - 0.5 * Bb2 * g[a]) * ib);
            }
        }
        if (p->boundary_flags & SPHX_BND_PENALTY_CUBE) { /* sph.rs:451-483 */
            double ext = p->half_extent, ad = p->activation_distance, K = p->stiffness;
            double cc[3] = {s->x[i], s->y[i], s->z[i]};
            for (int k = 0; k < 3; ++k) {
                double d_low = cc[k] - (-ext);
                if (d_low < ad && d_low >= 0.0) F[k] += K * (ad - d_low);
                double d_high = ext - cc[k];
                if (d_high < ad && d_high >= 0.0) F[k] -= K * (ad - d_high);
            }
        }
        s->ax[i] = F[0] / s->m[i];
        s->ay[i] = F[1] / s->m[i];
        s->az[i] = F[2] / s->m[i];
        if (energy) s->du[i] = du;
    }
}

/* nsteps velocity-Verlet ticks (main.rs:103-157) with lists rebuilt for each
 * evaluation; returns -1 when XSPH or the MHD passes are asked for. */
int orc_omp_step(void* h, OrcState* s, const SphxParams* p, uint32_t nsteps) {
    if (p->xsph_epsilon != 0.0 || (s->bx && (p->mhd_flags & (SPHX_MHD_INDUCTION | SPHX_MHD_CLEANING)))) return -1;
    for (uint32_t t = 0; t < nsteps; ++t) {
        orc_omp_build(h, s, p);
        orc_omp_eval(h, s, p);
        orc_vv_position(s, p);
        orc_omp_build(h, s, p);
        orc_omp_eval(h, s, p);
        orc_vv_velocity(s, p);
    }
    return 0;
}
