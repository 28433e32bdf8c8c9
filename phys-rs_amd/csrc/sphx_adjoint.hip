// sphx_adjoint.hip — reverse-mode gradient of the velocity-Verlet SPH tick.
//
// The tick (main.rs:103-157) is x_{n+1} = R(x_n + v_n dt + a_n dt^2/2),
// v' = S v_n, a_prev = S a_n, a* = A(x_{n+1}, v'), v_{n+1} = v' + (a_prev + a*) dt/2,
// with a_n = A(x_n, v_n), R the wall reflection and S = diag(+-1) its sign.
// A = gravity + pressure (sph.rs:177-228) + viscosity (sph.rs:260-322) on
// the summation density (sph.rs:72-116) and its EOS (sph.rs:149-169).
//
// The vector-Jacobian product of A at (x, v) with a cotangent abar is, per
// undirected pair (i, j) with d = x_j - x_i, u = abar_i/m_i - abar_j/m_j,
// c = u.d and G = (dW/dr)/r, the derivative of c * m_i m_j (T + Pi) G
// (T = P_i/rho_i^2 + P_j/rho_j^2) in d, v, rho and alpha; rhobar then flows
// through the density sum.  Both passes run in gather form (one thread per
// particle, each pair visited from both ends, each end adding its own share),
// so no atomics and every sum is deterministic.  The neighbour sets and the
// viscosity switch are those of the forward state (the map is piecewise
// smooth; this is the gradient of the active piece).
//
// Adjoint vectors live in upload order (index = SphxState slot), which does
// not change when the particles are re-sorted; kernels gather and scatter
// through Particles::orig.

#include "sphx_device.h"

namespace sphx {

// cotangent of the acceleration, gathered into sorted order (ab4: one
// 32-byte record per particle) for the pair pass; apb keeps hdt * vbar for
// the reflection step
__global__ void k_adj_seed(const uint32_t* __restrict__ orig, uint32_t n, double hdt, Vec3 vb, Vec3 apb,
                           double4* __restrict__ ab4) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t u = orig[i];
    const double x = hdt * vb.x[u], y = hdt * vb.y[u], z = hdt * vb.z[u];
    apb.x[u] = x; apb.y[u] = y; apb.z[u] = z;
    double2* r = reinterpret_cast<double2*>(ab4 + i);
    r[0] = make_double2(x, y);
    r[1] = make_double2(z, 0.0);
}
void launch_adj_seed(const uint32_t* orig, uint32_t n, double hdt, const Vec3& vb, const Vec3& apb, double4* ab4,
                     hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_adj_seed, dim3(nblocks(n)), dim3(kBlock), 0, s, orig, n, hdt, vb, apb, ab4);
}

// ab4 (sorted records) = src (upload order) through orig
__global__ void k_gather_ab(const uint32_t* __restrict__ orig, uint32_t n, Vec3 src, double4* __restrict__ ab4) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t u = orig[i];
    double2* r = reinterpret_cast<double2*>(ab4 + i);
    r[0] = make_double2(src.x[u], src.y[u]);
    r[1] = make_double2(src.z[u], 0.0);
}
void launch_gather_ab(const uint32_t* orig, uint32_t n, const Vec3& src, double4* ab4, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_gather_ab, dim3(nblocks(n)), dim3(kBlock), 0, s, orig, n, src, ab4);
}

// dst (upload order) = src (sorted) through orig; with src2.x set, also
// dst2 = src2 (positions and velocities of the tape in one pass)
__global__ void k_scatter3(const uint32_t* __restrict__ orig, uint32_t n, Vec3 src, Vec3 dst, Vec3 src2,
                           Vec3 dst2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t u = orig[i];
    dst.x[u] = src.x[i]; dst.y[u] = src.y[i]; dst.z[u] = src.z[i];
    if (src2.x) {
        dst2.x[u] = src2.x[i]; dst2.y[u] = src2.y[i]; dst2.z[u] = src2.z[i];
    }
}
void launch_scatter3(const uint32_t* orig, uint32_t n, const Vec3& src, const Vec3& dst, hipStream_t s) {
    const Vec3 none{nullptr, nullptr, nullptr};
    if (n) hipLaunchKernelGGL(k_scatter3, dim3(nblocks(n)), dim3(kBlock), 0, s, orig, n, src, dst, none, none);
}
void launch_scatter6(const uint32_t* orig, uint32_t n, const Vec3& src, const Vec3& dst, const Vec3& src2,
                     const Vec3& dst2, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_scatter3, dim3(nblocks(n)), dim3(kBlock), 0, s, orig, n, src, dst, src2, dst2);
}

// Reverse of the position update and wall reflection (integrators.rs:63-83):
// ybar = S xbar; vbar = S vbar' + dt ybar; abar = S abar_prev + dt^2/2 ybar;
// xbar = ybar.  flip bit k of particle u set: S_k = -1.
__global__ void k_adj_flip(uint32_t n, const uint8_t* __restrict__ flip, double dt, double hdt2, Vec3 xb, Vec3 vb,
                           Vec3 apb, Vec3 ab) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const uint32_t f = flip[u];
    double* xs[3] = {xb.x, xb.y, xb.z};
    double* vs[3] = {vb.x, vb.y, vb.z};
    double* ps[3] = {apb.x, apb.y, apb.z};
    double* as[3] = {ab.x, ab.y, ab.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double sg = ((f >> k) & 1u) ? -1.0 : 1.0;
        const double y = sg * xs[k][u];
        vs[k][u] = sg * vs[k][u] + dt * y;
        as[k][u] = sg * ps[k][u] + hdt2 * y;
        xs[k][u] = y;
    }
}
void launch_adj_flip(uint32_t n, const uint8_t* flip, double dt, double hdt2, const Vec3& xb, const Vec3& vb,
                     const Vec3& apb, const Vec3& ab, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_adj_flip, dim3(nblocks(n)), dim3(kBlock), 0, s, n, flip, dt, hdt2, xb, vb, apb, ab);
}

// Block sum of one double per thread; thread 0 gets the total.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    return lds[0];
}

// L = sum |x - target|^2 / (3n) seeds xbar = dL/dx_N (upload order) and
// leaves one partial sum of squares per block.
__global__ void k_adj_loss(Particles P, uint32_t n, Vec3 target, Vec3 xb, double* __restrict__ part) {
    __shared__ double lds[kBlock];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (i < n) {
        const uint32_t u = P.orig[i];
        const double s = 2.0 / (3.0 * (double)n);
        const double ex = P.x[i] - target.x[u], ey = P.y[i] - target.y[u], ez = P.z[i] - target.z[u];
        xb.x[u] = s * ex; xb.y[u] = s * ey; xb.z[u] = s * ez;
        e2 = (ex * ex + ey * ey) + ez * ez;
    }
    const double t = block_sum(e2, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
void launch_adj_loss(const Particles& P, uint32_t n, const Vec3& target, const Vec3& xb, double* part,
                     hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_adj_loss, dim3(nblocks(n)), dim3(kBlock), 0, s, P, n, target, xb, part);
}

// Four lanes per particle (the VJP passes): lanes l, l + 16, l + 32 and
// l + 48 of a wave share particle (wave * 16 + l % 16); lane quarter q =
// lane / 16 sums list entries [q qs, (q + 1) qs) (the last quarter to the
// end), qs a multiple of 4 near a quarter of the list, and the quarter sums
// are added as (q0 + q1) + (q2 + q3) into quarter 0.  At the sizes the
// adjoint runs (tens of thousands of particles) the passes are bound by the
// latency of their gathers, and four lanes give four times the waves of one.
__device__ __forceinline__ uint32_t vjp_slot(uint32_t t) { return ((t >> 6) << 4) | (t & 15u); }
__device__ __forceinline__ uint32_t vjp_quarter() { return (threadIdx.x >> 4) & 3u; }
__device__ __forceinline__ void vjp_range(const NbrList& L, uint32_t i, uint32_t q, uint32_t& kb, uint32_t& ke) {
    const uint32_t qs = (((uint32_t)L.ntot[i] + 15u) >> 4) << 2;
    kb = q * qs;
    ke = q == 3u ? 0xffffffffu : (q + 1u) * qs;
}
__device__ __forceinline__ double quarter_total(double v) {
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}

uint32_t vjp_blocks(uint32_t n) { return nblocks(4 * n); }

// VJP of the pressure + viscosity acceleration at the current (sorted)
// state.  ab4: abar records in sorted order; neighbours' x, rho, v, P/rho^2
// come from the f64 gather record.  Adds dL/dx and dL/dv into xb, vb
// (upload order), writes rhobar (sorted) and adds this block's share of
// dL/dalpha to alpha_acc[block]; with `overwrite` it stores them instead of
// adding (a pass that may be redone).
template <bool UNI>
__global__ void __launch_bounds__(kBlock) k_vjp_forces(Particles P, Derived D, uint32_t n, Grid g, NbrList L,
                                                        KernelParams kp, const double4* __restrict__ ab4, Vec3 xb,
                                                        Vec3 vb, double* __restrict__ rhob,
                                                        double* __restrict__ alpha_acc, int overwrite) {
    __shared__ double lds[kBlock];
    const uint32_t i = vjp_slot(particle_index()), h = vjp_quarter();
    double da = 0.0;
    if (i < n) {
        const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i], mi = P.m[i];
        const double vxi = P.vx[i], vyi = P.vy[i], vzi = P.vz[i];
        const double rhoi = D.rho[i], pr2i = D.pr2[i];
        const double inv_mi = 1.0 / mi;
        const double2* __restrict__ AB = reinterpret_cast<const double2*>(ab4);
        const double2* __restrict__ R = reinterpret_cast<const double2*>(D.r64);
        const double2 ab01 = AB[2 * i], ab2 = AB[2 * i + 1];
        const double uxi = ab01.x * inv_mi, uyi = ab01.y * inv_mi, uzi = ab2.x * inv_mi;
        // d(P_i/rho_i^2)/d(rho_i)
        const double dTi = eos_dpdrho(kp, rhoi) / (rhoi * rhoi) - 2.0 * pr2i / rhoi;
        const double cs = kp.speed_of_sound, alpha = kp.alpha, beta = kp.beta;
        const bool ref_conv = kp.visc_conv == SPHX_VISC_REFERENCE;
        const double sc = ref_conv ? 1.0 : -1.0;
        double gx = 0, gy = 0, gz = 0, gvx = 0, gvy = 0, gvz = 0, grho = 0;
        // Branch-free pair body with two refined reciprocals (1/Dn, 1/rho_bar)
        // in place of divisions; a pair outside the kernel support, at r <= 1e-10
        // or (for the viscosity terms) not approaching/separating contributes 0.
        uint32_t kb, ke;
        vjp_range(L, i, h, kb, ke);
        for_neighbours(L, i, [&](uint32_t e) {
            const uint32_t j = e & kIdxMask;
            const double2 xy = R[4 * j], zr = R[4 * j + 1], vxy = R[4 * j + 2], vzp = R[4 * j + 3];
            const double2 a01 = AB[2 * j], a2 = AB[2 * j + 1];
            double dx, dy, dz;
            pair_vec(g, xi, yi, zi, xy.x, xy.y, zr.x, dx, dy, dz);
            const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
            const bool ok = r2 > 1e-20;  // r <= 1e-10: no pressure force (sph.rs:52)
            double r, inv_r;
            sqrt_rsqrt1(ok ? r2 : 1.0, r, inv_r);
            const double h = UNI ? kp.h0 : 0.5 * (hi + P.h[j]);
            const double inv_h = UNI ? kp.inv_h0 : rcp_nr1(h);
            const double ih2 = inv_h * inv_h, sig3 = 3.0 * ((ih2 * ih2) * kInvPi);
            // dW/dr and d2W/dr2 of the cubic spline from its one-sided pieces
            // s = (1 - q)+ and t = (1 - q/2)+: 3 sigma_g (s^2 - t^2) and
            // 3 (sigma_g / h) (t - 2 s), zero from q = 2 on
            const double sk = fmax(fma(r, -inv_h, 1.0), 0.0), tk = fmax(fma(r, -0.5 * inv_h, 1.0), 0.0);
            const double dw = sig3 * fma(-tk, tk, sk * sk);
            const double d2w = (sig3 * inv_h) * fma(-2.0, sk, tk);
            const double okf = ok ? 1.0 : 0.0;
            const double G = (dw * inv_r) * okf;
            const double Gp = ((d2w - dw * inv_r) * inv_r) * okf;  // dG/dr
            const double mj = UNI ? kp.m0 : P.m[j], mm = mi * mj;
            const double inv_mj = UNI ? inv_mi : rcp_nr1(mj);
            const double ux = uxi - a01.x * inv_mj, uy = uyi - a01.y * inv_mj, uz = uzi - a2.x * inv_mj;
            const double c = fma(uz, dz, fma(uy, dy, ux * dx));
            const double rhoj = zr.y;
            const double T = pr2i + vzp.y;
            const double vx = vxi - vxy.x, vy = vyi - vxy.y, vz = vzi - vzp.x;
            const double sdot = fma(vz, dz, fma(vy, dy, vx * dx));
            const double actf = (ref_conv ? (sdot < 0.0) : (sdot > 0.0)) ? 1.0 : 0.0;
            const double Dn = fma(0.01 * h, h, r2);
            const double inv_Dn = rcp_nr1(Dn);
            const double inv_rb = 2.0 * rcp_nr1(rhoi + rhoj);  // 1 / rho_bar
            const double hsc = h * sc;
            const double mu = ((hsc * sdot) * inv_Dn) * actf;
            const double Pi = fma(beta * mu, mu, -alpha * cs * mu) * inv_rb;
            const double Pi_mu = (fma(2.0 * beta, mu, -alpha * cs) * inv_rb) * actf;
            const double TP = mm * (T + Pi);
            const double phi = TP * G;
            const double qd = ((c * TP) * Gp) * inv_r;
            const double cg = (c * mm) * G;
            const double w = (cg * Pi_mu) * inv_Dn;  // d(c phi)/d(mu) times dmu/d(.) / (...)
            const double tmu = -2.0 * mu;
            // dL/dd = phi u + qd d + w (h sc v - 2 mu d); particle i is the pair's "a": xbar_i -= dL/dd
            gx -= fma(w, fma(hsc, vx, tmu * dx), fma(phi, ux, qd * dx));
            gy -= fma(w, fma(hsc, vy, tmu * dy), fma(phi, uy, qd * dy));
            gz -= fma(w, fma(hsc, vz, tmu * dz), fma(phi, uz, qd * dz));
            // dL/dv_i
            const double wh = w * hsc;
            gvx = fma(wh, dx, gvx);
            gvy = fma(wh, dy, gvy);
            gvz = fma(wh, dz, gvz);
            grho = fma(cg, fma(-0.5 * Pi, inv_rb, dTi), grho);
            da = fma(0.5 * cg, (-cs * mu) * inv_rb, da);  // each pair is visited from both ends
        }, kb, ke);
        gx = quarter_total(gx); gy = quarter_total(gy); gz = quarter_total(gz);
        gvx = quarter_total(gvx); gvy = quarter_total(gvy); gvz = quarter_total(gvz);
        grho = quarter_total(grho);
        if (!h) {
            const uint32_t u = P.orig[i];
            if (overwrite) {
                xb.x[u] = gx; xb.y[u] = gy; xb.z[u] = gz;
                vb.x[u] = gvx; vb.y[u] = gvy; vb.z[u] = gvz;
            } else {
                xb.x[u] += gx; xb.y[u] += gy; xb.z[u] += gz;
                vb.x[u] += gvx; vb.y[u] += gvy; vb.z[u] += gvz;
            }
            rhob[i] = grho;
        }
    }
    const double t = block_sum(da, lds);
    if (threadIdx.x == 0) {
        if (overwrite) alpha_acc[xcd_block()] = t;
        else alpha_acc[xcd_block()] += t;
    }
}

// VJP of the summation density: rho_i = m_i W(0, h_i) + sum_j m_j W(r_ij, h_ij)
// gives xbar_i -= sum_j (rhobar_i m_j + rhobar_j m_i) G d.
template <bool UNI>
__global__ void __launch_bounds__(kBlock) k_vjp_density(Particles P, uint32_t n, Grid g, NbrList L,
                                                         KernelParams kp, const double* __restrict__ rhob, Vec3 xb) {
    const uint32_t i = vjp_slot(particle_index()), h = vjp_quarter();
    if (i >= n) return;
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i], mi = P.m[i], rbi = rhob[i];
    const double2* __restrict__ pos = reinterpret_cast<const double2*>(P.pos);
    double gx = 0, gy = 0, gz = 0;
    uint32_t kb, ke;
    vjp_range(L, i, h, kb, ke);
    for_neighbours(L, i, [&](uint32_t e) {
        const uint32_t j = e & kIdxMask;
        const double2 xy = pos[2 * j], zh = pos[2 * j + 1];
        double dx, dy, dz;
        pair_vec(g, xi, yi, zi, xy.x, xy.y, zh.x, dx, dy, dz);
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const bool ok = r2 > 0.0;
        double r, inv_r;
        sqrt_rsqrt1(ok ? r2 : 1.0, r, inv_r);
        const double inv_h = UNI ? kp.inv_h0 : rcp_nr1(0.5 * (hi + zh.y));
        const double ih2 = inv_h * inv_h;
        const double dwr = (((3.0 * ((ih2 * ih2) * kInvPi)) * dw3_shape_r(r, -inv_h, -0.5 * inv_h)) * inv_r) *
                           (ok ? 1.0 : 0.0);
        const double coef = UNI ? ((rbi + rhob[j]) * mi) * dwr : fma(rbi, P.m[j], rhob[j] * mi) * dwr;
        gx = fma(-coef, dx, gx);
        gy = fma(-coef, dy, gy);
        gz = fma(-coef, dz, gz);
    }, kb, ke);
    gx = quarter_total(gx); gy = quarter_total(gy); gz = quarter_total(gz);
    if (h) return;
    const uint32_t u = P.orig[i];
    xb.x[u] += gx; xb.y[u] += gy; xb.z[u] += gz;
}

void launch_vjp(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                const KernelParams& kp, const double4* ab4, const Vec3& xb, const Vec3& vb, double* rhob,
                double* alpha_acc, hipStream_t s, bool overwrite) {
    if (!n) return;
    const dim3 grid(vjp_blocks(n)), block(kBlock);  // four lanes per particle
    const int ow = overwrite ? 1 : 0;
    if (kp.uniform) {
        hipLaunchKernelGGL(k_vjp_forces<true>, grid, block, occupancy_lds(3), s, P, D, n, g, L, kp, ab4, xb, vb, rhob,
                           alpha_acc, ow);
        hipLaunchKernelGGL(k_vjp_density<true>, grid, block, occupancy_lds(4), s, P, n, g, L, kp,
                           (const double*)rhob, xb);
    } else {
        hipLaunchKernelGGL(k_vjp_forces<false>, grid, block, occupancy_lds(3), s, P, D, n, g, L, kp, ab4, xb, vb,
                           rhob, alpha_acc, ow);
        hipLaunchKernelGGL(k_vjp_density<false>, grid, block, occupancy_lds(4), s, P, n, g, L, kp,
                           (const double*)rhob, xb);
    }
}

__global__ void k_adj_accumulate(uint32_t n, Vec3 xs, Vec3 vs, Vec3 xb, Vec3 vb, const double* __restrict__ as,
                                 double* __restrict__ alpha_acc, uint32_t nb) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) {
        xb.x[u] += xs.x[u]; xb.y[u] += xs.y[u]; xb.z[u] += xs.z[u];
        vb.x[u] += vs.x[u]; vb.y[u] += vs.y[u]; vb.z[u] += vs.z[u];
    }
    if (u < nb) alpha_acc[u] += as[u];
}
void launch_adj_accumulate(uint32_t n, const Vec3& xs, const Vec3& vs, const Vec3& xb, const Vec3& vb,
                           const double* as, double* alpha_acc, uint32_t nb, hipStream_t s) {
    const uint32_t m = n > nb ? n : nb;
    if (m) hipLaunchKernelGGL(k_adj_accumulate, dim3(nblocks(m)), dim3(kBlock), 0, s, n, xs, vs, xb, vb, as, alpha_acc,
                              nb);
}

}  // namespace sphx
