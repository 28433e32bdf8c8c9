// sphx_particles.hip — per-particle kernels: velocity-Verlet halves and the
// energy/momentum reduction.  Streaming SoA f64 access; HBM-bound.

#include "sphx_device.h"

namespace sphx {

// After particle i's position changed to x: the packed position record and,
// in the mixed form, the f32 position relative to the particle's cell at the
// last neighbour build follow x.  Returns whether i has moved more than
// skin/2 since that build.
__device__ __forceinline__ bool track_position(const Particles& P, const Derived& D, const KernelParams& kp,
                                               const Grid& g, const uint32_t* __restrict__ skey, int mixed,
                                               const Mixed& X, uint32_t i, const double x[3]) {
    double2* pos = reinterpret_cast<double2*>(P.pos + i);
    pos[0] = make_double2(x[0], x[1]);
    pos[1] = make_double2(x[2], P.h[i]);
    if (mixed) {
        int cx, cy, cz;
        decode_cell(skey[i], g, cx, cy, cz);
        double d[3] = {x[0] - (g.lo[0] + cx * g.w[0]), x[1] - (g.lo[1] + cy * g.w[1]),
                       x[2] - (g.lo[2] + cz * g.w[2])};
        if (g.periodic)
            for (int k = 0; k < 3; ++k) min_image(d[k], g.L[k], g.halfL[k]);
        X.p4[i] = make_float4((float)d[0], (float)d[1], (float)d[2], (float)P.h[i]);
    }
    if (!(kp.skin > 0.0)) return false;  // no skin: the list is rebuilt every tick anyway
    const double2* ref = reinterpret_cast<const double2*>(D.xref + i);
    const double2 r01 = ref[0], r2_ = ref[1];
    double dx, dy, dz;
    pair_vec(g, r01.x, r01.y, r2_.x, x[0], x[1], x[2], dx, dy, dz);
    return (dx * dx + dy * dy) + dz * dz > kp.half_skin_sq;
}

// Raise the moved flag in epoch mf.epoch.  The device word is read first:
// once it holds the epoch (with a list rebuilt every tick every wave finds a
// particle beyond skin/2), waves skip the atomic instead of queueing on one
// address, and only the wave that moves it to the epoch writes the host word
// (one store across the bus per epoch).
__device__ __forceinline__ void raise_moved(const MovedFlag& mf) {
    if (*reinterpret_cast<volatile const uint32_t*>(mf.dev) != mf.epoch && atomicMax(mf.dev, mf.epoch) < mf.epoch)
        *reinterpret_cast<volatile uint32_t*>(mf.host) = mf.epoch;
}

// State replaced from outside (the adjoint's tape and checkpoints), in one
// pass: velocities (state and record), and positions with the bookkeeping of
// a position update.
__global__ void k_load_state(Particles P, Derived D, uint32_t n, KernelParams kp, Grid g,
                             const uint32_t* __restrict__ skey, int mixed, Mixed X, Vec3 xs, Vec3 vs, MovedFlag mf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool far = false;
    if (i < n) {
        const uint32_t u = P.orig[i];
        const double vx = vs.x[u], vy = vs.y[u], vz = vs.z[u];
        P.vx[i] = vx;
        P.vy[i] = vy;
        P.vz[i] = vz;
        set_record_velocity(D, kp, i, mixed != 0, vx, vy, vz);
        if (xs.x) {
            const double x[3] = {xs.x[u], xs.y[u], xs.z[u]};
            P.x[i] = x[0];
            P.y[i] = x[1];
            P.z[i] = x[2];
            far = track_position(P, D, kp, g, skey, mixed, X, i, x) && mf.dev;
        }
    }
    if (__ballot(far) && (threadIdx.x & 63u) == 0u) raise_moved(mf);
}
void launch_load_state(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp, const Grid& g,
                       const uint32_t* skey, bool mixed, const Mixed& X, const double* x3, const double* v3,
                       const MovedFlag& mf, hipStream_t s) {
    const size_t N = n;
    double* xp = const_cast<double*>(x3);
    double* vp = const_cast<double*>(v3);
    const Vec3 xs = x3 ? Vec3{xp, xp + N, xp + 2 * N} : Vec3{nullptr, nullptr, nullptr};
    const Vec3 vs{vp, vp + N, vp + 2 * N};
    if (n)
        hipLaunchKernelGGL(k_load_state, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, kp, g, skey, (int)mixed, X,
                           xs, vs, mf);
}

// velocity_verlet_position_system (integrators.rs:63-83), then the box:
// reflection mirrors x and flips v and the stored acceleration; periodic wraps.
// When `flip` is set, bit k of flip[orig[i]] (upload order) records a
// reflection on axis k (the adjoint needs the sign of the map).  A particle
// that has moved more than skin/2 since the last neighbour build raises
// *moved (see track_position).
__global__ void k_vv_position(Particles P, Derived D, uint32_t n, KernelParams kp, Grid g,
                              const uint32_t* __restrict__ skey, int mixed, Mixed X, MovedFlag mf,
                              uint8_t* flip) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool far = false;
    if (i < n) {
        double* xs[3] = {P.x, P.y, P.z};
        double* vs[3] = {P.vx, P.vy, P.vz};
        double* aps[3] = {P.apx, P.apy, P.apz};
        const double* as[3] = {D.ax, D.ay, D.az};
        double nx[3];
        uint32_t bits = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double a = as[k][i];
            const double v = vs[k][i];
            double ap = a;
            double x = xs[k][i] + (v * kp.dt + a * kp.half_dt_sq);
            if (kp.bnd & SPHX_BND_REFLECT) {
                const double lo = kp.box_lo[k], hi = kp.box_hi[k];
                const bool below = x < lo, above = x > hi;
                if (below || above) {  // v changes only here: no store otherwise
                    x = below ? lo + (lo - x) : hi - (x - hi);
                    vs[k][i] = -v;
                    ap = -ap;
                    bits |= 1u << k;
                }
            } else if (kp.bnd & SPHX_BND_PERIODIC) {
                const double lo = kp.box_lo[k], hi = kp.box_hi[k], L = hi - lo;
                if (x >= hi) x -= L;
                else if (x < lo) x += L;
            }
            xs[k][i] = x;
            aps[k][i] = ap;
            nx[k] = x;
        }
        if (flip) flip[P.orig[i]] = (uint8_t)bits;
        if (kp.eos == SPHX_EOS_ADIABATIC) P.dup[i] = D.du[i];  // du/dt of the tick's first evaluation
        far = track_position(P, D, kp, g, skey, mixed, X, i, nx);
    }
    if (__ballot(far) && (threadIdx.x & 63u) == 0u) raise_moved(mf);
}
void launch_vv_position(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp, const Grid& g,
                        const uint32_t* skey, bool mixed, const Mixed& X, const MovedFlag& mf, uint8_t* flip,
                        hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_vv_position, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, kp, g, skey, (int)mixed, X,
                           mf, flip);
}

// velocity_verlet_velocity_system (integrators.rs:92-104)
// The force pass's gather record is updated with the new velocity.  With the
// adiabatic EOS the internal energy is advanced the same way,
// u += (du_n + du*) dt / 2.
__global__ void k_vv_velocity(Particles P, Derived D, uint32_t n, KernelParams kp, int mixed) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vx = P.vx[i] + (P.apx[i] + D.ax[i]) * kp.half_dt;
    const double vy = P.vy[i] + (P.apy[i] + D.ay[i]) * kp.half_dt;
    const double vz = P.vz[i] + (P.apz[i] + D.az[i]) * kp.half_dt;
    P.vx[i] = vx;
    P.vy[i] = vy;
    P.vz[i] = vz;
    set_record_velocity(D, kp, i, mixed != 0, vx, vy, vz);
    if (kp.eos == SPHX_EOS_ADIABATIC) P.u[i] += (P.dup[i] + D.du[i]) * kp.half_dt;
}

// Adiabatic EOS after the internal energy changed at fixed positions: P,
// P/rho^2 and the gather records' P/rho^2 from the kept density and the new u
// (the same arithmetic as the density pass's epilogue).
__global__ void k_eos(Particles P, Derived D, uint32_t n, KernelParams kp, int mixed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double rho = D.rho[i];
    const double pr = eos_pressure(kp, rho, P.u[i]);
    const double pr2 = pr / (rho * rho);
    D.p[i] = pr;
    D.pr2[i] = pr2;
    if (mixed) D.r32[i].pr2 = (float)pr2;
    if (kp.rec64) D.r64[i].pr2 = pr2;
}
void launch_eos(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp, bool mixed,
                hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_eos, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, kp, (int)mixed);
}

// The Maxwell-stress records from the current B and density, for a force pass
// that follows a change of B without a cleaning pass (which writes them).
__global__ void k_stress_rec(Particles P, Derived D, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    D.recs[i] = stress_record(P.bx[i], P.by[i], P.bz[i], D.rho[i]);
}
void launch_stress_rec(const Particles& P, const Derived& D, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_stress_rec, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n);
}

void launch_vv_velocity(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp, bool mixed,
                        const Mixed& X, hipStream_t s) {
    (void)X;
    if (n) hipLaunchKernelGGL(k_vv_velocity, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, kp, (int)mixed);
}

// ---------------------------------------------------------- energy
// Per-block partial sums (deterministic), finished on the host.
static constexpr int kEnergyVals = 9;
uint32_t energy_blocks(uint32_t n) { return nblocks(n ? n : 1); }

__device__ double specific_internal(const KernelParams& kp, double rho) {
    if (kp.eos == SPHX_EOS_TAIT) {
        const double rho0 = kp.rest_density, g = kp.gamma;
        if (rho <= rho0) return 0.0;
        return kp.tait_B * ((pow(rho / rho0, g - 1.0) - 1.0) / ((g - 1.0) * rho0) + 1.0 / rho - 1.0 / rho0);
    }
    if (kp.eos == SPHX_EOS_POLYTROPIC) return kp.poly_k * pow(rho, kp.gamma - 1.0) / (kp.gamma - 1.0);
    return kp.iso_cs2 * log(rho / kp.rest_density);
}

__global__ void k_energy(Particles P, Derived D, uint32_t n, Grid g, NbrList L, KernelParams kp, int mhd,
                         double* __restrict__ partial) {
    __shared__ double lds[kEnergyVals][kBlock];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[kEnergyVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double m = P.m[i], vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
        v[0] = 0.5 * m * ((vx * vx + vy * vy) + vz * vz);
        v[1] = -m * ((kp.g[0] * P.x[i] + kp.g[1] * P.y[i]) + kp.g[2] * P.z[i]);
        v[2] = m * (kp.eos == SPHX_EOS_ADIABATIC ? P.u[i] : specific_internal(kp, D.rho[i]));
        if (mhd) {
            const double bx = P.bx[i], by = P.by[i], bz = P.bz[i];
            v[3] = 0.5 * (m / D.rho[i]) * ((bx * bx + by * by) + bz * bz);
        }
        v[4] = m * vx; v[5] = m * vy; v[6] = m * vz;
        const uint32_t c = exact_count(P, g, L, i);
        v[7] = (double)c;
        v[8] = (double)c;
    }
    for (int k = 0; k < kEnergyVals; ++k) lds[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int k = 0; k < 8; ++k) lds[k][threadIdx.x] += lds[k][threadIdx.x + off];
            lds[8][threadIdx.x] = fmax(lds[8][threadIdx.x], lds[8][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x < kEnergyVals) partial[(size_t)blockIdx.x * kEnergyVals + threadIdx.x] = lds[threadIdx.x][0];
}
void launch_energy(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                   const KernelParams& kp, bool mhd, double* partial, uint32_t nb, hipStream_t s) {
    hipLaunchKernelGGL(k_energy, dim3(nb), dim3(kBlock), 0, s, P, D, n, g, L, kp, (int)mhd, partial);
}

}  // namespace sphx
