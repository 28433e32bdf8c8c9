// sphx_kernels.hip — HIP kernels of the SPH particle step for gfx950 (MI355X).
//
// Data layout: every per-particle quantity is its own f64 (or u32) array in
// HBM, in cell-sorted order, so that a wavefront's 64 lanes load 64
// consecutive 8-byte words (one 512-byte coalesced access per array).
// Neighbour lists are column-major (idx[k * cap + i]) for the same reason.
// One thread owns one particle and gathers over its neighbours: no atomics
// on the floating-point path, so every sum is deterministic.
//
// Compiled with -ffp-contract=off: the neighbour predicate must round exactly
// like the reference's sqrt((dx*dx + dy*dy) + dz*dz) < 2*h_avg
// (src/physics/sph.rs:102-108).  Elsewhere fused multiply-adds are requested
// explicitly with fma().

#include <hip/hip_runtime.h>
#include <math.h>

#include "sphx_internal.h"

namespace sphx {

static constexpr int kBlock = 256;
static constexpr double kInvPi = 0.31830988618379067154;  // 1/pi
static constexpr double kPi = 3.141592653589793;           // std::f64::consts::PI

static inline uint32_t nblocks(size_t n, int b = kBlock) { return (uint32_t)((n + b - 1) / b); }

// blockIdx -> logical block so that each XCD (blocks are dealt round-robin to
// the 8 XCDs) works on one contiguous range of sorted particles and its L2
// holds that range's neighbours.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t xcd = b & 7u, local = b >> 3;
    const uint32_t base = nb >> 3, rem = nb & 7u;
    return xcd * base + min(xcd, rem) + local;
}

__device__ __forceinline__ void min_image(double& d, double L, double halfL) {
    if (d > halfL) d -= L;
    else if (d < -halfL) d += L;
}

__device__ __forceinline__ void pair_vec(const Grid& g, double xi, double yi, double zi, double xj,
                                         double yj, double zj, double& dx, double& dy, double& dz) {
    dx = xj - xi;
    dy = yj - yi;
    dz = zj - zi;
    if (g.periodic) {
        min_image(dx, g.L[0], g.halfL[0]);
        min_image(dy, g.L[1], g.halfL[1]);
        min_image(dz, g.L[2], g.halfL[2]);
    }
}

__device__ __forceinline__ int cell_coord(double x, double lo, double inv_w, int dim) {
    double u = (x - lo) * inv_w;
    u = fmin(fmax(u, 0.0), (double)(dim - 1));
    return (int)floor(u);
}

// Cubic spline W(r, h) of sph.rs:29-43 given 1/h (sigma = 1/(pi h^3)).
__device__ __forceinline__ double kernel_w(double r, double inv_h) {
    const double q = r * inv_h;
    const double sigma = (inv_h * inv_h) * inv_h * kInvPi;
    if (q < 1.0) return sigma * fma(q * q, fma(0.75, q, -1.5), 1.0);
    if (q < 2.0) {
        const double t = 2.0 - q;
        return (0.25 * sigma) * ((t * t) * t);
    }
    return 0.0;
}

// dW/dr of sph.rs:50-64 given 1/h (sigma = 1/(pi h^4)); caller handles r < 1e-10.
__device__ __forceinline__ double kernel_dw(double r, double inv_h) {
    const double q = r * inv_h;
    const double ih2 = inv_h * inv_h;
    const double sigma = (ih2 * ih2) * kInvPi;
    if (q < 1.0) return sigma * (q * fma(2.25, q, -3.0));
    if (q < 2.0) {
        const double t = 2.0 - q;
        return sigma * (-0.75 * (t * t));
    }
    return 0.0;
}

// ---------------------------------------------------------------- utilities
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(nblocks(n)), dim3(kBlock), 0, s, p, v, n);
}
__global__ void k_iota_u32(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}
void launch_iota_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_iota_u32, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n);
}

// ------------------------------------------------------ counting sort (a1)
// key = linear cell index (x fastest); rank = arrival order inside the cell.
__global__ void k_cell_hash(const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, uint32_t n, Grid g,
                            uint32_t* __restrict__ key, uint32_t* __restrict__ rank,
                            uint32_t* __restrict__ count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cx = cell_coord(x[i], g.lo[0], g.inv_w[0], g.dims[0]);
    int cy = cell_coord(y[i], g.lo[1], g.inv_w[1], g.dims[1]);
    int cz = cell_coord(z[i], g.lo[2], g.inv_w[2], g.dims[2]);
    uint32_t c = ((uint32_t)cz * (uint32_t)g.dims[1] + (uint32_t)cy) * (uint32_t)g.dims[0] + (uint32_t)cx;
    key[i] = c;
    rank[i] = atomicAdd(&count[c], 1u);
}
void launch_cell_hash(const Particles& P, uint32_t n, const Grid& g, const Sort& S, hipStream_t s) {
    launch_fill_u32(S.cell_count, 0u, g.ncell, s);
    if (n)
        hipLaunchKernelGGL(k_cell_hash, dim3(nblocks(n)), dim3(kBlock), 0, s, P.x, P.y, P.z, n, g,
                           S.key, S.rank, S.cell_count);
}

// Exclusive scan of cell_count into cell_start (ncell + 1 entries), three
// kernels: per-1024 block scans, one-block scan of the block sums, add.
static constexpr int kScanItems = 4;
static constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    // Hillis-Steele over 256 threads in LDS.
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        uint32_t add = (t >= off) ? lds[t - off] : 0u;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    total = lds[kBlock - 1];
    uint32_t incl = lds[t];
    __syncthreads();
    return incl - v;
}

__global__ void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
                             uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t lds[kBlock];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, lds, total);
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void k_scan_sums(uint32_t* __restrict__ sums, uint32_t m, uint32_t* __restrict__ grand) {
    __shared__ uint32_t lds[kBlock];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < m; base += kBlock) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = (i < m) ? sums[i] : 0u;
        uint32_t total;
        uint32_t ex = block_exclusive_scan(v, lds, total);
        if (i < m) sums[i] = ex + carry;
        carry += total;
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, uint32_t n, const uint32_t* __restrict__ sums) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const uint32_t add = sums[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) out[base + k] += add;
}

void launch_scan(const Sort& S, uint32_t ncell, hipStream_t s) {
    uint32_t tiles = (ncell + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kBlock), 0, s, S.cell_count, S.cell_start,
                       ncell, S.block_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, S.block_sums, tiles,
                       S.cell_start + ncell);
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kBlock), 0, s, S.cell_start, ncell, S.block_sums);
}

__global__ void k_scatter(const uint32_t* __restrict__ key, const uint32_t* __restrict__ rank,
                          const uint32_t* __restrict__ start, uint32_t n, uint32_t* __restrict__ perm) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[start[key[i]] + rank[i]] = i;
}
void launch_scatter(const Sort& S, uint32_t n, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_scatter, dim3(nblocks(n)), dim3(kBlock), 0, s, S.key, S.rank,
                           S.cell_start, n, S.perm);
}

// Atomic ranks arrive in any order; sorting each cell's slice of perm by the
// previous index makes the order deterministic (and stable with respect to
// the previous step, which keeps the reorder gather nearly sequential).
__global__ void k_cell_order(const uint32_t* __restrict__ start, uint32_t ncell,
                             uint32_t* __restrict__ perm) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const uint32_t b = start[c], e = start[c + 1];
    for (uint32_t k = b + 1; k < e; ++k) {
        uint32_t v = perm[k];
        uint32_t t = k;
        while (t > b && perm[t - 1] > v) {
            perm[t] = perm[t - 1];
            --t;
        }
        perm[t] = v;
    }
}
void launch_cell_order(const Sort& S, uint32_t ncell, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_order, dim3(nblocks(ncell)), dim3(kBlock), 0, s, S.cell_start, ncell,
                       S.perm);
}

__device__ __forceinline__ void decode_cell(uint32_t key, const Grid& g, int& cx, int& cy, int& cz) {
    const uint32_t nx = (uint32_t)g.dims[0], ny = (uint32_t)g.dims[1];
    cx = (int)(key % nx);
    const uint32_t t = key / nx;
    cy = (int)(t % ny);
    cz = (int)(t / ny);
}

__global__ void k_reorder(Particles src, Particles dst, const uint32_t* __restrict__ perm,
                          const uint32_t* __restrict__ key, uint32_t* __restrict__ skey, uint32_t n,
                          int mhd, int charge, int mixed, Mixed X, Grid g) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = perm[k];
    const uint32_t c = key[i];
    skey[k] = c;
    const double x = src.x[i], y = src.y[i], z = src.z[i];
    const double vx = src.vx[i], vy = src.vy[i], vz = src.vz[i];
    const double m = src.m[i], h = src.h[i];
    dst.x[k] = x;
    dst.y[k] = y;
    dst.z[k] = z;
    dst.vx[k] = vx;
    dst.vy[k] = vy;
    dst.vz[k] = vz;
    dst.apx[k] = src.apx[i];
    dst.apy[k] = src.apy[i];
    dst.apz[k] = src.apz[i];
    dst.m[k] = m;
    dst.h[k] = h;
    dst.id[k] = src.id[i];
    dst.orig[k] = src.orig[i];
    if (mhd) {
        dst.bx[k] = src.bx[i];
        dst.by[k] = src.by[i];
        dst.bz[k] = src.bz[i];
        dst.psi[k] = src.psi[i];
    }
    if (charge) dst.q[k] = src.q[i];
    if (mixed) {
        int cx, cy, cz;
        decode_cell(c, g, cx, cy, cz);
        X.xr[k] = (float)(x - (g.lo[0] + cx * g.w[0]));
        X.yr[k] = (float)(y - (g.lo[1] + cy * g.w[1]));
        X.zr[k] = (float)(z - (g.lo[2] + cz * g.w[2]));
        X.vx[k] = (float)vx;
        X.vy[k] = (float)vy;
        X.vz[k] = (float)vz;
        X.m[k] = (float)m;
        X.h[k] = (float)h;
    }
}
void launch_reorder(const Particles& src, const Particles& dst, const Sort& S, uint32_t n, bool mhd,
                    bool charge, bool mixed, const Mixed& X, const Grid& g, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_reorder, dim3(nblocks(n)), dim3(kBlock), 0, s, src, dst, S.perm, S.key,
                           S.skey, n, (int)mhd, (int)charge, (int)mixed, X, g);
}

// ---------------------------------------------- neighbour lists (a1, a2)
// Visits the 27 cells around particle i's cell in (z, y, x) order, calling
// f(j, code) for every particle j of each; code packs the cell offset
// (ox+1) | (oy+1) << 2 | (oz+1) << 4.  With x the fastest cell axis the
// candidates come in ascending sorted index within each (y, z) row.
template <typename F>
__device__ __forceinline__ void for_each_candidate(const Grid& g, const uint32_t* __restrict__ start,
                                                   uint32_t key, F&& f) {
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    int cx, cy, cz;
    decode_cell(key, g, cx, cy, cz);
    for (int oz = -1; oz <= 1; ++oz) {
        int z = cz + oz;
        if (g.periodic) z = (z + nz) % nz;
        else if (z < 0 || z >= nz) continue;
        for (int oy = -1; oy <= 1; ++oy) {
            int y = cy + oy;
            if (g.periodic) y = (y + ny) % ny;
            else if (y < 0 || y >= ny) continue;
            const uint32_t row = ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)nx;
            for (int ox = -1; ox <= 1; ++ox) {
                int x = cx + ox;
                if (g.periodic) x = (x + nx) % nx;
                else if (x < 0 || x >= nx) continue;
                const uint32_t code = (uint32_t)(ox + 1) | ((uint32_t)(oy + 1) << 2) | ((uint32_t)(oz + 1) << 4);
                const uint32_t b = start[row + x], e = start[row + x + 1];
                for (uint32_t j = b; j < e; ++j) f(j, code);
            }
        }
    }
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}

// Exact membership r < 2*h_avg with r = sqrt((dx*dx + dy*dy) + dz*dz) rounded
// as the reference rounds it.  r2 far from the threshold (relative 1e-12,
// about 5000 ulp) decides without the square root; the band in between takes
// the correctly rounded sqrt and the reference comparison.
__device__ __forceinline__ bool exact_neighbour(double dx, double dy, double dz, double hi, double hj) {
    const double r2 = (dx * dx + dy * dy) + dz * dz;
    const double h_avg = 0.5 * (hi + hj);
    const double t = 2.0 * h_avg;
    const double t2 = t * t;
    if (r2 < t2 * (1.0 - 1e-12)) return true;
    if (!(r2 <= t2 * (1.0 + 1e-12))) return false;
    return sqrt(r2) < t;
}

// Cell offset code -> f32 offset of the neighbour's cell origin.
__device__ __forceinline__ void code_offset(uint32_t code, float wx, float wy, float wz, float& ox,
                                            float& oy, float& oz) {
    ox = (float)((int)(code & 3u) - 1) * wx;
    oy = (float)((int)((code >> 2) & 3u) - 1) * wy;
    oz = (float)((int)((code >> 4) & 3u) - 1) * wz;
}

template <bool MIXED>
__global__ void __launch_bounds__(kBlock) k_build_nbr(Particles P, Mixed X, uint32_t n, uint32_t cap, Grid g,
                                                       const uint32_t* __restrict__ key,
                                                       const uint32_t* __restrict__ start,
                                                       uint32_t* __restrict__ idx, uint32_t* __restrict__ cnt,
                                                       uint32_t* __restrict__ ovf, int kmax) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    uint32_t count = 0;
    if (i < n) {
        auto push = [&](uint32_t j, uint32_t code) {
            if (count < (uint32_t)kmax) idx[(size_t)count * cap + i] = j | (code << kIdxBits);
            ++count;
        };
        if (!MIXED) {
            const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i];
            for_each_candidate(g, start, key[i], [&](uint32_t j, uint32_t code) {
                if (j == i) return;
                double dx, dy, dz;
                pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], dx, dy, dz);
                if (exact_neighbour(dx, dy, dz, hi, P.h[j])) push(j, code);
            });
        } else {
            // f32 predicate on cell-relative coordinates; d_ji = -d_ij exactly,
            // so the lists stay symmetric.
            const float xi = X.xr[i], yi = X.yr[i], zi = X.zr[i], hi = X.h[i];
            const float wx = (float)g.w[0], wy = (float)g.w[1], wz = (float)g.w[2];
            for_each_candidate(g, start, key[i], [&](uint32_t j, uint32_t code) {
                if (j == i) return;
                float ox, oy, oz;
                code_offset(code, wx, wy, wz, ox, oy, oz);
                const float dx = (X.xr[j] - xi) + ox, dy = (X.yr[j] - yi) + oy, dz = (X.zr[j] - zi) + oz;
                const float t = 2.0f * (0.5f * (hi + X.h[j]));
                if ((dx * dx + dy * dy) + dz * dz < t * t) push(j, code);
            });
        }
        cnt[i] = min(count, (uint32_t)kmax);
    }
    const uint32_t wmax = wave_max_u32(count);
    if ((threadIdx.x & 63) == 0 && wmax > 0) {
        atomicMax(&ovf[0], wmax);
        if (wmax > (uint32_t)kmax) atomicOr(&ovf[1], 1u);
    }
}
void launch_build_nbr(const Particles& P, uint32_t n, uint32_t cap, const Grid& g, const Sort& S,
                      const NbrList& L, bool mixed, const Mixed& X, hipStream_t s) {
    if (!n) return;
    if (mixed)
        hipLaunchKernelGGL(k_build_nbr<true>, dim3(nblocks(n)), dim3(kBlock), 0, s, P, X, n, cap, g, S.skey,
                           S.cell_start, L.idx, L.cnt, L.overflow, L.kmax);
    else
        hipLaunchKernelGGL(k_build_nbr<false>, dim3(nblocks(n)), dim3(kBlock), 0, s, P, X, n, cap, g, S.skey,
                           S.cell_start, L.idx, L.cnt, L.overflow, L.kmax);
}

// ------------------------------------------------------- density + EOS (a4, a5)
__device__ __forceinline__ double eos_pressure(const KernelParams& kp, double rho) {
    if (kp.eos == SPHX_EOS_TAIT) {
        double P = kp.tait_B * (pow(rho / kp.rest_density, kp.gamma) - 1.0);
        return P < 0.0 ? 0.0 : P;
    }
    if (kp.eos == SPHX_EOS_POLYTROPIC) return kp.poly_k * pow(rho, kp.gamma);
    return rho * kp.iso_cs2;
}

__device__ __forceinline__ double inv_h_pair(double hi, double hj, double inv_hi) {
    return (hi == hj) ? inv_hi : 1.0 / (0.5 * (hi + hj));
}

__global__ void __launch_bounds__(kBlock) k_density(Particles P, Derived D, uint32_t n, uint32_t cap, Grid g,
                                                     const uint32_t* __restrict__ idx,
                                                     const uint32_t* __restrict__ cnt, KernelParams kp) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i];
    const double inv_hi = 1.0 / hi;
    // self term m_i W(0, h_i) = m_i / (pi h_i^3)   (sph.rs:89-91)
    double rho = P.m[i] * (1.0 / (kPi * ((hi * hi) * hi)));
    const uint32_t c = cnt[i];
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t j = idx[(size_t)k * cap + i] & kIdxMask;
        double dx, dy, dz;
        pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], dx, dy, dz);
        const double r = sqrt((dx * dx + dy * dy) + dz * dz);
        rho = fma(P.m[j], kernel_w(r, inv_h_pair(hi, P.h[j], inv_hi)), rho);
    }
    const double p = eos_pressure(kp, rho);
    D.rho[i] = rho;
    D.p[i] = p;
    D.pr2[i] = p / (rho * rho);
}

// f32 cubic spline and derivative (FP32-mixed form)
__device__ __forceinline__ float kernel_w_f(float r, float inv_h) {
    const float q = r * inv_h;
    const float sigma = (inv_h * inv_h) * inv_h * (float)kInvPi;
    if (q < 1.0f) return sigma * fmaf(q * q, fmaf(0.75f, q, -1.5f), 1.0f);
    if (q < 2.0f) {
        const float t = 2.0f - q;
        return (0.25f * sigma) * ((t * t) * t);
    }
    return 0.0f;
}
__device__ __forceinline__ float kernel_dw_f(float r, float inv_h) {
    const float q = r * inv_h;
    const float ih2 = inv_h * inv_h;
    const float sigma = (ih2 * ih2) * (float)kInvPi;
    if (q < 1.0f) return sigma * (q * fmaf(2.25f, q, -3.0f));
    if (q < 2.0f) {
        const float t = 2.0f - q;
        return sigma * (-0.75f * (t * t));
    }
    return 0.0f;
}
__device__ __forceinline__ float inv_h_pair_f(float hi, float hj, float inv_hi) {
    return (hi == hj) ? inv_hi : __builtin_amdgcn_rcpf(0.5f * (hi + hj));
}

// FP32-mixed density: f32 pair terms, f64 sum; rho, P and P/rho^2 in f64
// with f32 copies for the force pass.
__global__ void __launch_bounds__(kBlock) k_density_mixed(Particles P, Derived D, Mixed X, uint32_t n,
                                                           uint32_t cap, Grid g,
                                                           const uint32_t* __restrict__ idx,
                                                           const uint32_t* __restrict__ cnt, KernelParams kp) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double hd = P.h[i];
    double rho = P.m[i] * (1.0 / (kPi * ((hd * hd) * hd)));
    const float xi = X.xr[i], yi = X.yr[i], zi = X.zr[i], hi = X.h[i];
    const float inv_hi = 1.0f / hi;
    const float wx = (float)g.w[0], wy = (float)g.w[1], wz = (float)g.w[2];
    const uint32_t c = cnt[i];
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t e = idx[(size_t)k * cap + i];
        const uint32_t j = e & kIdxMask;
        float ox, oy, oz;
        code_offset(e >> kIdxBits, wx, wy, wz, ox, oy, oz);
        const float dx = (X.xr[j] - xi) + ox, dy = (X.yr[j] - yi) + oy, dz = (X.zr[j] - zi) + oz;
        const float r = __builtin_amdgcn_sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
        rho += (double)(X.m[j] * kernel_w_f(r, inv_h_pair_f(hi, X.h[j], inv_hi)));
    }
    const double p = eos_pressure(kp, rho);
    const double pr2 = p / (rho * rho);
    D.rho[i] = rho;
    D.p[i] = p;
    D.pr2[i] = pr2;
    X.rho[i] = (float)rho;
    X.pr2[i] = (float)pr2;
}

void launch_density(const Particles& P, const Derived& D, uint32_t n, uint32_t cap, const Grid& g,
                    const NbrList& L, const KernelParams& kp, bool mixed, const Mixed& X, hipStream_t s) {
    if (!n) return;
    if (mixed)
        hipLaunchKernelGGL(k_density_mixed, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, X, n, cap, g, L.idx,
                           L.cnt, kp);
    else
        hipLaunchKernelGGL(k_density, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, cap, g, L.idx, L.cnt, kp);
}

// ------------------------- pressure + viscosity + Maxwell stress + walls (a6-a9, a13)
template <bool VISC_ONLY, bool STRESS, bool CHARGE>
__global__ void __launch_bounds__(kBlock) k_forces(Particles P, Derived D, uint32_t n, uint32_t cap, Grid g,
                                                    const uint32_t* __restrict__ idx,
                                                    const uint32_t* __restrict__ cnt, KernelParams kp) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i];
    const double vxi = P.vx[i], vyi = P.vy[i], vzi = P.vz[i];
    const double mi = P.m[i], rhoi = D.rho[i], pr2i = D.pr2[i];
    const double inv_hi = 1.0 / hi;
    const double ac = -kp.alpha * kp.speed_of_sound;
    const bool ref_conv = kp.visc_conv == SPHX_VISC_REFERENCE;
    double bxi = 0, byi = 0, bzi = 0, irho2i = 0, hb2i = 0;
    if (STRESS) {
        bxi = P.bx[i]; byi = P.by[i]; bzi = P.bz[i];
        irho2i = 1.0 / (rhoi * rhoi);
        hb2i = 0.5 * ((bxi * bxi + byi * byi) + bzi * bzi);
    }
    double fpx = 0, fpy = 0, fpz = 0;  // pressure
    double fvx = 0, fvy = 0, fvz = 0;  // viscosity
    double fmx = 0, fmy = 0, fmz = 0;  // Maxwell stress
    const uint32_t c = cnt[i];
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t j = idx[(size_t)k * cap + i] & kIdxMask;
        double dx, dy, dz;
        pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], dx, dy, dz);
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        const double r = sqrt(r2);
        if (!(r > 1e-10)) continue;  // dW/dr = 0 below 1e-10 (sph.rs:52): no force
        const double hj = P.h[j];
        const double inv_h = inv_h_pair(hi, hj, inv_hi);
        const double dw = kernel_dw(r, inv_h);
        const double inv_r = 1.0 / r;
        const double mm = mi * P.m[j];
        const double dwr = dw * inv_r;  // (dW/dr) / r: force = d * (scalar * dwr)
        if (!VISC_ONLY) {
            const double s = (mm * (pr2i + D.pr2[j])) * dwr;
            fpx = fma(dx, s, fpx);
            fpy = fma(dy, s, fpy);
            fpz = fma(dz, s, fpz);
        }
        // (v_i - v_j).(x_j - x_i) as sph.rs:292-295 forms it
        const double sdot = ((vxi - P.vx[j]) * dx + (vyi - P.vy[j]) * dy) + (vzi - P.vz[j]) * dz;
        const bool active = ref_conv ? (sdot < 0.0) : (sdot > 0.0);
        if (active) {
            const double h_avg = 0.5 * (hi + hj);
            const double rho_bar = 0.5 * (rhoi + D.rho[j]);
            const double ss = ref_conv ? sdot : -sdot;
            const double mu = (h_avg * ss) / fma(0.01 * h_avg, h_avg, r2);
            const double pi_ij = fma(kp.beta * mu, mu, ac * mu) / rho_bar;
            const double s = (mm * pi_ij) * dwr;
            fvx = fma(dx, s, fvx);
            fvy = fma(dy, s, fvy);
            fvz = fma(dz, s, fvz);
        }
        if (STRESS) {
            // grad_i W = -(dW/dr) d / r
            const double gx = -dx * dwr, gy = -dy * dwr, gz = -dz * dwr;
            const double bxj = P.bx[j], byj = P.by[j], bzj = P.bz[j];
            const double rhoj = D.rho[j];
            const double irho2j = 1.0 / (rhoj * rhoj);
            const double hb2j = 0.5 * ((bxj * bxj + byj * byj) + bzj * bzj);
            const double big = (bxi * gx + byi * gy) + bzi * gz;
            const double bjg = (bxj * gx + byj * gy) + bzj * gz;
            fmx += mm * ((bxi * big - hb2i * gx) * irho2i + (bxj * bjg - hb2j * gx) * irho2j);
            fmy += mm * ((byi * big - hb2i * gy) * irho2i + (byj * bjg - hb2j * gy) * irho2j);
            fmz += mm * ((bzi * big - hb2i * gz) * irho2i + (bzj * bjg - hb2j * gz) * irho2j);
        }
    }
    double sx, sy, sz;
    if (VISC_ONLY) {
        sx = D.spx[i]; sy = D.spy[i]; sz = D.spz[i];
    } else {
        sx = mi * kp.g[0] + fpx;
        sy = mi * kp.g[1] + fpy;
        sz = mi * kp.g[2] + fpz;
        D.spx[i] = sx; D.spy[i] = sy; D.spz[i] = sz;
    }
    if (CHARGE) {  // lorentz.rs:27-37
        const double q = P.q[i];
        const double bx = P.bx[i], by = P.by[i], bz = P.bz[i];
        sx += q * (kp.e[0] + (vyi * bz - vzi * by));
        sy += q * (kp.e[1] + (vzi * bx - vxi * bz));
        sz += q * (kp.e[2] + (vxi * by - vyi * bx));
    }
    double F[3] = {sx + fvx + fmx, sy + fvy + fmy, sz + fvz + fmz};
    if (kp.bnd & SPHX_BND_PENALTY_CUBE) {  // sph.rs:451-483
        const double c3[3] = {xi, yi, zi};
        for (int a = 0; a < 3; ++a) {
            const double d_low = c3[a] - (-kp.half_extent);
            if (d_low < kp.act_dist && d_low >= 0.0) F[a] += kp.stiffness * (kp.act_dist - d_low);
            const double d_high = kp.half_extent - c3[a];
            if (d_high < kp.act_dist && d_high >= 0.0) F[a] -= kp.stiffness * (kp.act_dist - d_high);
        }
    }
    D.ax[i] = F[0] / mi;  // integrators.rs:115
    D.ay[i] = F[1] / mi;
    D.az[i] = F[2] / mi;
}

// FP32-mixed force pass: pair terms in f32 on cell-relative coordinates, the
// pressure / viscosity / stress sums in f64; same structure and epilogue as
// k_forces, so VISC_ONLY reuse stays bit-identical within this form.
template <bool VISC_ONLY, bool STRESS, bool CHARGE>
__global__ void __launch_bounds__(kBlock) k_forces_mixed(Particles P, Derived D, Mixed X, uint32_t n,
                                                          uint32_t cap, Grid g,
                                                          const uint32_t* __restrict__ idx,
                                                          const uint32_t* __restrict__ cnt, KernelParams kp) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float xi = X.xr[i], yi = X.yr[i], zi = X.zr[i], hi = X.h[i];
    const float vxi = X.vx[i], vyi = X.vy[i], vzi = X.vz[i];
    const float mi = X.m[i], rhoi = X.rho[i], pr2i = X.pr2[i];
    const float inv_hi = 1.0f / hi;
    const float ac = (float)(-kp.alpha * kp.speed_of_sound);
    const float beta = (float)kp.beta;
    const bool ref_conv = kp.visc_conv == SPHX_VISC_REFERENCE;
    const float wx = (float)g.w[0], wy = (float)g.w[1], wz = (float)g.w[2];
    float bxi = 0, byi = 0, bzi = 0, irho2i = 0, hb2i = 0;
    if (STRESS) {
        bxi = (float)P.bx[i]; byi = (float)P.by[i]; bzi = (float)P.bz[i];
        irho2i = 1.0f / (rhoi * rhoi);
        hb2i = 0.5f * ((bxi * bxi + byi * byi) + bzi * bzi);
    }
    double fpx = 0, fpy = 0, fpz = 0;
    double fvx = 0, fvy = 0, fvz = 0;
    double fmx = 0, fmy = 0, fmz = 0;
    const uint32_t c = cnt[i];
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t e = idx[(size_t)k * cap + i];
        const uint32_t j = e & kIdxMask;
        float ox, oy, oz;
        code_offset(e >> kIdxBits, wx, wy, wz, ox, oy, oz);
        const float dx = (X.xr[j] - xi) + ox, dy = (X.yr[j] - yi) + oy, dz = (X.zr[j] - zi) + oz;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        if (!(r2 > 1e-20f)) continue;  // r <= 1e-10: no force
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float r = r2 * inv_r;
        const float hj = X.h[j];
        const float inv_h = inv_h_pair_f(hi, hj, inv_hi);
        const float dwr = kernel_dw_f(r, inv_h) * inv_r;
        const float mm = mi * X.m[j];
        if (!VISC_ONLY) {
            const float s = (mm * (pr2i + X.pr2[j])) * dwr;
            fpx += (double)(dx * s);
            fpy += (double)(dy * s);
            fpz += (double)(dz * s);
        }
        const float sdot = fmaf(vzi - X.vz[j], dz, fmaf(vyi - X.vy[j], dy, (vxi - X.vx[j]) * dx));
        const bool active = ref_conv ? (sdot < 0.0f) : (sdot > 0.0f);
        if (active) {
            const float h_avg = 0.5f * (hi + hj);
            const float rho_bar = 0.5f * (rhoi + X.rho[j]);
            const float ss = ref_conv ? sdot : -sdot;
            const float mu = (h_avg * ss) * __builtin_amdgcn_rcpf(fmaf(0.01f * h_avg, h_avg, r2));
            const float pi_ij = fmaf(beta * mu, mu, ac * mu) * __builtin_amdgcn_rcpf(rho_bar);
            const float s = (mm * pi_ij) * dwr;
            fvx += (double)(dx * s);
            fvy += (double)(dy * s);
            fvz += (double)(dz * s);
        }
        if (STRESS) {
            const float gx = -dx * dwr, gy = -dy * dwr, gz = -dz * dwr;
            const float bxj = (float)P.bx[j], byj = (float)P.by[j], bzj = (float)P.bz[j];
            const float rhoj = X.rho[j];
            const float irho2j = 1.0f / (rhoj * rhoj);
            const float hb2j = 0.5f * ((bxj * bxj + byj * byj) + bzj * bzj);
            const float big = (bxi * gx + byi * gy) + bzi * gz;
            const float bjg = (bxj * gx + byj * gy) + bzj * gz;
            fmx += (double)(mm * ((bxi * big - hb2i * gx) * irho2i + (bxj * bjg - hb2j * gx) * irho2j));
            fmy += (double)(mm * ((byi * big - hb2i * gy) * irho2i + (byj * bjg - hb2j * gy) * irho2j));
            fmz += (double)(mm * ((bzi * big - hb2i * gz) * irho2i + (bzj * bjg - hb2j * gz) * irho2j));
        }
    }
    const double mid = P.m[i];
    double sx, sy, sz;
    if (VISC_ONLY) {
        sx = D.spx[i]; sy = D.spy[i]; sz = D.spz[i];
    } else {
        sx = mid * kp.g[0] + fpx;
        sy = mid * kp.g[1] + fpy;
        sz = mid * kp.g[2] + fpz;
        D.spx[i] = sx; D.spy[i] = sy; D.spz[i] = sz;
    }
    const double xd = P.x[i], yd = P.y[i], zd = P.z[i];
    if (CHARGE) {
        const double q = P.q[i], vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
        const double bx = P.bx[i], by = P.by[i], bz = P.bz[i];
        sx += q * (kp.e[0] + (vy * bz - vz * by));
        sy += q * (kp.e[1] + (vz * bx - vx * bz));
        sz += q * (kp.e[2] + (vx * by - vy * bx));
    }
    double F[3] = {sx + fvx + fmx, sy + fvy + fmy, sz + fvz + fmz};
    if (kp.bnd & SPHX_BND_PENALTY_CUBE) {
        const double c3[3] = {xd, yd, zd};
        for (int a = 0; a < 3; ++a) {
            const double d_low = c3[a] - (-kp.half_extent);
            if (d_low < kp.act_dist && d_low >= 0.0) F[a] += kp.stiffness * (kp.act_dist - d_low);
            const double d_high = kp.half_extent - c3[a];
            if (d_high < kp.act_dist && d_high >= 0.0) F[a] -= kp.stiffness * (kp.act_dist - d_high);
        }
    }
    D.ax[i] = F[0] / mid;
    D.ay[i] = F[1] / mid;
    D.az[i] = F[2] / mid;
}

void launch_forces(const Particles& P, const Derived& D, uint32_t n, uint32_t cap, const Grid& g,
                   const NbrList& L, const KernelParams& kp, bool visc_only, bool stress, bool charge,
                   bool mixed, const Mixed& X, hipStream_t s) {
    if (!n) return;
    dim3 grid(nblocks(n)), block(kBlock);
#define SPHX_FORCES(V, M, C)                                                                           \
    if (visc_only == V && stress == M && charge == C) {                                               \
        if (mixed)                                                                                     \
            hipLaunchKernelGGL((k_forces_mixed<V, M, C>), grid, block, 0, s, P, D, X, n, cap, g, L.idx, \
                               L.cnt, kp);                                                             \
        else                                                                                           \
            hipLaunchKernelGGL((k_forces<V, M, C>), grid, block, 0, s, P, D, n, cap, g, L.idx, L.cnt, kp); \
        return;                                                                                        \
    }
    SPHX_FORCES(false, false, false)
    SPHX_FORCES(true, false, false)
    SPHX_FORCES(false, true, false)
    SPHX_FORCES(true, true, false)
    SPHX_FORCES(false, false, true)
    SPHX_FORCES(true, false, true)
    SPHX_FORCES(false, true, true)
    SPHX_FORCES(true, true, true)
#undef SPHX_FORCES
}

// ------------------------------------------------------------ integrator (a10)
// velocity_verlet_position_system (integrators.rs:63-83), then the box:
// reflection mirrors x and flips v and the stored acceleration; periodic wraps.
__global__ void k_vv_position(Particles P, Derived D, uint32_t n, KernelParams kp) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* xs[3] = {P.x, P.y, P.z};
    double* vs[3] = {P.vx, P.vy, P.vz};
    double* aps[3] = {P.apx, P.apy, P.apz};
    const double* as[3] = {D.ax, D.ay, D.az};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double a = as[k][i];
        double v = vs[k][i];
        double ap = a;
        double x = xs[k][i] + (v * kp.dt + a * kp.half_dt_sq);
        if (kp.bnd & SPHX_BND_REFLECT) {
            const double lo = kp.box_lo[k], hi = kp.box_hi[k];
            if (x < lo) { x = lo + (lo - x); v = -v; ap = -ap; }
            else if (x > hi) { x = hi - (x - hi); v = -v; ap = -ap; }
        } else if (kp.bnd & SPHX_BND_PERIODIC) {
            const double lo = kp.box_lo[k], hi = kp.box_hi[k], L = hi - lo;
            if (x >= hi) x -= L;
            else if (x < lo) x += L;
        }
        xs[k][i] = x;
        vs[k][i] = v;
        aps[k][i] = ap;
    }
}
void launch_vv_position(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp,
                        hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_vv_position, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, kp);
}

// velocity_verlet_velocity_system (integrators.rs:92-104)
__global__ void k_vv_velocity(Particles P, Derived D, uint32_t n, KernelParams kp, int mixed, Mixed X) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double vx = P.vx[i] + (P.apx[i] + D.ax[i]) * kp.half_dt;
    const double vy = P.vy[i] + (P.apy[i] + D.ay[i]) * kp.half_dt;
    const double vz = P.vz[i] + (P.apz[i] + D.az[i]) * kp.half_dt;
    P.vx[i] = vx;
    P.vy[i] = vy;
    P.vz[i] = vz;
    if (mixed) {
        X.vx[i] = (float)vx;
        X.vy[i] = (float)vy;
        X.vz[i] = (float)vz;
    }
}
void launch_vv_velocity(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp,
                        bool mixed, const Mixed& X, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_vv_velocity, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, kp, (int)mixed, X);
}

// ---------------------------------------------------------------- XSPH (a16)
// sph_xsph_system (sph.rs:349-405); Jacobi: writes the new velocity to tmp3.
__global__ void __launch_bounds__(kBlock) k_xsph(Particles P, Derived D, uint32_t n, uint32_t cap, Grid g,
                                                  const uint32_t* __restrict__ idx,
                                                  const uint32_t* __restrict__ cnt, KernelParams kp,
                                                  int mixed, Mixed X) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i], inv_hi = 1.0 / hi;
    const double vxi = P.vx[i], vyi = P.vy[i], vzi = P.vz[i], rhoi = D.rho[i];
    double ux = 0, uy = 0, uz = 0;
    const uint32_t c = cnt[i];
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t j = idx[(size_t)k * cap + i] & kIdxMask;
        double dx, dy, dz;
        pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], dx, dy, dz);
        const double r = sqrt((dx * dx + dy * dy) + dz * dz);
        const double w = kernel_w(r, inv_h_pair(hi, P.h[j], inv_hi));
        const double rho_bar = 0.5 * (rhoi + D.rho[j]);
        const double f = (kp.xsph_eps * (P.m[j] / rho_bar)) * w;
        ux = fma(f, P.vx[j] - vxi, ux);
        uy = fma(f, P.vy[j] - vyi, uy);
        uz = fma(f, P.vz[j] - vzi, uz);
    }
    D.tmp3[0][i] = vxi + ux;
    D.tmp3[1][i] = vyi + uy;
    D.tmp3[2][i] = vzi + uz;
    if (mixed) {  // f32 copies are read by the next force pass, not by this one
        X.vx[i] = (float)(vxi + ux);
        X.vy[i] = (float)(vyi + uy);
        X.vz[i] = (float)(vzi + uz);
    }
}
void launch_xsph(const Particles& P, const Derived& D, uint32_t n, uint32_t cap, const Grid& g,
                 const NbrList& L, const KernelParams& kp, bool mixed, const Mixed& X, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_xsph, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, cap, g, L.idx, L.cnt, kp,
                           (int)mixed, X);
}

// ------------------------------------------------------ MHD induction (a11)
// mhd_induction_system (mhd.rs:68-152) in gather form.  The reference adds
// the pair term to entity_a and its negation to entity_b (mhd.rs:121-122,
// 130-131), so particle i's term carries +1 when id_i < id_j, -1 otherwise.
__global__ void __launch_bounds__(kBlock) k_induction(Particles P, Derived D, uint32_t n, uint32_t cap, Grid g,
                                                       const uint32_t* __restrict__ idx,
                                                       const uint32_t* __restrict__ cnt, KernelParams kp) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i], inv_hi = 1.0 / hi;
    const double vxi = P.vx[i], vyi = P.vy[i], vzi = P.vz[i], rhoi = D.rho[i];
    const double bxi = P.bx[i], byi = P.by[i], bzi = P.bz[i];
    const uint32_t idi = P.id[i];
    double ux = 0, uy = 0, uz = 0;
    const uint32_t c = cnt[i];
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t j = idx[(size_t)k * cap + i] & kIdxMask;
        double dx, dy, dz;
        pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], dx, dy, dz);
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        const double r = sqrt(r2);
        if (!(r > 1e-10)) continue;
        const double hj = P.h[j];
        const double inv_h = inv_h_pair(hi, hj, inv_hi);
        const double dwr = kernel_dw(r, inv_h) / r;
        const double gx = dx * dwr, gy = dy * dwr, gz = dz * dwr;  // r_hat dW/dr
        const double vx = vxi - P.vx[j], vy = vyi - P.vy[j], vz = vzi - P.vz[j];
        const double rhoj = D.rho[j], mj = P.m[j];
        const double v_dot_g = (vx * gx + vy * gy) + vz * gz;
        const double b_dot_g = (bxi * gx + byi * gy) + bzi * gz;
        const double f = mj / (rhoi * rhoj);
        const double h_avg = 0.5 * (hi + hj);
        const double r_dot_g = (dx * gx + dy * gy) + dz * gz;
        const double res = (2.0 * kp.eta * (mj / rhoj)) * r_dot_g / fma(0.01 * h_avg, h_avg, r2);
        const double sg = (idi < P.id[j]) ? kp.dt : -kp.dt;
        ux += sg * (f * (bxi * v_dot_g - vx * b_dot_g) + res * (P.bx[j] - bxi));
        uy += sg * (f * (byi * v_dot_g - vy * b_dot_g) + res * (P.by[j] - byi));
        uz += sg * (f * (bzi * v_dot_g - vz * b_dot_g) + res * (P.bz[j] - bzi));
    }
    D.tmp3[0][i] = bxi + ux;
    D.tmp3[1][i] = byi + uy;
    D.tmp3[2][i] = bzi + uz;
}
void launch_induction(const Particles& P, const Derived& D, uint32_t n, uint32_t cap, const Grid& g,
                      const NbrList& L, const KernelParams& kp, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_induction, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, cap, g, L.idx,
                           L.cnt, kp);
}

// ---------------------------------------------- divergence cleaning (a12)
// divergence_cleaning_system (mhd.rs:172-259): Jacobi over B and psi, then
// psi *= exp(-sigma c_h dt / h) for every particle.
__global__ void __launch_bounds__(kBlock) k_div_clean(Particles P, Derived D, uint32_t n, uint32_t cap, Grid g,
                                                       const uint32_t* __restrict__ idx,
                                                       const uint32_t* __restrict__ cnt, KernelParams kp) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i], inv_hi = 1.0 / hi;
    const double bxi = P.bx[i], byi = P.by[i], bzi = P.bz[i], psii = P.psi[i];
    double div = 0, gpx = 0, gpy = 0, gpz = 0;
    const uint32_t c = cnt[i];
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t j = idx[(size_t)k * cap + i] & kIdxMask;
        double dx, dy, dz;
        pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], dx, dy, dz);
        const double r = sqrt((dx * dx + dy * dy) + dz * dz);
        if (!(r > 1e-10)) continue;
        const double dwr = kernel_dw(r, inv_h_pair(hi, P.h[j], inv_hi)) / r;
        const double gx = dx * dwr, gy = dy * dwr, gz = dz * dwr;
        const double vol = P.m[j] / D.rho[j];
        div += vol * (((P.bx[j] - bxi) * gx + (P.by[j] - byi) * gy) + (P.bz[j] - bzi) * gz);
        const double dp = vol * (P.psi[j] - psii);
        gpx = fma(dp, gx, gpx);
        gpy = fma(dp, gy, gpy);
        gpz = fma(dp, gz, gpz);
    }
    const double dt = kp.dt;
    D.tmp3[0][i] = bxi - gpx * dt;
    D.tmp3[1][i] = byi - gpy * dt;
    D.tmp3[2][i] = bzi - gpz * dt;
    const double decay = exp(-kp.sigma_c * kp.ch * dt / hi);
    D.tmp1[i] = (psii - (kp.ch * kp.ch) * div * dt) * decay;
}
void launch_div_clean(const Particles& P, const Derived& D, uint32_t n, uint32_t cap, const Grid& g,
                      const NbrList& L, const KernelParams& kp, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_div_clean, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, n, cap, g, L.idx,
                           L.cnt, kp);
}

// ---------------------------------------------------------- pair export
__global__ void k_export_pairs(const uint32_t* __restrict__ id, uint32_t n, uint32_t cap,
                               const uint32_t* __restrict__ idx, const uint32_t* __restrict__ cnt,
                               uint32_t* __restrict__ out, unsigned long long* counter, uint64_t out_cap) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t idi = id[i];
    const uint32_t c = cnt[i];
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t idj = id[idx[(size_t)k * cap + i] & kIdxMask];
        if (idi < idj) {
            unsigned long long pos = atomicAdd(counter, 1ull);
            if (pos < out_cap) {
                out[2 * pos] = idi;
                out[2 * pos + 1] = idj;
            }
        }
    }
}
void launch_export_pairs(const Particles& P, uint32_t n, uint32_t cap, const NbrList& L, uint32_t* out,
                         unsigned long long* counter, uint64_t out_cap, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_export_pairs, dim3(nblocks(n)), dim3(kBlock), 0, s, P.id, n, cap, L.idx,
                           L.cnt, out, counter, out_cap);
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

__global__ void k_energy(Particles P, Derived D, uint32_t n, const uint32_t* __restrict__ cnt,
                         KernelParams kp, int mhd, double* __restrict__ partial) {
    __shared__ double lds[kEnergyVals][kBlock];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[kEnergyVals] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double m = P.m[i], vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
        v[0] = 0.5 * m * ((vx * vx + vy * vy) + vz * vz);
        v[1] = -m * ((kp.g[0] * P.x[i] + kp.g[1] * P.y[i]) + kp.g[2] * P.z[i]);
        v[2] = m * specific_internal(kp, D.rho[i]);
        if (mhd) {
            const double bx = P.bx[i], by = P.by[i], bz = P.bz[i];
            v[3] = 0.5 * (m / D.rho[i]) * ((bx * bx + by * by) + bz * bz);
        }
        v[4] = m * vx; v[5] = m * vy; v[6] = m * vz;
        v[7] = (double)cnt[i];
        v[8] = (double)cnt[i];
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
void launch_energy(const Particles& P, const Derived& D, uint32_t n, uint32_t cap, const NbrList& L,
                   const KernelParams& kp, bool mhd, double* partial, uint32_t nb, hipStream_t s) {
    (void)cap;
    hipLaunchKernelGGL(k_energy, dim3(nb), dim3(kBlock), 0, s, P, D, n, L.cnt, kp, (int)mhd, partial);
}

}  // namespace sphx
