// sphx_sort.hip — cell hashing, counting sort, reorder and tile descriptors.
//
// This replaces the reference's K-D tree broadphase (broadphase.rs:33-96):
// a uniform grid of cells >= 2*max(h) wide, a counting sort of particles by
// cell (histogram, exclusive scan, scatter), and a gather of every state
// array into the new order.  All integer work; HBM-bound.

#include <algorithm>

#include "sphx_device.h"

namespace sphx {

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

// key = linear cell index (x fastest); rank = arrival order inside the cell.
// Particles arrive in the previous step's sorted order, so neighbouring
// lanes mostly share a cell: each run of equal keys in a wave takes one
// atomic for the whole run.
__global__ void k_cell_hash(const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, uint32_t n, Grid g,
                            uint32_t* __restrict__ key, uint32_t* __restrict__ rank,
                            uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    uint32_t c = 0xffffffffu;
    if (valid) {
        const int cx = cell_coord(x[i], g.lo[0], g.inv_w[0], g.dims[0]);
        const int cy = cell_coord(y[i], g.lo[1], g.inv_w[1], g.dims[1]);
        const int cz = cell_coord(z[i], g.lo[2], g.inv_w[2], g.dims[2]);
        c = ((uint32_t)cz * (uint32_t)g.dims[1] + (uint32_t)cy) * (uint32_t)g.dims[0] + (uint32_t)cx;
        key[i] = c;
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t prev = (uint32_t)__shfl_up((int)c, 1);
    const bool head = lane == 0u || prev != c;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
    const uint32_t h = 63u - (uint32_t)__clzll(heads & upto);            // this lane's run head
    const unsigned long long after = h == 63u ? 0ull : (heads & ~((2ull << h) - 1ull));
    const uint32_t end = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;  // one past the run
    uint32_t base = 0;
    if (valid && lane == h) base = atomicAdd(&count[c], end - h);
    base = (uint32_t)__shfl((int)base, (int)h);
    if (valid) rank[i] = base + (lane - h);
}
void launch_cell_hash(const Particles& P, uint32_t n, const Grid& g, const Sort& S, hipStream_t s) {
    launch_fill_u32(S.cell_count, 0u, g.ncell, s);
    if (n)
        hipLaunchKernelGGL(k_cell_hash, dim3(nblocks(n)), dim3(kBlock), 0, s, P.x, P.y, P.z, n, g,
                           S.key, S.rank, S.cell_count);
}

// Exclusive scan of n u32 into out[0..n] (out[n] = total): per-1024 tile scans,
// one-block scan of the tile sums, add.
static constexpr int kScanItems = 4;
static constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
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

void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums, hipStream_t s) {
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles) hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kBlock), 0, s, in, out, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, sums, tiles, out + n);
    if (tiles) hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kBlock), 0, s, out, n, sums);
}

__global__ void k_scatter(const uint32_t* __restrict__ key, const uint32_t* __restrict__ rank,
                          const uint32_t* __restrict__ start, uint32_t n, uint32_t* __restrict__ perm) {
    // perm here is the arrival-order scatter; k_cell_rank makes it deterministic
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[start[key[i]] + rank[i]] = i;
}
void launch_scatter(const Sort& S, uint32_t n, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_scatter, dim3(nblocks(n)), dim3(kBlock), 0, s, S.key, S.rank,
                           S.cell_start, n, S.tmp);
}

// Atomic ranks arrive in any order.  Each particle's final place in its
// cell is the number of cell members before it by (x, previous index): a
// deterministic order, sorted by x so that the neighbour build can cut each
// row to the reach in x.  One thread per particle scans its own cell
// (members are L1-resident, the cell's threads read the same lines).
__global__ void k_cell_rank(const uint32_t* __restrict__ key, const uint32_t* __restrict__ start,
                            const uint32_t* __restrict__ arrival, const double* __restrict__ x, uint32_t n,
                            uint32_t* __restrict__ perm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = key[i];
    const uint32_t b = start[c], e = start[c + 1];
    const double xv = x[i];
    uint32_t r = 0;
    for (uint32_t m = b; m < e; ++m) {
        const uint32_t j = arrival[m];
        const double xj = x[j];
        r += (xj < xv || (xj == xv && j < i)) ? 1u : 0u;
    }
    perm[b + r] = i;
}
void launch_cell_order(const Sort& S, uint32_t n, const double* x, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_cell_rank, dim3(nblocks(n)), dim3(kBlock), 0, s, S.key, S.cell_start, S.tmp, x, n,
                           S.perm);
}

// Gather of all persistent state into the new order, plus the sorted cell
// key, the packed position record and the f32 copies (positions relative to
// the cell origin).
__global__ void k_reorder(Particles src, Particles dst, const uint32_t* __restrict__ perm,
                          const uint32_t* __restrict__ key, uint32_t* __restrict__ skey, uint32_t n,
                          int mhd, int charge, int mixed, int energy, Mixed X, Grid g,
                          double4* __restrict__ recb, double4* __restrict__ xref) {
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
    double2* pos = reinterpret_cast<double2*>(dst.pos + k);
    pos[0] = make_double2(x, y);
    pos[1] = make_double2(z, h);
    double2* ref = reinterpret_cast<double2*>(xref + k);
    ref[0] = make_double2(x, y);
    ref[1] = make_double2(z, 0.0);
    dst.id[k] = src.id[i];
    dst.orig[k] = src.orig[i];
    if (mhd) {
        const double bx = src.bx[i], by = src.by[i], bz = src.bz[i], psi = src.psi[i];
        dst.bx[k] = bx;
        dst.by[k] = by;
        dst.bz[k] = bz;
        dst.psi[k] = psi;
        double2* rb = reinterpret_cast<double2*>(recb + k);
        rb[0] = make_double2(bx, by);
        rb[1] = make_double2(bz, psi);
    }
    if (charge) dst.q[k] = src.q[i];
    if (energy) {
        dst.u[k] = src.u[i];
        dst.dup[k] = src.dup[i];
    }
    // f32 cell-relative coordinates and h: the neighbour build's prefilter in
    // both forms, and the pair-side data of the mixed form
    int cx, cy, cz;
    decode_cell(c, g, cx, cy, cz);
    X.p4[k] = make_float4((float)(x - (g.lo[0] + cx * g.w[0])), (float)(y - (g.lo[1] + cy * g.w[1])),
                          (float)(z - (g.lo[2] + cz * g.w[2])), (float)h);
    if (mixed) X.m[k] = (float)m;
}
void launch_reorder(const Particles& src, const Particles& dst, const Sort& S, uint32_t n, bool mhd,
                    bool charge, bool mixed, bool energy, const Mixed& X, const Grid& g, double4* recb,
                    double4* xref, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_reorder, dim3(nblocks(n)), dim3(kBlock), 0, s, src, dst, S.perm, S.key,
                           S.skey, n, (int)mhd, (int)charge, (int)mixed, (int)energy, X, g, recb, xref);
}

// ---- tiles: each row cut into runs of <= kTile sorted particles ---------
__global__ void k_row_tiles(const uint32_t* __restrict__ start, uint32_t nrows, uint32_t nx,
                            uint32_t* __restrict__ count) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const uint32_t a = start[r * nx], b = start[r * nx + nx];
    count[r] = (b - a + kTile - 1) / kTile;
}

__global__ void k_row_tiles_write(const uint32_t* __restrict__ start, uint32_t nrows, uint32_t nx,
                                  const uint32_t* __restrict__ off, uint32_t* __restrict__ p0,
                                  uint32_t* __restrict__ np, uint32_t* __restrict__ row) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const uint32_t a = start[r * nx], b = start[r * nx + nx];
    const uint32_t t0 = off[r], t1 = off[r + 1];
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t first = a + (t - t0) * kTile;
        p0[t] = first;
        np[t] = min((uint32_t)kTile, b - first);
        row[t] = r;
    }
}

void launch_tiles(const Grid& g, const Sort& S, const Tiles& T, hipStream_t s) {
    const uint32_t nx = (uint32_t)g.dims[0];
    hipLaunchKernelGGL(k_row_tiles, dim3(nblocks(T.nrows)), dim3(kBlock), 0, s, S.cell_start, T.nrows, nx,
                       T.row_count);
    launch_scan(T.row_count, T.row_off, T.nrows, T.scan_sums, s);
    hipLaunchKernelGGL(k_row_tiles_write, dim3(nblocks(T.nrows)), dim3(kBlock), 0, s, S.cell_start, T.nrows,
                       nx, T.row_off, T.p0, T.np, T.row);
}

// ---- bricks: the ids of the bricks that hold particles, in brick order ----
__global__ void k_brick_flag(const uint32_t* __restrict__ start, Grid g, Bricks B) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B.count) return;
    const uint32_t nbx = (uint32_t)B.nb[0], nby = (uint32_t)B.nb[1];
    const int x0 = (int)(b % nbx) * B.b[0], y0 = (int)((b / nbx) % nby) * B.b[1], z0 = (int)(b / (nbx * nby)) * B.b[2];
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    const int x1 = min(x0 + B.b[0], nx);
    uint32_t any = 0;
    for (int z = z0; z < min(z0 + B.b[2], nz); ++z)
        for (int y = y0; y < min(y0 + B.b[1], ny); ++y) {
            const uint32_t row = ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)nx;
            any |= start[row + x1] - start[row + x0];
        }
    B.flag[b] = any ? 1u : 0u;
}

__global__ void k_brick_compact(Bricks B) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B.count && B.flag[b]) B.active[B.pos[b]] = b;
}

void launch_bricks(const Grid& g, const Sort& S, const Bricks& B, hipStream_t s) {
    if (!B.on) return;
    hipLaunchKernelGGL(k_brick_flag, dim3(nblocks(B.count)), dim3(kBlock), 0, s, S.cell_start, g, B);
    launch_scan(B.flag, B.pos, B.count, B.sums, s);
    hipLaunchKernelGGL(k_brick_compact, dim3(nblocks(B.count)), dim3(kBlock), 0, s, B);
}

// Upper bound of the tile count: sum over rows of ceil(n_r / kTile).
// Blocks of a launch over the tiles: a bound on their count (a row of m
// particles has at most m / kTile + 1 tiles) plus the 8 that the XCD dealing
// of xcd_block_of needs; blocks past the device-side count return at once.
uint32_t tile_grid(uint32_t n, const Tiles& T) {
    return (uint32_t)((uint64_t)n / kTile + std::min<uint64_t>(T.nrows, n) + 8);
}

}  // namespace sphx
