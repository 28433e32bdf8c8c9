// sphx_device.h — device helpers shared by the kernel translation units:
// SPH kernel functions, pair vectors, the EOS, the tile/stage framework of
// the neighbour build and the list walk of the neighbour passes.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include <type_traits>

#include "sphx_internal.h"

namespace sphx {

static constexpr int kBlock = kTile;            // one thread per particle
static constexpr int kStagedBlock = 512;        // workgroup of the brick-staged passes
static constexpr int kStageCap = 384;           // candidates staged in LDS per chunk
static constexpr double kInvPi = 0.31830988618379067154;  // 1/pi
static constexpr double kPi = 3.141592653589793;           // std::f64::consts::PI


// Dynamic LDS that caps a gather pass at `blocks` workgroups per CU (160 KiB
// of LDS per CU; the LDS itself is not used).  The gather passes are bound
// by L1 misses of their neighbour loads: fewer waves per CU keep more of each
// wave's neighbour lines in L1 between consecutive list entries (measured on
// the 1M dam break: FP64 force pass 0.72 -> 0.62 ms at 3 workgroups per CU).
static constexpr size_t kLdsPerCU = 160 * 1024;
static inline size_t occupancy_lds(int blocks) { return (kLdsPerCU / (size_t)blocks) & ~(size_t)255; }

// blockIdx -> logical block so that each XCD (blocks are dealt round-robin to
// the 8 XCDs) works on one contiguous range of tiles / particles and its L2
// holds that range's neighbours.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t xcd = b & 7u, local = b >> 3;
    const uint32_t base = nb >> 3, rem = nb & 7u;
    return xcd * base + min(xcd, rem) + local;
}

// The same dealing for a grid launched larger than the count of work items
// (count known only on the device): each XCD gets an equal contiguous share
// of the count items; blocks left over return 0xffffffff.  The grid must hold
// at least count + 8 blocks.
__device__ __forceinline__ uint32_t xcd_block_of(uint32_t count) {
    const uint32_t b = blockIdx.x;
    const uint32_t xcd = b & 7u, local = b >> 3;
    const uint32_t base = count >> 3, rem = count & 7u;
    if (local >= base + (xcd < rem ? 1u : 0u)) return 0xffffffffu;
    return xcd * base + min(xcd, rem) + local;
}

// key % nx for a cell key (< 2^27) and nx (< 2^16): (key + 1/2) / nx in f64
// lies at least 1/(2 nx) from an integer, far beyond the rounding of the
// product, so its floor is the exact quotient (a few instructions instead of
// an integer division).
__device__ __forceinline__ uint32_t cell_x_of(uint32_t key, uint32_t nx, double inv_nx) {
    const uint32_t q = (uint32_t)(((double)key + 0.5) * inv_nx);
    return key - q * nx;
}

__device__ __forceinline__ void min_image(double& d, double L, double halfL) {
    if (d > halfL) d -= L;
    else if (d < -halfL) d += L;
}

// The Maxwell-stress record of a particle: b = B / rho and |b|^2 / 2, so that
// the stress term of a pair, (B (B.g) - |B|^2 g / 2) / rho^2 for either end,
// is b (b.g) - (|b|^2 / 2) g with no reciprocal per pair.  Every writer
// (density and cleaning epilogues, k_stress_rec) forms it here, so the bits do
// not depend on which one wrote it.
__device__ __forceinline__ double4 stress_record(double bx, double by, double bz, double rho) {
    const double ir = 1.0 / rho;
    const double sx = bx * ir, sy = by * ir, sz = bz * ir;
    return make_double4(sx, sy, sz, 0.5 * ((sx * sx + sy * sy) + sz * sz));
}

// The minimum image of a neighbour-list pair, branch-free: d - k L with
// k = rint(d / L), one rounding for k = +-1 as in d -= L.  A periodic axis
// holds at least 3 cells no narrower than a list pair's reach, so a list pair
// has |d| < L / 3 or |d| > 2 L / 3 (positions stay in the box), where k is
// exactly the shift min_image applies: the same bits.
__device__ __forceinline__ void min_image_list(double& d, double L, double invL) {
    d = fma(-rint(d * invL), L, d);
}

// x_j - x_i as core/coordinates.rs:51-65 forms it in one sector (0 + (b - a)),
// with the minimum image in a periodic box.
// PER: 1 periodic (neighbour-list pairs only: min_image_list), 0 closed box,
// -1 read from the grid (a compile-time PER keeps a pair body free of the
// branch).
template <int PER = -1>
__device__ __forceinline__ void pair_vec(const Grid& g, double xi, double yi, double zi, double xj,
                                         double yj, double zj, double& dx, double& dy, double& dz) {
    dx = xj - xi;
    dy = yj - yi;
    dz = zj - zi;
    if (PER == 1) {
        min_image_list(dx, g.L[0], g.invL[0]);
        min_image_list(dy, g.L[1], g.invL[1]);
        min_image_list(dz, g.L[2], g.invL[2]);
    } else if (PER < 0 && g.periodic) {
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

__device__ __forceinline__ void decode_cell(uint32_t key, const Grid& g, int& cx, int& cy, int& cz) {
    const uint32_t nx = (uint32_t)g.dims[0], ny = (uint32_t)g.dims[1];
    cx = (int)(key % nx);
    const uint32_t t = key / nx;
    cy = (int)(t % ny);
    cz = (int)(t / ny);
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

// ---- cubic spline (sph.rs:29-64), f64 given 1/h --------------------------
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

// d^2 W / dr^2 (for the adjoint of the force pass)
__device__ __forceinline__ double kernel_d2w(double r, double inv_h) {
    const double q = r * inv_h;
    const double ih2 = inv_h * inv_h;
    const double sigma = (ih2 * ih2) * inv_h * kInvPi;
    if (q < 1.0) return sigma * fma(4.5, q, -3.0);
    if (q < 2.0) return sigma * (1.5 * (2.0 - q));
    return 0.0;
}

// ---- f64 arithmetic for the pair passes ----------------------------------
// The passes match the reference to a relative tolerance (the neighbour
// predicate, which must round exactly, does not use these).  1/x: hardware
// estimate and two Newton steps (within ~1 ulp).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
// sqrt(x) and 1/sqrt(x) for normal x > 0: Goldschmidt iterations on the
// hardware rsqrt estimate (two steps take its ~2^-23 error below 1 ulp).
__device__ __forceinline__ void sqrt_rsqrt(double x, double& sq, double& rs) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    const double d = fma(-g, g, x);
    sq = fma(d, h, g);
    rs = 2.0 * h;
}
// One Goldschmidt step: sqrt(x) and 1/sqrt(x) to about 2^-44 relative, for
// the force passes (their results are compared relative to the largest
// acceleration, 1e-11).
__device__ __forceinline__ void sqrt_rsqrt1(double x, double& sq, double& rs) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double e = fma(-h, g, 0.5);
    sq = fma(g, e, g);
    rs = 2.0 * fma(h, e, h);
}
// As sqrt_rsqrt1, with 1 / (2 sqrt(x)) in place of 1 / sqrt(x): one add
// fewer, and the factor 2 is exact wherever a caller folds it back in.
__device__ __forceinline__ void sqrt_rhalf1(double x, double& sq, double& rh) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double e = fma(-h, g, 0.5);
    sq = fma(g, e, g);
    rh = fma(h, e, h);
}
// One Newton step on the hardware reciprocal: about 2^-46 relative.
__device__ __forceinline__ double rcp_nr1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}
// Cubic-spline shape functions without the normalisation: W = sigma_w * w_shape(q)
// with sigma_w = 1/(pi h^3); dW/dr = sigma_g * dw_shape(q) with sigma_g = 1/(pi h^4).
// For q < 1, 1 - 1.5 q^2 + 0.75 q^3 = 1 - 0.75 q^2 (2 - q), formed from t = 2 - q
// (no constant to load into a register per pair).
__device__ __forceinline__ double w_shape(double q) {
    const double t = fmax(2.0 - q, 0.0);
    const double a = fma((q * q) * t, -0.75, 1.0);
    const double b = 0.25 * ((t * t) * t);
    return q < 1.0 ? a : b;
}
// 4 w_shape(q), one multiply fewer; the scaling by 4 is exact, so a sum of
// these times 0.25 is bitwise the sum of w_shape
__device__ __forceinline__ double w4_shape(double q) {
    const double t = fmax(2.0 - q, 0.0);
    const double a = fma((q * q) * t, -3.0, 4.0);
    return q < 1.0 ? a : (t * t) * t;
}
__device__ __forceinline__ double dw_shape(double q) {
    const double t = fmax(2.0 - q, 0.0);
    const double a = q * fma(2.25, q, -3.0);
    const double b = -0.75 * (t * t);
    return q < 1.0 ? a : b;
}
// The same shapes from the two one-sided pieces s = max(1 - q, 0) and
// t = max(1 - q/2, 0), formed from r and nih = -1/h, nih2 = -1/(2h) (no q, no
// select): for q < 1, (2 - q)^3 / 4 - (1 - q)^3 = 1 - 1.5 q^2 + 0.75 q^3 and
// 3 (1 - q)^2 - 0.75 (2 - q)^2 = -3 q + 2.25 q^2; for 1 <= q < 2 only the
// t piece is left, (2 - q)^3 / 4 and -0.75 (2 - q)^2; both are 0 from q = 2 on.
// w_shape(q) / 2 = t^3 - s^3 / 2
__device__ __forceinline__ double wh_shape_r(double r, double nih, double nih2) {
    const double s = fmax(fma(r, nih, 1.0), 0.0);
    const double t = fmax(fma(r, nih2, 1.0), 0.0);
    return fma((s * s) * s, -0.5, (t * t) * t);
}
// dw_shape(q) / 3 = s^2 - t^2 (0 at r = 0 exactly: s = t = 1)
__device__ __forceinline__ double dw3_shape_r(double r, double nih, double nih2) {
    const double s = fmax(fma(r, nih, 1.0), 0.0);
    const double t = fmax(fma(r, nih2, 1.0), 0.0);
    return fma(-t, t, s * s);
}
// dw_shape(q) / 0.75 (the negation of t^2 is free)
__device__ __forceinline__ float dw43_shape_f(float q) {
    const float t = fmaxf(2.0f - q, 0.0f);
    const float a = q * fmaf(3.0f, q, -4.0f);
    return q < 1.0f ? a : -(t * t);
}

// ---- equations of state (sph.rs:157-167; thermodynamics.rs:232-238) -----
// u: the particle's specific internal energy (read by the adiabatic EOS only)
__device__ __forceinline__ double eos_pressure(const KernelParams& kp, double rho, double u = 0.0) {
    if (kp.eos == SPHX_EOS_TAIT) {
        double P = kp.tait_B * (pow(rho / kp.rest_density, kp.gamma) - 1.0);
        return P < 0.0 ? 0.0 : P;
    }
    if (kp.eos == SPHX_EOS_POLYTROPIC) return kp.poly_k * pow(rho, kp.gamma);
    if (kp.eos == SPHX_EOS_ADIABATIC) return (kp.gamma - 1.0) * rho * u;
    return rho * kp.iso_cs2;
}
// dP/drho (0 where the Tait clamp is active)
__device__ __forceinline__ double eos_dpdrho(const KernelParams& kp, double rho) {
    if (kp.eos == SPHX_EOS_TAIT) {
        const double ratio = rho / kp.rest_density;
        if (kp.tait_B * (pow(ratio, kp.gamma) - 1.0) < 0.0) return 0.0;
        return kp.tait_B * kp.gamma * pow(ratio, kp.gamma - 1.0) / kp.rest_density;
    }
    if (kp.eos == SPHX_EOS_POLYTROPIC) return kp.poly_k * kp.gamma * pow(rho, kp.gamma - 1.0);
    return kp.iso_cs2;
}

// v of particle i in the gather records (the f32 one in the mixed form, the
// f64 one whenever it is kept current)
__device__ __forceinline__ void set_record_velocity(const Derived& D, const KernelParams& kp, uint32_t i, bool mixed,
                                                    double vx, double vy, double vz) {
    if (mixed) {
        D.r32[i].vx = (float)vx;
        D.r32[i].vy = (float)vy;
        D.r32[i].vz = (float)vz;
    }
    if (kp.rec64) {  // (vx, vy) as one 16-byte store (offset 32 of the 64-byte record)
        *reinterpret_cast<double2*>(&D.r64[i].vx) = make_double2(vx, vy);
        D.r64[i].vz = vz;
    }
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}


// =========================================================================
// Tile / stage framework.
//
// A tile is <= kTile consecutive sorted particles of one cell row (fixed
// cell y, z) whose cells span [cx_lo, cx_hi].  Stage s = (oy, oz) in
// {-1,0,1}^2 covers the neighbour row (cy + oy, cz + oz) over cells
// [cx_lo - 1, cx_hi + 1]: one contiguous slice of the sorted arrays, or two
// when a periodic row wraps.  Its particles are staged into LDS in chunks of
// kStageCap; "slot" = position in that staged sequence.  Every value here
// is uniform across the workgroup.
// =========================================================================
struct TileCtx {
    uint32_t p0, np;
    int cy, cz, cx_lo, cx_hi;
};

struct Stage {
    int nseg;
    int c00, c01;               // physical cell x range of segment 0 (inclusive)
    int c10, c11;               // ... of segment 1
    uint32_t ps0, ps1;          // first particle of each segment
    uint32_t off1;              // slot of segment 1's first particle (segment 0 starts at 0)
    uint32_t total;             // staged candidates
    uint32_t rowbase;           // linear index of cell (0, y, z)
    float oy, oz;               // stage cell offsets times the cell width (mixed form)
    // logical cell x of a staged particle = physical + lshift of its segment:
    // the stage's cells L0 .. L1 around the tile, unwrapped across a
    // periodic boundary, ascend with the slots; lorigin = the first one
    int lshift0, lshift1, lorigin;
};

__device__ __forceinline__ uint32_t tile_count(const Tiles& T) { return T.row_off[T.nrows]; }

__device__ __forceinline__ TileCtx load_tile(uint32_t b, const Tiles& T, const Grid& g,
                                             const uint32_t* __restrict__ skey) {
    TileCtx tc;
    tc.p0 = T.p0[b];
    tc.np = T.np[b];
    const uint32_t row = T.row[b];
    const uint32_t ny = (uint32_t)g.dims[1], nx = (uint32_t)g.dims[0];
    tc.cy = (int)(row % ny);
    tc.cz = (int)(row / ny);
    tc.cx_lo = (int)(skey[tc.p0] % nx);
    tc.cx_hi = (int)(skey[tc.p0 + tc.np - 1] % nx);
    return tc;
}

__device__ __forceinline__ void stage_segments(const Grid& g, const uint32_t* __restrict__ start,
                                               const TileCtx& tc, int s, Stage& st) {
    const int oy = s % 3 - 1, oz = s / 3 - 1;
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    st.nseg = 0;
    st.total = 0;
    st.off1 = 0;
    st.ps0 = st.ps1 = 0;
    st.c00 = st.c01 = st.c10 = st.c11 = 0;
    st.lshift0 = st.lshift1 = st.lorigin = 0;
    st.oy = (float)(oy * g.w[1]);
    st.oz = (float)(oz * g.w[2]);
    st.rowbase = 0;
    int y = tc.cy + oy, z = tc.cz + oz;
    if (g.periodic) {
        y = (y + ny) % ny;
        z = (z + nz) % nz;
    } else if (y < 0 || y >= ny || z < 0 || z >= nz) {
        return;
    }
    st.rowbase = ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)nx;
    const int L0 = tc.cx_lo - 1, L1 = tc.cx_hi + 1;
    if (!g.periodic) {
        st.nseg = 1; st.c00 = max(L0, 0); st.c01 = min(L1, nx - 1);
    } else if (L1 - L0 + 1 >= nx) {
        st.nseg = 1; st.c00 = 0; st.c01 = nx - 1;
    } else if (L0 < 0) {
        st.nseg = 2; st.c00 = L0 + nx; st.c01 = nx - 1; st.c10 = 0; st.c11 = L1;
        st.lshift0 = -nx;
    } else if (L1 > nx - 1) {
        st.nseg = 2; st.c00 = L0; st.c01 = nx - 1; st.c10 = 0; st.c11 = L1 - nx;
        st.lshift1 = nx;
    } else {
        st.nseg = 1; st.c00 = L0; st.c01 = L1;
    }
    st.lorigin = st.c00 + st.lshift0;
    st.ps0 = start[st.rowbase + st.c00];
    st.total = start[st.rowbase + st.c01 + 1] - st.ps0;
    if (st.nseg == 2) {
        st.ps1 = start[st.rowbase + st.c10];
        st.off1 = st.total;
        st.total += start[st.rowbase + st.c11 + 1] - st.ps1;
    }
}

// slot -> sorted particle index
__device__ __forceinline__ uint32_t slot_particle(const Stage& st, uint32_t slot) {
    return (st.nseg == 2 && slot >= st.off1) ? st.ps1 + (slot - st.off1) : st.ps0 + slot;
}

// Slots of the cell x = cx + ox of a stage, as [a, b); empty if the cell is
// outside a closed box.
__device__ __forceinline__ void cell_slots(const Grid& g, const uint32_t* __restrict__ start, const Stage& st,
                                           int cx, int ox, uint32_t& a, uint32_t& b) {
    a = b = 0;
    if (st.nseg == 0) return;
    const int nx = g.dims[0];
    int c = cx + ox;
    if (g.periodic) c = c < 0 ? c + nx : (c >= nx ? c - nx : c);
    else if (c < 0 || c >= nx) return;
    const bool second = st.nseg == 2 && c >= st.c10 && c <= st.c11;
    const uint32_t base = second ? st.off1 - st.ps1 : 0u - st.ps0;  // slot = particle + base (mod 2^32)
    a = start[st.rowbase + c] + base;
    b = start[st.rowbase + c + 1] + base;
}

// =========================================================================
// Neighbour lists (built per tile, walked per particle).
// =========================================================================
// Entry k of sorted particle i (layout in NbrList): group k / 4 of the
// particle's list at nbr_base + (k / 4) * 256, entry k % 4 within it.
__device__ __forceinline__ size_t nbr_index(const NbrList& L, uint32_t i) {
    return ((((size_t)(i >> 6) * (uint32_t)(L.ks >> 2)) << 6) + (i & 63u)) * 4u;
}
__device__ __forceinline__ uint32_t* nbr_base(const NbrList& L, uint32_t i) { return L.idx + nbr_index(L, i); }
__device__ __forceinline__ size_t nbr_off(uint32_t k) { return ((size_t)(k >> 2) << 8) + (k & 3u); }
// the four entries of the group holding entry k (k a multiple of 4)
__device__ __forceinline__ uint4 nbr_group(const uint32_t* __restrict__ e, uint32_t k) {
    return *reinterpret_cast<const uint4*>(e + ((size_t)(k >> 2) << 8));
}

// f(e) for every entry of sorted particle i (entries kbeg .. kend - 1, kbeg a
// multiple of 4), a group of four per iteration (one 16-byte load) so that
// the bodies overlap each other's gathers.
template <typename F>
__device__ __forceinline__ void for_neighbours(const NbrList& L, uint32_t i, F&& f, uint32_t kbeg = 0,
                                               uint32_t kend = 0xffffffffu) {
    const uint32_t* __restrict__ e = nbr_base(L, i);
    const uint32_t cnt = min((uint32_t)L.ntot[i], kend);
    uint32_t k = kbeg;
    for (; k + 3 < cnt; k += 4) {
        const uint4 q = nbr_group(e, k);
        f(q.x);
        f(q.y);
        f(q.z);
        f(q.w);
    }
    if (k < cnt) {
        const uint4 q = nbr_group(e, k);
        f(q.x);
        if (k + 1 < cnt) f(q.y);
        if (k + 2 < cnt) f(q.z);
    }
}

// f(e) for every entry, and flush() after each group of four consecutive
// entries (0-3, 4-7, ...) and after the last one: the FP32-mixed passes sum a
// group in f32 and add it to their f64 sums at the flush.
// Entries kbeg .. kend - 1 only (kbeg a multiple of 4).
template <typename F, typename G>
__device__ __forceinline__ void for_neighbour_groups(const NbrList& L, uint32_t i, F&& f, G&& flush,
                                                     uint32_t kbeg = 0, uint32_t kend = 0xffffffffu) {
    const uint32_t* __restrict__ e = nbr_base(L, i);
    const uint32_t cnt = min((uint32_t)L.ntot[i], kend);
    uint32_t k = kbeg;
    for (; k + 3 < cnt; k += 4) {
        const uint4 q = nbr_group(e, k);
        f(q.x);
        f(q.y);
        f(q.z);
        f(q.w);
        flush();
    }
    if (k < cnt) {
        const uint4 q = nbr_group(e, k);
        f(q.x);
        if (k + 1 < cnt) f(q.y);
        if (k + 2 < cnt) f(q.z);
        flush();
    }
}

struct NoFlush {
    __device__ __forceinline__ void operator()() const {}
};

// Sorted particle i's exact neighbour count r < 2*h_avg (its list may also
// hold pairs out to 2*h_avg + skin).
__device__ __forceinline__ uint32_t exact_count(const Particles& P, const Grid& g, const NbrList& L, uint32_t i) {
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i];
    uint32_t c = 0;
    for_neighbours(L, i, [&](uint32_t e) {
        const uint32_t j = e & kIdxMask;
        double dx, dy, dz;
        pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], dx, dy, dz);
        c += exact_neighbour(dx, dy, dz, hi, P.h[j]) ? 1u : 0u;
    });
    return c;
}

// Cell offset of an entry's neighbour times the cell widths (FP32-mixed pair
// vectors are (xr_j - xr_i) + offset).  code = stage * 3 + ox + 1 < 27 and
// stage = (oz + 1) * 3 + oy + 1; (c * 11) >> 5 == c / 3 for c < 27.
__device__ __forceinline__ void entry_offset(uint32_t e, float wx, float wy, float wz, float& ox, float& oy,
                                             float& oz) {
    const uint32_t c = e >> kIdxBits;
    const uint32_t s = (c * 11u) >> 5;
    const uint32_t sz = (s * 11u) >> 5;
    ox = (float)((int)(c - 3u * s) - 1) * wx;
    oy = (float)((int)(s - 3u * sz) - 1) * wy;
    oz = (float)((int)sz - 1) * wz;
}

// Sorted particle index of the calling thread in a one-thread-per-particle pass.
__device__ __forceinline__ uint32_t particle_index() { return xcd_block() * kBlock + threadIdx.x; }

// Two lanes per particle: lanes l and l + 32 of a wave share one particle,
// lane l < 32 walking the first half of its list and lane l + 32 the second,
// so that each half-wave walks 32 consecutive particles (their list groups
// are contiguous).  pair_slot(t): the particle slot of thread t (t / 2 of
// 32 slots per wave); lane_half(): 0 or 1; other_half(v): the value of the
// lane that shares the particle.
__device__ __forceinline__ uint32_t pair_slot(uint32_t t) { return ((t >> 6) << 5) | (t & 31u); }
__device__ __forceinline__ uint32_t lane_half() { return (threadIdx.x >> 5) & 1u; }
__device__ __forceinline__ double other_half(double v) { return __shfl_xor(v, 32); }

// The entry code of a particle's own cell: stage 4 (oy = oz = 0), x offset 0.
static constexpr uint32_t kOwnCode = 13u << kIdxBits;

// Gather record of list entry e: NQ 16-byte quads of type Q at record j of
// glob, then (NQ2 > 0) NQ2 quads at record j of glob2.
template <class Q, int NQ, int NQ2 = 0>
struct GlobalFetch {
    static constexpr int NR = NQ + NQ2;  // quads per record: NQ from glob, then NQ2 from glob2
    static constexpr bool BREL = false;  // positions cell-relative: pair vectors add the entry's cell offset
    const Q* __restrict__ glob;
    const Q* __restrict__ glob2 = nullptr;
    uint32_t kbeg = 0, kend = 0xffffffffu;  // the walk's range of list entries
    __device__ __forceinline__ void offset(uint32_t e, float wx, float wy, float wz, float& ox, float& oy,
                                          float& oz) const {
        entry_offset(e, wx, wy, wz, ox, oy, oz);
    }
    __device__ __forceinline__ void operator()(uint32_t e, Q (&q)[NR]) const {
        const size_t j = e & kIdxMask;
        const Q* p = glob + j * NQ;
#pragma unroll
        for (int k = 0; k < NQ; ++k) q[k] = p[k];
        if constexpr (NQ2 > 0) {
            const Q* p2 = glob2 + j * NQ2;
#pragma unroll
            for (int k = 0; k < NQ2; ++k) q[NQ + k] = p2[k];
        }
    }
    // the record of the walking particle i itself
    __device__ __forceinline__ void own(uint32_t i, Q (&q)[NR]) const { (*this)(i | kOwnCode, q); }
    // the length of particle i's list
    __device__ __forceinline__ uint32_t count(const NbrList& L, uint32_t i) const { return L.ntot[i]; }
};

// body(e, record) for every list entry of sorted particle i, records from
// global memory.
template <class Q, int NQ, int NQ2, class Body, class Flush = NoFlush>
__device__ __forceinline__ void walk(const NbrList& L, uint32_t i, const GlobalFetch<Q, NQ, NQ2>& F, Body&& body,
                                     Flush&& flush = Flush()) {
    auto one = [&](uint32_t e) {
        Q q[NQ + NQ2];
        F(e, q);
        body(e, q);
    };
    if (std::is_same<typename std::decay<Flush>::type, NoFlush>::value)
        for_neighbours(L, i, one, F.kbeg, F.kend);
    else
        for_neighbour_groups(L, i, one, flush, F.kbeg, F.kend);
}

// body(e, record) for the entries kb .. ke - 1 of sorted particle i (kb a
// multiple of 4), records from global memory, W records loaded before the W
// bodies run (their gathers in flight together); flush() after each group of
// four.  Entries past the list's end in the last group are replaced by
// particle i itself in its own cell (r = 0), so the bodies must contribute
// nothing at r = 0.
template <int W, class Q, int NQ, int NQ2, class Body, class Flush = NoFlush>
__device__ __forceinline__ void walk_ahead(const NbrList& L, uint32_t i, uint32_t kb, uint32_t ke,
                                           const GlobalFetch<Q, NQ, NQ2>& F, Body&& body, Flush&& flush = Flush()) {
    static_assert(4 % W == 0, "W divides the group of four");
    constexpr uint32_t kOwnCell = kOwnCode;
    const uint32_t* __restrict__ ep = nbr_base(L, i);
    const uint32_t cnt = min((uint32_t)L.ntot[i], ke);
    for (uint32_t k = kb; k < cnt; k += 4) {
        const uint4 grp = nbr_group(ep, k);
        const uint32_t ent[4] = {grp.x, grp.y, grp.z, grp.w};
#pragma unroll
        for (int u0 = 0; u0 < 4; u0 += W) {
            uint32_t e[W];
            Q q[W][NQ + NQ2];
#pragma unroll
            for (int w = 0; w < W; ++w) {
                e[w] = k + (uint32_t)(u0 + w) < cnt ? ent[u0 + w] : (i | kOwnCell);
                F(e[w], q[w]);
            }
#pragma unroll
            for (int w = 0; w < W; ++w) body(e[w], q[w]);
        }
        flush();
    }
}

// =========================================================================
// Brick-staged passes (the Bricks decomposition).
// =========================================================================
struct BrickHdr {
    uint4 halo[kMaxHaloRows];           // per halo row: {first sorted index, staged count, LDS record base, 0}
    uint4 code[27];                     // per entry cell code: {ox, oy, oz (f32 bits, times the cell widths), 0}
    uint32_t own_start[kMaxOwnRows];    // per brick row: its first sorted particle
    uint32_t own_cum[kMaxOwnRows + 1];  // exclusive prefix of the brick rows' particle counts
    uint32_t whole;                     // 1: every halo row is staged whole (no list entry is read from global)
};

// Halo row `lane` (< (b[1] + 2) (b[2] + 2)) of brick `id`: cells x0 - 1 ..
// x0 + b[0] of row y0 - 1 + lane % (b[1] + 2), z0 - 1 + lane / (b[1] + 2),
// wrapped in a periodic box, clamped in a closed one (a row outside it is
// empty).  gs: its first sorted particle, gl: its particles; cut: a periodic
// row that wraps in x (only its unwrapped part is staged); whole_x: the row
// holds the whole period in x (a periodic box at most b[0] + 2 cells wide).
// Called by one wave per brick, lane = row; the slot base of a row is the
// exclusive prefix of gl over the rows (brick_row_base).
__device__ __forceinline__ void brick_row(uint32_t id, const Bricks& B, const Grid& g,
                                          const uint32_t* __restrict__ start, int lane, uint32_t& gs,
                                          uint32_t& gl, bool& cut, bool* whole_x = nullptr) {
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    const int BX = B.b[0], BY = B.b[1], BZ = B.b[2];
    const uint32_t nbx = (uint32_t)B.nb[0], nby = (uint32_t)B.nb[1];
    const int x0 = (int)(id % nbx) * BX, y0 = (int)((id / nbx) % nby) * BY, z0 = (int)(id / (nbx * nby)) * BZ;
    const int hy = BY + 2, nh = hy * (BZ + 2);
    gs = 0;
    gl = 0;
    cut = false;
    if (lane >= nh) return;
    int y = y0 - 1 + lane % hy, z = z0 - 1 + lane / hy;
    bool ok = true;
    if (g.periodic) {
        y = (y % ny + ny) % ny;
        z = (z % nz + nz) % nz;
    } else {
        ok = y >= 0 && y < ny && z >= 0 && z < nz;
    }
    if (!ok) return;
    int xa = x0 - 1, xe = x0 + BX;
    const bool wx = g.periodic && xe - xa + 1 >= nx;  // the row holds the whole period in x
    if (wx) { xa = 0; xe = nx - 1; }
    cut = g.periodic && (xa < 0 || xe > nx - 1);
    xa = max(xa, 0);
    xe = min(xe, nx - 1);
    const uint32_t row = ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)nx;
    gs = start[row + (uint32_t)xa];
    gl = start[row + (uint32_t)xe + 1u] - gs;
    if (whole_x) *whole_x = wx;
}
// exclusive prefix of gl over the lanes of the wave: the LDS slot of the
// row's first record
__device__ __forceinline__ uint32_t brick_row_base(uint32_t gl) {
    const int lane = (int)(threadIdx.x & 63u);
    uint32_t incl = gl;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
    }
    return incl - gl;
}

// Brick `id`'s halo rows (brick_row; the first `cap` records are staged),
// its own rows and the table of entry codes.  Needs blockDim.x >= 155.
// BREL (records staged brick-relative): a row that holds the whole period in
// x is not staged, since its cells would need two images.
__device__ __forceinline__ void brick_setup(uint32_t id, const Bricks& B, const Grid& g,
                                            const uint32_t* __restrict__ start, uint32_t cap, BrickHdr& H,
                                            bool brel = false) {
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    const int BX = B.b[0], BY = B.b[1], BZ = B.b[2];
    const uint32_t nbx = (uint32_t)B.nb[0], nby = (uint32_t)B.nb[1];
    const int x0 = (int)(id % nbx) * BX, y0 = (int)((id / nbx) % nby) * BY, z0 = (int)(id / (nbx * nby)) * BZ;
    const uint32_t t = threadIdx.x;
    const int lane = (int)(t & 63u);
    if (t < 64) {
        const int nh = (BY + 2) * (BZ + 2);
        uint32_t gs, gl;
        bool cut, wx;
        brick_row(id, B, g, start, lane, gs, gl, cut, &wx);
        const uint32_t base = brick_row_base(gl);
        const uint32_t staged = base >= cap || (brel && wx) ? 0u : min(gl, cap - base);
        if (lane < nh) H.halo[lane] = make_uint4(gs, staged, base, 0u);
        const bool whole = __all(staged == gl && !cut) != 0;
        if (lane == 0) H.whole = whole ? 1u : 0u;
    } else if (t >= 128 && t < 128 + 27) {
        // entry code c = stage * 3 + x offset + 1, stage = (oz + 1) * 3 + oy + 1
        const uint32_t c = t - 128u, st = c / 3u, sy = st % 3u, sz = st / 3u;
        const float ox = (float)((int)(c - 3u * st) - 1) * (float)g.w[0];
        const float oy = (float)((int)sy - 1) * (float)g.w[1], oz = (float)((int)sz - 1) * (float)g.w[2];
        H.code[c] = make_uint4(__float_as_uint(ox), __float_as_uint(oy), __float_as_uint(oz), 0u);
    }
    if (t >= 64 && t < 128) {
        const int no = BY * BZ;
        uint32_t s0 = 0, cnt = 0;
        if (lane < no) {
            const int y = y0 + lane % BY, z = z0 + lane / BY;
            if (y < ny && z < nz) {
                const uint32_t row = ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)nx;
                s0 = start[row + (uint32_t)x0];
                cnt = start[row + (uint32_t)min(x0 + BX, nx)] - s0;
            }
        }
        uint32_t incl = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        if (lane < no) {
            H.own_start[lane] = s0;
            H.own_cum[lane] = incl - cnt;
        }
        if (lane == no - 1) H.own_cum[no] = incl;
    }
    __syncthreads();
}

// LDS pointers carry their address space so that reads through them are
// ds_read (a generic pointer would turn them into flat loads).
#define SPHX_LDS __attribute__((address_space(3)))
// 16-byte LDS word, read as a builtin vector (HIP's vector classes cannot be
// copied out of an address-space-qualified object)
typedef uint32_t lds_quad_t __attribute__((ext_vector_type(4)));
template <class Q>
__device__ __forceinline__ Q lds_quad(const SPHX_LDS lds_quad_t* p) {
    const lds_quad_t v = *p;
    Q q;
    __builtin_memcpy(&q, &v, sizeof q);
    return q;
}
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return (uint32_t)(size_t)(const SPHX_LDS void*)p;
}

// Row-code table: one 16-byte entry per (brick row k, entry code c), at
// rc[k * 27 + c], for list entries e = j | c << kIdxBits of the particles of
// brick row k.  Entry {gs | c << kIdxBits, staged, adj, 0} of the halo row the
// code points to (first sorted index gs, `staged` records from LDS slot lb):
// e is staged when e - (gs | c << kIdxBits) = j - gs < staged, and quad 0 of
// its record is at LDS byte address e * 16 + adj = rec + (lb + j - gs) * 16
// (adj = rec + (lb - gs) * 16 - (c << 31), all modulo 2^32).
static constexpr int kRowCodes = 27;

// A cell-relative f32 position moved to the brick origin: x + k_x w_x etc.
__device__ __forceinline__ void brel_shift(float4& q, int kx, int ky, int kz, float wx, float wy, float wz) {
    q.x = q.x + (float)kx * wx;
    q.y = q.y + (float)ky * wy;
    q.z = q.z + (float)kz * wz;
}

// What the staging needs to store FP32-mixed positions brick-relative
// (BrickFetch BREL): halo row r holds cells y = ky0 + r % hy, z = kz0 + r / hy
// from the origin; a record's x cell comes from its sorted cell key.
struct BrelStage {
    const uint32_t* __restrict__ skey;
    int ox0, ky0, kz0, hy;             // origin cell x; cell offsets of halo row 0
    uint32_t nx;
    double inv_nx;
    float wx, wy, wz;
};

// Copy the staged records (NQ 16-byte quads each) into LDS, quad-major: quad
// k of staged record r at dst[k * cap + r], so that one quad of 64 random
// records spreads over all LDS banks (record-major, a 64-byte record stride
// would put quad k of every record in the same 16 of the 64 banks).  Wave w
// copies halo rows w, w + waves, ..., each one contiguous range, eight quads
// per lane in flight: the eight loads (past the row's end: its last quad
// again, stored again to its own slot) and then the eight stores, with no
// branch between them (a branch would let the compiler sink each load to
// its store).  BREL: quad 0 of each record (a cell-relative float4
// position) is stored brick-relative, its record's cell key loaded beside it.
template <int NQ, bool BREL = false>
__device__ __forceinline__ void brick_copy(const BrickHdr& H, int nh, const uint4* __restrict__ src,
                                           uint4* __restrict__ dst, uint32_t cap, const BrelStage& bs) {
    const SPHX_LDS BrickHdr* HL = (const SPHX_LDS BrickHdr*)&H;
    const int w = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    for (int r = w; r < nh; r += nw) {
        const uint32_t gs = HL->halo[r].x, nq = HL->halo[r].y * (uint32_t)NQ, base = HL->halo[r].z;
        const uint4* __restrict__ s = src + (size_t)gs * NQ;
        for (uint32_t q0 = lane; q0 < nq; q0 += 512u) {
            uint32_t qq[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) qq[u] = min(q0 + 64u * (uint32_t)u, nq - 1u);
            uint4 v[8];
            uint32_t key[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                v[u] = s[qq[u]];
                if constexpr (BREL) key[u] = bs.skey[gs + qq[u] / NQ];
            }
            if constexpr (BREL) {
                const int ky = bs.ky0 + r % bs.hy, kz = bs.kz0 + r / bs.hy;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    float4 f;
                    __builtin_memcpy(&f, &v[u], sizeof f);
                    const int kx = (int)cell_x_of(key[u], bs.nx, bs.inv_nx) - bs.ox0;
                    brel_shift(f, kx, ky, kz, bs.wx, bs.wy, bs.wz);
                    uint4 t;
                    __builtin_memcpy(&t, &f, sizeof t);
                    if (qq[u] % NQ == 0) v[u] = t;  // quad 0 only
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) dst[(qq[u] % NQ) * cap + base + qq[u] / NQ] = v[u];
        }
    }
}
template <int NQ, int NQ2 = 0, bool BREL = false>
__device__ __forceinline__ void brick_stage(const BrickHdr& H, int nh, const uint4* __restrict__ src,
                                            const uint4* __restrict__ src2, uint4* __restrict__ dst, uint32_t cap,
                                            uint4* __restrict__ rc, uint32_t no, uint32_t by, uint32_t hy,
                                            const BrelStage& bs) {
    for (uint32_t t = threadIdx.x; t < no * (uint32_t)kRowCodes; t += blockDim.x) {
        const uint32_t k = t / (uint32_t)kRowCodes, c = t - k * (uint32_t)kRowCodes;
        const uint32_t st = c / 3u, sy = st % 3u, sz = st / 3u;
        const uint4 h = H.halo[(k % by + sy) + hy * (k / by + sz)];
        const uint32_t adj = lds_addr(dst) + (h.z - h.x) * 16u - (c << 31);
        rc[t] = make_uint4(h.x | (c << kIdxBits), h.y, adj, 0u);
    }
    brick_copy<NQ, BREL>(H, nh, src, dst, cap, bs);
    if constexpr (NQ2 > 0) brick_copy<NQ2>(H, nh, src2, dst + (size_t)NQ * cap, cap, bs);
    __syncthreads();
}

// Record of list entry e of a particle in brick row k: the row-code table of
// row k (rc) gives its LDS address when it is staged.
// PIN_: hold the walk's list prefetch at the top of each step with a full
// scheduling barrier (always done in two-entry walks; in four-entry walks
// only where it measured faster).
// T64_: the pass uses only the first 8 bytes of a record's last quad, read
// from LDS as 8 bytes (its second half is left 0): fewer bytes per random
// gather, the LDS cost that bounds the staged passes after their VALU work.
// S16_: in a brick whose halo is staged whole the walk reads the particle's
// slot list (NbrList::s16) instead of its entries, and the body receives the
// slot in place of the entry: only for bodies that use neither the
// neighbour's sorted index nor its cell code.
// BREL_ (FP32-mixed records, float4 quad 0 = cell-relative x, y, z): every
// record the fetch returns holds its position relative to the brick's origin
// cell (x0 + b[0]/2, y0 + b[1]/2, z0 + b[2]/2): x + k_x w_x (f32) with k the
// cell's offset from the origin, converted once when staged (brick_copy) and
// on the fly when read from global memory (from the entry's cell code), the
// same f32 operations either way.  A pair vector is then x_j - x_i, with no
// cell offset.
template <class Q, int NQ, int W_ = 2, bool PF_ = false, int NQ2 = 0, bool PIN_ = false, bool T64_ = false,
          bool S16_ = false, bool BREL_ = false>
struct BrickFetch {
    static_assert(sizeof(Q) == 16, "records are made of 16-byte quads");
    static_assert(!BREL_ || std::is_same<Q, float4>::value, "brick-relative records are FP32-mixed records");
    static constexpr int W = W_;
    static constexpr bool PF = PF_;
    static constexpr bool PIN = PIN_ || W_ == 2;
    static constexpr bool T64 = T64_;
    static constexpr bool S16 = S16_;
    static constexpr bool BREL = BREL_;
    static constexpr int NR = NQ + NQ2;  // quads per record: NQ from glob, then NQ2 from glob2
    const SPHX_LDS lds_quad_t* rc;    // the row-code table of the particle's brick row
    const SPHX_LDS lds_quad_t* code;  // per entry code: the cell offset
    const Q* __restrict__ glob;
    const Q* __restrict__ glob2;
    uint32_t qstride;                 // bytes from quad k to quad k + 1 of a record (16 cap)
    uint32_t whole;                   // every list entry of the brick's particles is staged (uniform)
    uint32_t rec0 = 0;                // S16: LDS byte address of slot 0's quad 0
    const uint16_t* __restrict__ s16 = nullptr;  // S16: the particle's slot list (null: walk the entries)
    // S16, four-entry walks of a brick staged whole: the particle's slot list
    // to fill while the entries are walked (the first pass after a build)
    uint16_t* __restrict__ w16 = nullptr;
    // BREL: the walking particle's cell offset from the brick origin,
    // (kx + 128) | (ky + 128) << 8 | (kz + 128) << 16 (one register)
    uint32_t kxyz = 0;
    float wx = 0, wy = 0, wz = 0;     // BREL: cell widths (f32)
    uint32_t kbeg = 0, kend = 0xffffffffu;  // the walk's range of list entries (kbeg a multiple of 4)
    uint32_t n_list = 0;              // the length of the walking particle's list (loaded by brick_pass)
    __device__ __forceinline__ uint32_t count(const NbrList&, uint32_t) const { return n_list; }
    // LDS byte address of entry e's record; false if it is not staged
    __device__ __forceinline__ bool locate(uint32_t e, uint32_t& addr) const {
        const lds_quad_t h = rc[e >> kIdxBits];
        addr = e * 16u + h.z;
        return e - h.x < h.y;
    }
    // LDS byte address of a staged entry's record
    __device__ __forceinline__ uint32_t address(uint32_t e) const { return e * 16u + rc[e >> kIdxBits].z; }
    // cell offset of entry e's neighbour times the cell widths (as
    // entry_offset; BREL: none, the records are brick-relative)
    __device__ __forceinline__ void offset(uint32_t e, float, float, float, float& ox, float& oy, float& oz) const {
        if constexpr (BREL) {
            ox = oy = oz = 0.0f;
        } else {
            const lds_quad_t t = code[e >> kIdxBits];
            ox = __uint_as_float(t.x);
            oy = __uint_as_float(t.y);
            oz = __uint_as_float(t.z);
        }
    }
    __device__ __forceinline__ void from_lds(uint32_t addr, Q (&q)[NR]) const {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const uint32_t a = addr + (uint32_t)k * qstride;
            if (T64 && k == NR - 1) {
                typedef uint32_t lds_pair_t __attribute__((ext_vector_type(2)));
                const lds_pair_t v = *(const SPHX_LDS lds_pair_t*)(size_t)a;
                const lds_quad_t w = {v.x, v.y, 0u, 0u};
                __builtin_memcpy(&q[k], &w, sizeof w);
            } else {
                q[k] = lds_quad<Q>((const SPHX_LDS lds_quad_t*)(size_t)a);
            }
        }
    }
    // the record of the walking particle i itself (from its own cell's halo
    // row: LDS when staged; the same bits as the global record)
    __device__ __forceinline__ void own(uint32_t i, Q (&q)[NR]) const { (*this)(i | kOwnCode, q); }
    __device__ __forceinline__ void operator()(uint32_t e, Q (&q)[NR]) const {
        uint32_t addr;
        if (locate(e, addr)) {
            from_lds(addr, q);
        } else {
            const size_t j = e & kIdxMask;
            const Q* p = glob + j * NQ;
#pragma unroll
            for (int k = 0; k < NQ; ++k) q[k] = p[k];
            if constexpr (NQ2 > 0) {
                const Q* p2 = glob2 + j * NQ2;
#pragma unroll
                for (int k = 0; k < NQ2; ++k) q[NQ + k] = p2[k];
            }
            if constexpr (BREL) {  // the cell of entry e: code c = stage * 3 + x offset + 1
                const uint32_t c = e >> kIdxBits, st = (c * 11u) >> 5, sz = (st * 11u) >> 5;
                const int ox = (int)(c - 3u * st) - 1, oy = (int)(st - 3u * sz) - 1, oz = (int)sz - 1;
                const int kx = (int)(kxyz & 255u) - 128, ky = (int)((kxyz >> 8) & 255u) - 128,
                          kz = (int)(kxyz >> 16) - 128;
                brel_shift(q[0], kx + ox, ky + oy, kz + oz, wx, wy, wz);
            }
        }
    }
};

// body(e, record) for every list entry of sorted particle i, records from the
// brick's LDS: W (2 or 4) entries per iteration, with PF the next W loaded
// ahead.  When all W neighbours of every active lane are staged (a
// wave-uniform test) the records are read with no per-lane branch, so the W
// bodies, written branch-free, form one basic block the compiler can
// interleave.  In a brick whose halo rows are all staged whole (F.whole)
// every entry is staged and the test is skipped.  flush() as in for_neighbour_groups (it needs W = 4).
template <class Q, int NQ, int W, bool PF, int NQ2, bool PIN, bool T64, bool S16, bool BREL, class Body,
          class Flush = NoFlush>
__device__ __forceinline__ void walk(const NbrList& L, uint32_t i,
                                     const BrickFetch<Q, NQ, W, PF, NQ2, PIN, T64, S16, BREL>& F, Body&& body,
                                     Flush&& flush = Flush()) {
    constexpr int NR = NQ + NQ2;
    static_assert(W == 2 || W == 4, "two or four entries per iteration");
    static_assert(W == 4 || std::is_same<typename std::decay<Flush>::type, NoFlush>::value,
                  "groups of four entries need four entries per iteration");
    const uint32_t cnt = min(F.count(L, i), F.kend);
    if constexpr (S16) {
        if (F.whole && F.s16) {
            // the slot list: a group of four 16-bit slots per 8-byte load,
            // the next group loaded ahead (the last group again past the
            // end), walked W entries at a time; record address rec0 + 16 slot
            const uint16_t* __restrict__ s = F.s16;
            uint32_t k = F.kbeg;
            if (k >= cnt) return;
            const uint32_t last = (cnt - 1u) & ~3u;
            auto load = [&](uint32_t kk) { return *reinterpret_cast<const uint2*>(s + ((size_t)(kk >> 2) << 8)); };
            uint2 a = load(k);
            for (; k + 3u < cnt; k += 4u) {
                const uint2 c = a;
                a = load(min(k + 4u, last));
                if constexpr (PIN) __builtin_amdgcn_sched_barrier(0);
                const uint32_t sl[4] = {c.x & 0xffffu, c.x >> 16, c.y & 0xffffu, c.y >> 16};
#pragma unroll
                for (int u0 = 0; u0 < 4; u0 += W) {
                    Q q[W][NR];
#pragma unroll
                    for (int w = 0; w < W; ++w) F.from_lds(F.rec0 + sl[u0 + w] * 16u, q[w]);
#pragma unroll
                    for (int w = 0; w < W; ++w) body(sl[u0 + w], q[w]);
                    // two-entry walks: the second pair's reads stay after the first pair
                    if constexpr (W == 2) if (u0 == 0) __builtin_amdgcn_sched_barrier(0);
                }
                flush();
            }
            if (k < cnt) {
                const uint32_t sl[3] = {a.x & 0xffffu, a.x >> 16, a.y & 0xffffu};
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    if (k + (uint32_t)u < cnt) {
                        Q q[NR];
                        F.from_lds(F.rec0 + sl[u] * 16u, q);
                        body(sl[u], q);
                    }
                }
                flush();
            }
            return;
        }
    }
    const uint32_t* __restrict__ e = nbr_base(L, i);
    // W entries of the list from position k (k a multiple of W): one 16- or
    // 8-byte load
    auto load = [&](uint32_t k, uint32_t (&c)[W]) {
        if constexpr (W == 4) {
            const uint4 q = nbr_group(e, k);
            c[0] = q.x; c[1] = q.y; c[2] = q.z; c[3] = q.w;
        } else {
            const uint2 q = *reinterpret_cast<const uint2*>(e + nbr_off(k));
            c[0] = q.x; c[W - 1] = q.y;
        }
    };
    uint32_t a[W];
    uint32_t k = F.kbeg;
    if (PF && cnt > k) load(k, a);
    // WHOLE (the brick's halo is staged whole): no staged test per entry
    auto steps = [&](auto whole) {
        for (; k + (W - 1) < cnt; k += W) {
            uint32_t c[W];
            if (PF) {
                // the next W entries are loaded unconditionally (the group past
                // the end is the last one again) so that the load has its own
                // registers and nothing waits for it this iteration
#pragma unroll
                for (int u = 0; u < W; ++u) c[u] = a[u];
                load(min(k + W, (cnt - 1u) & ~(uint32_t)(W - 1)), a);
                // keep the load at the top of the iteration (the scheduler
                // may otherwise sink it next to its wait)
                if constexpr (PIN || W == 2) __builtin_amdgcn_sched_barrier(0);
            } else {
                load(k, c);
            }
            Q q[W][NR];
            if constexpr (decltype(whole)::value) {
                uint32_t ad[W];
#pragma unroll
                for (int u = 0; u < W; ++u) {
                    ad[u] = F.address(c[u]);
                    F.from_lds(ad[u], q[u]);
                }
                if constexpr (S16 && W == 4) {
                    if (F.w16) {  // the group's four slots, one 8-byte store
                        const uint32_t s0 = (ad[0] - F.rec0) >> 4, s1 = (ad[1] - F.rec0) >> 4;
                        const uint32_t s2 = (ad[2] - F.rec0) >> 4, s3 = (ad[3] - F.rec0) >> 4;
                        *reinterpret_cast<uint2*>(F.w16 + ((size_t)(k >> 2) << 8)) =
                            make_uint2(s0 | (s1 << 16), s2 | (s3 << 16));
                    }
                }
            } else {
                uint32_t slot[W];
                bool staged = true;
#pragma unroll
                for (int u = 0; u < W; ++u) staged = F.locate(c[u], slot[u]) && staged;
                if (__all(staged)) {
#pragma unroll
                    for (int u = 0; u < W; ++u) F.from_lds(slot[u], q[u]);
                } else {
#pragma unroll
                    for (int u = 0; u < W; ++u) F(c[u], q[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < W; ++u) body(c[u], q[u]);
            flush();
        }
    };
    if (F.whole) steps(std::true_type{});
    else steps(std::false_type{});
    Q q[NR];
    if (k < cnt) {
        uint32_t c[W];
        if (PF) {
#pragma unroll
            for (int u = 0; u < W; ++u) c[u] = a[u];
        } else {
            load(k, c);
        }
#pragma unroll
        for (int u = 0; u < W - 1; ++u) {
            if (k + u < cnt) {
                F(c[u], q);
                body(c[u], q);
            }
        }
        if constexpr (S16 && W == 4) {
            if (F.w16) {  // the last, partial group (its entries are staged: the brick is whole)
                uint32_t sl[3];
#pragma unroll
                for (int u = 0; u < 3; ++u) sl[u] = k + (uint32_t)u < cnt ? (F.address(c[u]) - F.rec0) >> 4 : 0u;
                *reinterpret_cast<uint2*>(F.w16 + ((size_t)(k >> 2) << 8)) = make_uint2(sl[0] | (sl[1] << 16), sl[2]);
            }
        }
        flush();
    }
}

// The workgroup's brick (one per active brick, dealt so that each XCD takes a
// contiguous range, see xcd_block): set up, stage the records of `src` (NQ
// quads each) and, with NQ2 > 0, of `src2` (NQ2 quads each, read as quads NQ
// .. NQ + NQ2 - 1), then body(i, fetch) for each of the brick's particles.
// LANES = 2: two neighbouring lanes share each particle (body sees both).
template <class Q, int NQ, int W, bool PF, int LANES = 1, int NQ2 = 0, bool PIN = false, bool T64 = false,
          bool S16 = false, bool BREL = false, class Body>
__device__ __forceinline__ void brick_pass(const Bricks& B, const Grid& g, const uint32_t* __restrict__ start,
                                           uint32_t cap, const NbrList& L, const Q* __restrict__ src, BrickHdr& H,
                                           uint4* __restrict__ sh, Body&& body, const Q* __restrict__ src2 = nullptr,
                                           bool write16 = false) {
    const uint32_t hy = (uint32_t)B.b[1] + 2u, nh = hy * ((uint32_t)B.b[2] + 2u);
    const uint32_t BY = (uint32_t)B.b[1], no = BY * (uint32_t)B.b[2];
    const SPHX_LDS BrickHdr* HL = (const SPHX_LDS BrickHdr*)&H;
    // dynamic LDS: the row-code table (brick_rc_quads), then the records
    uint4* rc = sh;
    uint4* rec = sh + brick_rc_quads(B);
    const uint32_t id = B.active[xcd_block()];
    brick_setup(id, B, g, start, cap, H, BREL);
    // the list length of the thread's first particle, loaded while the halo
    // is staged (its latency then overlaps the staging's)
    const uint32_t nown = HL->own_cum[no], u0 = LANES == 2 ? pair_slot(threadIdx.x) : threadIdx.x;
    uint32_t k = 0, n0 = 0;
    if (u0 < nown) {
        while (u0 >= HL->own_cum[k + 1]) ++k;
        n0 = L.ntot[HL->own_start[k] + (u0 - HL->own_cum[k])];
    }
    BrelStage bs{};
    if constexpr (BREL) {
        bs.skey = B.skey;
        bs.ox0 = (int)(id % (uint32_t)B.nb[0]) * B.b[0] + B.b[0] / 2;
        bs.ky0 = -1 - (int)BY / 2;
        bs.kz0 = -1 - B.b[2] / 2;
        bs.hy = (int)hy;
        bs.nx = (uint32_t)g.dims[0];
        bs.inv_nx = g.inv_nx;
        bs.wx = (float)g.w[0];
        bs.wy = (float)g.w[1];
        bs.wz = (float)g.w[2];
    }
    brick_stage<NQ, NQ2, BREL>(H, (int)nh, reinterpret_cast<const uint4*>(src), reinterpret_cast<const uint4*>(src2),
                               rec, cap, rc, no, BY, hy, bs);
    using Fetch = BrickFetch<Q, NQ, W, PF, NQ2, PIN, T64, S16, BREL>;
    Fetch base_fetch{(const SPHX_LDS lds_quad_t*)rc, (const SPHX_LDS lds_quad_t*)HL->code, src, src2, cap * 16u,
                     (uint32_t)__builtin_amdgcn_readfirstlane(HL->whole)};
    base_fetch.rec0 = lds_addr(rec);
    if constexpr (BREL) {
        base_fetch.wx = (float)g.w[0];
        base_fetch.wy = (float)g.w[1];
        base_fetch.wz = (float)g.w[2];
    }
    // the slot lists are read (or, write16, filled) only in a brick whose
    // halo is staged whole
    const bool slots = S16 && L.s16 != nullptr && base_fetch.whole;
    for (uint32_t u = u0; u < nown; u += blockDim.x / LANES) {
        while (u >= HL->own_cum[k + 1]) ++k;
        const uint32_t i = HL->own_start[k] + (u - HL->own_cum[k]);
        Fetch f = base_fetch;
        f.rc += k * (uint32_t)kRowCodes;
        f.n_list = u == u0 ? n0 : (uint32_t)L.ntot[i];
        if (slots && write16) f.w16 = L.s16 + nbr_index(L, i);
        else if (slots) f.s16 = L.s16 + nbr_index(L, i);
        if constexpr (BREL) {
            // the particle's cell offset from the brick origin (its x only
            // where some record comes from global memory)
            if (!base_fetch.whole) {
                const int kx = (int)cell_x_of(B.skey[i], bs.nx, bs.inv_nx) - bs.ox0;
                const int ky = (int)(k % BY) - (int)BY / 2, kz = (int)(k / BY) - B.b[2] / 2;
                f.kxyz = (uint32_t)(kx + 128) | ((uint32_t)(ky + 128) << 8) | ((uint32_t)(kz + 128) << 16);
            }
        }
        body(i, f);
    }
}

}  // namespace sphx
