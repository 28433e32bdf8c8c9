// sphx_internal.h — device buffers, grid description and kernel launchers
// shared by the kernel translation units (sphx_sort.hip, sphx_pairs.hip,
// sphx_particles.hip, sphx_adjoint.hip) and the runtime (sphx_runtime.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/sphx.h"

namespace sphx {

// Uniform cell grid: cell (cx,cy,cz) holds positions with
// floor((x - lo) * inv_w) == cx after clamping to [0, dims).  Cell width is
// >= 2*max(h) so every neighbour (r < 2*h_avg <= 2*max(h)) lies in the 27 cells
// around a particle's own; clamping is monotone and 1-Lipschitz in cell
// units, so particles outside the grid box stay correct (just slower).
struct Grid {
    double lo[3];
    double w[3];      // cell width per axis
    double inv_w[3];
    int dims[3];
    int periodic;     // 1: periodic box, cells wrap, minimum image
    double L[3];      // periodic box length
    double halfL[3];
    double invL[3];   // 1 / L (periodic), 0 otherwise
    uint32_t ncell;
    double inv_nx;    // 1 / dims[0] (cell_x_of)
};

// Persistent per-particle state, in sorted (cell) order.  Two copies: the
// counting sort gathers from one into the other.
struct Particles {
    double4* pos;              // (x, y, z, h) packed, kept current with x: one 32-byte gather per neighbour
    double *x, *y, *z;
    double *vx, *vy, *vz;
    double *apx, *apy, *apz;   // PreviousAcceleration
    double *m, *h;
    double *bx, *by, *bz, *psi;
    double *q;
    double *u, *dup;           // adiabatic EOS: specific internal energy, du/dt of the tick's first evaluation
    uint32_t *id, *orig;       // rank of the entity (pair orientation); upload slot
};

// Gather records of the force pass: one particle's neighbour-side fields in
// one 64-byte (f64) or 32-byte (FP32-mixed) block, so a neighbour costs 4 or
// 2 wide loads from one cache line instead of 8 gathers from 8 arrays.
// Written by the density pass (x, rho, P/rho^2 and v) and kept current by
// every later writer of v (vv_velocity, XSPH).
struct alignas(64) Rec64 {
    double x, y, z, rho;
    double vx, vy, vz, pr2;
};
struct alignas(32) Rec32 {
    float xr, yr, zr, rho;     // cell-relative position
    float vx, vy, vz, pr2;
};

// Per-particle quantities that are recomputed after every sort.
struct Derived {
    Rec64* r64;
    Rec32* r32;
    double4* recb;             // (bx, by, bz, psi) in sorted order, written by the reorder and the MHD passes
    double4* recb_next;        // the MHD passes' output (Jacobi), swapped with recb afterwards
    double4* recs;             // Maxwell stress: (B / rho, |B|^2 / (2 rho^2)) in sorted order (stress_record)
    double4* xref;             // positions at the last neighbour build (sorted order)
    double *rho, *p, *pr2;     // density, pressure, P / rho^2
    double *ax, *ay, *az;      // acceleration
    double *spx, *spy, *spz;   // m g + sum of pressure terms (reused by VISC_ONLY)
    double* du;                // adiabatic EOS: du/dt of the last full force pass
    double *tmp3[3];           // Jacobi scratch (xsph / induction / cleaning)
    double *tmp1;
};

struct Sort {
    uint32_t *key, *rank, *perm;
    uint32_t* tmp;             // arrival-order scatter (cell order before the deterministic ranking)
    uint32_t* skey;            // key in sorted order (written by the reorder)
    uint32_t *cell_count;      // ncell
    uint32_t *cell_start;      // ncell + 1
    uint32_t *block_sums;      // scan scratch
    uint32_t cell_cap;
};

// Tiles: every cell row (fixed y, z) is cut into runs of at most kTile
// consecutive sorted particles.  The neighbour build runs one workgroup per
// tile: the candidates of one neighbour row ("stage", 9 per tile) are one or
// two contiguous slices of the sorted arrays, staged through LDS.
static constexpr int kTile = 256;
static constexpr int kStages = 9;
// blocks of a one-thread-per-particle launch (kTile threads per block)
static inline uint32_t nblocks(size_t n, int b = kTile) { return (uint32_t)((n + b - 1) / b); }
struct Tiles {
    uint32_t* row_count;       // tiles per row (nrows)
    uint32_t* row_off;         // exclusive scan (nrows + 1); row_off[nrows] = tile count
    uint32_t* scan_sums;
    uint32_t *p0, *np, *row;   // per tile: first particle, particles, row index
    uint32_t nrows, cap;       // rows of the grid; tile arrays capacity
    uint32_t nrows_cap;        // rows the row arrays can hold
};

// Neighbour list of sorted particle i: ntot[i] entries, entry k at
// idx[(((i / 64) * (ks / 4) + k / 4) * 64 + i % 64) * 4 + k % 4] (ks a multiple
// of 4): each lane reads or writes four consecutive entries as one 16-byte
// word, and the 64 lanes of a wave at equal k / 4 read 1 KiB contiguously.
// An entry is the neighbour's
// sorted index (bits 0..26) and its cell's offset from particle i's cell
// (bits 27..31: stage * 3 + x offset + 1, stage = (oz + 1) * 3 + oy + 1).
static constexpr uint32_t kIdxBits = 27;
static constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1u;
static constexpr uint32_t kMaxParticles = 1u << kIdxBits;

struct NbrList {
    uint32_t* idx;
    // the same entries as LDS record slots of the particle's brick halo (16
    // bits each, same index as idx), written by the first density pass after
    // each build in the bricks it stages whole; later staged passes read them
    // in a brick whose halo rows are all staged whole (null: not written)
    uint16_t* s16;
    uint16_t* ntot;            // entries per particle (<= ks)
    uint32_t* overflow;        // [0] largest list seen, [1] some list overflowed, [2] a particle moved
                               // more than skin/2 since the build
    int ks;                    // list capacity per particle
};

// Bricks of b[0] x b[1] x b[2] cells for the staged passes.  A workgroup
// takes one brick: it copies the gather records of the (b[1]+2) x (b[2]+2)
// cell rows around the brick, cells b.x - 1 .. b.x + b[0], into LDS (each row
// one contiguous range of the sorted order) and walks its particles' lists
// against LDS.  Every list entry of a brick particle lies in those rows; the
// entries that are not staged (a periodic row that wraps in x, LDS full) are
// read from global memory, so the result never depends on the staging.
static constexpr int kMaxBrickYZ = 4;
static constexpr int kMaxHaloRows = (kMaxBrickYZ + 2) * (kMaxBrickYZ + 2);
static constexpr int kMaxOwnRows = kMaxBrickYZ * kMaxBrickYZ;
struct Bricks {
    int b[3];                  // cells per brick along x, y, z (b[1], b[2] <= kMaxBrickYZ)
    int nb[3];                 // bricks along x, y, z
    uint32_t count;            // nb[0] * nb[1] * nb[2]
    uint32_t cap;              // capacity of the arrays below
    uint32_t* flag;            // 1 if the brick holds particles [cap]
    uint32_t* pos;             // exclusive scan of flag [cap + 1]; pos[count] = active bricks
    uint32_t* sums;            // scan scratch
    uint32_t* active;          // ids of the bricks that hold particles, in brick order [cap]
    const uint32_t* skey;      // cell key of each sorted particle (the sort's)
    int cus;                   // compute units: the staged passes run cus * k workgroups
    int on;                    // staged passes enabled
    int threads;               // workgroup size of the staged passes (256 or 512)
    uint32_t nactive;          // active bricks of the current sort (host copy): one workgroup each
    double ppc;                // mean particles per non-empty cell at upload (LDS sizing)
};
// 16-byte words of dynamic LDS the staged passes keep before the records: the
// row-code table, 27 entries per brick row
__host__ __device__ inline uint32_t brick_rc_quads(const Bricks& B) {
    return (uint32_t)(B.b[1] * B.b[2]) * 27u;
}

// f32 copies in sorted order: p4 = (xr, yr, zr, h), the position as an offset
// from the particle's own cell origin (computed in f64, rounded once), so a
// pair vector is (xr_j - xr_i) + (cell offset) * w, never a difference of
// large numbers.  The neighbour build prefilters with them in both forms; the
// FP32-mixed passes compute with them (one 16-byte record per neighbour).
struct Mixed {
    float4* p4;
    float* m;
};

struct Vec3 {
    double *x, *y, *z;
};

struct KernelParams {          // what the kernels need from SphxParams, by value
    double dt, half_dt, half_dt_sq;
    double rest_density, speed_of_sound, gamma, tait_B;
    double poly_k, iso_cs2;
    double alpha, beta;
    double xsph_eps;
    double half_extent, act_dist, stiffness;
    double box_lo[3], box_hi[3];
    double g[3], e[3];
    double eta, ch, sigma_c;
    int eos, visc_conv, bnd, mhd;
    // set at upload when every particle has the same mass / smoothing radius:
    // the pair passes then use these instead of staging m and h
    int uniform;
    double m0, h0, inv_h0;
    int rec64;                 // keep the f64 gather record current (the FP64 form)
    int srec;                  // Maxwell stress on: density and cleaning passes write D.recs
    double skin;               // neighbour-list skin (list radius 2 h_avg + skin)
    double half_skin_sq;       // (skin / 2)^2: displacement since the last build that forces a rebuild
};

// ---- kernel launchers -------------------------------------------------------
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void launch_iota_u32(uint32_t* p, uint32_t n, hipStream_t s);
void launch_cell_hash(const Particles& P, uint32_t n, const Grid& g, const Sort& S, hipStream_t s);
void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums, hipStream_t s);
void launch_scatter(const Sort& S, uint32_t n, hipStream_t s);
void launch_cell_order(const Sort& S, uint32_t n, const double* x, hipStream_t s);
void launch_reorder(const Particles& src, const Particles& dst, const Sort& S, uint32_t n,
                    bool mhd, bool charge, bool mixed, bool energy, const Mixed& X, const Grid& g, double4* recb,
                    double4* xref, hipStream_t s);
void launch_tiles(const Grid& g, const Sort& S, const Tiles& T, hipStream_t s);
void launch_stress_rec(const Particles& P, const Derived& D, uint32_t n, hipStream_t s);
void launch_bricks(const Grid& g, const Sort& S, const Bricks& B, hipStream_t s);
uint32_t tile_grid(uint32_t n, const Tiles& T);
void launch_build_nbr(const Particles& P, uint32_t n, const Grid& g, const Sort& S, const Tiles& T,
                      const NbrList& L, bool uniform, const Mixed& X, double h_max, double skin, hipStream_t s);
// write_slots: fill L.s16 (the first density pass after a build; L.s16 set)
void launch_density(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                    const KernelParams& kp, bool mixed, const Mixed& X, const Bricks& B,
                    const uint32_t* cell_start, bool write_slots, hipStream_t s);
// energy: also du/dt of the adiabatic EOS (full passes only)
void launch_forces(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                   const KernelParams& kp, bool visc_only, bool energy, bool mhd_stress, bool charge, bool mixed,
                   const Mixed& X, const Bricks& B, const uint32_t* cell_start, hipStream_t s);
// P, P/rho^2 and the records' P/rho^2 from the kept density and the current
// internal energy (the adiabatic EOS after the internal energy changed)
void launch_eos(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp, bool mixed,
                hipStream_t s);
// The moved-skin/2 flag of the position updates: the device word and its
// copy in pinned host memory, both holding the epoch in which a particle was
// last found beyond skin/2.  The host starts a new epoch at each list build
// and each replacement of the positions, so nothing needs clearing.
struct MovedFlag {
    uint32_t* dev;
    uint32_t* host;
    uint32_t epoch;
};
void launch_vv_position(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp, const Grid& g,
                        const uint32_t* skey, bool mixed, const Mixed& X, const MovedFlag& mf, uint8_t* flip,
                        hipStream_t s);
// State (upload order, through orig) -> sorted arrays: velocities and their
// record copies, and, with x3, positions with the position bookkeeping of a
// position update (the moved flag raised in epoch mf.epoch when mf.dev).
void launch_load_state(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp, const Grid& g,
                       const uint32_t* skey, bool mixed, const Mixed& X, const double* x3, const double* v3,
                       const MovedFlag& mf, hipStream_t s);
void launch_vv_velocity(const Particles& P, const Derived& D, uint32_t n, const KernelParams& kp,
                        bool mixed, const Mixed& X, hipStream_t s);
void launch_xsph(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                 const KernelParams& kp, bool mixed, const Mixed& X, hipStream_t s);
void launch_induction(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                      const KernelParams& kp, bool mixed, const Mixed& X, const Bricks& B,
                      const uint32_t* cell_start, hipStream_t s);
void launch_div_clean(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                      const KernelParams& kp, bool mixed, const Mixed& X, const Bricks& B,
                      const uint32_t* cell_start, hipStream_t s);
void launch_export_candidates(const Particles& P, uint32_t n, const Grid& g, const uint32_t* cell_start,
                              const double* R, double half_skin, uint32_t* out, unsigned long long* counter,
                              uint64_t out_cap, hipStream_t s);
void launch_export_pairs(const Particles& P, uint32_t n, const Grid& g, const NbrList& L, uint32_t* out,
                         unsigned long long* counter, uint64_t out_cap, hipStream_t s);
void launch_energy(const Particles& P, const Derived& D, uint32_t n, const Grid& g,
                   const NbrList& L, const KernelParams& kp, bool mhd, double* partial,
                   uint32_t nblocks, hipStream_t s);
uint32_t energy_blocks(uint32_t n);

// adjoint (sphx_adjoint.hip); Vec3 arrays named *_s are in sorted order,
// the others in upload order
void launch_adj_seed(const uint32_t* orig, uint32_t n, double hdt, const Vec3& vb, const Vec3& apb,
                     double4* ab4, hipStream_t s);
void launch_gather_ab(const uint32_t* orig, uint32_t n, const Vec3& src, double4* ab4, hipStream_t s);
void launch_scatter3(const uint32_t* orig, uint32_t n, const Vec3& src_s, const Vec3& dst, hipStream_t s);
void launch_scatter6(const uint32_t* orig, uint32_t n, const Vec3& src, const Vec3& dst, const Vec3& src2,
                     const Vec3& dst2, hipStream_t s);
void launch_adj_flip(uint32_t n, const uint8_t* flip, double dt, double hdt2, const Vec3& xb, const Vec3& vb,
                     const Vec3& apb, const Vec3& ab, hipStream_t s);
void launch_adj_loss(const Particles& P, uint32_t n, const Vec3& target, const Vec3& xb, double* part,
                     hipStream_t s);
// workgroups of the VJP passes (two lanes per particle): the size of the
// per-block dL/dalpha array
uint32_t vjp_blocks(uint32_t n);
void launch_vjp(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                const KernelParams& kp, const double4* ab4, const Vec3& xb, const Vec3& vb, double* rhob,
                double* alpha_acc, hipStream_t s,
                bool overwrite = false);
// xb += xs, vb += vs (upload order, n each) and alpha_acc[b] += as[b] for the
// nb VJP blocks: the VJP of a tick written with overwrite = true, applied
// once its neighbour list is known to be valid
void launch_adj_accumulate(uint32_t n, const Vec3& xs, const Vec3& vs, const Vec3& xb, const Vec3& vb,
                           const double* as, double* alpha_acc, uint32_t nb, hipStream_t s);

}  // namespace sphx
