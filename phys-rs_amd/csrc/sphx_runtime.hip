// sphx_runtime.hip — context, device memory, pass sequencing and the C ABI.
//
// The context owns all device memory.  State stays resident in HBM between
// calls; sphx_upload / sphx_download are the only host<->device copies.
// No entry point throws or aborts: every failure becomes a status code and a
// message for sphx_last_error().

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "sphx_internal.h"

using namespace sphx;

namespace {

constexpr int kDefaultNeighbours = 192;  // list capacity per particle when max_neighbours <= 0
constexpr int kMaxNeighbours = 4096;

enum Slot { kSort = 0, kNbr, kDensity, kForce, kViscOnly, kIntegrate, kMhd, kNumSlots };

// ---- scratch of the entry point that allocates nothing -----------------
// sphx_export_pairs_noalloc stages the pairs (two u32 a pair) and their
// 64-bit counter in this file-scope device array instead of memory the
// library allocates: 2^21 pairs (16 MiB) and the counter, held by every GPU
// the library runs on until the process ends.  Host code never passes the
// name `scratch` to a kernel or copy: its device address on each device is
// resolved once, at sphx_create with that device current, into
// g_scratch[device].
constexpr uint64_t kNoAllocPairs = 1ull << 21;
constexpr size_t kScratchWords = 2 * kNoAllocPairs + 64;  // the pairs, then the counter (256-byte aligned)
static __device__ __align__(256) uint32_t scratch[kScratchWords];
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (once per device).
int resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return SPHX_E_NODEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return SPHX_OK;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return SPHX_E_HIP;
    }
    g_scratch[device] = ptr;
    return SPHX_OK;
}

struct TimedRange {
    hipEvent_t a, b;
    int slot;
};

}  // namespace

// Device memory of sphx_adjoint, kept between calls and grown on demand.
// Tapes and adjoint vectors are in upload order.
struct AdjBuffers {
    uint32_t n = 0, k = 0, m = 0;  // particles, ticks per segment, checkpoints
    std::vector<void*> allocs;
    double* ckpt = nullptr;        // m x 6n: x, y, z, vx, vy, vz at ticks 0, k, 2k, ...
    double* tx = nullptr;          // (k + 1) x 3n: x_t of the current segment
    double* tv = nullptr;          // (k + 1) x 3n: v_t
    double* tvp = nullptr;         // k x 3n: v after the position update (v')
    uint8_t* flip = nullptr;       // k x n: wall reflections of each tick
    double* fin = nullptr;         // 6n: state after the last tick
    double* adj = nullptr;         // 4 x 3n: xbar, vbar, abar, abar_prev
    double4* ab4 = nullptr;        // n: abar records in sorted order
    double* rhob = nullptr;        // n: rhobar in sorted order
    double* target = nullptr;      // 3n
    double* part = nullptr;        // per-block loss partial sums
    double* alpha_acc = nullptr;   // per-block dL/dalpha
    double* spec = nullptr;        // 2 x 3n: a tick's VJP (xbar, vbar shares) before it is added
    double* alpha_spec = nullptr;  // per-block dL/dalpha share of that VJP
    void release() {
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        n = k = m = 0;
    }
};

struct SphxCtx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t cap = 0, n = 0;
    SphxParams prm{};
    KernelParams kp{};
    Grid grid{};
    bool grid_set = false;
    Particles A[2]{};
    int cur = 0;
    Derived D{};
    Sort S{};
    Tiles T{};
    Bricks B{};
    uint32_t* bricks_host = nullptr;       // pinned: active brick count of the last sort
    hipEvent_t bricks_ready = nullptr;     // recorded after that count's copy
    bool bricks_pending = false;           // count not read yet
    bool slots = false;                    // the list gets slot lists (staged passes that read them)
    bool slots_fresh = false;              // the first density pass after the build has written them
    hipEvent_t moved_ready = nullptr;      // recorded after the kernel that raises the moved flag
    uint32_t* moved_host = nullptr;        // pinned, fine-grained: the moved flag as the kernels leave it
    uint32_t epoch = 1;                    // epoch of the moved flag (see MovedFlag); the words start at 0
    bool moved_pending = false;            // flag not read yet (passes enqueued on the kept list)
    NbrList L{};
    Mixed X{};
    double* energy_partial = nullptr;
    uint32_t* pair_buf = nullptr;
    unsigned long long* pair_counter = nullptr;
    uint64_t pair_buf_cap = 0;
    double* cand_r = nullptr;  // bounding radii in sorted order (sphx_export_candidates)
    uint32_t cand_r_cap = 0;
    bool has_mhd = false, has_charge = false;
    bool has_u = false;                    // an internal energy was uploaded
    bool pres = false;                     // P (adiabatic: from the current u) matches the density
    // the Maxwell-stress records (D.recs) match B and the density; the density
    // and cleaning passes write them, the induction pass makes them stale
    bool srec = false;
    bool uniform = false;                  // all m equal and all h equal
    double m0 = 0.0, h0 = 0.0;
    double hmax = 0.0;                     // largest smoothing radius (h is fixed per particle)
    // validity of derived data for the current positions and sorted order
    bool sorted = false, listed = false, dens = false, sp = false, acc = false;
    // list lifetime in ticks: position updates since the current list's
    // build, and the count at which the previous list was found stale (0: none yet)
    uint32_t list_age = 0, last_life = 0;
    uint64_t ticks = 0;                    // sphx_step ticks since the upload (kept by checkpoints)
    int spec_mode = 0;                     // SPHX_SPECULATE: 0 by list lifetime, 1 always, 2 never
    // timing: the slots bracketed with events (bit slot + 1, SPHX_TIMING_*)
    uint32_t timing = 0;
    std::vector<TimedRange> pending;
    std::vector<hipEvent_t> free_events;
    double ms[kNumSlots] = {0};
    uint64_t launches[kNumSlots] = {0};
    std::vector<void*> allocs;
    std::string err;
    AdjBuffers adj;

    Particles& P() { return A[cur]; }
    bool mixed() const { return prm.precision == SPHX_PREC_FP32_MIXED; }
    // the adiabatic EOS: u is evolved with du/dt from the full force passes
    bool energy() const { return has_u && prm.eos == SPHX_EOS_ADIABATIC; }
};

namespace {

int fail(SphxCtx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                  \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(ctx, SPHX_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

template <typename T>
int dalloc(SphxCtx* c, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, SPHX_E_OOM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                    hipGetErrorString(e));
    }
    c->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return SPHX_OK;
}

void free_all(SphxCtx* c) {
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
}

int check_params(SphxCtx* c, const SphxParams* p) {
    if (!p) return fail(c, SPHX_E_INVALID, "params is NULL");
    if (!(p->dt > 0.0) || !std::isfinite(p->dt)) return fail(c, SPHX_E_INVALID, "dt must be > 0");
    if (p->eos < SPHX_EOS_TAIT || p->eos > SPHX_EOS_ADIABATIC)
        return fail(c, SPHX_E_INVALID, "unknown eos %d", p->eos);
    if (p->eos == SPHX_EOS_TAIT && (!(p->rest_density > 0.0) || !(p->gamma != 0.0)))
        return fail(c, SPHX_E_INVALID, "Tait EOS needs rest_density > 0 and gamma != 0");
    if (p->visc_convention != SPHX_VISC_REFERENCE && p->visc_convention != SPHX_VISC_MONAGHAN)
        return fail(c, SPHX_E_INVALID, "unknown viscosity convention %d", p->visc_convention);
    if ((p->boundary_flags & SPHX_BND_REFLECT) && (p->boundary_flags & SPHX_BND_PERIODIC))
        return fail(c, SPHX_E_INVALID, "reflective and periodic boxes are exclusive");
    if (p->boundary_flags & (SPHX_BND_REFLECT | SPHX_BND_PERIODIC))
        for (int k = 0; k < 3; ++k)
            if (!(p->box_hi[k] > p->box_lo[k]))
                return fail(c, SPHX_E_INVALID, "box_hi must exceed box_lo on every axis");
    if (p->precision != SPHX_PREC_FP64 && p->precision != SPHX_PREC_FP32_MIXED)
        return fail(c, SPHX_E_INVALID, "unknown precision %d", p->precision);
    return SPHX_OK;
}

void make_kernel_params(SphxCtx* c) {
    const SphxParams& p = c->prm;
    KernelParams& k = c->kp;
    k.dt = p.dt;
    k.half_dt = 0.5 * p.dt;                 // integrators.rs:97
    k.half_dt_sq = 0.5 * p.dt * p.dt;       // integrators.rs:73
    k.rest_density = p.rest_density;
    k.speed_of_sound = p.speed_of_sound;
    k.gamma = p.gamma;
    k.tait_B = p.rest_density * (p.speed_of_sound * p.speed_of_sound) / p.gamma;  // sph.rs:157
    k.poly_k = p.poly_k;
    k.iso_cs2 = p.iso_cs2;
    k.alpha = p.visc_alpha;
    k.beta = p.visc_beta;
    k.xsph_eps = p.xsph_epsilon;
    k.half_extent = p.half_extent;
    k.act_dist = p.activation_distance;
    k.stiffness = p.stiffness;
    for (int a = 0; a < 3; ++a) {
        k.box_lo[a] = p.box_lo[a];
        k.box_hi[a] = p.box_hi[a];
        k.g[a] = p.gravity[a];
        k.e[a] = p.e_field[a];
    }
    k.eta = p.resistivity;
    k.ch = p.cleaning_speed;
    k.sigma_c = p.cleaning_damping;
    k.eos = p.eos;
    k.visc_conv = p.visc_convention;
    k.bnd = p.boundary_flags;
    k.mhd = p.mhd_flags;
    k.uniform = c->uniform ? 1 : 0;
    k.rec64 = c->mixed() ? 0 : 1;
    k.srec = (c->has_mhd && (p.mhd_flags & SPHX_MHD_MAXWELL_STRESS)) ? 1 : 0;
    c->srec = false;
    k.skin = p.neighbour_skin < 0.0 ? 0.04 * c->hmax : p.neighbour_skin;
    k.half_skin_sq = 0.25 * k.skin * k.skin;
    k.m0 = c->m0;
    k.h0 = c->h0;
    k.inv_h0 = c->uniform ? 1.0 / c->h0 : 0.0;
}

void brick_defaults(SphxCtx* c) {
    Bricks& B = c->B;
    B.b[0] = B.b[1] = B.b[2] = 3;
    B.on = 0;
    B.threads = 512;
    B.ppc = 1.0;
}

// Brick shape and workgroup size of the staged passes from ppc, the mean
// number of particles in a non-empty cell (counted here from the uploaded
// positions).  Among the shapes whose ppc * cells particles fill 50-100 % of a
// 256- or 512-thread workgroup, whose f64 force halo (64-byte records) fits
// the LDS at 8 waves per CU, and which leave at least 8 bricks per CU, the
// one with the most brick particles per staged record, weighted by the lane
// fill (the 1M dam break at ~19 per cell: 3 x 3 x 3 with 512 threads).  When
// none fits (few particles) the passes run unstaged.  SPHX_BRICKS=off or
// SPHX_BRICKS=bx,by,bz (by, bz <= 4) overrides the choice.
void choose_bricks(SphxCtx* c, const SphxState* host, const Grid& g) {
    Bricks& B = c->B;
    B.on = 0;
    const uint32_t n = host->n;
    if (!n) return;
    std::vector<uint64_t> seen(((size_t)g.ncell + 63) / 64, 0);
    uint64_t cells = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const double xs[3] = {host->x[i], host->y[i], host->z[i]};
        uint64_t key = 0;
        for (int k = 2; k >= 0; --k) {
            double u = (xs[k] - g.lo[k]) * g.inv_w[k];
            if (!(u >= 0.0)) u = 0.0;
            u = std::min(u, (double)(g.dims[k] - 1));
            key = key * (uint64_t)g.dims[k] + (uint64_t)u;
        }
        const uint64_t bit = 1ull << (key & 63);
        if (!(seen[key >> 6] & bit)) {
            seen[key >> 6] |= bit;
            ++cells;
        }
    }
    const double ppc = (double)n / (double)std::max<uint64_t>(cells, 1);
    B.ppc = ppc;
    const char* e = getenv("SPHX_BRICKS");
    int bx, by, bz;
    if (e && !strcmp(e, "off")) return;
    if (e && sscanf(e, "%d,%d,%d", &bx, &by, &bz) == 3 && bx >= 1 && by >= 1 && bz >= 1 && by <= kMaxBrickYZ &&
        bz <= kMaxBrickYZ) {
        B.b[0] = bx; B.b[1] = by; B.b[2] = bz;
        B.threads = ppc * bx * by * bz <= 256 ? 256 : 512;
        B.on = 1;
        return;
    }
    // the f64 force record (64 bytes); in the FP64 form with a magnetic field
    // the MHD passes and the Maxwell-stress force pass stage the B or stress
    // record beside it (96 bytes; 4M MHD box: 4x3x3 cells instead of 4x4x3,
    // 8.34 -> 8.27 ms per step)
    const SphxParams& p = c->prm;
    const bool b_rec = host->bx && p.precision == SPHX_PREC_FP64 &&
                       (p.mhd_flags & (SPHX_MHD_MAXWELL_STRESS | SPHX_MHD_INDUCTION | SPHX_MHD_CLEANING));
    const double rec = b_rec ? 96.0 : 64.0;
    // at least 8 bricks per CU: a brick is one workgroup's work, and with
    // fewer the last round of workgroups leaves much of the GPU idle
    double best = 0.0;
    for (int z = 1; z <= kMaxBrickYZ; ++z)
        for (int y = 1; y <= kMaxBrickYZ; ++y)
            for (int x = 1; x <= 8; ++x) {
                // records of one 512-thread group's LDS, less the header and
                // the row-code table; a halo that overflows it is read from
                // global memory, which takes its waves off the LDS fast path
                const double lds_records = (double)(160 * 1024 - 2048 - 16 * 27 * y * z) / rec;
                const double own = ppc * x * y * z, halo = ppc * (x + 2) * (y + 2) * (z + 2);
                if ((double)n / own < 8.0 * B.cus) continue;
                for (int T : {256, 512}) {
                    if (own > T || own < 0.5 * T) continue;
                    // two standard deviations of a Poisson count above the mean
                    // halo (256 threads: two groups per CU)
                    if (halo + 2.0 * std::sqrt(halo) > lds_records * (T / 512.0)) continue;
                    const double score = (own / halo) * std::min(1.0, own / (0.85 * T));
                    if (score > best) {
                        best = score;
                        B.b[0] = x; B.b[1] = y; B.b[2] = z;
                        B.threads = T;
                        B.on = 1;
                    }
                }
            }
}

// Grid from the parameters and (for open boxes) the uploaded positions, or
// the given one (a checkpoint's), then its cell, tile and brick arrays.
int setup_grid(SphxCtx* c, const SphxState* host, const Grid* fixed = nullptr) {
    const SphxParams& p = c->prm;
    double hmax = 0.0;
    for (uint32_t i = 0; i < host->n; ++i) hmax = std::max(hmax, host->h[i]);
    if (host->n && !(hmax > 0.0)) return fail(c, SPHX_E_INVALID, "smoothing radius must be > 0");
    c->hmax = hmax;
    make_kernel_params(c);  // the skin follows h_max
    Grid g{};
    if (fixed) {
        g = *fixed;
    } else {
        double cs = (2.0 * hmax + c->kp.skin) * (1.0 + 1e-9);
        if (p.cell_size > cs) cs = p.cell_size;
        if (!(cs > 0.0)) cs = 1.0;
        const bool periodic = (p.boundary_flags & SPHX_BND_PERIODIC) != 0;
        double lo[3], hi[3];
        bool explicit_grid = true;
        for (int k = 0; k < 3; ++k)
            if (!(p.grid_hi[k] > p.grid_lo[k])) explicit_grid = false;
        if (periodic) {
            for (int k = 0; k < 3; ++k) { lo[k] = p.box_lo[k]; hi[k] = p.box_hi[k]; }
        } else if (explicit_grid) {
            for (int k = 0; k < 3; ++k) { lo[k] = p.grid_lo[k]; hi[k] = p.grid_hi[k]; }
        } else if (p.boundary_flags & SPHX_BND_REFLECT) {
            for (int k = 0; k < 3; ++k) { lo[k] = p.box_lo[k]; hi[k] = p.box_hi[k]; }
        } else {
            // bounding box of the uploaded particles, padded by a quarter on each side
            if (host->n == 0) { for (int k = 0; k < 3; ++k) { lo[k] = 0; hi[k] = 1; } }
            else {
                const double* xs[3] = {host->x, host->y, host->z};
                for (int k = 0; k < 3; ++k) {
                    lo[k] = hi[k] = xs[k][0];
                    for (uint32_t i = 1; i < host->n; ++i) {
                        lo[k] = std::min(lo[k], xs[k][i]);
                        hi[k] = std::max(hi[k], xs[k][i]);
                    }
                    double pad = 0.25 * (hi[k] - lo[k]) + cs;
                    lo[k] -= pad;
                    hi[k] += pad;
                }
            }
        }
        // Cells never get narrower than 2*max(h); a grid above kMaxCells cells is
        // made coarser (wider cells stay correct, the walk just tests more).
        const double kMaxCells = (double)(1u << 24);
        double ext[3];
        for (int k = 0; k < 3; ++k) {
            ext[k] = hi[k] - lo[k];
            if (!std::isfinite(ext[k]) || !(ext[k] > 0.0)) return fail(c, SPHX_E_NONFINITE, "bad grid extent on axis %d", k);
        }
        for (int iter = 0;; ++iter) {
            double dd[3], prod = 1.0;
            for (int k = 0; k < 3; ++k) {
                dd[k] = periodic ? std::floor(ext[k] / cs) : std::ceil(ext[k] / cs);
                dd[k] = std::max(dd[k], 1.0);
                prod *= dd[k];
            }
            if (prod <= kMaxCells || iter > 64) {
                for (int k = 0; k < 3; ++k) g.dims[k] = (int)std::min(dd[k], 65536.0);
                break;
            }
            cs *= std::max(1.01, std::cbrt(prod / kMaxCells));
        }
        uint64_t ncell = 1;
        for (int k = 0; k < 3; ++k) {
            const int d = g.dims[k];
            if (periodic) {
                if (d < 3)
                    return fail(c, SPHX_E_INVALID,
                                "periodic box axis %d holds %d cells of width 2*h_max; at least 3 are needed", k, d);
                g.w[k] = ext[k] / d;
                g.L[k] = ext[k];
                g.halfL[k] = 0.5 * ext[k];
                g.invL[k] = 1.0 / ext[k];
            } else {
                g.w[k] = cs;
            }
            g.lo[k] = lo[k];
            g.inv_w[k] = 1.0 / g.w[k];
            ncell *= (uint64_t)d;
        }
        if (ncell > (1ull << 27)) return fail(c, SPHX_E_CAPACITY, "grid of %llu cells exceeds 2^27", (unsigned long long)ncell);
        g.periodic = periodic ? 1 : 0;
        g.ncell = (uint32_t)ncell;
        g.inv_nx = 1.0 / (double)g.dims[0];
    }
    if (g.ncell > c->S.cell_cap) {
        int rc;
        if ((rc = dalloc(c, &c->S.cell_count, g.ncell))) return rc;
        if ((rc = dalloc(c, &c->S.cell_start, (size_t)g.ncell + 1))) return rc;
        size_t tiles = ((size_t)g.ncell + 1023) / 1024 + 1;
        if ((rc = dalloc(c, &c->S.block_sums, tiles))) return rc;
        c->S.cell_cap = g.ncell;
    }
    const uint32_t nrows = (uint32_t)g.dims[1] * (uint32_t)g.dims[2];
    const uint32_t tcap = c->cap / kTile + nrows + 2;
    if (nrows > c->T.nrows_cap || tcap > c->T.cap) {
        int rc;
        if ((rc = dalloc(c, &c->T.row_count, nrows))) return rc;
        if ((rc = dalloc(c, &c->T.row_off, (size_t)nrows + 1))) return rc;
        if ((rc = dalloc(c, &c->T.scan_sums, ((size_t)nrows + 1023) / 1024 + 1))) return rc;
        if ((rc = dalloc(c, &c->T.p0, tcap)) || (rc = dalloc(c, &c->T.np, tcap)) || (rc = dalloc(c, &c->T.row, tcap)))
            return rc;
        c->T.cap = tcap;
        c->T.nrows_cap = nrows;
    }
    c->T.nrows = nrows;
    choose_bricks(c, host, g);
    {
        Bricks& B = c->B;
        uint64_t count = 1;
        for (int k = 0; k < 3; ++k) {
            B.nb[k] = (g.dims[k] + B.b[k] - 1) / B.b[k];
            count *= (uint64_t)B.nb[k];
        }
        B.count = (uint32_t)count;
        if (B.on && B.count > B.cap) {
            int rc;
            if ((rc = dalloc(c, &B.flag, B.count)) || (rc = dalloc(c, &B.pos, (size_t)B.count + 1)) ||
                (rc = dalloc(c, &B.sums, ((size_t)B.count + 1023) / 1024 + 1)) || (rc = dalloc(c, &B.active, B.count)))
                return rc;
            B.cap = B.count;
        }
    }
    c->grid = g;
    c->grid_set = true;
    return SPHX_OK;
}

// ---- timing -----------------------------------------------------------
hipEvent_t take_event(SphxCtx* c) {
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct Timer {
    SphxCtx* c;
    int slot;
    hipEvent_t a = nullptr;
    Timer(SphxCtx* ctx, int s) : c(ctx), slot(s) {
        if (c->timing & (2u << s)) {
            a = take_event(c);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    ~Timer() {
        if (!a) return;
        hipEvent_t b = take_event(c);
        if (!b) return;
        (void)hipEventRecord(b, c->stream);
        c->pending.push_back({a, b, slot});
    }
};

int drain_timing(SphxCtx* c) {
    for (auto& t : c->pending) {
        HIP_TRY(c, hipEventSynchronize(t.b));
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, t.a, t.b));
        c->ms[t.slot] += ms;
        c->launches[t.slot] += 1;
        c->free_events.push_back(t.a);
        c->free_events.push_back(t.b);
    }
    c->pending.clear();
    return SPHX_OK;
}

int launch_check(SphxCtx* c) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, SPHX_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return SPHX_OK;
}

// ---- passes ------------------------------------------------------------
// The list as the pair passes see it: the slot lists only once the first
// density pass after the build has written them (write: that pass itself).
NbrList pass_list(const SphxCtx* c, bool write = false) {
    NbrList L = c->L;
    if (!c->slots || (!c->slots_fresh && !write)) L.s16 = nullptr;
    return L;
}

int do_neighbours(SphxCtx* c) {
    const uint32_t n = c->n;
    {
        Timer t(c, kSort);
        Particles& src = c->A[c->cur];
        Particles& dst = c->A[1 - c->cur];
        launch_cell_hash(src, n, c->grid, c->S, c->stream);
        launch_scan(c->S.cell_count, c->S.cell_start, c->grid.ncell, c->S.block_sums, c->stream);
        launch_scatter(c->S, n, c->stream);
        launch_cell_order(c->S, n, src.x, c->stream);
        launch_reorder(src, dst, c->S, n, c->has_mhd, c->has_charge, c->mixed(), c->energy(), c->X, c->grid, c->D.recb,
                       c->D.xref, c->stream);
        c->cur = 1 - c->cur;
        launch_tiles(c->grid, c->S, c->T, c->stream);
        if (c->B.on) {
            // the staged passes launch one workgroup per active brick: the
            // count travels to the host while the neighbour build runs
            launch_bricks(c->grid, c->S, c->B, c->stream);
            HIP_TRY(c, hipMemcpyAsync(c->bricks_host, c->B.pos + c->B.count, sizeof(uint32_t),
                                      hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipEventRecord(c->bricks_ready, c->stream));
            c->bricks_pending = true;
        }
    }
    {
        Timer t(c, kNbr);
        ++c->epoch;  // displacements count from this build
        // the slot lists are read by the staged density and force passes
        // (FP64 and FP32-mixed) with uniform m and h; not with the Maxwell
        // stress, whose 96-byte FP64 records leave few bricks staged whole
        // (4M MHD box: no gain in its passes)
        const bool stress = c->has_mhd && (c->prm.mhd_flags & SPHX_MHD_MAXWELL_STRESS);
        c->slots = c->B.on && c->kp.uniform && !stress;
        c->slots_fresh = false;
        launch_build_nbr(c->P(), n, c->grid, c->S, c->T, c->L, c->kp.uniform != 0, c->X, c->hmax, c->kp.skin,
                         c->stream);
    }
    c->sorted = c->listed = true;
    c->list_age = 0;
    c->dens = c->sp = c->acc = false;
    return launch_check(c);
}

int ensure_list(SphxCtx* c) {
    int rc;
    if (!c->listed && (rc = do_neighbours(c))) return rc;
    if (c->bricks_pending) {
        HIP_TRY(c, hipEventSynchronize(c->bricks_ready));
        c->B.nactive = *c->bricks_host;
        c->bricks_pending = false;
    }
    return SPHX_OK;
}

int do_density(SphxCtx* c) {
    int rc;
    if ((rc = ensure_list(c))) return rc;
    {
        Timer t(c, kDensity);
        const bool write = c->slots && !c->slots_fresh;  // the first density pass on this list
        launch_density(c->P(), c->D, c->n, c->grid, pass_list(c, write), c->kp, c->mixed(), c->X, c->B,
                       c->S.cell_start, write, c->stream);
        c->slots_fresh = c->slots;
    }
    c->dens = c->pres = true;
    c->srec = c->kp.srec != 0;
    return launch_check(c);
}

int do_forces(SphxCtx* c, bool visc_only) {
    int rc;
    if (!c->dens && (rc = do_density(c))) return rc;
    if (visc_only && !c->sp)
        return fail(c, SPHX_E_STATE, "SPHX_FORCES_VISC_ONLY needs a full force pass at the current positions");
    if (visc_only && c->energy())
        return fail(c, SPHX_E_STATE, "SPHX_FORCES_VISC_ONLY is not available with the adiabatic EOS (the energy "
                                     "equation needs the pressure term of a full pass)");
    const bool stress = c->has_mhd && (c->prm.mhd_flags & SPHX_MHD_MAXWELL_STRESS);
    const bool charge = c->has_charge && c->has_mhd && (c->prm.mhd_flags & SPHX_MHD_LORENTZ_CHARGE);
    if (!c->pres) {  // the internal energy changed at these positions: P from the kept density
        Timer t(c, kDensity);
        launch_eos(c->P(), c->D, c->n, c->kp, c->mixed(), c->stream);
        c->pres = true;
    }
    if (stress && !c->srec) {  // B changed since the density and cleaning passes
        Timer t(c, kMhd);
        launch_stress_rec(c->P(), c->D, c->n, c->stream);
        c->srec = true;
    }
    {
        Timer t(c, visc_only ? kViscOnly : kForce);
        launch_forces(c->P(), c->D, c->n, c->grid, pass_list(c), c->kp, visc_only, c->energy(), stress, charge,
                      c->mixed(), c->X, c->B, c->S.cell_start, c->stream);
    }
    c->sp = true;
    c->acc = true;
    return launch_check(c);
}

// Position update; the neighbour list stays valid while no particle has
// moved more than skin/2 since it was built (the kernel raises a flag).
// defer: the flag travels to pinned memory and is read by settle_moved, so
// that passes can be enqueued on the kept list before the host knows.
// The moved flag as the last kernel that raises it left it (read after
// moved_ready has completed).
bool moved_flag(const SphxCtx* c) {
    return *reinterpret_cast<volatile const uint32_t*>(c->moved_host) == c->epoch;
}
MovedFlag moved_flag_args(SphxCtx* c) { return MovedFlag{c->L.overflow + 2, c->moved_host, c->epoch}; }

int do_vv_position(SphxCtx* c, uint8_t* flip = nullptr, bool defer = false) {
    if (!c->acc) return fail(c, SPHX_E_STATE, "vv_position needs accelerations (call sphx_forces)");
    {
        Timer t(c, kIntegrate);
        launch_vv_position(c->P(), c->D, c->n, c->kp, c->grid, c->S.skey, c->mixed(), c->X, moved_flag_args(c), flip,
                           c->stream);
    }
    int rc;
    if ((rc = launch_check(c))) return rc;
    c->dens = c->sp = c->acc = false;
    ++c->list_age;
    bool rebuild = c->kp.skin <= 0.0 || !c->listed;
    if (!rebuild) {
        HIP_TRY(c, hipEventRecord(c->moved_ready, c->stream));
        if (defer) {
            c->moved_pending = true;
            return SPHX_OK;
        }
        HIP_TRY(c, hipEventSynchronize(c->moved_ready));
        rebuild = moved_flag(c);
        if (rebuild) c->last_life = c->list_age;
    }
    if (rebuild) c->sorted = c->listed = false;
    return SPHX_OK;
}

// Whether to enqueue the passes of a tick before its moved flag is read: a
// stale list costs a discarded density and force pass, so only while the
// list is young compared with the previous one's lifetime (younger than it
// by more than two ticks), or, before any list has gone stale, once it has
// survived 4 ticks.
bool speculate(const SphxCtx* c) {
    if (!c->listed || c->kp.skin <= 0.0 || c->spec_mode == 2) return false;
    if (c->spec_mode == 1) return true;
    const uint32_t age = c->list_age + 1;  // after this tick's position update
    return c->last_life ? age + 2 < c->last_life : age >= 4;
}

// Reads the flag of a deferred position update.  If a particle moved more
// than skin/2, what was computed on the kept list since is discarded (redo).
int settle_moved(SphxCtx* c, bool& redo) {
    redo = false;
    if (!c->moved_pending) return SPHX_OK;
    c->moved_pending = false;
    HIP_TRY(c, hipEventSynchronize(c->moved_ready));
    if (moved_flag(c)) {
        c->last_life = c->list_age;
        c->sorted = c->listed = false;
        c->dens = c->sp = c->acc = false;
        redo = true;
    }
    return SPHX_OK;
}

int do_vv_velocity(SphxCtx* c) {
    if (!c->acc) return fail(c, SPHX_E_STATE, "vv_velocity needs accelerations (call sphx_forces)");
    {
        Timer t(c, kIntegrate);
        launch_vv_velocity(c->P(), c->D, c->n, c->kp, c->mixed(), c->X, c->stream);
    }
    c->acc = false;  // a is a function of v, which just changed
    if (c->energy()) c->pres = c->sp = false;  // and P of u, which did too
    return launch_check(c);
}

void swap3(double*& a, double*& b, double*& cc, double** t) {
    std::swap(a, t[0]);
    std::swap(b, t[1]);
    std::swap(cc, t[2]);
}

// The Jacobi passes write into D.tmp3 / tmp1 and swap those with the state
// arrays; the other copy of Particles keeps pointing at buffers the next
// reorder overwrites, so both copies are kept consistent by swapping in place.
int do_xsph(SphxCtx* c) {
    int rc;
    if (!c->dens && (rc = do_density(c))) return rc;
    {
        Timer t(c, kMhd);
        launch_xsph(c->P(), c->D, c->n, c->grid, c->L, c->kp, c->mixed(), c->X, c->stream);
    }
    Particles& P = c->P();
    swap3(P.vx, P.vy, P.vz, c->D.tmp3);
    c->acc = false;
    return launch_check(c);
}

int do_induction(SphxCtx* c) {
    int rc;
    if (!c->has_mhd) return fail(c, SPHX_E_STATE, "no magnetic field was uploaded");
    if (!c->dens && (rc = do_density(c))) return rc;
    {
        Timer t(c, kMhd);
        launch_induction(c->P(), c->D, c->n, c->grid, c->L, c->kp, c->mixed(), c->X, c->B, c->S.cell_start,
                         c->stream);
    }
    Particles& P = c->P();
    swap3(P.bx, P.by, P.bz, c->D.tmp3);
    std::swap(c->D.recb, c->D.recb_next);
    c->acc = false;
    c->srec = false;
    return launch_check(c);
}

int do_div_clean(SphxCtx* c) {
    int rc;
    if (!c->has_mhd) return fail(c, SPHX_E_STATE, "no magnetic field was uploaded");
    if (!c->dens && (rc = do_density(c))) return rc;
    {
        Timer t(c, kMhd);
        launch_div_clean(c->P(), c->D, c->n, c->grid, c->L, c->kp, c->mixed(), c->X, c->B, c->S.cell_start,
                         c->stream);
    }
    Particles& P = c->P();
    swap3(P.bx, P.by, P.bz, c->D.tmp3);
    std::swap(P.psi, c->D.tmp1);
    std::swap(c->D.recb, c->D.recb_next);
    c->acc = false;
    c->srec = c->kp.srec != 0;  // the pass wrote the records of the new B
    return launch_check(c);
}

// ---- adjoint ---------------------------------------------------------------
Vec3 vec3(double* base, uint32_t n) { return Vec3{base, base + n, base + 2 * (size_t)n}; }
Vec3 pos3(Particles& P) { return Vec3{P.x, P.y, P.z}; }
Vec3 vel3(Particles& P) { return Vec3{P.vx, P.vy, P.vz}; }

int adj_alloc(SphxCtx* c, uint32_t n, uint32_t k, uint32_t m) {
    AdjBuffers& A = c->adj;
    if (n <= A.n && k <= A.k && m <= A.m) return SPHX_OK;
    A.release();
    const size_t N = n, nb = nblocks(n);
    auto get = [&](auto** p, size_t count) -> int {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(**p));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(c, SPHX_E_OOM, "adjoint: hipMalloc of %zu bytes failed", count * sizeof(**p));
        }
        A.allocs.push_back(q);
        *p = static_cast<std::remove_reference_t<decltype(*p)>>(q);
        return SPHX_OK;
    };
    int rc;
    if ((rc = get(&A.ckpt, (size_t)m * 6 * N)) || (rc = get(&A.tx, (size_t)(k + 1) * 3 * N)) ||
        (rc = get(&A.tv, (size_t)(k + 1) * 3 * N)) || (rc = get(&A.tvp, (size_t)k * 3 * N)) ||
        (rc = get(&A.flip, (size_t)k * N)) || (rc = get(&A.fin, 6 * N)) || (rc = get(&A.adj, 12 * N)) ||
        (rc = get(&A.ab4, N)) || (rc = get(&A.rhob, N)) || (rc = get(&A.target, 3 * N)) ||
        (rc = get(&A.part, nb)) || (rc = get(&A.alpha_acc, vjp_blocks(n))) || (rc = get(&A.spec, 6 * N)) ||
        (rc = get(&A.alpha_spec, vjp_blocks(n)))) {
        A.release();
        return rc;
    }
    A.n = n;
    A.k = k;
    A.m = m;
    return SPHX_OK;
}

// state (upload order) -> working arrays (current sorted order); x may be
// NULL.  The neighbour list is kept while no particle is further than skin/2
// from where it was at the list's build.
// defer: as for do_vv_position, the moved flag is read by settle_moved.
int adj_load(SphxCtx* c, const double* x3, const double* v3, bool defer = false) {
    const uint32_t n = c->n;
    Particles& P = c->P();
    c->acc = false;
    // the moved flag is tested against the list's build positions in a new
    // epoch (the replaced positions start it afresh)
    const bool track = x3 && c->listed && c->kp.skin > 0.0;
    if (track) ++c->epoch;
    MovedFlag mf = moved_flag_args(c);
    if (!track) mf.dev = nullptr;
    launch_load_state(P, c->D, n, c->kp, c->grid, c->S.skey, c->mixed(), c->X, x3, v3, mf, c->stream);
    if (x3) {
        c->dens = c->sp = c->acc = false;
        if (track) {
            HIP_TRY(c, hipEventRecord(c->moved_ready, c->stream));
            if (defer) {
                c->moved_pending = true;
            } else {
                HIP_TRY(c, hipEventSynchronize(c->moved_ready));
                if (moved_flag(c)) c->sorted = c->listed = false;
            }
        } else {
            c->sorted = c->listed = false;
        }
    }
    return launch_check(c);
}
void adj_save(SphxCtx* c, double* x3, double* v3) {
    const uint32_t n = c->n;
    Particles& P = c->P();
    if (x3 && v3) launch_scatter6(P.orig, n, pos3(P), vec3(x3, n), vel3(P), vec3(v3, n), c->stream);
    else if (x3) launch_scatter3(P.orig, n, pos3(P), vec3(x3, n), c->stream);
    else if (v3) launch_scatter3(P.orig, n, vel3(P), vec3(v3, n), c->stream);
}

// Density and the second force evaluation of a tick at the new positions,
// a* = A(x_{n+1}, v_n).  After a deferred position update they run on the
// kept list before the moved flag is read, and again on the rebuilt list if
// a particle did move skin/2 (they write only derived arrays).
int second_evaluation(SphxCtx* c) {
    int rc;
    for (bool redo = true; redo;) {
        if ((rc = ensure_list(c)) || (rc = do_density(c)) || (rc = do_forces(c, false))) {
            c->moved_pending = false;
            c->sorted = c->listed = false;
            c->dens = c->sp = c->acc = false;
            return rc;
        }
        if ((rc = settle_moved(c, redo))) return rc;
    }
    return SPHX_OK;
}

// One tick of sphx_step (no XSPH / MHD); records the reflections and v' when asked.
int adj_tick(SphxCtx* c, uint8_t* flip, double* vp3) {
    int rc;
    if ((rc = do_forces(c, c->sp && c->dens && !c->energy()))) return rc;
    if ((rc = do_vv_position(c, flip, speculate(c)))) return rc;
    if (vp3) adj_save(c, nullptr, vp3);
    if ((rc = second_evaluation(c))) return rc;
    return do_vv_velocity(c);
}

int check_overflow(SphxCtx* c) {
    uint32_t ovf[2] = {0, 0};
    HIP_TRY(c, hipMemcpyAsync(ovf, c->L.overflow, sizeof ovf, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (ovf[1])
        return fail(c, SPHX_E_CAPACITY,
                    "a particle has %u neighbours, more than max_neighbours = %d; create the context with a "
                    "larger max_neighbours", ovf[0], c->L.ks);
    return SPHX_OK;
}

}  // namespace

// =========================================================================
extern "C" {

void sphx_default_params(SphxParams* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->dt = 1.0 / 60.0;
    p->sector_size = 1e12;
    p->rest_density = 1000.0;
    p->speed_of_sound = 1500.0;
    p->gamma = 7.0;
    p->eos = SPHX_EOS_TAIT;
    p->visc_convention = SPHX_VISC_REFERENCE;
    p->poly_k = 1.0;
    p->iso_cs2 = 1.0;
    p->visc_alpha = 0.1;
    p->visc_beta = 0.2;
    p->xsph_epsilon = 0.5;
    p->half_extent = 10.0;
    p->activation_distance = 0.5;
    p->stiffness = 1e6;
    p->resistivity = 1e-4;
    p->cleaning_speed = 100.0;
    p->cleaning_damping = 0.5;
    p->precision = SPHX_PREC_FP64;
    p->max_neighbours = 0;
    p->neighbour_skin = -1.0;
}

int sphx_params_from_uniform(const SphxUniformData* u, SphxParams* p) {
    if (!u || !p) return SPHX_E_INVALID;
    if (!(u->dt > 0.0f) || !(u->rest_density > 0.0f) || !(u->speed_of_sound > 0.0f) || !(u->gamma > 0.0f))
        return SPHX_E_INVALID;
    p->dt = (double)u->dt;
    p->rest_density = (double)u->rest_density;
    p->speed_of_sound = (double)u->speed_of_sound;
    p->gamma = (double)u->gamma;
    p->visc_alpha = (double)u->alpha_visc;
    p->visc_beta = (double)u->beta_visc;
    return SPHX_OK;
}

int sphx_state_from_vec4(const SphxUniformData* u, const float* pos4, const float* vel4, SphxState* st) {
    if (!u || !st) return SPHX_E_INVALID;
    const uint32_t n = u->num_particles;
    if (!(u->mass > 0.0f) || !(u->h > 0.0f)) return SPHX_E_INVALID;
    if (n && (!pos4 || !vel4 || !st->x || !st->y || !st->z || !st->vx || !st->vy || !st->vz || !st->m || !st->h))
        return SPHX_E_INVALID;
    for (uint32_t i = 0; i < n; ++i) {
        const float* a = pos4 + 4 * (size_t)i;
        const float* b = vel4 + 4 * (size_t)i;
        st->x[i] = a[0]; st->y[i] = a[1]; st->z[i] = a[2];
        st->vx[i] = b[0]; st->vy[i] = b[1]; st->vz[i] = b[2];
        st->m[i] = u->mass;
        st->h[i] = u->h;
        if (st->id) st->id[i] = i;
    }
    st->n = n;
    return SPHX_OK;
}

int sphx_device_count(int* count) {
    if (!count) return SPHX_E_INVALID;
    *count = 0;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return SPHX_E_NODEVICE;
    }
    *count = c;
    return SPHX_OK;
}

const char* sphx_last_error(const SphxCtx* ctx) {
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

int sphx_create(int device, uint32_t capacity, const SphxParams* params, SphxCtx** out) {
    if (!out) return SPHX_E_INVALID;
    *out = nullptr;
    int rc = check_params(nullptr, params);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return SPHX_E_NODEVICE;
    }
    if (device < 0 || device >= ndev) return SPHX_E_NODEVICE;
    if (capacity > kMaxParticles) return SPHX_E_CAPACITY;
    SphxCtx* c = new (std::nothrow) SphxCtx();
    if (!c) return SPHX_E_OOM;
    c->device = device;
    c->cap = capacity ? capacity : 1;
    c->prm = *params;
    make_kernel_params(c);
    // a multiple of 4: lists are read and written in groups of four entries
    c->L.ks = params->max_neighbours > 0 ? (params->max_neighbours + 3) / 4 * 4 : kDefaultNeighbours;
    if (c->L.ks > kMaxNeighbours) {
        delete c;
        return SPHX_E_INVALID;
    }
    auto bail = [&](int code) {
        free_all(c);
        if (c->bricks_ready) (void)hipEventDestroy(c->bricks_ready);
        if (c->moved_ready) (void)hipEventDestroy(c->moved_ready);
        if (c->bricks_host) (void)hipHostFree(c->bricks_host);
        if (c->moved_host) (void)hipHostFree(c->moved_host);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(SPHX_E_HIP);
    if (resolve_scratch(device) != SPHX_OK) return bail(SPHX_E_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(SPHX_E_HIP);
    if (hipDeviceGetAttribute(&c->B.cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return bail(SPHX_E_HIP);
    brick_defaults(c);
    // SPHX_SPECULATE=always|never overrides when a tick's second evaluation
    // is enqueued before its moved flag is read (the results do not change)
    if (const char* e = getenv("SPHX_SPECULATE")) c->spec_mode = !strcmp(e, "always") ? 1 : !strcmp(e, "never") ? 2 : 0;
    if (hipHostMalloc((void**)&c->bricks_host, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        c->bricks_host = nullptr;
        return bail(SPHX_E_OOM);
    }
    if (hipEventCreateWithFlags(&c->bricks_ready, hipEventDisableTiming) != hipSuccess) {
        c->bricks_ready = nullptr;
        return bail(SPHX_E_HIP);
    }
    if (hipEventCreateWithFlags(&c->moved_ready, hipEventDisableTiming) != hipSuccess) {
        c->moved_ready = nullptr;
        return bail(SPHX_E_HIP);
    }
    if (hipHostMalloc((void**)&c->moved_host, sizeof(uint32_t), hipHostMallocCoherent) != hipSuccess) {
        c->moved_host = nullptr;
        return bail(SPHX_E_OOM);
    }
    *c->moved_host = 0u;
    const size_t N = c->cap;
    for (int b = 0; b < 2; ++b) {
        Particles& P = c->A[b];
        double** dd[] = {&P.x, &P.y, &P.z, &P.vx, &P.vy, &P.vz, &P.apx, &P.apy, &P.apz,
                         &P.m, &P.h, &P.bx, &P.by, &P.bz, &P.psi, &P.q, &P.u, &P.dup};
        for (double** d : dd)
            if (dalloc(c, d, N)) return bail(SPHX_E_OOM);
        if (dalloc(c, &P.id, N) || dalloc(c, &P.orig, N) || dalloc(c, &P.pos, N)) return bail(SPHX_E_OOM);
    }
    Derived& D = c->D;
    double** dd[] = {&D.rho, &D.p, &D.pr2, &D.ax, &D.ay, &D.az, &D.spx, &D.spy, &D.spz,
                     &D.tmp3[0], &D.tmp3[1], &D.tmp3[2], &D.tmp1, &D.du};
    for (double** d : dd)
        if (dalloc(c, d, N)) return bail(SPHX_E_OOM);
    if (dalloc(c, &D.r64, N) || dalloc(c, &D.r32, N) || dalloc(c, &D.recb, N) || dalloc(c, &D.recb_next, N) ||
        dalloc(c, &D.recs, N) ||
        dalloc(c, &D.xref, N))
        return bail(SPHX_E_OOM);
    {
        Mixed& X = c->X;
        if (dalloc(c, &X.p4, N) || dalloc(c, &X.m, N)) return bail(SPHX_E_OOM);
    }
    if (dalloc(c, &c->S.key, N) || dalloc(c, &c->S.rank, N) || dalloc(c, &c->S.perm, N) ||
        dalloc(c, &c->S.skey, N) || dalloc(c, &c->S.tmp, N))
        return bail(SPHX_E_OOM);
    c->B.skey = c->S.skey;
    if (dalloc(c, &c->L.ntot, N) || dalloc(c, &c->L.overflow, 4) ||
        dalloc(c, &c->L.idx, ((N + 63) / 64) * 64 * (size_t)c->L.ks) ||
        dalloc(c, &c->L.s16, ((N + 63) / 64) * 64 * (size_t)c->L.ks))
        return bail(SPHX_E_OOM);
    if (dalloc(c, &c->energy_partial, (size_t)energy_blocks(c->cap) * 9)) return bail(SPHX_E_OOM);
    if (dalloc(c, &c->pair_counter, 1)) return bail(SPHX_E_OOM);
    launch_fill_u32(c->L.overflow, 0u, 4, c->stream);
    launch_fill_u32(reinterpret_cast<uint32_t*>(c->L.ntot), 0u, (N + 1) / 2, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(SPHX_E_HIP);
    *out = c;
    return SPHX_OK;
}

void sphx_destroy(SphxCtx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& t : c->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (hipEvent_t e : c->free_events) (void)hipEventDestroy(e);
    if (c->bricks_ready) (void)hipEventDestroy(c->bricks_ready);
    if (c->moved_ready) (void)hipEventDestroy(c->moved_ready);
    if (c->bricks_host) (void)hipHostFree(c->bricks_host);
    if (c->moved_host) (void)hipHostFree(c->moved_host);
    c->adj.release();
    free_all(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int sphx_set_params(SphxCtx* c, const SphxParams* p) {
    if (!c) return SPHX_E_INVALID;
    int rc = check_params(c, p);
    if (rc) return rc;
    // the list capacity is fixed at sphx_create, where max_neighbours was
    // rounded up to a multiple of 4 (0 or less here keeps it); a failed call
    // leaves the context as it was
    if (p->max_neighbours > 0 && (p->max_neighbours + 3) / 4 * 4 != c->L.ks)
        return fail(c, SPHX_E_INVALID, "max_neighbours is fixed at sphx_create (capacity %d, asked %d)", c->L.ks,
                    p->max_neighbours);
    if (p->eos == SPHX_EOS_ADIABATIC && c->n && !c->has_u)
        return fail(c, SPHX_E_STATE, "the adiabatic EOS needs the internal energy u (upload it)");
    const SphxParams old = c->prm;
    c->prm = *p;
    make_kernel_params(c);
    // a change of grid-relevant parameters needs a new upload
    bool grid_change = old.boundary_flags != p->boundary_flags || old.cell_size != p->cell_size ||
                       old.neighbour_skin != p->neighbour_skin;
    for (int k = 0; k < 3; ++k)
        grid_change |= old.box_lo[k] != p->box_lo[k] || old.box_hi[k] != p->box_hi[k] ||
                       old.grid_lo[k] != p->grid_lo[k] || old.grid_hi[k] != p->grid_hi[k];
    if (grid_change && c->n) c->grid_set = false;
    // derived quantities depend on the EOS, viscosity and walls; the
    // precision form also changes the neighbour predicate and f32 copies
    c->dens = c->sp = c->acc = false;
    if (old.precision != p->precision) c->sorted = c->listed = false;
    return SPHX_OK;
}

}  // extern "C"

namespace {

// What a checkpoint load sets beside the state of sphx_upload: the grid of the
// saved context, the upload slot of each particle and the previous
// accelerations (all in the order of the given state).
struct UploadExtras {
    const Grid* grid = nullptr;
    const uint32_t* orig = nullptr;
    const double* ap[3] = {nullptr, nullptr, nullptr};
};

int do_upload(SphxCtx* c, const SphxState* h, const UploadExtras* ex) {
    if (!c || !h) return SPHX_E_INVALID;
    if (h->n > c->cap) return fail(c, SPHX_E_CAPACITY, "n = %u exceeds capacity %u", h->n, c->cap);
    if (h->n && (!h->x || !h->y || !h->z || !h->vx || !h->vy || !h->vz || !h->m || !h->h))
        return fail(c, SPHX_E_INVALID, "x, y, z, vx, vy, vz, m and h are required");
    for (uint32_t i = 0; i < h->n; ++i) {
        const double vals[8] = {h->x[i], h->y[i], h->z[i], h->vx[i], h->vy[i], h->vz[i], h->m[i], h->h[i]};
        for (double v : vals)
            if (!std::isfinite(v)) return fail(c, SPHX_E_NONFINITE, "non-finite input at particle %u", i);
        if (!(h->m[i] > 0.0)) return fail(c, SPHX_E_INVALID, "mass must be > 0 (particle %u)", i);
        if (!(h->h[i] > 0.0)) return fail(c, SPHX_E_INVALID, "smoothing radius must be > 0 (particle %u)", i);
    }
    if (c->prm.eos == SPHX_EOS_ADIABATIC && h->n && !h->u)
        return fail(c, SPHX_E_INVALID, "the adiabatic EOS needs the internal energy u");
    if (h->u)
        for (uint32_t i = 0; i < h->n; ++i)
            if (!std::isfinite(h->u[i])) return fail(c, SPHX_E_NONFINITE, "non-finite u at particle %u", i);
    const bool mhd = h->bx && h->by && h->bz;
    if ((h->bx || h->by || h->bz) && !mhd) return fail(c, SPHX_E_INVALID, "bx, by, bz go together");
    if (h->q && !mhd) return fail(c, SPHX_E_INVALID, "charge needs a magnetic field");
    HIP_TRY(c, hipSetDevice(c->device));
    c->n = h->n;
    c->uniform = h->n > 0;
    for (uint32_t i = 1; i < h->n && c->uniform; ++i)
        c->uniform = h->m[i] == h->m[0] && h->h[i] == h->h[0];
    c->m0 = h->n ? h->m[0] : 0.0;
    c->h0 = h->n ? h->h[0] : 0.0;
    c->has_mhd = mhd;
    make_kernel_params(c);
    int rc = setup_grid(c, h, ex ? ex->grid : nullptr);
    if (rc) return rc;
    c->has_charge = h->q != nullptr;
    c->has_u = h->u != nullptr;
    c->cur = 0;
    Particles& P = c->A[0];
    const size_t bytes = sizeof(double) * h->n;
    if (h->n) {
        const double* src[] = {h->x, h->y, h->z, h->vx, h->vy, h->vz, h->m, h->h};
        double* dst[] = {P.x, P.y, P.z, P.vx, P.vy, P.vz, P.m, P.h};
        for (int k = 0; k < 8; ++k) HIP_TRY(c, hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyHostToDevice, c->stream));
        std::vector<double> zero(h->n, 0.0);
        double* z[] = {P.apx, P.apy, P.apz};
        for (int k = 0; k < 3; ++k)
            HIP_TRY(c, hipMemcpyAsync(z[k], ex && ex->ap[k] ? ex->ap[k] : zero.data(), bytes, hipMemcpyHostToDevice,
                                      c->stream));
        if (mhd) {
            HIP_TRY(c, hipMemcpyAsync(P.bx, h->bx, bytes, hipMemcpyHostToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(P.by, h->by, bytes, hipMemcpyHostToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(P.bz, h->bz, bytes, hipMemcpyHostToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(P.psi, h->psi ? h->psi : zero.data(), bytes, hipMemcpyHostToDevice, c->stream));
        }
        if (h->q) HIP_TRY(c, hipMemcpyAsync(P.q, h->q, bytes, hipMemcpyHostToDevice, c->stream));
        if (h->u) {
            HIP_TRY(c, hipMemcpyAsync(P.u, h->u, bytes, hipMemcpyHostToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(P.dup, zero.data(), bytes, hipMemcpyHostToDevice, c->stream));
        }
        if (h->id) HIP_TRY(c, hipMemcpyAsync(P.id, h->id, sizeof(uint32_t) * h->n, hipMemcpyHostToDevice, c->stream));
        else launch_iota_u32(P.id, h->n, c->stream);
        if (ex && ex->orig)
            HIP_TRY(c, hipMemcpyAsync(P.orig, ex->orig, sizeof(uint32_t) * h->n, hipMemcpyHostToDevice, c->stream));
        else
            launch_iota_u32(P.orig, h->n, c->stream);
        HIP_TRY(c, hipStreamSynchronize(c->stream));  // `zero` goes out of scope
    }
    launch_fill_u32(c->L.overflow, 0u, 4, c->stream);
    c->sorted = c->listed = c->dens = c->sp = c->acc = false;
    c->ticks = 0;
    return launch_check(c);
}

}  // namespace

extern "C" {

int sphx_upload(SphxCtx* c, const SphxState* h) { return do_upload(c, h, nullptr); }

int sphx_download(SphxCtx* c, SphxState* h, uint32_t fields) {
    if (!c || !h) return SPHX_E_INVALID;
    if (h->n != c->n) return fail(c, SPHX_E_INVALID, "state.n = %u but the context holds %u", h->n, c->n);
    HIP_TRY(c, hipSetDevice(c->device));
    if ((fields & (SPHX_F_RHO | SPHX_F_P)) && !c->dens)
        return fail(c, SPHX_E_STATE, "density is not valid for the current positions (call sphx_density)");
    if ((fields & SPHX_F_ACC) && !c->acc)
        return fail(c, SPHX_E_STATE, "accelerations are not valid for the current state (call sphx_forces)");
    if ((fields & (SPHX_F_B | SPHX_F_PSI)) && !c->has_mhd)
        return fail(c, SPHX_E_STATE, "no magnetic field was uploaded");
    if ((fields & SPHX_F_U) && !c->has_u) return fail(c, SPHX_E_STATE, "no internal energy was uploaded");
    const uint32_t n = c->n;
    if (!n) return SPHX_OK;
    Particles& P = c->P();
    if ((fields & SPHX_F_P) && !c->pres) {  // P of the current internal energy
        launch_eos(P, c->D, n, c->kp, c->mixed(), c->stream);
        int rc0;
        if ((rc0 = launch_check(c))) return rc0;
        c->pres = true;
    }
    std::vector<uint32_t> orig(n);
    HIP_TRY(c, hipMemcpyAsync(orig.data(), P.orig, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    std::vector<double> buf(n);
    auto fetch = [&](const double* dev, double* dst) -> int {
        if (!dst) return SPHX_OK;
        HIP_TRY(c, hipMemcpyAsync(buf.data(), dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (uint32_t k = 0; k < n; ++k) dst[orig[k]] = buf[k];
        return SPHX_OK;
    };
    int rc = SPHX_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (fields & SPHX_F_POS) { if ((rc = fetch(P.x, h->x)) || (rc = fetch(P.y, h->y)) || (rc = fetch(P.z, h->z))) return rc; }
    if (fields & SPHX_F_VEL) { if ((rc = fetch(P.vx, h->vx)) || (rc = fetch(P.vy, h->vy)) || (rc = fetch(P.vz, h->vz))) return rc; }
    if (fields & SPHX_F_RHO) { if ((rc = fetch(c->D.rho, h->rho))) return rc; }
    if (fields & SPHX_F_P) { if ((rc = fetch(c->D.p, h->p))) return rc; }
    if (fields & SPHX_F_ACC) { if ((rc = fetch(c->D.ax, h->ax)) || (rc = fetch(c->D.ay, h->ay)) || (rc = fetch(c->D.az, h->az))) return rc; }
    if (fields & SPHX_F_B) { if ((rc = fetch(P.bx, h->bx)) || (rc = fetch(P.by, h->by)) || (rc = fetch(P.bz, h->bz))) return rc; }
    if (fields & SPHX_F_PSI) { if ((rc = fetch(P.psi, h->psi))) return rc; }
    if (fields & SPHX_F_U) { if ((rc = fetch(P.u, h->u))) return rc; }
    return SPHX_OK;
}

int sphx_update(SphxCtx* c, const SphxState* h, uint32_t fields) {
    if (!c || !h) return SPHX_E_INVALID;
    if (!c->grid_set) return fail(c, SPHX_E_STATE, "upload particles first");
    if (h->n != c->n) return fail(c, SPHX_E_INVALID, "state.n = %u but the context holds %u", h->n, c->n);
    if (fields & ~(uint32_t)(SPHX_F_VEL | SPHX_F_B | SPHX_F_PSI | SPHX_F_U))
        return fail(c, SPHX_E_INVALID, "sphx_update takes SPHX_F_VEL, SPHX_F_B, SPHX_F_PSI and SPHX_F_U (positions, "
                                       "masses and radii change through sphx_upload)");
    const bool vel = fields & SPHX_F_VEL, b = fields & SPHX_F_B, psi = fields & SPHX_F_PSI, ue = fields & SPHX_F_U;
    if ((b || psi) && !c->has_mhd) return fail(c, SPHX_E_STATE, "no magnetic field was uploaded");
    if (ue && !c->has_u) return fail(c, SPHX_E_STATE, "no internal energy was uploaded");
    if ((vel && (!h->vx || !h->vy || !h->vz)) || (b && (!h->bx || !h->by || !h->bz)) || (psi && !h->psi) ||
        (ue && !h->u))
        return fail(c, SPHX_E_INVALID, "an array of the requested fields is missing");
    const uint32_t n = c->n;
    for (uint32_t i = 0; i < n; ++i) {
        const bool ok = (!vel || (std::isfinite(h->vx[i]) && std::isfinite(h->vy[i]) && std::isfinite(h->vz[i]))) &&
                        (!b || (std::isfinite(h->bx[i]) && std::isfinite(h->by[i]) && std::isfinite(h->bz[i]))) &&
                        (!psi || std::isfinite(h->psi[i])) && (!ue || std::isfinite(h->u[i]));
        if (!ok) return fail(c, SPHX_E_NONFINITE, "non-finite input at particle %u", i);
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (!n || !fields) return SPHX_OK;
    Particles& P = c->P();
    std::vector<uint32_t> orig(n);
    HIP_TRY(c, hipMemcpyAsync(orig.data(), P.orig, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // sorted slot k holds the particle uploaded at orig[k]
    std::vector<double> buf(n);
    auto put = [&](const double* src, double* dev) -> int {
        for (uint32_t k = 0; k < n; ++k) buf[k] = src[orig[k]];
        HIP_TRY(c, hipMemcpyAsync(dev, buf.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return SPHX_OK;
    };
    int rc = SPHX_OK;
    if (vel && ((rc = put(h->vx, P.vx)) || (rc = put(h->vy, P.vy)) || (rc = put(h->vz, P.vz)))) return rc;
    if (b && ((rc = put(h->bx, P.bx)) || (rc = put(h->by, P.by)) || (rc = put(h->bz, P.bz)))) return rc;
    if (psi && (rc = put(h->psi, P.psi))) return rc;
    if (ue && (rc = put(h->u, P.u))) return rc;
    if (b || psi) {
        // the passes read B and psi from the (bx, by, bz, psi) records
        std::vector<double> bx(n), by(n), bz(n), ps(n);
        HIP_TRY(c, hipMemcpyAsync(bx.data(), P.bx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(by.data(), P.by, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(bz.data(), P.bz, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(ps.data(), P.psi, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        std::vector<double> rec(4 * (size_t)n);
        for (uint32_t k = 0; k < n; ++k) {
            rec[4 * (size_t)k] = bx[k];
            rec[4 * (size_t)k + 1] = by[k];
            rec[4 * (size_t)k + 2] = bz[k];
            rec[4 * (size_t)k + 3] = ps[k];
        }
        HIP_TRY(c, hipMemcpyAsync(c->D.recb, rec.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    // the gather records hold velocities, and a, the pressure sums and the
    // MHD terms depend on v and B: all recomputed at the next use
    c->dens = c->sp = c->acc = false;
    return SPHX_OK;
}

#define SPHX_ENTER(c)                                        \
    if (!(c)) return SPHX_E_INVALID;                         \
    if (!(c)->grid_set) return fail(c, SPHX_E_STATE, "upload particles first"); \
    HIP_TRY(c, hipSetDevice((c)->device))

int sphx_neighbours(SphxCtx* c) {
    SPHX_ENTER(c);
    return do_neighbours(c);
}
int sphx_density(SphxCtx* c) {
    SPHX_ENTER(c);
    return do_density(c);
}
int sphx_forces(SphxCtx* c, uint32_t flags) {
    SPHX_ENTER(c);
    if (flags & ~1u) return fail(c, SPHX_E_INVALID, "unknown force flags 0x%x", flags);
    return do_forces(c, (flags & SPHX_FORCES_VISC_ONLY) != 0);
}
int sphx_vv_position(SphxCtx* c) {
    SPHX_ENTER(c);
    return do_vv_position(c);
}
int sphx_vv_velocity(SphxCtx* c) {
    SPHX_ENTER(c);
    return do_vv_velocity(c);
}
int sphx_xsph(SphxCtx* c) {
    SPHX_ENTER(c);
    return do_xsph(c);
}
int sphx_mhd_induction(SphxCtx* c) {
    SPHX_ENTER(c);
    return do_induction(c);
}
int sphx_div_clean(SphxCtx* c) {
    SPHX_ENTER(c);
    return do_div_clean(c);
}

int sphx_step(SphxCtx* c, uint32_t nsteps) {
    SPHX_ENTER(c);
    int rc;
    for (uint32_t t = 0; t < nsteps; ++t) {
        // a_n = A(x_n, v_n): the pressure sum at x_n is still valid from the
        // previous tick's second evaluation unless positions were replaced
        // (never with the adiabatic EOS: P follows u, which vv_velocity moved).
        if ((rc = do_forces(c, c->sp && c->dens && !c->energy()))) return rc;
        // While the list is expected to stay valid, density and the second
        // force pass are enqueued before the host has read the moved flag,
        // so the GPU does not wait for the host; they only write derived
        // arrays, and if a particle did move skin/2 they are redone on the
        // rebuilt list.
        if ((rc = do_vv_position(c, nullptr, speculate(c)))) return rc;
        if ((rc = second_evaluation(c))) return rc;  // a* = A(x_{n+1}, v_n)
        if ((rc = do_vv_velocity(c))) return rc;
        if (c->prm.xsph_epsilon != 0.0 && (rc = do_xsph(c))) return rc;
        if (c->has_mhd && (c->prm.mhd_flags & SPHX_MHD_INDUCTION) && (rc = do_induction(c))) return rc;
        if (c->has_mhd && (c->prm.mhd_flags & SPHX_MHD_CLEANING) && (rc = do_div_clean(c))) return rc;
        ++c->ticks;
    }
    return check_overflow(c);
}

int sphx_adjoint(SphxCtx* c, uint32_t nsteps, uint32_t every, const double* tgx, const double* tgy,
                 const double* tgz, double* loss, double* gx, double* gy, double* gz, double* gvx, double* gvy,
                 double* gvz, double* galpha) {
    SPHX_ENTER(c);
    if (!tgx || !tgy || !tgz || !loss || !gvx || !gvy || !gvz || !galpha)
        return fail(c, SPHX_E_INVALID, "target, loss, grad_v and grad_alpha are required");
    if ((gx || gy || gz) && !(gx && gy && gz)) return fail(c, SPHX_E_INVALID, "grad_x, grad_y, grad_z go together");
    if (nsteps == 0 || c->n == 0) return fail(c, SPHX_E_INVALID, "sphx_adjoint needs particles and nsteps >= 1");
    if (c->mixed()) return fail(c, SPHX_E_INVALID, "sphx_adjoint runs in the FP64 form only");
    if (c->prm.boundary_flags & (SPHX_BND_PENALTY_CUBE | SPHX_BND_PERIODIC))
        return fail(c, SPHX_E_INVALID, "sphx_adjoint supports open and reflective boxes only");
    if (c->has_mhd || c->prm.xsph_epsilon != 0.0)
        return fail(c, SPHX_E_INVALID, "sphx_adjoint does not differentiate XSPH or MHD");
    if (c->energy()) return fail(c, SPHX_E_INVALID, "sphx_adjoint does not differentiate the adiabatic EOS");
    const uint32_t n = c->n;
    const uint32_t k = (every == 0 || every > nsteps) ? nsteps : every;
    const uint32_t M = (nsteps + k - 1) / k;
    int rc;
    if ((rc = adj_alloc(c, n, k, M))) return rc;
    AdjBuffers& A = c->adj;
    const size_t n3 = 3 * (size_t)n;
    HIP_TRY(c, hipMemcpyAsync(A.target, tgx, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(A.target + n, tgy, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(A.target + 2 * (size_t)n, tgz, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    const Vec3 xb = vec3(A.adj, n), vb = vec3(A.adj + n3, n), ab = vec3(A.adj + 2 * n3, n),
               apb = vec3(A.adj + 3 * n3, n);
    const Vec3 xs = vec3(A.spec, n), vs = vec3(A.spec + n3, n);
    const uint32_t last0 = (M - 1) * k;  // first tick of the last segment
    // forward: checkpoints every k ticks, the last segment's tape kept
    for (uint32_t t = 0; t < nsteps; ++t) {
        const uint32_t s = t % k;
        if (s == 0) adj_save(c, A.ckpt + (size_t)(t / k) * 6 * n, A.ckpt + (size_t)(t / k) * 6 * n + n3);
        if (t >= last0) {
            adj_save(c, A.tx + (t - last0) * n3, A.tv + (t - last0) * n3);
            rc = adj_tick(c, A.flip + (size_t)(t - last0) * n, A.tvp + (t - last0) * n3);
        } else {
            rc = adj_tick(c, nullptr, nullptr);
        }
        if (rc) return rc;
    }
    if ((rc = check_overflow(c))) return rc;
    adj_save(c, A.fin, A.fin + n3);
    // loss and dL/dx_N
    launch_adj_loss(c->P(), n, vec3(A.target, n), xb, A.part, c->stream);
    launch_fill_u32(reinterpret_cast<uint32_t*>(vb.x), 0u, 2 * n3, c->stream);
    launch_fill_u32(reinterpret_cast<uint32_t*>(A.alpha_acc), 0u, 2 * (size_t)vjp_blocks(n), c->stream);
    const double dt = c->kp.dt, hdt = c->kp.half_dt, hdt2 = c->kp.half_dt_sq;
    // reverse, one segment at a time
    for (uint32_t m = M; m-- > 0;) {
        const uint32_t t0 = m * k, len = std::min(k, nsteps - t0);
        if (m + 1 < M) {  // recompute the segment's tape from its checkpoint
            if ((rc = adj_load(c, A.ckpt + (size_t)m * 6 * n, A.ckpt + (size_t)m * 6 * n + n3))) return rc;
            for (uint32_t s = 0; s < len; ++s) {
                adj_save(c, A.tx + s * n3, A.tv + s * n3);
                if ((rc = adj_tick(c, A.flip + (size_t)s * n, A.tvp + s * n3))) return rc;
            }
        }
        for (uint32_t s = len; s-- > 0;) {
            // at (x_{t+1}, v'_t): positions sorted, listed and dense already
            if ((rc = adj_load(c, nullptr, A.tvp + s * n3))) return rc;
            launch_adj_seed(c->P().orig, n, hdt, vb, apb, A.ab4, c->stream);
            launch_vjp(c->P(), c->D, n, c->grid, c->L, c->kp, A.ab4, xb, vb, A.rhob, A.alpha_acc, c->stream);
            launch_adj_flip(n, A.flip + (size_t)s * n, dt, hdt2, xb, vb, apb, ab, c->stream);
            // at (x_t, v_t): enqueued before the moved flag of x_t is read,
            // into A.spec (stored, not added), and redone on a rebuilt list
            // if a particle moved skin/2; then added to xbar, vbar, alpha
            if ((rc = adj_load(c, A.tx + s * n3, A.tv + s * n3, true))) return rc;
            for (bool redo = true; redo;) {
                if ((rc = ensure_list(c)) || (rc = do_density(c))) {
                    c->moved_pending = false;
                    return rc;
                }
                launch_gather_ab(c->P().orig, n, ab, A.ab4, c->stream);
                launch_vjp(c->P(), c->D, n, c->grid, c->L, c->kp, A.ab4, xs, vs, A.rhob, A.alpha_spec, c->stream,
                           true);
                if ((rc = launch_check(c)) || (rc = settle_moved(c, redo))) return rc;
            }
            launch_adj_accumulate(n, xs, vs, xb, vb, A.alpha_spec, A.alpha_acc, vjp_blocks(n), c->stream);
        }
    }
    if ((rc = check_overflow(c))) return rc;
    // results, then the state after the last tick back in place
    const uint32_t nb = nblocks(n), nv = vjp_blocks(n);
    std::vector<double> part(nb), acc(nv);
    HIP_TRY(c, hipMemcpyAsync(part.data(), A.part, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(acc.data(), A.alpha_acc, sizeof(double) * nv, hipMemcpyDeviceToHost, c->stream));
    double* outs[6] = {gvx, gvy, gvz, gx, gy, gz};
    const double* srcs[6] = {vb.x, vb.y, vb.z, xb.x, xb.y, xb.z};
    for (int q = 0; q < 6; ++q)
        if (outs[q]) HIP_TRY(c, hipMemcpyAsync(outs[q], srcs[q], sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if ((rc = adj_load(c, A.fin, A.fin + n3))) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    double l = 0.0, ga = 0.0;
    for (uint32_t b = 0; b < nb; ++b) l += part[b];
    for (uint32_t b = 0; b < nv; ++b) ga += acc[b];
    *loss = l / (3.0 * (double)n);
    *galpha = ga;
    return launch_check(c);
}

int sphx_synchronize(SphxCtx* c) {
    if (!c) return SPHX_E_INVALID;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_overflow(c);
}

int sphx_export_pairs(SphxCtx* c, uint32_t* pairs, uint64_t cap_pairs, uint64_t* count) {
    SPHX_ENTER(c);
    if (!count) return fail(c, SPHX_E_INVALID, "count is NULL");
    int rc;
    if ((rc = ensure_list(c))) return rc;
    if ((rc = check_overflow(c))) return rc;
    // the lists hold every neighbour on both sides (and skin pairs): total/2
    // bounds the undirected pairs
    std::vector<uint16_t> cnt(c->n);
    if (c->n) HIP_TRY(c, hipMemcpy(cnt.data(), c->L.ntot, sizeof(uint16_t) * c->n, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (uint16_t v : cnt) total += v;
    const uint64_t np = total / 2;
    if (np > c->pair_buf_cap) {
        if ((rc = dalloc(c, &c->pair_buf, 2 * np))) return rc;
        c->pair_buf_cap = np;
    }
    launch_fill_u32(reinterpret_cast<uint32_t*>(c->pair_counter), 0u, 2, c->stream);
    launch_export_pairs(c->P(), c->n, c->grid, c->L, c->pair_buf, c->pair_counter,
                        c->pair_buf_cap, c->stream);
    if ((rc = launch_check(c))) return rc;
    unsigned long long got = 0;
    HIP_TRY(c, hipMemcpyAsync(&got, c->pair_counter, sizeof got, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint64_t ng = got;  // pairs with r < 2 h_avg (<= np)
    *count = ng;
    if (ng > cap_pairs) return fail(c, SPHX_E_CAPACITY, "%llu pairs do not fit in %llu", (unsigned long long)ng, (unsigned long long)cap_pairs);
    if (!ng) return SPHX_OK;
    if (!pairs) return fail(c, SPHX_E_INVALID, "pairs is NULL");
    std::vector<uint64_t> packed(ng);
    std::vector<uint32_t> raw(2 * ng);
    HIP_TRY(c, hipMemcpy(raw.data(), c->pair_buf, sizeof(uint32_t) * 2 * ng, hipMemcpyDeviceToHost));
    for (uint64_t k = 0; k < ng; ++k) packed[k] = ((uint64_t)raw[2 * k] << 32) | raw[2 * k + 1];
    std::sort(packed.begin(), packed.end());
    for (uint64_t k = 0; k < ng; ++k) {
        pairs[2 * k] = (uint32_t)(packed[k] >> 32);
        pairs[2 * k + 1] = (uint32_t)packed[k];
    }
    return SPHX_OK;
}

// ---- checkpoint / resume ---------------------------------------------------
// Layout (native little-endian, 8-byte aligned): CkptHeader, then n-element
// arrays in the context's sorted order: f64 x, y, z, the positions at the
// last list build xr, yr, zr, vx, vy, vz, apx, apy, apz, m, h, [bx, by, bz,
// psi], [q], [u], then u32 id and the upload slot.
namespace {
constexpr char kCkptMagic[8] = {'S', 'P', 'H', 'X', 'C', 'K', 'P', '1'};
constexpr uint32_t kCkptVersion = 1;
enum : uint32_t { kCkptMhd = 1, kCkptCharge = 2, kCkptU = 4 };
struct CkptHeader {
    char magic[8];
    uint32_t version, header_bytes, params_bytes, n;
    uint32_t flags, ks, list_age, last_life;
    uint64_t ticks;
    double lo[3], w[3], L[3];
    int32_t dims[3], periodic;
    SphxParams params;
};
uint32_t ckpt_flags(const SphxCtx* c) {
    return (c->has_mhd ? kCkptMhd : 0u) | (c->has_charge && c->has_mhd ? kCkptCharge : 0u) | (c->has_u ? kCkptU : 0u);
}
uint64_t ckpt_arrays(uint32_t flags) {  // f64 arrays per particle
    return 14u + ((flags & kCkptMhd) ? 4u : 0u) + ((flags & kCkptCharge) ? 1u : 0u) + ((flags & kCkptU) ? 1u : 0u);
}
uint64_t ckpt_bytes(uint32_t n, uint32_t flags) {
    return sizeof(CkptHeader) + (uint64_t)n * (8u * ckpt_arrays(flags) + 8u);
}
}  // namespace

int sphx_checkpoint_size(SphxCtx* c, uint64_t* bytes) {
    if (!c || !bytes) return SPHX_E_INVALID;
    if (!c->grid_set) return fail(c, SPHX_E_STATE, "upload particles first");
    *bytes = ckpt_bytes(c->n, ckpt_flags(c));
    return SPHX_OK;
}

int sphx_save(SphxCtx* c, void* buf, uint64_t cap, uint64_t* written) {
    SPHX_ENTER(c);
    const uint32_t n = c->n, flags = ckpt_flags(c);
    const uint64_t need = ckpt_bytes(n, flags);
    if (written) *written = need;
    if (!buf || cap < need) return fail(c, SPHX_E_CAPACITY, "a checkpoint needs %llu bytes", (unsigned long long)need);
    int rc;
    // the list and the sorted order are part of what a resume reproduces
    if ((rc = ensure_list(c))) return rc;
    CkptHeader h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = kCkptVersion;
    h.header_bytes = sizeof(CkptHeader);
    h.params_bytes = sizeof(SphxParams);
    h.n = n;
    h.flags = flags;
    h.ks = (uint32_t)c->L.ks;
    h.list_age = c->list_age;
    h.last_life = c->last_life;
    h.ticks = c->ticks;
    for (int k = 0; k < 3; ++k) {
        h.lo[k] = c->grid.lo[k];
        h.w[k] = c->grid.w[k];
        h.L[k] = c->grid.L[k];
        h.dims[k] = c->grid.dims[k];
    }
    h.periodic = c->grid.periodic;
    h.params = c->prm;
    unsigned char* out = static_cast<unsigned char*>(buf);
    std::memcpy(out, &h, sizeof h);
    double* f = reinterpret_cast<double*>(out + sizeof h);
    const size_t N = n;
    Particles& P = c->P();
    std::vector<double4> xr(n);
    if (n) HIP_TRY(c, hipMemcpyAsync(xr.data(), c->D.xref, sizeof(double4) * N, hipMemcpyDeviceToHost, c->stream));
    std::vector<const double*> src = {P.x, P.y, P.z};
    std::vector<const double*> rest = {P.vx, P.vy, P.vz, P.apx, P.apy, P.apz, P.m, P.h};
    if (flags & kCkptMhd) rest.insert(rest.end(), {P.bx, P.by, P.bz, P.psi});
    if (flags & kCkptCharge) rest.push_back(P.q);
    if (flags & kCkptU) rest.push_back(P.u);
    for (size_t k = 0; k < 3; ++k)
        if (n) HIP_TRY(c, hipMemcpyAsync(f + k * N, src[k], sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    for (size_t k = 0; k < rest.size(); ++k)
        if (n) HIP_TRY(c, hipMemcpyAsync(f + (6 + k) * N, rest[k], sizeof(double) * N, hipMemcpyDeviceToHost,
                                         c->stream));
    uint32_t* ids = reinterpret_cast<uint32_t*>(f + ckpt_arrays(flags) * N);
    if (n) {
        HIP_TRY(c, hipMemcpyAsync(ids, P.id, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(ids + N, P.orig, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < n; ++i) {
        f[3 * N + i] = xr[i].x;
        f[4 * N + i] = xr[i].y;
        f[5 * N + i] = xr[i].z;
    }
    return SPHX_OK;
}

int sphx_load(SphxCtx* c, const void* buf, uint64_t size) {
    if (!c) return SPHX_E_INVALID;
    if (!buf || size < sizeof(CkptHeader)) return fail(c, SPHX_E_INVALID, "not a checkpoint (too short)");
    CkptHeader h;
    std::memcpy(&h, buf, sizeof h);
    if (std::memcmp(h.magic, kCkptMagic, 8) || h.version != kCkptVersion || h.header_bytes != sizeof(CkptHeader) ||
        h.params_bytes != sizeof(SphxParams))
        return fail(c, SPHX_E_INVALID, "not a checkpoint of this library version");
    if (h.flags & ~(uint32_t)(kCkptMhd | kCkptCharge | kCkptU) || size != ckpt_bytes(h.n, h.flags))
        return fail(c, SPHX_E_INVALID, "checkpoint size does not match its header");
    if (h.n > c->cap) return fail(c, SPHX_E_CAPACITY, "checkpoint of %u particles exceeds capacity %u", h.n, c->cap);
    if ((int)h.ks != c->L.ks)
        return fail(c, SPHX_E_INVALID, "checkpoint lists hold %u neighbours, this context %d", h.ks, c->L.ks);
    int rc = check_params(c, &h.params);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    const uint32_t n = h.n;
    const size_t N = n;
    const unsigned char* in = static_cast<const unsigned char*>(buf);
    // the arrays may not be 8-byte aligned in the caller's buffer: copied out
    std::vector<double> f(ckpt_arrays(h.flags) * N);
    std::vector<uint32_t> ids(2 * N);
    std::memcpy(f.data(), in + sizeof h, sizeof(double) * f.size());
    std::memcpy(ids.data(), in + sizeof h + sizeof(double) * f.size(), sizeof(uint32_t) * ids.size());
    const uint32_t* orig = ids.data() + N;
    {
        std::vector<uint8_t> seen(N, 0);
        for (uint32_t i = 0; i < n; ++i) {
            if (orig[i] >= n || seen[orig[i]]) return fail(c, SPHX_E_INVALID, "checkpoint slot map is not a permutation");
            seen[orig[i]] = 1;
        }
    }
    Grid g{};
    for (int k = 0; k < 3; ++k) {
        g.lo[k] = h.lo[k];
        g.w[k] = h.w[k];
        g.inv_w[k] = 1.0 / h.w[k];
        g.L[k] = h.L[k];
        g.halfL[k] = 0.5 * h.L[k];
        g.invL[k] = h.periodic ? 1.0 / h.L[k] : 0.0;
        g.dims[k] = h.dims[k];
        if (!(h.w[k] > 0.0) || h.dims[k] < (h.periodic ? 3 : 1) || (h.periodic && !(h.L[k] > 0.0)))
            return fail(c, SPHX_E_INVALID, "checkpoint grid is invalid");
    }
    g.periodic = h.periodic;
    g.ncell = (uint32_t)((uint64_t)h.dims[0] * (uint64_t)h.dims[1] * (uint64_t)h.dims[2]);
    g.inv_nx = 1.0 / (double)g.dims[0];
    const SphxParams old = c->prm;
    c->prm = h.params;
    const double* a = f.data();
    // the state at the positions of the last list build, in the saved sorted
    // order: the sort reproduces that order and the build the saved list
    SphxState st{};
    st.n = n;
    st.x = const_cast<double*>(a + 3 * N); st.y = const_cast<double*>(a + 4 * N); st.z = const_cast<double*>(a + 5 * N);
    st.vx = const_cast<double*>(a + 6 * N); st.vy = const_cast<double*>(a + 7 * N); st.vz = const_cast<double*>(a + 8 * N);
    st.m = const_cast<double*>(a + 12 * N);
    st.h = const_cast<double*>(a + 13 * N);
    size_t k = 14;
    if (h.flags & kCkptMhd) {
        st.bx = const_cast<double*>(a + k++ * N); st.by = const_cast<double*>(a + k++ * N);
        st.bz = const_cast<double*>(a + k++ * N); st.psi = const_cast<double*>(a + k++ * N);
    }
    if (h.flags & kCkptCharge) st.q = const_cast<double*>(a + k++ * N);
    if (h.flags & kCkptU) st.u = const_cast<double*>(a + k++ * N);
    st.id = ids.data();
    UploadExtras ex;
    ex.grid = &g;
    ex.orig = orig;
    ex.ap[0] = a + 9 * N; ex.ap[1] = a + 10 * N; ex.ap[2] = a + 11 * N;
    if ((rc = do_upload(c, &st, &ex))) {
        c->prm = old;
        make_kernel_params(c);
        return rc;
    }
    if (!n) return SPHX_OK;
    if ((rc = do_neighbours(c)) || (rc = ensure_list(c))) return rc;
    // then the current positions (and velocities) through the upload slots,
    // with the moved-skin/2 test against the build positions
    std::vector<double> x3(3 * N), v3(3 * N);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t u = orig[i];
        for (int d = 0; d < 3; ++d) {
            x3[d * N + u] = a[d * N + i];
            v3[d * N + u] = a[(6 + d) * N + i];
        }
    }
    double* dx3 = nullptr;
    if ((rc = dalloc(c, &dx3, 6 * N))) return rc;
    HIP_TRY(c, hipMemcpyAsync(dx3, x3.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(dx3 + 3 * N, v3.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
    rc = adj_load(c, dx3, dx3 + 3 * N);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->allocs.pop_back();
    (void)hipFree(dx3);
    if (rc) return rc;
    c->list_age = h.list_age;
    c->last_life = h.last_life;
    c->ticks = h.ticks;
    return SPHX_OK;
}

int sphx_export_pairs_noalloc(SphxCtx* c, uint32_t* pairs, uint64_t cap_pairs, uint64_t* count) {
    SPHX_ENTER(c);
    if (!count) return fail(c, SPHX_E_INVALID, "count is NULL");
    void* base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
    if (!base) return fail(c, SPHX_E_STATE, "no scratch on device %d", c->device);
    int rc;
    if ((rc = ensure_list(c))) return rc;
    if ((rc = check_overflow(c))) return rc;
    uint32_t* out = static_cast<uint32_t*>(base);
    unsigned long long* counter = reinterpret_cast<unsigned long long*>(out + 2 * kNoAllocPairs);
    // the counter keeps the last call's value: cleared by a kernel first
    launch_fill_u32(reinterpret_cast<uint32_t*>(counter), 0u, 2, c->stream);
    launch_export_pairs(c->P(), c->n, c->grid, c->L, out, counter, kNoAllocPairs, c->stream);
    if ((rc = launch_check(c))) return rc;
    unsigned long long got = 0;
    HIP_TRY(c, hipMemcpyAsync(&got, counter, sizeof got, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    *count = got;
    if (got > kNoAllocPairs)
        return fail(c, SPHX_E_CAPACITY, "%llu pairs exceed the %llu the no-allocation export holds (use "
                    "sphx_export_pairs)", got, (unsigned long long)kNoAllocPairs);
    if (got > cap_pairs)
        return fail(c, SPHX_E_CAPACITY, "%llu pairs do not fit in %llu", got, (unsigned long long)cap_pairs);
    if (!got) return SPHX_OK;
    if (!pairs) return fail(c, SPHX_E_INVALID, "pairs is NULL");
    HIP_TRY(c, hipMemcpyAsync(pairs, out, sizeof(uint32_t) * 2 * got, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // ascending (id_a, id_b), sorted in place in the caller's array
    struct IdPair {
        uint32_t a, b;
    };
    IdPair* pp = reinterpret_cast<IdPair*>(pairs);
    std::sort(pp, pp + got, [](const IdPair& x, const IdPair& y) { return x.a != y.a ? x.a < y.a : x.b < y.b; });
    return SPHX_OK;
}

int sphx_export_candidates(SphxCtx* c, const double* radius, uint32_t* pairs, uint64_t cap_pairs,
                           uint64_t* count) {
    SPHX_ENTER(c);
    if (!count) return fail(c, SPHX_E_INVALID, "count is NULL");
    const uint32_t n = c->n;
    if (n && !radius) return fail(c, SPHX_E_INVALID, "bounding radii are required");
    for (uint32_t i = 0; i < n; ++i)
        if (!std::isfinite(radius[i]) || radius[i] < 0.0)
            return fail(c, SPHX_E_INVALID, "bounding radius %u is negative or not finite", i);
    const Grid& g = c->grid;
    int rc;
    if ((rc = ensure_list(c))) return rc;
    *count = 0;
    if (!n) return SPHX_OK;
    Particles& P = c->P();
    if (n > c->cand_r_cap) {
        if ((rc = dalloc(c, &c->cand_r, n))) return rc;
        c->cand_r_cap = n;
    }
    {
        std::vector<uint32_t> orig(n);
        HIP_TRY(c, hipMemcpyAsync(orig.data(), P.orig, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        std::vector<double> rs(n);
        for (uint32_t k = 0; k < n; ++k) rs[k] = radius[orig[k]];
        HIP_TRY(c, hipMemcpyAsync(c->cand_r, rs.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    unsigned long long got = 0;
    for (int pass = 0; pass < 2; ++pass) {
        launch_fill_u32(reinterpret_cast<uint32_t*>(c->pair_counter), 0u, 2, c->stream);
        launch_export_candidates(P, n, g, c->S.cell_start, c->cand_r, 0.5 * c->kp.skin, c->pair_buf,
                                 c->pair_counter, c->pair_buf_cap, c->stream);
        if ((rc = launch_check(c))) return rc;
        HIP_TRY(c, hipMemcpyAsync(&got, c->pair_counter, sizeof got, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (got <= c->pair_buf_cap) break;
        // the first pass counted them: grow the buffer and write them all
        if ((rc = dalloc(c, &c->pair_buf, 2 * (size_t)got))) return rc;
        c->pair_buf_cap = got;
    }
    const uint64_t ng = got;
    *count = ng;
    if (ng > cap_pairs)
        return fail(c, SPHX_E_CAPACITY, "%llu pairs do not fit in %llu", (unsigned long long)ng,
                    (unsigned long long)cap_pairs);
    if (!ng) return SPHX_OK;
    if (!pairs) return fail(c, SPHX_E_INVALID, "pairs is NULL");
    std::vector<uint32_t> raw(2 * ng);
    HIP_TRY(c, hipMemcpy(raw.data(), c->pair_buf, sizeof(uint32_t) * 2 * ng, hipMemcpyDeviceToHost));
    std::vector<uint64_t> packed(ng);
    for (uint64_t k = 0; k < ng; ++k) packed[k] = ((uint64_t)raw[2 * k] << 32) | raw[2 * k + 1];
    std::sort(packed.begin(), packed.end());
    for (uint64_t k = 0; k < ng; ++k) {
        pairs[2 * k] = (uint32_t)(packed[k] >> 32);
        pairs[2 * k + 1] = (uint32_t)packed[k];
    }
    return SPHX_OK;
}

int sphx_enable_timing(SphxCtx* c, int enable) {
    if (!c) return SPHX_E_INVALID;
    if (enable != 0 && enable != SPHX_TIMING_ALL && (enable & ~0xfe))
        return fail(c, SPHX_E_INVALID, "unknown timing slots 0x%x", (unsigned)enable);
    c->timing = enable == SPHX_TIMING_ALL ? 0xfeu : (uint32_t)enable;
    return SPHX_OK;
}

int sphx_kernel_times(SphxCtx* c, SphxTimes* out, int reset) {
    if (!c || !out) return SPHX_E_INVALID;
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = drain_timing(c);
    if (rc) return rc;
    out->sort_ms = c->ms[kSort];
    out->neighbour_ms = c->ms[kNbr];
    out->density_ms = c->ms[kDensity];
    out->force_ms = c->ms[kForce];
    out->visc_only_ms = c->ms[kViscOnly];
    out->integrate_ms = c->ms[kIntegrate];
    out->mhd_ms = c->ms[kMhd];
    for (int k = 0; k < kNumSlots; ++k) out->launches[k] = c->launches[k];
    if (reset) {
        for (int k = 0; k < kNumSlots; ++k) { c->ms[k] = 0; c->launches[k] = 0; }
    }
    return SPHX_OK;
}

int sphx_energy(SphxCtx* c, SphxEnergy* out) {
    SPHX_ENTER(c);
    if (!out) return fail(c, SPHX_E_INVALID, "out is NULL");
    int rc;
    if (!c->dens && (rc = do_density(c))) return rc;
    const uint32_t nb = energy_blocks(c->n);
    if (!c->listed) return fail(c, SPHX_E_STATE, "energy needs the neighbour list (call sphx_neighbours)");
    launch_energy(c->P(), c->D, c->n, c->grid, c->L, c->kp, c->has_mhd, c->energy_partial, nb, c->stream);
    if ((rc = launch_check(c))) return rc;
    std::vector<double> part((size_t)nb * 9);
    HIP_TRY(c, hipMemcpyAsync(part.data(), c->energy_partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    double s[9] = {0};
    for (uint32_t b = 0; b < nb; ++b) {
        for (int k = 0; k < 8; ++k) s[k] += part[(size_t)b * 9 + k];
        s[8] = std::max(s[8], part[(size_t)b * 9 + 8]);
    }
    out->kinetic = s[0];
    out->potential = s[1];
    out->internal = s[2];
    out->magnetic = s[3];
    out->momentum[0] = s[4];
    out->momentum[1] = s[5];
    out->momentum[2] = s[6];
    out->total_neighbours = (uint64_t)s[7];
    out->max_neighbours = (uint64_t)s[8];
    return SPHX_OK;
}

}  // extern "C"
