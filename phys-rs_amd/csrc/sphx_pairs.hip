// sphx_pairs.hip — the neighbour build and the neighbour passes of the SPH
// step on gfx950.
//
// Build: one workgroup per tile (<= 256 sorted particles of one cell row).
// For each of the 9 neighbour rows ("stages") the workgroup copies the
// candidates' f32 cell-relative coordinates into LDS with contiguous loads
// and each thread tests its 3-cell window, appending accepted neighbours to
// one list per particle (all stages together).
//
// Passes: one thread per sorted particle walks its whole list (one loop over
// all stages keeps lanes busy; a per-stage loop would idle every lane whose
// particle has no neighbour in that row).  The density, force and mixed force
// passes run brick-staged: a workgroup per brick of cells copies the
// neighbours' gather records around the brick into LDS with contiguous loads
// and reads them from there (gathering them from L2 was bound by L1 misses,
// ~0.7 L2 requests per pair).  The other passes gather from global memory.
// One thread owns one particle and gathers: no atomics on the floating-point
// path, every sum is deterministic, staged or not.  No MFMA: the path is not a
// dense contraction.
//
// Compiled with -ffp-contract=off: the neighbour predicate rounds exactly
// like the reference's sqrt((dx*dx + dy*dy) + dz*dz) < 2*h_avg
// (src/physics/sph.rs:102-108).  Fused multiply-adds elsewhere are fma().

#include <algorithm>

#include "sphx_device.h"

namespace sphx {

#define SPHX_TILE_PROLOGUE()                                             \
    const uint32_t b_ = xcd_block_of(tile_count(T));                     \
    if (b_ == 0xffffffffu) return;                                       \
    const TileCtx tc = load_tile(b_, T, g, skey);                        \
    const uint32_t t = threadIdx.x;                                      \
    const bool valid = t < tc.np;                                        \
    const uint32_t p = tc.p0 + (valid ? t : 0u)

// Stage the chunk [cb, ce) of stage st: body(u, j) stores particle j's
// fields at LDS index u.
#define SPHX_STAGE_CHUNK(body)                                           \
    for (uint32_t u = t; u < ce - cb; u += kBlock) {                     \
        const uint32_t j = slot_particle(st, cb + u);                    \
        body;                                                            \
    }

// ------------------------------------------------ neighbour lists (a1, a2)

// The list holds every pair with r < 2 h_avg + skin (a superset: each pass
// gets exactly the pairs r < 2 h_avg, because the cubic spline and its
// derivatives are exactly zero from q = 2 on, and the pair export and the
// neighbour counts apply the reference predicate itself).  The test is f32
// on staged cell-relative coordinates with a 1e-5 relative margin on r^2:
// |xr| <= 2w for the owner bounds the f32 error of r^2 by ~2e-6 relative,
// so no pair inside the radius is missed.  Owners further than 2 cells
// from their cell origin (particles clamped into the grid) test in f64.
//
// Cells are sorted by x and a row's cells are consecutive in the sorted
// order, so a thread's candidates in one neighbour row -- the three cells
// around its own, cut to |dx| < reach = 2 h_max (1 + 1e-4) + skin -- are one
// contiguous slot range, found by binary search on the staged f32 x at both
// ends.  All threads of a wave then walk ranges of about two cells in
// lockstep.  The f32 error of dx (<~1e-7 w for |xr| <= 2w) is far inside the
// 1e-4 margin, so no neighbour is cut; owners outside that bound do not cut.
// Rows whose three cells are not consecutive (a periodic wrap) walk the
// cells one at a time.
template <bool UNI>
__global__ void __launch_bounds__(kBlock) k_build(Particles P, Mixed X, Tiles T, Grid g,
                                                   const uint32_t* __restrict__ skey,
                                                   const uint32_t* __restrict__ start, NbrList L,
                                                   float reach, double skin) {
    // staged candidates: cell-relative x, y, z and h; three spare records so
    // that a four-slot step may read past the chunk (those slots are masked)
    __shared__ float4 lds[kStageCap + 3];
    SPHX_TILE_PROLOGUE();
    const int cx = valid ? (int)(skey[p] % (uint32_t)g.dims[0]) : 0;
    float xf = 0, yf = 0, zf = 0, hf = 0;
    double xi = 0, yi = 0, zi = 0, hi = 0;
    bool exact_all = false;
    if (valid) {
        const float4 o = X.p4[p];
        xf = o.x; yf = o.y; zf = o.z; hf = o.w;
        const float lim = 2.0f * (float)fmax(g.w[0], fmax(g.w[1], g.w[2]));
        exact_all = fabsf(xf) > lim || fabsf(yf) > lim || fabsf(zf) > lim;
        xi = P.x[p]; yi = P.y[p]; zi = P.z[p]; hi = P.h[p];
    }
    const float skin_f = (float)skin;
    const float wy = (float)g.w[1], wz = (float)g.w[2];
    const bool trim = !exact_all;
    const bool any_exact = __any(exact_all) != 0;  // wave-uniform
    const float wx = (float)g.w[0];
    uint32_t total = 0;
    uint32_t* __restrict__ out = nbr_base(L, p);
    // the last four accepted entries (b3 newest): the list is written a group
    // of four (16 bytes) at a time
    uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    for (int s = 0; s < kStages; ++s) {
        Stage st;
        stage_segments(g, start, tc, s, st);
        uint32_t ra[3], rb[3];
        for (int o = 0; o < 3; ++o) cell_slots(g, start, st, cx, o - 1, ra[o], rb[o]);
        uint32_t self = 0xffffffffu;
        if (s == 4 && valid) self = ra[1] + (p - start[st.rowbase + cx]);
        const bool e0 = ra[0] == rb[0], e2 = ra[2] == rb[2];
        const bool merged = (e0 || rb[0] == ra[1]) && (e2 || rb[1] == ra[2]);
        const bool single = st.total <= (uint32_t)kStageCap;
        // The reach in x left for this row: a candidate of row (oy, oz) is at
        // least dy (dz) away, the distance from i to the cell face towards
        // that row (0 for the own row), so |dx| < sqrt(reach^2 - dy^2 - dz^2);
        // a row with nothing in reach is skipped.  Shrunk distances keep the
        // bound conservative in f32.
        float rreach = reach;
        bool row_empty = false;
        if (trim) {
            const int oy = s % 3 - 1, oz = s / 3 - 1;
            const float dyl = oy == 0 ? 0.0f : fmaxf(oy < 0 ? yf : wy - yf, 0.0f) * (1.0f - 1e-5f);
            const float dzl = oz == 0 ? 0.0f : fmaxf(oz < 0 ? zf : wz - zf, 0.0f) * (1.0f - 1e-5f);
            const float rr2 = fmaf(-dzl, dzl, fmaf(-dyl, dyl, reach * reach));
            row_empty = rr2 <= 0.0f;
            rreach = __builtin_amdgcn_sqrtf(fmaxf(rr2, 0.0f)) * (1.0f + 1e-5f);
        }
        // first slot of [a, e) (ascending x, whole stage staged) with
        // (x_slot - x_i) + off beyond the row's reach: below -rreach (left
        // cell) or at least rreach (right cell)
        // the row's y and z offsets folded into the owner's coordinates once
        // (x_j - (x_i - off) differs from (x_j - x_i) + off by a few ulp of
        // the cell width, far inside the 1e-5 margin of the superset test)
        const float yo = yf - st.oy, zo = zf - st.oz;
        auto cut = [&](uint32_t a, uint32_t e, float off, bool left) {
            while (a < e) {
                const uint32_t mid = (a + e) >> 1;
                const float d = (lds[mid].x - xf) + off;
                if (left ? d <= -rreach : d < rreach) a = mid + 1;
                else e = mid;
            }
            return a;
        };
        // the slots [a, e) of one cell (x offset o - 1), chunk starting at cb:
        // the candidate tests of up to 32 slots go into a bit mask first, then
        // each lane appends its accepted slots; EX: some lane of the wave
        // tests in f64 (exact_all), decided once per wave so that the usual
        // f32-only loop has no branch for it
        auto test_cell = [&](uint32_t a, uint32_t e, uint32_t o, uint32_t cb, auto ex) {
            const float ox = (float)((int)o - 1) * wx;
            const float tu = (hf + hf) + skin_f;
            const float lim_u = (tu * tu) * (1.0f + 1e-5f);  // (2 h + skin)^2 with a 1e-5 margin
            const uint32_t code = ((uint32_t)s * 3u + o) << kIdxBits;
            for (uint32_t base = a; base < e; base += 32u) {
                const uint32_t cnt = min(32u, e - base);
                uint32_t mask = 0;
                // four slots a step from one LDS address (their reads issued
                // together); slots past e read spare records and are masked
                const float4* __restrict__ row = lds + (base - cb);
                for (uint32_t k = 0; k < cnt; k += 4u) {
                    float4 q[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) q[u] = row[k + (uint32_t)u];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float dx = (q[u].x - xf) + ox, dy = q[u].y - yo, dz = q[u].z - zo;
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        float lim = lim_u;
                        if (!UNI) {
                            const float tt = (hf + q[u].w) + skin_f;  // 2 h_avg + skin
                            lim = (tt * tt) * (1.0f + 1e-5f);
                        }
                        bool in = r2 < lim;
                        if constexpr (decltype(ex)::value) {
                            if (exact_all) {
                                const uint32_t j = slot_particle(st, min(base + k + (uint32_t)u, e - 1u));
                                double ddx, ddy, ddz;
                                pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], ddx, ddy, ddz);
                                const double t = (hi + P.h[j]) + skin;
                                in = (ddx * ddx + ddy * ddy) + ddz * ddz < (t * t) * (1.0 + 1e-12);
                            }
                        }
                        mask |= in ? (1u << (k + (uint32_t)u)) : 0u;
                    }
                }
                if (cnt < 32u) mask &= (1u << cnt) - 1u;
                if (self - base < cnt) mask &= ~(1u << (self - base));
                while (mask) {
                    const uint32_t k = (uint32_t)__builtin_ctz(mask);
                    mask &= mask - 1u;
                    if (total < (uint32_t)L.ks) {
                        b0 = b1;
                        b1 = b2;
                        b2 = b3;
                        b3 = slot_particle(st, base + k) | code;
                        if ((total & 3u) == 3u)
                            *reinterpret_cast<uint4*>(out + nbr_off(total - 3u)) = make_uint4(b0, b1, b2, b3);
                    }
                    ++total;
                }
            }
        };
        // Uniform h: the lane's whole window [a, e) over its three cells in
        // one loop.  The staged w holds the candidate's logical cell (see
        // Stage), so the cell offset of the pair vector is (w - own) wx with
        // w - own in {-1, 0, 1}: fma((w - own), wx, x_j - x_i) rounds exactly
        // as (x_j - x_i) + offset does.
        auto test_window = [&](uint32_t a, uint32_t e, uint32_t cb, auto ex) {
            const float own = (float)(cx - st.lorigin);
            const float tu = (hf + hf) + skin_f;
            const float lim = (tu * tu) * (1.0f + 1e-5f);
            for (uint32_t base = a; base < e; base += 32u) {
                const uint32_t cnt = min(32u, e - base);
                uint32_t mask = 0;
                const float4* __restrict__ row = lds + (base - cb);
                for (uint32_t k = 0; k < cnt; k += 4u) {
                    float4 q[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) q[u] = row[k + (uint32_t)u];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float dx = fmaf(q[u].w - own, wx, q[u].x - xf);
                        const float dy = q[u].y - yo, dz = q[u].z - zo;
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        bool in = r2 < lim;
                        if constexpr (decltype(ex)::value) {
                            if (exact_all) {
                                const uint32_t j = slot_particle(st, min(base + k + (uint32_t)u, e - 1u));
                                double ddx, ddy, ddz;
                                pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], ddx, ddy, ddz);
                                const double t = (hi + P.h[j]) + skin;
                                in = (ddx * ddx + ddy * ddy) + ddz * ddz < (t * t) * (1.0 + 1e-12);
                            }
                        }
                        mask |= in ? (1u << (k + (uint32_t)u)) : 0u;
                    }
                }
                if (cnt < 32u) mask &= (1u << cnt) - 1u;
                if (self - base < cnt) mask &= ~(1u << (self - base));
                while (mask) {
                    const uint32_t k = (uint32_t)__builtin_ctz(mask);
                    mask &= mask - 1u;
                    if (total < (uint32_t)L.ks) {
                        const uint32_t slot = base + k;
                        const uint32_t o = (slot >= ra[1] ? 1u : 0u) + (slot >= rb[1] ? 1u : 0u);
                        b0 = b1;
                        b1 = b2;
                        b2 = b3;
                        b3 = slot_particle(st, slot) | (((uint32_t)s * 3u + o) << kIdxBits);
                        if ((total & 3u) == 3u)
                            *reinterpret_cast<uint4*>(out + nbr_off(total - 3u)) = make_uint4(b0, b1, b2, b3);
                    }
                    ++total;
                }
            }
        };
        // non-merged rows: the cells in slot order (a periodic row spanning
        // the whole grid stages in physical, not logical, cell order)
        uint32_t qa[3] = {ra[0], ra[1], ra[2]}, qb[3] = {rb[0], rb[1], rb[2]}, qo[3] = {0, 1, 2};
        if (!merged) {
#define SPHX_SWAP_RANGE(a_, b_)                                                            \
            if (qa[b_] < qa[a_]) {                                                         \
                uint32_t t0 = qa[a_]; qa[a_] = qa[b_]; qa[b_] = t0;                        \
                t0 = qb[a_]; qb[a_] = qb[b_]; qb[b_] = t0;                                 \
                t0 = qo[a_]; qo[a_] = qo[b_]; qo[b_] = t0;                                 \
            }
            SPHX_SWAP_RANGE(0, 1)
            SPHX_SWAP_RANGE(1, 2)
            SPHX_SWAP_RANGE(0, 1)
#undef SPHX_SWAP_RANGE
        }
        for (uint32_t cb = 0; cb < st.total; cb += kStageCap) {
            const uint32_t ce = min(st.total, cb + kStageCap);
            __syncthreads();
            SPHX_STAGE_CHUNK({
                float4 v = X.p4[j];
                if (UNI)  // h is uniform: w carries the candidate's logical cell
                    v.w = (float)((int)cell_x_of(skey[j], (uint32_t)g.dims[0], g.inv_nx) +
                                  (st.nseg == 2 && cb + u >= st.off1 ? st.lshift1 : st.lshift0) - st.lorigin);
                lds[u] = v;
            });
            __syncthreads();
            if (!valid || row_empty) continue;
            auto run = [&](auto ex) {
                if (merged) {  // the three cells are consecutive slots: left, own, right
                    uint32_t A = e0 ? ra[1] : ra[0], E = e2 ? rb[1] : rb[2];
                    if (trim && single) {
                        if (!e0) A = cut(ra[0], rb[0], -wx, true);
                        if (!e2) E = cut(ra[2], rb[2], wx, false);
                    }
                    if (UNI) {
                        test_window(max(A, cb), min(E, ce), cb, ex);
                    } else {
                        test_cell(max(A, cb), min(ra[1], ce), 0u, cb, ex);
                        test_cell(max(ra[1], cb), min(rb[1], ce), 1u, cb, ex);
                        test_cell(max(rb[1], cb), min(E, ce), 2u, cb, ex);
                    }
                } else {
                    for (int q = 0; q < 3; ++q) {
                        uint32_t a = qa[q], e = qb[q];
                        const uint32_t o = qo[q];
                        if (trim && single && o != 1u && a < e) {
                            if (o == 0u) a = cut(a, e, -wx, true);
                            else e = cut(a, e, wx, false);
                        }
                        test_cell(max(a, cb), min(e, ce), o, cb, ex);
                    }
                }
            };
            if (any_exact) run(std::true_type{});
            else run(std::false_type{});
        }
    }
    if (valid) {
        // the last, partial group (ks is a multiple of 4, so a full list has none)
        const uint32_t kept = min(total, (uint32_t)L.ks), r = kept & 3u;
        if (r) {
            const uint4 v = r == 1u ? make_uint4(b3, 0u, 0u, 0u)
                                    : (r == 2u ? make_uint4(b2, b3, 0u, 0u) : make_uint4(b1, b2, b3, 0u));
            *reinterpret_cast<uint4*>(out + nbr_off(kept - r)) = v;
        }
        L.ntot[p] = (uint16_t)kept;
    }
    const uint32_t wmax = wave_max_u32(total);
    if ((t & 63) == 0 && wmax) {
        atomicMax(&L.overflow[0], wmax);
        if (wmax > (uint32_t)L.ks) atomicOr(&L.overflow[1], 1u);
    }
}

void launch_build_nbr(const Particles& P, uint32_t n, const Grid& g, const Sort& S, const Tiles& T,
                      const NbrList& L, bool uniform, const Mixed& X, double h_max, double skin, hipStream_t s) {
    if (!n) return;
    const dim3 grid(tile_grid(n, T)), block(kBlock);
    const float reach = (float)(2.0 * h_max * (1.0 + 1e-4) + skin);
    if (uniform)
        hipLaunchKernelGGL(k_build<true>, grid, block, 0, s, P, X, T, g, S.skey, S.cell_start, L, reach, skin);
    else
        hipLaunchKernelGGL(k_build<false>, grid, block, 0, s, P, X, T, g, S.skey, S.cell_start, L, reach, skin);
}

// ------------------------------------------------------- density + EOS (a4, a5)
// Pair geometry shared by the f64 passes: the smoothing length of the pair
// (1/h_avg) and the kernel normalisations, uniform-h case folded to constants.
struct PairH {
    double inv_h, sig_w, sig_g, h_avg;
};
template <bool UNI>
__device__ __forceinline__ PairH pair_h(const KernelParams& kp, double hi, double hj) {
    PairH q;
    if (UNI) {
        q.inv_h = kp.inv_h0;
        q.h_avg = kp.h0;
    } else {
        q.h_avg = 0.5 * (hi + hj);
        q.inv_h = rcp_nr(q.h_avg);
    }
    const double ih2 = q.inv_h * q.inv_h;
    q.sig_w = (ih2 * q.inv_h) * kInvPi;
    q.sig_g = (ih2 * ih2) * kInvPi;
    return q;
}

#define SPHX_PASS_PROLOGUE()                \
    const uint32_t i = particle_index();    \
    if (i >= n) return

// The list is summed in two halves, entries 0 .. half - 1 and half .. end
// (half a multiple of 4 near the middle), added in that order; with LANES = 2
// the calling lane and its neighbour (lane ^ 1) take one half each.
template <bool UNI, int PER, int LANES, class Fetch>
__device__ __forceinline__ void density_particle(const Particles& P, const Derived& D, const Grid& g,
                                                 const NbrList& L, const KernelParams& kp, uint32_t i,
                                                 const Fetch& fetch) {
    double2 own[2];  // the particle's own position record (x, y | z, h)
    fetch.own(i, own);
    const double xi = own[0].x, yi = own[0].y, zi = own[1].x, hi = UNI ? kp.h0 : own[1].y;
    double rho = (UNI ? kp.m0 : P.m[i]) * (1.0 / (kPi * ((hi * hi) * hi)));  // m_i W(0, h_i)  (sph.rs:89-91)
    double acc = 0.0;
    const double nih0 = -kp.inv_h0, nih20 = -0.5 * kp.inv_h0;
    auto pair = [&](uint32_t e, const double2 (&q)[2]) {
        const double2 xy = q[0], zh = q[1];
        double dx, dy, dz;
        pair_vec<PER>(g, xi, yi, zi, xy.x, xy.y, zh.x, dx, dy, dz);
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        // r2 = 0 (coincident particles) as 1e-300: q ~ 1e-148 gives W(0) exactly
        double r, inv_r;
        sqrt_rsqrt(fmax(r2, 1e-300), r, inv_r);
        const PairH ph = pair_h<UNI>(kp, hi, UNI ? hi : zh.y);
        if (UNI) {  // m, sigma_w and the factor 2 of wh_shape are folded out of the sum
            acc += wh_shape_r(r, nih0, nih20);
        } else {
            const double nih = -ph.inv_h;
            acc = fma(P.m[e & kIdxMask], (2.0 * ph.sig_w) * wh_shape_r(r, nih, 0.5 * nih), acc);
        }
    };
    const uint32_t half = ((fetch.count(L, i) + 7u) >> 3) << 2;
    auto sum_range = [&](uint32_t kb, uint32_t ke) {
        Fetch f = fetch;
        f.kbeg = kb;
        f.kend = ke;
        walk(L, i, f, pair);
    };
    if (LANES == 2) {
        const uint32_t h = lane_half();
        sum_range(h ? half : 0u, h ? 0xffffffffu : half);
        acc += other_half(acc);  // lane 0: first half + second half
        if (h) return;
    } else {
        sum_range(0u, half);
        const double a0 = acc;
        acc = 0.0;
        sum_range(half, 0xffffffffu);
        acc = a0 + acc;
    }
    rho = UNI ? fma(kp.m0 * ((((kp.inv_h0 * kp.inv_h0) * kp.inv_h0) * kInvPi) * 2.0), acc, rho) : rho + acc;
    const double vxi = P.vx[i], vyi = P.vy[i], vzi = P.vz[i];
    const double ui = kp.eos == SPHX_EOS_ADIABATIC ? P.u[i] : 0.0;
    const double pr = eos_pressure(kp, rho, ui);
    const double pr2 = pr / (rho * rho);
    D.rho[i] = rho;
    D.p[i] = pr;
    D.pr2[i] = pr2;
    double2* rec = reinterpret_cast<double2*>(D.r64 + i);
    rec[0] = make_double2(xi, yi);
    rec[1] = make_double2(zi, rho);
    rec[2] = make_double2(vxi, vyi);
    rec[3] = make_double2(vzi, pr2);
    if (kp.srec) D.recs[i] = stress_record(P.bx[i], P.by[i], P.bz[i], rho);
}

template <bool UNI>
__global__ void __launch_bounds__(kBlock) k_density(Particles P, Derived D, uint32_t n, Grid g, NbrList L,
                                                     KernelParams kp) {
    const uint32_t i = pair_slot(particle_index());  // two lanes per particle
    if (i >= n) return;
    density_particle<UNI, -1, 2>(P, D, g, L, kp, i,
                                 GlobalFetch<double2, 2>{reinterpret_cast<const double2*>(P.pos)});
}

// Brick-staged forms: the neighbours' records come from LDS (bit-identical
// results: same records, same order, same arithmetic).
// One lane per particle summing the two halves in turn (two lanes per
// particle measured slower here: one workgroup per CU instead of two, 0.21 ->
// 0.25 ms on the 1M dam break).
// S16: the slot lists are written (the launch picks the form without the
// slot walk when they are not, which saves its registers).
// write16: the first density pass after a build fills the slot lists as it
// walks the entries.
template <bool UNI, int PER, int W, bool PF, bool S16>
__global__ void __launch_bounds__(kStagedBlock) __attribute__((amdgpu_waves_per_eu(4))) k_density_staged(
    Particles P, Derived D, Grid g, NbrList L, KernelParams kp, Bricks B, const uint32_t* __restrict__ start,
    uint32_t cap, int write16) {
    __shared__ BrickHdr H;
    extern __shared__ uint4 sh_dyn[];
    // uniform h: the neighbour's h (the last 8 bytes of its record) is not
    // read, and the pair body needs only the record (the slot lists); the
    // epilogue's velocity loads wait for the walk (ahead of it they held
    // registers that then spilled: 1M density 0.173 -> 0.168 ms)
    brick_pass<double2, 2, W, PF, 1, 0, false, UNI, UNI && S16>(
        B, g, start, cap, L, reinterpret_cast<const double2*>(P.pos), H, sh_dyn,
        [&](uint32_t i, const auto& f) { density_particle<UNI, PER, 1>(P, D, g, L, kp, i, f); }, nullptr,
        write16 != 0);
}

// FP32-mixed density: f32 pair terms on cell-relative coordinates, f64 sum.
template <bool UNI, class Fetch>
__device__ __forceinline__ void density_mixed_particle(const Particles& P, const Derived& D, const Mixed& X,
                                                       const Grid& g, const NbrList& L, const KernelParams& kp,
                                                       uint32_t i, const Fetch& fetch) {
    float4 ownr[1];  // the particle's own position record (brick-relative with Fetch::BREL)
    fetch.own(i, ownr);
    const float4 own = ownr[0];
    const float xi = own.x, yi = own.y, zi = own.z, hi = own.w;
    const double hd = UNI ? kp.h0 : P.h[i];
    double rho = (UNI ? kp.m0 : P.m[i]) * (1.0 / (kPi * ((hd * hd) * hd)));
    // the epilogue's per-particle loads, issued ahead of the walk
    const double vxi = P.vx[i], vyi = P.vy[i], vzi = P.vz[i];
    const double ui = kp.eos == SPHX_EOS_ADIABATIC ? P.u[i] : 0.0;
    const float inv_h0 = (float)kp.inv_h0;
    const float wx = (float)g.w[0], wy = (float)g.w[1], wz = (float)g.w[2];
    double acc = 0.0;
    float gacc = 0.0f;  // f32 sum over each group of four list entries
    walk(L, i, fetch, [&](uint32_t e, const float4 (&q)[1]) {
        float dx, dy, dz;
        if constexpr (Fetch::BREL) {  // brick-relative records: no cell offset
            dx = q[0].x - xi;
            dy = q[0].y - yi;
            dz = q[0].z - zi;
        } else {
            float ox, oy, oz;
            fetch.offset(e, wx, wy, wz, ox, oy, oz);
            dx = (q[0].x - xi) + ox;
            dy = (q[0].y - yi) + oy;
            dz = (q[0].z - zi) + oz;
        }
        const float r = __builtin_amdgcn_sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
        if (UNI) {  // m, sigma_w and the factor 2 of the one-sided shape are folded out of the sum
            // W / (2 sigma) = t^3 - s^3 / 2 from the one-sided pieces (as
            // the f64 pass: 1M mixed density 0.128 -> 0.125 ms)
            const float ps = fmaxf(fmaf(r, -inv_h0, 1.0f), 0.0f), pt = fmaxf(fmaf(r, -0.5f * inv_h0, 1.0f), 0.0f);
            gacc += fmaf((ps * ps) * ps, -0.5f, (pt * pt) * pt);
        } else {
            const float inv_h = __builtin_amdgcn_rcpf(0.5f * (hi + q[0].w));
            const float ps = fmaxf(fmaf(r, -inv_h, 1.0f), 0.0f), pt = fmaxf(fmaf(r, -0.5f * inv_h, 1.0f), 0.0f);
            const float w = ((inv_h * inv_h) * inv_h * (float)(2.0 * kInvPi)) * fmaf((ps * ps) * ps, -0.5f, (pt * pt) * pt);
            gacc += X.m[e & kIdxMask] * w;
        }
    }, [&]() {
        acc += (double)gacc;
        gacc = 0.0f;
    });
    rho = UNI ? fma(kp.m0 * ((((kp.inv_h0 * kp.inv_h0) * kp.inv_h0) * kInvPi) * 2.0), acc, rho) : rho + acc;
    const double pr = eos_pressure(kp, rho, ui);
    const double pr2 = pr / (rho * rho);
    D.rho[i] = rho;
    D.p[i] = pr;
    D.pr2[i] = pr2;
    float4* rec = reinterpret_cast<float4*>(D.r32 + i);
    // the force record keeps the cell-relative position
    const float4 p4i = Fetch::BREL ? X.p4[i] : own;
    rec[0] = make_float4(p4i.x, p4i.y, p4i.z, (float)rho);
    rec[1] = make_float4((float)vxi, (float)vyi, (float)vzi, (float)pr2);
    if (kp.srec) D.recs[i] = stress_record(P.bx[i], P.by[i], P.bz[i], rho);
}

template <bool UNI>
__global__ void __launch_bounds__(kBlock) k_density_mixed(Particles P, Derived D, Mixed X, uint32_t n, Grid g,
                                                           NbrList L, KernelParams kp) {
    SPHX_PASS_PROLOGUE();
    density_mixed_particle<UNI>(P, D, X, g, L, kp, i, GlobalFetch<float4, 1>{X.p4});
}

template <bool UNI, int W, bool PF, bool S16>
__global__ void __launch_bounds__(kStagedBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_density_mixed_staged(Particles P, Derived D, Mixed X, Grid g, NbrList L, KernelParams kp, Bricks B,
                       const uint32_t* __restrict__ start, uint32_t cap, int write16) {
    __shared__ BrickHdr H;
    extern __shared__ uint4 sh_dyn[];
    // brick-relative records; uniform m and h: the pair body needs only the
    // record (the slot lists)
    brick_pass<float4, 1, W, PF, 1, 0, false, false, UNI && S16, true>(
        B, g, start, cap, L, X.p4, H, sh_dyn,
        [&](uint32_t i, const auto& f) { density_mixed_particle<UNI>(P, D, X, g, L, kp, i, f); }, nullptr,
        write16 != 0);
}

// Staged launch of a pass with records of rec_bytes that runs at most
// max_waves waves per CU: LDS for as many workgroups per CU as the waves
// allow and the mean halo fits (at least one), one workgroup per active
// brick, and the record capacity.
struct StagedLaunch {
    dim3 grid, block;
    size_t dyn;
    uint32_t cap;
};
static StagedLaunch staged_launch(const Bricks& B, int max_waves, size_t rec_bytes) {
    StagedLaunch l;
    const double halo = B.ppc * (B.b[0] + 2) * (B.b[1] + 2) * (B.b[2] + 2);
    const size_t rc = (size_t)brick_rc_quads(B) * 16u;  // the row-code table (dynamic LDS)
    const size_t fixed = sizeof(BrickHdr) + 512 + rc;
    int per_cu = std::min(max_waves / (B.threads / 64), (int)((double)kLdsPerCU / (halo * rec_bytes + fixed)));
    per_cu = std::max(per_cu, 1);
    const size_t per_block = kLdsPerCU / (size_t)per_cu - fixed;
    l.cap = (uint32_t)(per_block / rec_bytes);
    l.dyn = (size_t)l.cap * rec_bytes + rc;
    l.grid = dim3(B.nactive);
    l.block = dim3((uint32_t)B.threads);
    return l;
}
// The record capacity of the density pass, which writes the slot lists in
// the bricks it stages whole: a pass that reads them stages at most this
// many records, so that every brick it stages whole has them written (the
// FP64 force pass alone could stage a few more: its fixed LDS share is
// split over half as many workgroups).
static uint32_t slot_cap(const Bricks& B, bool mixed) { return staged_launch(B, 16, mixed ? 16 : 32).cap; }

void launch_density(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                    const KernelParams& kp, bool mixed, const Mixed& X, const Bricks& B,
                    const uint32_t* cell_start, bool write_slots, hipStream_t s) {
    const int w16 = write_slots ? 1 : 0;
    if (!n) return;
    const dim3 grid(nblocks(n)), block(kBlock);
    const bool uni = kp.uniform != 0;
    if (!mixed && B.on && B.nactive) {
        const StagedLaunch sl = staged_launch(B, 16, 32);
        // walk widths measured on the 1M dam break: four entries with the next
        // four loaded ahead for the uniform density (0.27 -> 0.24 ms)
#define SPHX_DENS_STAGED(U, PER, W, PF, S)                                                                  \
        hipLaunchKernelGGL((k_density_staged<U, PER, W, PF, S>), sl.grid, sl.block, sl.dyn, s, P, D, g, L, kp, B, \
                           cell_start, sl.cap, w16)
        const bool slots = L.s16 != nullptr;
        if (uni && g.periodic && slots) SPHX_DENS_STAGED(true, 1, 4, true, true);
        else if (uni && g.periodic) SPHX_DENS_STAGED(true, 1, 4, true, false);
        else if (uni && slots) SPHX_DENS_STAGED(true, 0, 4, true, true);
        else if (uni) SPHX_DENS_STAGED(true, 0, 4, true, false);
        else if (g.periodic) SPHX_DENS_STAGED(false, 1, 2, false, false);
        else SPHX_DENS_STAGED(false, 0, 2, false, false);
#undef SPHX_DENS_STAGED
        return;
    }
    if (mixed && B.on && B.nactive) {
        const StagedLaunch sl = staged_launch(B, 16, 16);
        if (uni && L.s16)
            hipLaunchKernelGGL((k_density_mixed_staged<true, 4, true, true>), sl.grid, sl.block, sl.dyn, s, P, D, X,
                               g, L, kp, B, cell_start, sl.cap, w16);
        else if (uni)
            hipLaunchKernelGGL((k_density_mixed_staged<true, 4, true, false>), sl.grid, sl.block, sl.dyn, s, P, D, X,
                               g, L, kp, B, cell_start, sl.cap, 0);
        else
            hipLaunchKernelGGL((k_density_mixed_staged<false, 4, false, false>), sl.grid, sl.block, sl.dyn, s, P, D,
                               X, g, L, kp, B, cell_start, sl.cap, 0);
        return;
    }
    if (mixed && uni)
        hipLaunchKernelGGL(k_density_mixed<true>, grid, block, 0, s, P, D, X, n, g, L, kp);
    else if (mixed)
        hipLaunchKernelGGL(k_density_mixed<false>, grid, block, 0, s, P, D, X, n, g, L, kp);
    else if (uni)
        hipLaunchKernelGGL(k_density<true>, dim3(nblocks(2 * n)), block, occupancy_lds(4), s, P, D, n, g, L, kp);
    else
        hipLaunchKernelGGL(k_density<false>, dim3(nblocks(2 * n)), block, occupancy_lds(4), s, P, D, n, g, L, kp);
}

// ------------------------- pressure + viscosity + Maxwell stress + walls (a6-a9, a13)
// Shared epilogue: S = m g + sum(pressure) is stored by the full pass and
// reloaded by VISC_ONLY, so the reuse is bit-identical.  Then Lorentz
// q(E + v x B) (lorentz.rs:27-37), viscosity and stress sums, penalty walls
// (sph.rs:451-483) and a = F / m (integrators.rs:115).
// VISC_ONLY: sp holds the stored sum of particle p, loaded by the caller
// before its walk (so that the load's latency is spent in the walk).
template <bool VISC_ONLY>
__device__ __forceinline__ void force_epilogue(const Particles& P, const Derived& D, const KernelParams& kp,
                                               bool charge, uint32_t p, const double fp[3], const double fv[3],
                                               const double fm[3], const double sp[3]) {
    const double mi = kp.uniform ? kp.m0 : P.m[p];
    double S[3];
    if (VISC_ONLY) {
        S[0] = sp[0]; S[1] = sp[1]; S[2] = sp[2];
    } else {
        S[0] = mi * kp.g[0] + fp[0];
        S[1] = mi * kp.g[1] + fp[1];
        S[2] = mi * kp.g[2] + fp[2];
        D.spx[p] = S[0]; D.spy[p] = S[1]; D.spz[p] = S[2];
    }
    if (charge) {
        const double q = P.q[p], vx = P.vx[p], vy = P.vy[p], vz = P.vz[p];
        const double bx = P.bx[p], by = P.by[p], bz = P.bz[p];
        S[0] += q * (kp.e[0] + (vy * bz - vz * by));
        S[1] += q * (kp.e[1] + (vz * bx - vx * bz));
        S[2] += q * (kp.e[2] + (vx * by - vy * bx));
    }
    double F[3] = {S[0] + fv[0] + fm[0], S[1] + fv[1] + fm[1], S[2] + fv[2] + fm[2]};
    if (kp.bnd & SPHX_BND_PENALTY_CUBE) {
        const double c3[3] = {P.x[p], P.y[p], P.z[p]};
        for (int a = 0; a < 3; ++a) {
            const double d_low = c3[a] - (-kp.half_extent);
            if (d_low < kp.act_dist && d_low >= 0.0) F[a] += kp.stiffness * (kp.act_dist - d_low);
            const double d_high = kp.half_extent - c3[a];
            if (d_high < kp.act_dist && d_high >= 0.0) F[a] -= kp.stiffness * (kp.act_dist - d_high);
        }
    }
    D.ax[p] = F[0] / mi;
    D.ay[p] = F[1] / mi;
    D.az[p] = F[2] / mi;
}

// The i term of the Maxwell-stress sum (see force_particle):
// fm <- -(b_i (b_i . gs) - e_i gs + fm).
__device__ __forceinline__ void stress_close(const double2* __restrict__ RS, uint32_t i, const double gs[3],
                                             double fm[3]) {
    const double2 sxy = RS[2 * i], sze = RS[2 * i + 1];
    const double bg = fma(sze.x, gs[2], fma(sxy.y, gs[1], sxy.x * gs[0]));
    const double b[3] = {sxy.x, sxy.y, sze.x};
    for (int a = 0; a < 3; ++a) fm[a] = -(fma(b[a], bg, -sze.y * gs[a]) + fm[a]);
}

// LANES = 2: the calling lane and its neighbour (lane ^ 1) share particle i.
// ENERGY (adiabatic EOS, full pass only): du_i/dt = sum_j m_j (P_i/rho_i^2 +
// Pi_ij/2) v_ij . grad_i W_ij, from the same pair terms (v_ij . grad_i W =
// -s dW/r with s = (v_i - v_j).(x_j - x_i)), summed beside the force.
// B0 (beta = 0): the viscosity term is Pi = cv hs' / (den (rho_i + rho_j))
// with cv = -2 alpha c (and h, when uniform) folded out of the sum.
template <bool VISC_ONLY, bool ENERGY, bool STRESS, bool UNI, bool B0, int PER, int LANES, class Fetch>
__device__ __forceinline__ void force_particle(const Particles& P, const Derived& D, const Grid& g,
                                               const NbrList& L, const KernelParams& kp, int charge, uint32_t i,
                                               const Fetch& fetch) {
    // the particle's own gather record (x, y | z, rho | vx, vy | vz, P/rho^2)
    double2 own[Fetch::NR];
    fetch.own(i, own);
    const double xi = own[0].x, yi = own[0].y, zi = own[1].x, hi = UNI ? kp.h0 : P.h[i];
    const double vxi = own[2].x, vyi = own[2].y, vzi = own[3].x;
    const double mi = UNI ? kp.m0 : P.m[i], rhoi = own[1].y, pr2i = own[3].y;
    double sp[3] = {0, 0, 0};  // VISC_ONLY: the stored sum, loaded ahead of the walk
    if (VISC_ONLY) {
        sp[0] = D.spx[i]; sp[1] = D.spy[i]; sp[2] = D.spz[i];
    }
    // Pi = mu (beta mu - alpha c) / rho_bar with mu = h s / den, den = r^2 + 0.01 h^2
    // (sph.rs:281-300), formed with one reciprocal:
    // Pi = hs (2 beta hs - 2 alpha c den) / (den^2 (rho_i + rho_j)), hs = h s,
    // and for beta = 0, Pi = -2 alpha c hs / (den (rho_i + rho_j)).
    const double ac2 = -2.0 * kp.alpha * kp.speed_of_sound, beta2 = 2.0 * kp.beta;
    // the factor left out of the viscosity sum (B0), and half of it for the energy sum
    const double cv = B0 ? (UNI ? ac2 * kp.h0 : ac2) : 1.0, hcv = 0.5 * cv;
    const double nih0 = -kp.inv_h0, nih20 = -0.5 * kp.inv_h0;
    // s = (v_i - v_j).(x_j - x_i); the pair has a viscosity term for s < 0
    // (REFERENCE) or s > 0 (MONAGHAN): hs = h min(sv s, 0) with sv = 1 or -1
    // is then h s (REFERENCE), -h s (MONAGHAN) or 0, so Pi = 0 needs no select
    const double sv = kp.visc_conv == SPHX_VISC_REFERENCE ? 1.0 : -1.0;
    double fp[3] = {0, 0, 0}, fv[3] = {0, 0, 0}, fm[3] = {0, 0, 0};
    double de = 0.0;  // ENERGY: sum of (P_i/rho_i^2 + Pi/2) s dW/r
    // STRESS (Maxwell stress M = B B - |B|^2 I / 2, SPMHD):
    //   F_i = m_i m_j sum_j (M_i / rho_i^2 + M_j / rho_j^2) . grad_i W
    // with b = B / rho and e = |b|^2 / 2 from the stress records and
    // grad_i W = -(dW/r) d, d = x_j - x_i: the j terms are summed per pair,
    //   fm = sum_j (dW/r) (b_j (b_j . d) - e_j d),
    // and the i term leaves the sum through gs = sum_j (dW/r) d:
    //   F_i = -(b_i (b_i . gs) - e_i gs + fm).
    double gs[3] = {0, 0, 0};
    const double2* __restrict__ RS = reinterpret_cast<const double2*>(D.recs);
    // MERGE (full pass with the adiabatic EOS and the stress): the pressure
    // term's j part joins the stress sum, sum_j (dW/r) (b_j (b_j . d) -
    // (e_j + P_j/rho_j^2) d), and its i part is P_i/rho_i^2 gs; the stored
    // pressure sum then holds pressure and stress, which no pass reuses (the
    // viscosity-only pass is not available with the adiabatic EOS)
    constexpr bool MERGE = ENERGY && STRESS && !VISC_ONLY;
    // The pair body is branch-free: a pair without a term adds +-0, which
    // leaves a sum unchanged (the sums start at +0 and never become -0).
    // With uniform m and h, m_i m_j sigma_g is folded out of the sums.
    // q: the f64 record (x, y | z, rho | vx, vy | vz, P/rho^2), then, when the
    // fetch stages it, the B record (bx, by | bz, psi)
    auto pair = [&](uint32_t e, const double2 (&q)[Fetch::NR]) {
        const uint32_t j = e & kIdxMask;
        const double2 xy = q[0], zr = q[1], vxy = q[2], vzp = q[3];
        double dx, dy, dz;
        pair_vec<PER>(g, xi, yi, zi, xy.x, xy.y, zr.x, dx, dy, dz);
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        // r <= 1e-10: dW/dr = 0 (sph.rs:52), no force.  The clamp keeps 1/r
        // finite; below the cut r keeps only its low word, a denormal below
        // 2^-1042, so that s = t = 1 exactly and the shape is exactly 0.
        double r, inv_2r;
        sqrt_rhalf1(fmax(r2, 1e-300), r, inv_2r);
        r = __hiloint2double(r2 > 1e-20 ? __double2hiint(r) : 0, __double2loint(r));
        const PairH ph = pair_h<UNI>(kp, hi, UNI ? hi : P.h[j]);
        const double nih = UNI ? nih0 : -ph.inv_h, nih2 = UNI ? nih20 : 0.5 * nih;
        // (dW/dr) / (6 r): every sum below is a sixth of its value, and the
        // factor 6 is folded into the constants after the walk (uniform m and
        // h: into m_i m_j sigma_g, otherwise into sigma_g here)
        const double dwr0 = dw3_shape_r(r, nih, nih2) * inv_2r;
        const double dwr = UNI ? dwr0 : (mi * P.m[j]) * ((6.0 * ph.sig_g) * dwr0);
        if (!VISC_ONLY && !MERGE) {  // sph.rs:208-221
            const double sc = (pr2i + vzp.y) * dwr;
            fp[0] = fma(dx, sc, fp[0]);
            fp[1] = fma(dy, sc, fp[1]);
            fp[2] = fma(dz, sc, fp[2]);
        }
        // sph.rs:292-314: (v_i - v_j).(x_j - x_i); REFERENCE convention active when < 0
        const double sdot = fma(vzi - vzp.x, dz, fma(vyi - vxy.y, dy, (vxi - vxy.x) * dx));
        const double rhoj = zr.y;
        {
            const double ms = fmin(sv * sdot, 0.0);
            const double den = fma(0.01 * ph.h_avg, ph.h_avg, r2);
            double pi_ij;  // Pi / cv
            if (B0) {
                pi_ij = (UNI ? ms : ph.h_avg * ms) * rcp_nr1(den * (rhoi + rhoj));
            } else {
                const double hs = ph.h_avg * ms;
                pi_ij = (hs * fma(beta2, hs, ac2 * den)) * rcp_nr1((den * den) * (rhoi + rhoj));
            }
            const double sc = pi_ij * dwr;
            fv[0] = fma(dx, sc, fv[0]);
            fv[1] = fma(dy, sc, fv[1]);
            fv[2] = fma(dz, sc, fv[2]);
            if (ENERGY) de = fma(fma(hcv, pi_ij, pr2i), sdot * dwr, de);
        }
        if (STRESS) {
            double2 sxy, sze;  // (b_x, b_y), (b_z, e) of j
            if constexpr (Fetch::NR >= 6) {
                sxy = q[4];
                sze = q[5];
            } else {
                sxy = RS[2 * j];
                sze = RS[2 * j + 1];
            }
            gs[0] = fma(dwr, dx, gs[0]);
            gs[1] = fma(dwr, dy, gs[1]);
            gs[2] = fma(dwr, dz, gs[2]);
            const double bd = fma(sze.x, dz, fma(sxy.y, dy, sxy.x * dx));
            const double ej = MERGE ? sze.y + vzp.y : sze.y;
            fm[0] = fma(dwr, fma(sxy.x, bd, -ej * dx), fm[0]);
            fm[1] = fma(dwr, fma(sxy.y, bd, -ej * dy), fm[1]);
            fm[2] = fma(dwr, fma(sze.x, bd, -ej * dz), fm[2]);
        }
    };
    // The list is summed in two halves, entries 0 .. half - 1 and half .. end
    // (half a multiple of 4 near the middle), and the halves are added in
    // that order; the staged pass gives each half a lane of its own.
    const uint32_t cnt = fetch.count(L, i), half = ((cnt + 7u) >> 3) << 2;
    auto sum_range = [&](uint32_t kb, uint32_t ke) {
        Fetch f = fetch;
        f.kbeg = kb;
        f.kend = ke;
        walk(L, i, f, pair);
    };
    if (LANES == 2) {
        const uint32_t h = lane_half();
        sum_range(h ? half : 0u, h ? 0xffffffffu : half);
        for (int a = 0; a < 3; ++a) {  // lane h = 0: first half + second half
            if (!MERGE) fp[a] += other_half(fp[a]);
            fv[a] += other_half(fv[a]);
            if (STRESS) {
                fm[a] += other_half(fm[a]);
                gs[a] += other_half(gs[a]);
            }
        }
        if (ENERGY) de += other_half(de);
        if (h) return;
    } else {
        sum_range(0u, half);
        const double p0[3] = {fp[0], fp[1], fp[2]}, v0[3] = {fv[0], fv[1], fv[2]}, m0[3] = {fm[0], fm[1], fm[2]};
        const double g0[3] = {gs[0], gs[1], gs[2]};
        const double e0 = de;
        for (int a = 0; a < 3; ++a) fp[a] = fv[a] = fm[a] = gs[a] = 0.0;
        de = 0.0;
        sum_range(half, 0xffffffffu);
        for (int a = 0; a < 3; ++a) {
            fp[a] = p0[a] + fp[a];
            fv[a] = v0[a] + fv[a];
            fm[a] = m0[a] + fm[a];
            gs[a] = g0[a] + gs[a];
        }
        de = e0 + de;
    }
    if (STRESS) stress_close(RS, i, gs, fm);
    if (MERGE) {
        for (int a = 0; a < 3; ++a) {
            fp[a] = fma(pr2i, gs[a], fm[a]);
            fm[a] = 0.0;
        }
    }
    if (UNI) {
        const double ih2 = kp.inv_h0 * kp.inv_h0;
        const double k = 6.0 * ((kp.m0 * kp.m0) * ((ih2 * ih2) * kInvPi));  // 6: the sixth dW/(6r)
        const double kv = B0 ? k * cv : k;
        for (int a = 0; a < 3; ++a) {
            fp[a] *= k;
            fv[a] *= kv;
            fm[a] *= k;
        }
        if (ENERGY) D.du[i] = -de * (6.0 * (kp.m0 * ((ih2 * ih2) * kInvPi)));
    } else {
        if (B0)
            for (int a = 0; a < 3; ++a) fv[a] *= cv;
        if (ENERGY) D.du[i] = -de / mi;
    }
    force_epilogue<VISC_ONLY>(P, D, kp, charge != 0, i, fp, fv, fm, sp);
}

// Two lanes per particle, as the staged pass (same halves, same sums): at
// the sizes that run unstaged, one thread per particle is few waves per CU.
template <bool VISC_ONLY, bool ENERGY, bool STRESS, bool UNI>
__global__ void __launch_bounds__(kBlock) k_forces(Particles P, Derived D, uint32_t n, Grid g, NbrList L,
                                                    KernelParams kp, int charge) {
    const uint32_t i = pair_slot(particle_index());
    if (i >= n) return;
    const GlobalFetch<double2, 4> f{reinterpret_cast<const double2*>(D.r64)};
    if (kp.beta == 0.0)
        force_particle<VISC_ONLY, ENERGY, STRESS, UNI, true, -1, 2>(P, D, g, L, kp, charge, i, f);
    else
        force_particle<VISC_ONLY, ENERGY, STRESS, UNI, false, -1, 2>(P, D, g, L, kp, charge, i, f);
}

// Two lanes per particle (each walks half of its list): twice the waves per CU
// with the one halo that fits the LDS (measured: force 0.41 -> 0.39 ms,
// visc_only 0.40 -> 0.36 ms on the 1M dam break).  With the Maxwell stress
// the B records are staged beside the f64 records (96 bytes a record).
template <bool VISC_ONLY, bool ENERGY, bool STRESS, bool UNI, int PER>
__global__ void __launch_bounds__(2 * kStagedBlock) k_forces_staged(Particles P, Derived D, Grid g, NbrList L,
                                                                    KernelParams kp, int charge, Bricks B,
                                                                    const uint32_t* __restrict__ start,
                                                                    uint32_t cap) {
    __shared__ BrickHdr H;
    extern __shared__ uint4 sh_dyn[];
    // two entries per step (four measured slower once bricks with a whole
    // staged halo skip the staged test: VISC_ONLY 0.322 against 0.315 ms)
    constexpr int W = 2;
    constexpr int NQ2 = STRESS ? 2 : 0;
    // the next two list entries loaded ahead (with the stress: 4M MHD box
    // 1.69 -> 1.64 ms per launch despite a few spilled registers)
    constexpr bool PF = true;
    // VISC_ONLY without the stress: the neighbour's P/rho^2 (the last 8 bytes
    // of its record) is not read
    // uniform m and h: the pair body needs only the record (the slot lists;
    // not written with the Maxwell stress, whose 96-byte records leave few
    // bricks staged whole)
    brick_pass<double2, 4, W, PF, 2, NQ2, false, VISC_ONLY && !STRESS, UNI && !STRESS>(
        B, g, start, cap, L, reinterpret_cast<const double2*>(D.r64), H, sh_dyn,
        [&](uint32_t i, const auto& f) {
            if (kp.beta == 0.0)
                force_particle<VISC_ONLY, ENERGY, STRESS, UNI, true, PER, 2>(P, D, g, L, kp, charge, i, f);
            else
                force_particle<VISC_ONLY, ENERGY, STRESS, UNI, false, PER, 2>(P, D, g, L, kp, charge, i, f);
        },
        reinterpret_cast<const double2*>(D.recs));
}

// FP32-mixed force pass: f32 pair terms, f64 sums, same epilogue (and the
// ENERGY sum as in the f64 pass, f32 per group of four entries).
// With STRESS the list is summed in two halves, entries 0 .. half - 1 and
// half .. end (half a multiple of 4 near the middle), added in that order:
// with LANES = 2 the lanes l and l + 32 of a wave take one half each.
template <bool VISC_ONLY, bool ENERGY, bool STRESS, bool UNI, bool B0, int LANES, class Fetch>
__device__ __forceinline__ void force_mixed_particle(const Particles& P, const Derived& D, const Mixed& X,
                                                     const Grid& g, const NbrList& L, const KernelParams& kp,
                                                     int charge, uint32_t i, const Fetch& fetch) {
    const double2* __restrict__ RS = reinterpret_cast<const double2*>(D.recs);
    float4 own[Fetch::NR];  // the particle's own gather record
    fetch.own(i, own);
    const float4 own0 = own[0], own1 = own[1];
    double sp[3] = {0, 0, 0};  // VISC_ONLY: the stored sum, loaded ahead of the walk
    if (VISC_ONLY) {
        sp[0] = D.spx[i]; sp[1] = D.spy[i]; sp[2] = D.spz[i];
    }
    const float xi = own0.x, yi = own0.y, zi = own0.z, rhoi = own0.w;
    const float vxi = own1.x, vyi = own1.y, vzi = own1.z, pr2i = own1.w;
    const float hi = X.p4[i].w, mi = UNI ? 0.0f : X.m[i];
    const float inv_h0 = (float)kp.inv_h0, h0 = (float)kp.h0;
    // viscosity with one reciprocal, as in the f64 pass:
    // Pi = hs (2 beta hs - 2 alpha c den) / (den^2 (rho_i + rho_j)), hs = h s
    const float ac2 = (float)(-2.0 * kp.alpha * kp.speed_of_sound), beta2 = (float)(2.0 * kp.beta);
    const bool ref_conv = kp.visc_conv == SPHX_VISC_REFERENCE;
    const float ssign = ref_conv ? 1.0f : -1.0f;
    // s clamped to [s_lo, s_hi] (one med3) is 0 for a pair without a viscosity term
    const float s_lo = ref_conv ? -INFINITY : 0.0f, s_hi = ref_conv ? 0.0f : INFINITY;
    // B0 (beta = 0): Pi = cv s' / (den (rho_i + rho_j)) with cv = -2 alpha c
    // sv (and h, when uniform) folded out of the sum, as in the f64 pass
    const double cv = B0 ? (UNI ? (-2.0 * kp.alpha * kp.speed_of_sound) * (double)ssign * kp.h0
                                : (-2.0 * kp.alpha * kp.speed_of_sound) * (double)ssign)
                         : 1.0;
    const float hcv = (float)(0.5 * cv);
    const float wx = (float)g.w[0], wy = (float)g.w[1], wz = (float)g.w[2];
    double fp[3] = {0, 0, 0}, fv[3] = {0, 0, 0}, fm[3] = {0, 0, 0};
    // f32 sums over each group of four list entries, added to the f64 sums
    // once per group; with uniform m and h, m_i m_j sigma_g is folded out
    float gp[3] = {0, 0, 0}, gv[3] = {0, 0, 0};
    double de = 0.0;
    float ge = 0.0f;
    // STRESS: the j terms and gs = sum (dW/r) d as in the f64 pass, f32 per
    // group of four entries
    double gs[3] = {0, 0, 0};
    float gm[3] = {0, 0, 0}, gg[3] = {0, 0, 0};
    // branch-free pair body, as in the f64 pass (factors 1 or 0, not selects)
    auto pair = [&](uint32_t e, const float4 (&q)[Fetch::NR]) {
        const uint32_t j = e & kIdxMask;
        const float4 a = q[0], b = q[1];
        float dx, dy, dz;
        if constexpr (Fetch::BREL) {  // brick-relative records: no cell offset
            dx = a.x - xi;
            dy = a.y - yi;
            dz = a.z - zi;
        } else {
            float ox, oy, oz;
            fetch.offset(e, wx, wy, wz, ox, oy, oz);
            dx = (a.x - xi) + ox;
            dy = (a.y - yi) + oy;
            dz = (a.z - zi) + oz;
        }
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        // r <= 1e-10: no force; 1/r = 0 gives r = 0, q = 0 and dW/dq(0) = 0
        const float inv_r = r2 > 1e-20f ? __builtin_amdgcn_rsqf(r2) : 0.0f;
        const float r = r2 * inv_r;
        float inv_h = inv_h0, h_avg = h0;
        if (!UNI) {
            h_avg = 0.5f * (hi + X.p4[j].w);
            inv_h = __builtin_amdgcn_rcpf(h_avg);
        }
        // dW/dq / 3 = s^2 - t^2 from the one-sided pieces s = (1 - q)+, t =
        // (1 - q/2)+ formed from r (no q, no select: 1M mixed force 0.155 ->
        // 0.152 ms); r = 0 gives s = t = 1, so 0
        const float ps = fmaxf(fmaf(r, -inv_h, 1.0f), 0.0f), pt = fmaxf(fmaf(r, -0.5f * inv_h, 1.0f), 0.0f);
        const float dwr0 = fmaf(-pt, pt, ps * ps) * inv_r;  // its 3 is folded in below
        float dwr = dwr0;
        if (!UNI) {
            const float ih2 = inv_h * inv_h;
            dwr = (mi * X.m[j]) * (((ih2 * ih2) * (float)(3.0 * kInvPi)) * dwr0);
        }
        if (!VISC_ONLY) {
            const float sc = (pr2i + b.w) * dwr;
            gp[0] = fmaf(dx, sc, gp[0]);
            gp[1] = fmaf(dy, sc, gp[1]);
            gp[2] = fmaf(dz, sc, gp[2]);
        }
        const float sdot = fmaf(vzi - b.z, dz, fmaf(vyi - b.y, dy, (vxi - b.x) * dx));
        const float rhoj = a.w;
        {
            const float ms = __builtin_amdgcn_fmed3f(sdot, s_lo, s_hi);
            const float den = fmaf(0.01f * h_avg, h_avg, r2);
            float pi_ij;  // Pi / cv
            if (B0) {
                pi_ij = (UNI ? ms : h_avg * ms) * __builtin_amdgcn_rcpf(den * (rhoi + rhoj));
            } else {
                const float hs = (ssign * h_avg) * ms;
                pi_ij = (hs * fmaf(beta2, hs, ac2 * den)) * __builtin_amdgcn_rcpf((den * den) * (rhoi + rhoj));
            }
            const float sc = pi_ij * dwr;
            gv[0] = fmaf(dx, sc, gv[0]);
            gv[1] = fmaf(dy, sc, gv[1]);
            gv[2] = fmaf(dz, sc, gv[2]);
            if (ENERGY) ge = fmaf(fmaf(hcv, pi_ij, pr2i), sdot * dwr, ge);
        }
        if (STRESS) {
            // the stress record of j: quads 2 and 3 of a staged record hold
            // the bits of (b_x, b_y) and (b_z, e); unstaged it is read from
            // the stress records
            double2 sxy, sze;
            if constexpr (Fetch::NR >= 4) {
                const float4 u = q[2], v = q[3];
                sxy = make_double2(__hiloint2double(__float_as_int(u.y), __float_as_int(u.x)),
                                   __hiloint2double(__float_as_int(u.w), __float_as_int(u.z)));
                sze = make_double2(__hiloint2double(__float_as_int(v.y), __float_as_int(v.x)),
                                   __hiloint2double(__float_as_int(v.w), __float_as_int(v.z)));
            } else {
                sxy = RS[2 * j];
                sze = RS[2 * j + 1];
            }
            const float bx = (float)sxy.x, by = (float)sxy.y, bz = (float)sze.x, ej = (float)sze.y;
            gg[0] = fmaf(dwr, dx, gg[0]);
            gg[1] = fmaf(dwr, dy, gg[1]);
            gg[2] = fmaf(dwr, dz, gg[2]);
            const float bd = fmaf(bz, dz, fmaf(by, dy, bx * dx));
            gm[0] = fmaf(dwr, fmaf(bx, bd, -ej * dx), gm[0]);
            gm[1] = fmaf(dwr, fmaf(by, bd, -ej * dy), gm[1]);
            gm[2] = fmaf(dwr, fmaf(bz, bd, -ej * dz), gm[2]);
        }
    };
    auto flush = [&]() {
        for (int a = 0; a < 3; ++a) {
            if (!VISC_ONLY) fp[a] += (double)gp[a];
            fv[a] += (double)gv[a];
            gp[a] = gv[a] = 0.0f;
            if (STRESS) {
                fm[a] += (double)gm[a];
                gs[a] += (double)gg[a];
                gm[a] = gg[a] = 0.0f;
            }
        }
        if (ENERGY) {
            de += (double)ge;
            ge = 0.0f;
        }
    };
    auto sum_range = [&](uint32_t kb, uint32_t ke) {
        Fetch f = fetch;
        f.kbeg = kb;
        f.kend = ke;
        walk(L, i, f, pair, flush);
    };
    if (!STRESS) {
        walk(L, i, fetch, pair, flush);
    } else {
        const uint32_t cnt = fetch.count(L, i), half = ((cnt + 7u) >> 3) << 2;
        if (LANES == 2) {
            const uint32_t h = lane_half();
            sum_range(h ? half : 0u, h ? 0xffffffffu : half);
            for (int a = 0; a < 3; ++a) {  // lane half 0: first half + second half
                if (!VISC_ONLY) fp[a] += other_half(fp[a]);
                fv[a] += other_half(fv[a]);
                fm[a] += other_half(fm[a]);
                gs[a] += other_half(gs[a]);
            }
            if (ENERGY) de += other_half(de);
            if (h) return;
        } else {
            sum_range(0u, half);
            const double p0[3] = {fp[0], fp[1], fp[2]}, v0[3] = {fv[0], fv[1], fv[2]};
            const double m0[3] = {fm[0], fm[1], fm[2]}, g0[3] = {gs[0], gs[1], gs[2]};
            const double e0 = de;
            for (int a = 0; a < 3; ++a) fp[a] = fv[a] = fm[a] = gs[a] = 0.0;
            de = 0.0;
            sum_range(half, 0xffffffffu);
            for (int a = 0; a < 3; ++a) {
                fp[a] = p0[a] + fp[a];
                fv[a] = v0[a] + fv[a];
                fm[a] = m0[a] + fm[a];
                gs[a] = g0[a] + gs[a];
            }
            de = e0 + de;
        }
    }
    if (STRESS) stress_close(RS, i, gs, fm);
    if (UNI) {
        const double ih2 = kp.inv_h0 * kp.inv_h0;
        const double k = ((kp.m0 * kp.m0) * ((ih2 * ih2) * kInvPi)) * 3.0;
        const double kv = B0 ? k * cv : k;
        for (int a = 0; a < 3; ++a) {
            fp[a] *= k;
            fv[a] *= kv;
            fm[a] *= k;
        }
        if (ENERGY) D.du[i] = -de * ((kp.m0 * ((ih2 * ih2) * kInvPi)) * 3.0);
    } else {
        if (B0)
            for (int a = 0; a < 3; ++a) fv[a] *= cv;
        if (ENERGY) D.du[i] = -de / P.m[i];
    }
    force_epilogue<VISC_ONLY>(P, D, kp, charge != 0, i, fp, fv, fm, sp);
}

template <bool VISC_ONLY, bool ENERGY, bool STRESS, bool UNI>
__global__ void __launch_bounds__(kBlock) k_forces_mixed(Particles P, Derived D, Mixed X, uint32_t n, Grid g,
                                                          NbrList L, KernelParams kp, int charge) {
    SPHX_PASS_PROLOGUE();
    const GlobalFetch<float4, 2> f{reinterpret_cast<const float4*>(D.r32)};
    if (kp.beta == 0.0)
        force_mixed_particle<VISC_ONLY, ENERGY, STRESS, UNI, true, 1>(P, D, X, g, L, kp, charge, i, f);
    else
        force_mixed_particle<VISC_ONLY, ENERGY, STRESS, UNI, false, 1>(P, D, X, g, L, kp, charge, i, f);
}

// Staged FP32-mixed force pass without the Maxwell stress: one lane per
// particle, 32-byte records.
template <bool VISC_ONLY, bool ENERGY, bool UNI, int W, bool PF>
__global__ void __launch_bounds__(kStagedBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_forces_mixed_staged(
    Particles P, Derived D, Mixed X, Grid g, NbrList L, KernelParams kp, int charge, Bricks B,
    const uint32_t* __restrict__ start, uint32_t cap) {
    __shared__ BrickHdr H;
    extern __shared__ uint4 sh_dyn[];
    // the full pass holds its list prefetch at the top of each step (faster
    // there; slower in the viscosity-only pass)
    // brick-relative records; uniform m and h: the pair body needs only the
    // record (the slot lists)
    brick_pass<float4, 2, W, PF, 1, 0, !VISC_ONLY, false, UNI, true>(
        B, g, start, cap, L, reinterpret_cast<const float4*>(D.r32), H, sh_dyn,
        [&](uint32_t i, const auto& f) {
            if (kp.beta == 0.0)
                force_mixed_particle<VISC_ONLY, ENERGY, false, UNI, true, 1>(P, D, X, g, L, kp, charge, i, f);
            else
                force_mixed_particle<VISC_ONLY, ENERGY, false, UNI, false, 1>(P, D, X, g, L, kp, charge, i, f);
        });
}

// With the Maxwell stress the B records (bx, by, bz, psi) in f64 are staged
// beside the f32 records as quads 2 and 3 (64 bytes a record); its halo fills
// one CU's LDS, so two lanes per particle (lanes l and l + 32, each walking
// half of the list) give the CU four waves per SIMD.
template <bool VISC_ONLY, bool ENERGY, bool UNI>
__global__ void __launch_bounds__(2 * kStagedBlock) k_forces_mixed_stress_staged(
    Particles P, Derived D, Mixed X, Grid g, NbrList L, KernelParams kp, int charge, Bricks B,
    const uint32_t* __restrict__ start, uint32_t cap) {
    __shared__ BrickHdr H;
    extern __shared__ uint4 sh_dyn[];
    brick_pass<float4, 2, 4, false, 2, 2>(
        B, g, start, cap, L, reinterpret_cast<const float4*>(D.r32), H, sh_dyn,
        [&](uint32_t i, const BrickFetch<float4, 2, 4, false, 2>& f) {
            if (kp.beta == 0.0)
                force_mixed_particle<VISC_ONLY, ENERGY, true, UNI, true, 2>(P, D, X, g, L, kp, charge, i, f);
            else
                force_mixed_particle<VISC_ONLY, ENERGY, true, UNI, false, 2>(P, D, X, g, L, kp, charge, i, f);
        },
        reinterpret_cast<const float4*>(D.recs));
}

void launch_forces(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                   const KernelParams& kp, bool visc_only, bool energy, bool stress, bool charge, bool mixed,
                   const Mixed& X, const Bricks& B, const uint32_t* cell_start, hipStream_t s) {
    if (!n) return;
    const dim3 grid(nblocks(n)), block(kBlock);
    const int ch = charge ? 1 : 0;
    const bool uni = kp.uniform != 0;
    energy = energy && !visc_only;  // the energy equation needs the pressure term of a full pass
    if (B.on && B.nactive) {
#define SPHX_FORCES_STAGED(V, E, M, U)                                                                      \
        if (visc_only == V && energy == E && stress == M && uni == U) {                                     \
            if (mixed && M) {                                                                               \
                StagedLaunch sl = staged_launch(B, 8, 64);                                                  \
                sl.block = dim3(2u * (uint32_t)B.threads); /* two lanes per particle */                      \
                hipLaunchKernelGGL((k_forces_mixed_stress_staged<V, E, U>), sl.grid, sl.block, sl.dyn, s, P, D, X, \
                                   g, L, kp, ch, B, cell_start, sl.cap);                                    \
            } else if (mixed) {                                                                             \
                StagedLaunch sl = staged_launch(B, 16, 32);                                                 \
                if (L.s16) sl.cap = std::min(sl.cap, slot_cap(B, true));                                    \
                hipLaunchKernelGGL((k_forces_mixed_staged<V, E, U, 4, true>), sl.grid, sl.block, sl.dyn, s, P, D, \
                                   X, g, L, kp, ch, B, cell_start, sl.cap);                                 \
            } else {                                                                                        \
                StagedLaunch sl = staged_launch(B, 8, M ? 96 : 64);                                         \
                sl.block = dim3(2u * (uint32_t)B.threads); /* two lanes per particle */                      \
                if (L.s16) sl.cap = std::min(sl.cap, slot_cap(B, false));                                   \
                if (g.periodic)                                                                             \
                    hipLaunchKernelGGL((k_forces_staged<V, E, M, U, 1>), sl.grid, sl.block, sl.dyn, s, P, D, g, \
                                       L, kp, ch, B, cell_start, sl.cap);                                   \
                else                                                                                        \
                    hipLaunchKernelGGL((k_forces_staged<V, E, M, U, 0>), sl.grid, sl.block, sl.dyn, s, P, D, g, \
                                       L, kp, ch, B, cell_start, sl.cap);                                   \
            }                                                                                               \
            return;                                                                                         \
        }
#define SPHX_FORCES_STAGED_MU(M, U)        \
        SPHX_FORCES_STAGED(false, false, M, U) \
        SPHX_FORCES_STAGED(true, false, M, U)  \
        SPHX_FORCES_STAGED(false, true, M, U)
        SPHX_FORCES_STAGED_MU(false, false)
        SPHX_FORCES_STAGED_MU(false, true)
        SPHX_FORCES_STAGED_MU(true, false)
        SPHX_FORCES_STAGED_MU(true, true)
#undef SPHX_FORCES_STAGED_MU
#undef SPHX_FORCES_STAGED
    }
#define SPHX_FORCES(V, E, M, U)                                                                           \
    if (visc_only == V && energy == E && stress == M && uni == U) {                                      \
        if (mixed)                                                                                        \
            hipLaunchKernelGGL((k_forces_mixed<V, E, M, U>), grid, block, occupancy_lds(4), s, P, D, X, n, g, L, \
                               kp, ch);                                                                   \
        else                                                                                              \
            hipLaunchKernelGGL((k_forces<V, E, M, U>), dim3(nblocks(2 * n)), block, M ? 0 : occupancy_lds(3), s, \
                               P, D, n, g, L, kp, ch);                                                    \
        return;                                                                                           \
    }
#define SPHX_FORCES_MU(M, U)        \
    SPHX_FORCES(false, false, M, U) \
    SPHX_FORCES(true, false, M, U)  \
    SPHX_FORCES(false, true, M, U)
    SPHX_FORCES_MU(false, false)
    SPHX_FORCES_MU(true, false)
    SPHX_FORCES_MU(false, true)
    SPHX_FORCES_MU(true, true)
#undef SPHX_FORCES_MU
#undef SPHX_FORCES
}

// ---------------------------------------------------------------- XSPH (a16)
// sph_xsph_system (sph.rs:349-405); Jacobi: the new velocity goes to tmp3.
__global__ void __launch_bounds__(kBlock) k_xsph(Particles P, Derived D, Mixed X, uint32_t n, Grid g, NbrList L,
                                                  KernelParams kp, int mixed) {
    SPHX_PASS_PROLOGUE();
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i];
    const double vxi = P.vx[i], vyi = P.vy[i], vzi = P.vz[i], rhoi = D.rho[i];
    double ux = 0, uy = 0, uz = 0;
    for_neighbours(L, i, [&](uint32_t e) {
        const uint32_t j = e & kIdxMask;
        double dx, dy, dz;
        pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], dx, dy, dz);
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        double r = 0.0, inv_r;
        if (r2 > 0.0) sqrt_rsqrt(r2, r, inv_r);
        const PairH ph = pair_h<false>(kp, hi, P.h[j]);
        const double w = ph.sig_w * w_shape(r * ph.inv_h);
        const double rho_bar = 0.5 * (rhoi + D.rho[j]);
        const double f = (kp.xsph_eps * (P.m[j] * rcp_nr(rho_bar))) * w;
        ux = fma(f, P.vx[j] - vxi, ux);
        uy = fma(f, P.vy[j] - vyi, uy);
        uz = fma(f, P.vz[j] - vzi, uz);
    });
    D.tmp3[0][i] = vxi + ux;
    D.tmp3[1][i] = vyi + uy;
    D.tmp3[2][i] = vzi + uz;
    // the record is read by the next force pass, not by this one (Jacobi)
    set_record_velocity(D, kp, i, mixed != 0, vxi + ux, vyi + uy, vzi + uz);
}
void launch_xsph(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                 const KernelParams& kp, bool mixed, const Mixed& X, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_xsph, dim3(nblocks(n)), dim3(kBlock), 0, s, P, D, X, n, g, L, kp, (int)mixed);
}

// ------------------------------------------------------ MHD induction (a11)
// mhd_induction_system (mhd.rs:68-152) in gather form.  The reference adds
// the pair term to entity_a and its negation to entity_b (mhd.rs:121-122,
// 130-131), so particle i's term carries +1 when id_i < id_j, -1 otherwise.
// Neighbours come from the f64 record (x, rho, v) and the B record (one
// 6-quad record); the new B goes to tmp3 and, with psi, to recb_next (Jacobi).
// (dW/dr) / r of a pair of the MHD passes from r^2, 0 for r <= 1e-10
// (mhd.rs:100, where the reference skips the pair): the clamp keeps 1/r
// finite, and below the cut r keeps only its low word (a denormal) so that
// the spline's one-sided pieces give exactly 0.  One refinement step on the
// hardware rsqrt (the terms are compared relative to the largest change).
template <bool UNI>
__device__ __forceinline__ double mhd_dwr(const KernelParams& kp, double r2, double hi, double hj) {
    double r, inv_r;
    sqrt_rsqrt1(fmax(r2, 1e-300), r, inv_r);
    r = __hiloint2double(r2 > 1e-20 ? __double2hiint(r) : 0, __double2loint(r));
    const PairH ph = pair_h<UNI>(kp, hi, hj);
    const double nih = -ph.inv_h;
    return ((3.0 * ph.sig_g) * dw3_shape_r(r, nih, 0.5 * nih)) * inv_r;
}

// Two lanes per particle: lane h (lanes l and l + 32 of a wave) walks half h of the list,
// two neighbour records loaded ahead of their bodies, and the halves are added
// in order.  The pair body is branch-free (a factor 0 for r <= 1e-10, where
// the reference skips the pair: mhd.rs:100).  RES: the resistive term (eta != 0).
// range(kb, ke, body) walks list entries kb .. ke - 1 of particle i with
// records of 6 quads (the f64 record, then the B record).
template <bool UNI, bool RES, int PER, class Own, class Range>
__device__ __forceinline__ void induction_particle(const Particles& P, const Derived& D, const Grid& g,
                                                   const NbrList& L, const KernelParams& kp, uint32_t i,
                                                   Own&& own_record, Range&& range) {
    const uint32_t h = lane_half();
    double2 own[6];  // the particle's own records: (x, y | z, rho | vx, vy | vz, P/rho^2 | bx, by | bz, psi)
    own_record(own);
    const double xi = own[0].x, yi = own[0].y, zi = own[1].x, hi = UNI ? kp.h0 : P.h[i];
    const double vxi = own[2].x, vyi = own[2].y, vzi = own[3].x, rhoi = own[1].y;
    const double bxi = own[4].x, byi = own[4].y, bzi = own[5].x;
    const uint32_t idi = P.id[i];
    // m_j / (rho_i rho_j) = mi_rho / rho_j (uniform m: mi_rho = m0 / rho_i)
    const double mi_rho = (UNI ? kp.m0 : 1.0) * rcp_nr(rhoi);
    double ux = 0, uy = 0, uz = 0;
    const uint32_t cnt = L.ntot[i], half = ((cnt + 7u) >> 3) << 2;
    // per pair, with g = (dW/r) d (d = x_j - x_i) and v = v_i - v_j:
    //   dB_i += +-dt m_j / (rho_i rho_j) (B_i (v . g) - v (B_i . g))
    // = f (B_i (v . d) - v (B_i . d)),  f = +-dt (m_j / (rho_i rho_j)) dW/r,
    // one hardware-estimate refinement step for 1/r and 1/rho_j (the terms
    // are compared relative to the largest change of B)
    range(h ? half : 0u, h ? 0xffffffffu : half, [&](uint32_t e, const double2 (&q)[6]) {
        const uint32_t j = e & kIdxMask;
        const double2 xy = q[0], zr = q[1], vxy = q[2], vzp = q[3], bxy = q[4], bzp = q[5];
        double dx, dy, dz;
        pair_vec<PER>(g, xi, yi, zi, xy.x, xy.y, zr.x, dx, dy, dz);
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double dwr = mhd_dwr<UNI>(kp, r2, hi, UNI ? hi : P.h[j]);
        const double vx = vxi - vxy.x, vy = vyi - vxy.y, vz = vzi - vzp.x;
        const double rhoj = zr.y;
        const double v_dot_d = fma(vz, dz, fma(vy, dy, vx * dx));
        const double b_dot_d = fma(bzi, dz, fma(byi, dy, bxi * dx));
        const double inv_rhoj = rcp_nr1(rhoj);
        const double sdt = (idi < P.id[j]) ? kp.dt : -kp.dt;
        const double f = (sdt * dwr) * ((UNI ? mi_rho : P.m[j] * mi_rho) * inv_rhoj);
        if (RES) {
            const double mj = UNI ? kp.m0 : P.m[j];
            const double r_dot_g = r2 * dwr;
            const double h_avg = UNI ? kp.h0 : 0.5 * (hi + P.h[j]);
            const double res = sdt * ((2.0 * kp.eta * (mj * inv_rhoj)) * r_dot_g *
                                      rcp_nr(fma(0.01 * h_avg, h_avg, r2)));
            ux = fma(f, fma(bxi, v_dot_d, -vx * b_dot_d), fma(res, bxy.x - bxi, ux));
            uy = fma(f, fma(byi, v_dot_d, -vy * b_dot_d), fma(res, bxy.y - byi, uy));
            uz = fma(f, fma(bzi, v_dot_d, -vz * b_dot_d), fma(res, bzp.x - bzi, uz));
        } else {
            ux = fma(f, fma(bxi, v_dot_d, -vx * b_dot_d), ux);
            uy = fma(f, fma(byi, v_dot_d, -vy * b_dot_d), uy);
            uz = fma(f, fma(bzi, v_dot_d, -vz * b_dot_d), uz);
        }
    });
    ux += other_half(ux);  // lane 0: first half + second half
    uy += other_half(uy);
    uz += other_half(uz);
    if (h) return;
    const double bx = bxi + ux, by = byi + uy, bz = bzi + uz;
    D.tmp3[0][i] = bx;
    D.tmp3[1][i] = by;
    D.tmp3[2][i] = bz;
    double2* rb = reinterpret_cast<double2*>(D.recb_next + i);
    rb[0] = make_double2(bx, by);
    rb[1] = make_double2(bz, own[5].y);
}

template <bool UNI, bool RES, int PER>
__global__ void __launch_bounds__(kBlock) k_induction(Particles P, Derived D, uint32_t n, Grid g, NbrList L,
                                                       KernelParams kp) {
    const uint32_t i = pair_slot(particle_index());
    if (i >= n) return;
    const GlobalFetch<double2, 4, 2> F{reinterpret_cast<const double2*>(D.r64),
                                       reinterpret_cast<const double2*>(D.recb)};
    induction_particle<UNI, RES, PER>(P, D, g, L, kp, i, [&](double2 (&q)[6]) { F.own(i, q); },
                                      [&](uint32_t kb, uint32_t ke, auto&& body) {
                                          walk_ahead<4>(L, i, kb, ke, F, body);
                                      });
}

// Brick-staged: the neighbours' f64 and B records (96 bytes) from LDS, the
// same halves, order and arithmetic as k_induction (bit-identical); two
// entries per step with the next two loaded ahead (4M MHD box: 1.48 ->
// 1.37 ms per launch; unstaged 1.50 ms, bound by the gathers' address path).
template <bool UNI, bool RES, int PER>
__global__ void __launch_bounds__(2 * kStagedBlock) k_induction_staged(Particles P, Derived D, Grid g, NbrList L,
                                                                       KernelParams kp, Bricks B,
                                                                       const uint32_t* __restrict__ start,
                                                                       uint32_t cap) {
    __shared__ BrickHdr H;
    extern __shared__ uint4 sh_dyn[];
    using Fetch = BrickFetch<double2, 4, 2, true, 2>;
    brick_pass<double2, 4, 2, true, 2, 2>(
        B, g, start, cap, L, reinterpret_cast<const double2*>(D.r64), H, sh_dyn,
        [&](uint32_t i, const Fetch& f) {
            induction_particle<UNI, RES, PER>(P, D, g, L, kp, i, [&](double2 (&q)[6]) { f.own(i, q); },
                                              [&](uint32_t kb, uint32_t ke, auto&& body) {
                Fetch ff = f;
                ff.kbeg = kb;
                ff.kend = ke;
                walk(L, i, ff, body);
            });
        },
        reinterpret_cast<const double2*>(D.recb));
}

// FP32-mixed MHD passes: f32 pair terms on the cell-relative f32 record (as
// the mixed force pass), the neighbour's B converted to f32, each group of
// four list entries summed in f32 and added to f64 sums; B, psi and the
// update stay f64.  dW/dr / r with its 0.75 folded into sigma_g.
struct MixedPairG {
    float dx, dy, dz, r2, dwr, h_avg;  // dwr = (dW/dr) / r
};
template <bool UNI, class Fetch>
__device__ __forceinline__ MixedPairG mixed_pair_g(const Mixed& X, const Fetch& fetch, float hi, float inv_h0,
                                                   float h0, float sig0, uint32_t e, const float4& a, float xi,
                                                   float yi, float zi, float wx, float wy, float wz) {
    MixedPairG p;
    float ox, oy, oz;
    fetch.offset(e, wx, wy, wz, ox, oy, oz);
    p.dx = (a.x - xi) + ox;
    p.dy = (a.y - yi) + oy;
    p.dz = (a.z - zi) + oz;
    p.r2 = fmaf(p.dz, p.dz, fmaf(p.dy, p.dy, p.dx * p.dx));
    // r <= 1e-10: no term (1/r = 0 gives q = 0 and dW/dq(0) = 0)
    const float inv_r = p.r2 > 1e-20f ? __builtin_amdgcn_rsqf(p.r2) : 0.0f;
    float inv_h = inv_h0, sig = sig0;
    p.h_avg = h0;
    if (!UNI) {
        p.h_avg = 0.5f * (hi + X.p4[e & kIdxMask].w);
        inv_h = __builtin_amdgcn_rcpf(p.h_avg);
        const float ih2 = inv_h * inv_h;
        sig = (ih2 * ih2) * (float)(0.75 * kInvPi);
    }
    p.dwr = (sig * dw43_shape_f((p.r2 * inv_r) * inv_h)) * inv_r;
    return p;
}

// (b_x, b_y, b_z, psi) of a neighbour from quads 2 and 3 of its record (the
// bits of the f64 B record), in f32
__device__ __forceinline__ float4 mixed_b_record(const float4& u, const float4& v) {
    return make_float4((float)__hiloint2double(__float_as_int(u.y), __float_as_int(u.x)),
                       (float)__hiloint2double(__float_as_int(u.w), __float_as_int(u.z)),
                       (float)__hiloint2double(__float_as_int(v.y), __float_as_int(v.x)),
                       (float)__hiloint2double(__float_as_int(v.w), __float_as_int(v.z)));
}

// FP32-mixed induction of particle i: walk(fetch) gives records of NR quads,
// the f32 record (x, rho | v, P/rho^2) and, with RES, the B record.  Per pair
// dB_i += f (B_i (v . d) - v (B_i . d)), f = +-dt m_j / (rho_i rho_j) (dW/r).
template <bool UNI, bool RES, class Fetch>
__device__ __forceinline__ void induction_mixed_particle(const Particles& P, const Derived& D, const Mixed& X,
                                                         const Grid& g, const NbrList& L, const KernelParams& kp,
                                                         uint32_t i, const Fetch& fetch) {
    const float4* __restrict__ R = reinterpret_cast<const float4*>(D.r32);
    const float4 own0 = R[2 * i], own1 = R[2 * i + 1];
    const float xi = own0.x, yi = own0.y, zi = own0.z, rhoi = own0.w;
    const float vxi = own1.x, vyi = own1.y, vzi = own1.z;
    const double bxd = P.bx[i], byd = P.by[i], bzd = P.bz[i];
    const float bxi = (float)bxd, byi = (float)byd, bzi = (float)bzd;
    const uint32_t idi = P.id[i];
    const float hi = X.p4[i].w, inv_rhoi = __builtin_amdgcn_rcpf(rhoi);
    const float inv_h0 = (float)kp.inv_h0, h0 = (float)kp.h0, m0 = (float)kp.m0;
    const float sig0 = (float)((kp.inv_h0 * kp.inv_h0) * (kp.inv_h0 * kp.inv_h0) * (0.75 * kInvPi));
    const float wx = (float)g.w[0], wy = (float)g.w[1], wz = (float)g.w[2];
    const float eta2 = (float)(2.0 * kp.eta), dt = (float)kp.dt;
    const float mi_rho = (UNI ? m0 : 1.0f) * inv_rhoi;
    double u[3] = {0, 0, 0};
    float gu[3] = {0, 0, 0};
    walk(L, i, fetch, [&](uint32_t e, const float4 (&q)[Fetch::NR]) {
        const uint32_t j = e & kIdxMask;
        const MixedPairG p = mixed_pair_g<UNI>(X, fetch, hi, inv_h0, h0, sig0, e, q[0], xi, yi, zi, wx, wy, wz);
        const float vx = vxi - q[1].x, vy = vyi - q[1].y, vz = vzi - q[1].z;
        const float inv_rhoj = __builtin_amdgcn_rcpf(q[0].w);
        const float v_dot_d = fmaf(vz, p.dz, fmaf(vy, p.dy, vx * p.dx));
        const float b_dot_d = fmaf(bzi, p.dz, fmaf(byi, p.dy, bxi * p.dx));
        const float sdt = (idi < P.id[j]) ? dt : -dt;
        const float f = (sdt * p.dwr) * ((UNI ? mi_rho : X.m[j] * mi_rho) * inv_rhoj);
        if constexpr (RES) {
            const float4 bj = mixed_b_record(q[2], q[3]);
            const float mj = UNI ? m0 : X.m[j];
            const float res = sdt * (((eta2 * (mj * inv_rhoj)) * (p.r2 * p.dwr)) *
                                     __builtin_amdgcn_rcpf(fmaf(0.01f * p.h_avg, p.h_avg, p.r2)));
            gu[0] = fmaf(f, fmaf(bxi, v_dot_d, -vx * b_dot_d), fmaf(res, bj.x - bxi, gu[0]));
            gu[1] = fmaf(f, fmaf(byi, v_dot_d, -vy * b_dot_d), fmaf(res, bj.y - byi, gu[1]));
            gu[2] = fmaf(f, fmaf(bzi, v_dot_d, -vz * b_dot_d), fmaf(res, bj.z - bzi, gu[2]));
        } else {
            gu[0] = fmaf(f, fmaf(bxi, v_dot_d, -vx * b_dot_d), gu[0]);
            gu[1] = fmaf(f, fmaf(byi, v_dot_d, -vy * b_dot_d), gu[1]);
            gu[2] = fmaf(f, fmaf(bzi, v_dot_d, -vz * b_dot_d), gu[2]);
        }
    }, [&]() {
        for (int a = 0; a < 3; ++a) {
            u[a] += (double)gu[a];
            gu[a] = 0.0f;
        }
    });
    const double bx = bxd + u[0], by = byd + u[1], bz = bzd + u[2];
    D.tmp3[0][i] = bx;
    D.tmp3[1][i] = by;
    D.tmp3[2][i] = bz;
    double2* rb = reinterpret_cast<double2*>(D.recb_next + i);
    rb[0] = make_double2(bx, by);
    rb[1] = make_double2(bz, P.psi[i]);
}

template <bool UNI, bool RES>
__global__ void __launch_bounds__(kBlock) k_induction_mixed(Particles P, Derived D, Mixed X, uint32_t n, Grid g,
                                                             NbrList L, KernelParams kp) {
    SPHX_PASS_PROLOGUE();
    induction_mixed_particle<UNI, RES>(
        P, D, X, g, L, kp, i,
        GlobalFetch<float4, 2, RES ? 2 : 0>{reinterpret_cast<const float4*>(D.r32),
                                            reinterpret_cast<const float4*>(D.recb)});
}

// Brick-staged: the f32 records (and, with RES, the B records) from LDS, four
// entries a step with the next four loaded ahead (bit-identical to the
// gather form).
template <bool UNI, bool RES>
__global__ void __launch_bounds__(kStagedBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_induction_mixed_staged(Particles P, Derived D, Mixed X, Grid g, NbrList L, KernelParams kp, Bricks B,
                         const uint32_t* __restrict__ start, uint32_t cap) {
    constexpr int NQ2 = RES ? 2 : 0;
    __shared__ BrickHdr H;
    extern __shared__ uint4 sh_dyn[];
    brick_pass<float4, 2, 4, true, 1, NQ2>(
        B, g, start, cap, L, reinterpret_cast<const float4*>(D.r32), H, sh_dyn,
        [&](uint32_t i, const BrickFetch<float4, 2, 4, true, NQ2>& f) {
            induction_mixed_particle<UNI, RES>(P, D, X, g, L, kp, i, f);
        },
        reinterpret_cast<const float4*>(D.recb));
}

// Brick-staged wherever the force passes are; the gather forms below that.
void launch_induction(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                      const KernelParams& kp, bool mixed, const Mixed& X, const Bricks& B,
                      const uint32_t* cell_start, hipStream_t s) {
    if (!n) return;
    const bool res = kp.eta != 0.0, staged = B.on && B.nactive;
    if (mixed) {
#define SPHX_INDUCTION_MIXED(U, R)                                                                             \
        if (kp.uniform == (U ? 1 : 0) && res == R) {                                                           \
            if (staged) {                                                                                      \
                const StagedLaunch sl = staged_launch(B, 16, R ? 64 : 32);                                     \
                hipLaunchKernelGGL((k_induction_mixed_staged<U, R>), sl.grid, sl.block, sl.dyn, s, P, D, X, g, L, \
                                   kp, B, cell_start, sl.cap);                                                 \
            } else {                                                                                           \
                hipLaunchKernelGGL((k_induction_mixed<U, R>), dim3(nblocks(n)), dim3(kBlock), occupancy_lds(3), s, \
                                   P, D, X, n, g, L, kp);                                                      \
            }                                                                                                  \
            return;                                                                                            \
        }
        SPHX_INDUCTION_MIXED(true, false)
        SPHX_INDUCTION_MIXED(true, true)
        SPHX_INDUCTION_MIXED(false, false)
        SPHX_INDUCTION_MIXED(false, true)
#undef SPHX_INDUCTION_MIXED
        return;
    }
    const dim3 grid2(nblocks(2 * n));  // two lanes per particle
    StagedLaunch sl{};
    if (staged) {
        sl = staged_launch(B, 8, 96);
        sl.block = dim3(2u * (uint32_t)B.threads);  // two lanes per particle
    }
#define SPHX_INDUCTION(U, R, PER)                                                                              \
    if (kp.uniform == (U ? 1 : 0) && res == R && g.periodic == PER) {                                          \
        if (staged)                                                                                           \
            hipLaunchKernelGGL((k_induction_staged<U, R, PER>), sl.grid, sl.block, sl.dyn, s, P, D, g, L, kp, B, \
                               cell_start, sl.cap);                                                           \
        else                                                                                                  \
            hipLaunchKernelGGL((k_induction<U, R, PER>), grid2, dim3(kBlock), occupancy_lds(4), s, P, D, n, g, L, \
                               kp);                                                                           \
        return;                                                                                               \
    }
    SPHX_INDUCTION(true, false, 1)
    SPHX_INDUCTION(true, false, 0)
    SPHX_INDUCTION(true, true, 1)
    SPHX_INDUCTION(true, true, 0)
    SPHX_INDUCTION(false, false, 1)
    SPHX_INDUCTION(false, false, 0)
    SPHX_INDUCTION(false, true, 1)
    SPHX_INDUCTION(false, true, 0)
#undef SPHX_INDUCTION
}

// ---------------------------------------------- divergence cleaning (a12)
// divergence_cleaning_system (mhd.rs:172-259): Jacobi over B and psi, then
// psi *= exp(-sigma c_h dt / h) for every particle.
// divergence_cleaning_system (mhd.rs:172-259) in gather form, two lanes per
// particle as the induction pass (halves added in order, branch-free body);
// range as for induction_particle.
template <bool UNI, int PER, class Own, class Range>
__device__ __forceinline__ void clean_particle(const Particles& P, const Derived& D, const Grid& g,
                                               const NbrList& L, const KernelParams& kp, uint32_t i,
                                               Own&& own_record, Range&& range) {
    const uint32_t h = lane_half();
    double2 own[6];  // the particle's own records, as in the induction pass
    own_record(own);
    const double xi = own[0].x, yi = own[0].y, zi = own[1].x, hi = UNI ? kp.h0 : P.h[i];
    const double bxi = own[4].x, byi = own[4].y, bzi = own[5].x, psii = own[5].y;
    double div = 0, gpx = 0, gpy = 0, gpz = 0;
    const uint32_t cnt = L.ntot[i], half = ((cnt + 7u) >> 3) << 2;
    // per pair, with g = (dW/r) d (d = x_j - x_i) and w = (m_j / rho_j) dW/r:
    //   div B_i += w (B_j - B_i) . d,   grad psi_i += w (psi_j - psi_i) d,
    // one refinement step for 1/r and 1/rho_j (as the induction pass)
    range(h ? half : 0u, h ? 0xffffffffu : half, [&](uint32_t e, const double2 (&q)[6]) {
        const uint32_t j = e & kIdxMask;
        const double2 xy = q[0], zr = q[1], bxy = q[4], bzp = q[5];
        double dx, dy, dz;
        pair_vec<PER>(g, xi, yi, zi, xy.x, xy.y, zr.x, dx, dy, dz);
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double dwr = mhd_dwr<UNI>(kp, r2, hi, UNI ? hi : P.h[j]);
        const double w = ((UNI ? kp.m0 : P.m[j]) * rcp_nr1(zr.y)) * dwr;
        div = fma(w, fma(bzp.x - bzi, dz, fma(bxy.y - byi, dy, (bxy.x - bxi) * dx)), div);
        const double dp = w * (bzp.y - psii);
        gpx = fma(dp, dx, gpx);
        gpy = fma(dp, dy, gpy);
        gpz = fma(dp, dz, gpz);
    });
    div += other_half(div);
    gpx += other_half(gpx);
    gpy += other_half(gpy);
    gpz += other_half(gpz);
    if (h) return;
    const double dt = kp.dt;
    const double bx = bxi - gpx * dt, by = byi - gpy * dt, bz = bzi - gpz * dt;
    D.tmp3[0][i] = bx;
    D.tmp3[1][i] = by;
    D.tmp3[2][i] = bz;
    const double decay = exp(-kp.sigma_c * kp.ch * dt / hi);
    const double psi = (psii - (kp.ch * kp.ch) * div * dt) * decay;
    D.tmp1[i] = psi;
    double2* rb = reinterpret_cast<double2*>(D.recb_next + i);
    rb[0] = make_double2(bx, by);
    rb[1] = make_double2(bz, psi);
    if (kp.srec) D.recs[i] = stress_record(bx, by, bz, own[1].y);
}

template <bool UNI, int PER>
__global__ void __launch_bounds__(kBlock) k_div_clean(Particles P, Derived D, uint32_t n, Grid g, NbrList L,
                                                       KernelParams kp) {
    const uint32_t i = pair_slot(particle_index());
    if (i >= n) return;
    const GlobalFetch<double2, 4, 2> F{reinterpret_cast<const double2*>(D.r64),
                                       reinterpret_cast<const double2*>(D.recb)};
    clean_particle<UNI, PER>(P, D, g, L, kp, i, [&](double2 (&q)[6]) { F.own(i, q); },
                             [&](uint32_t kb, uint32_t ke, auto&& body) {
                                 walk_ahead<4>(L, i, kb, ke, F, body);
                             });
}

// Brick-staged, as k_induction_staged (bit-identical to k_div_clean).
template <bool UNI, int PER>
__global__ void __launch_bounds__(2 * kStagedBlock) k_div_clean_staged(Particles P, Derived D, Grid g, NbrList L,
                                                                       KernelParams kp, Bricks B,
                                                                       const uint32_t* __restrict__ start,
                                                                       uint32_t cap) {
    __shared__ BrickHdr H;
    extern __shared__ uint4 sh_dyn[];
    using Fetch = BrickFetch<double2, 4, 2, true, 2>;
    brick_pass<double2, 4, 2, true, 2, 2>(
        B, g, start, cap, L, reinterpret_cast<const double2*>(D.r64), H, sh_dyn,
        [&](uint32_t i, const Fetch& f) {
            clean_particle<UNI, PER>(P, D, g, L, kp, i, [&](double2 (&q)[6]) { f.own(i, q); },
                                     [&](uint32_t kb, uint32_t ke, auto&& body) {
                Fetch ff = f;
                ff.kbeg = kb;
                ff.kend = ke;
                walk(L, i, ff, body);
            });
        },
        reinterpret_cast<const double2*>(D.recb));
}

// FP32-mixed cleaning of particle i: records of the f32 record (x, rho | v,
// P/rho^2; the velocity quad is not read) and the B record.  Per pair, with
// w = (m_j / rho_j) (dW/r): div B += w (B_j - B_i) . d, grad psi += w (psi_j -
// psi_i) d, f32 per group of four entries, f64 sums.
template <bool UNI, class Fetch>
__device__ __forceinline__ void clean_mixed_particle(const Particles& P, const Derived& D, const Mixed& X,
                                                     const Grid& g, const NbrList& L, const KernelParams& kp,
                                                     uint32_t i, const Fetch& fetch) {
    const float4* __restrict__ R = reinterpret_cast<const float4*>(D.r32);
    const float4 own0 = R[2 * i];
    const float xi = own0.x, yi = own0.y, zi = own0.z;
    const double bxd = P.bx[i], byd = P.by[i], bzd = P.bz[i], psid = P.psi[i];
    const float bxi = (float)bxd, byi = (float)byd, bzi = (float)bzd, psii = (float)psid;
    const float hi = X.p4[i].w;
    const float inv_h0 = (float)kp.inv_h0, h0 = (float)kp.h0, m0 = (float)kp.m0;
    const float sig0 = (float)((kp.inv_h0 * kp.inv_h0) * (kp.inv_h0 * kp.inv_h0) * (0.75 * kInvPi));
    const float wx = (float)g.w[0], wy = (float)g.w[1], wz = (float)g.w[2];
    double div = 0, gp[3] = {0, 0, 0};
    float gdiv = 0, ggp[3] = {0, 0, 0};
    walk(L, i, fetch, [&](uint32_t e, const float4 (&q)[Fetch::NR]) {
        const uint32_t j = e & kIdxMask;
        const MixedPairG p = mixed_pair_g<UNI>(X, fetch, hi, inv_h0, h0, sig0, e, q[0], xi, yi, zi, wx, wy, wz);
        const float w = ((UNI ? m0 : X.m[j]) * __builtin_amdgcn_rcpf(q[0].w)) * p.dwr;
        const float4 bj = mixed_b_record(q[2], q[3]);
        gdiv = fmaf(w, fmaf(bj.z - bzi, p.dz, fmaf(bj.y - byi, p.dy, (bj.x - bxi) * p.dx)), gdiv);
        const float dp = w * (bj.w - psii);
        ggp[0] = fmaf(dp, p.dx, ggp[0]);
        ggp[1] = fmaf(dp, p.dy, ggp[1]);
        ggp[2] = fmaf(dp, p.dz, ggp[2]);
    }, [&]() {
        div += (double)gdiv;
        gdiv = 0.0f;
        for (int a = 0; a < 3; ++a) {
            gp[a] += (double)ggp[a];
            ggp[a] = 0.0f;
        }
    });
    const double dt = kp.dt;
    const double bx = bxd - gp[0] * dt, by = byd - gp[1] * dt, bz = bzd - gp[2] * dt;
    D.tmp3[0][i] = bx;
    D.tmp3[1][i] = by;
    D.tmp3[2][i] = bz;
    const double decay = exp(-kp.sigma_c * kp.ch * dt / P.h[i]);
    const double psi = (psid - (kp.ch * kp.ch) * div * dt) * decay;
    D.tmp1[i] = psi;
    double2* rb = reinterpret_cast<double2*>(D.recb_next + i);
    rb[0] = make_double2(bx, by);
    rb[1] = make_double2(bz, psi);
    if (kp.srec) D.recs[i] = stress_record(bx, by, bz, D.rho[i]);
}

template <bool UNI>
__global__ void __launch_bounds__(kBlock) k_div_clean_mixed(Particles P, Derived D, Mixed X, uint32_t n, Grid g,
                                                             NbrList L, KernelParams kp) {
    SPHX_PASS_PROLOGUE();
    clean_mixed_particle<UNI>(P, D, X, g, L, kp, i,
                              GlobalFetch<float4, 2, 2>{reinterpret_cast<const float4*>(D.r32),
                                                        reinterpret_cast<const float4*>(D.recb)});
}

// Brick-staged, as k_induction_mixed_staged.
template <bool UNI>
__global__ void __launch_bounds__(kStagedBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_div_clean_mixed_staged(Particles P, Derived D, Mixed X, Grid g, NbrList L, KernelParams kp, Bricks B,
                         const uint32_t* __restrict__ start, uint32_t cap) {
    __shared__ BrickHdr H;
    extern __shared__ uint4 sh_dyn[];
    brick_pass<float4, 2, 4, true, 1, 2>(
        B, g, start, cap, L, reinterpret_cast<const float4*>(D.r32), H, sh_dyn,
        [&](uint32_t i, const BrickFetch<float4, 2, 4, true, 2>& f) {
            clean_mixed_particle<UNI>(P, D, X, g, L, kp, i, f);
        },
        reinterpret_cast<const float4*>(D.recb));
}

void launch_div_clean(const Particles& P, const Derived& D, uint32_t n, const Grid& g, const NbrList& L,
                      const KernelParams& kp, bool mixed, const Mixed& X, const Bricks& B,
                      const uint32_t* cell_start, hipStream_t s) {
    if (!n) return;
    const bool staged = B.on && B.nactive;
    if (mixed) {
        const StagedLaunch sl = staged ? staged_launch(B, 16, 64) : StagedLaunch{};
        if (staged && kp.uniform)
            hipLaunchKernelGGL(k_div_clean_mixed_staged<true>, sl.grid, sl.block, sl.dyn, s, P, D, X, g, L, kp, B,
                               cell_start, sl.cap);
        else if (staged)
            hipLaunchKernelGGL(k_div_clean_mixed_staged<false>, sl.grid, sl.block, sl.dyn, s, P, D, X, g, L, kp, B,
                               cell_start, sl.cap);
        else if (kp.uniform)
            hipLaunchKernelGGL(k_div_clean_mixed<true>, dim3(nblocks(n)), dim3(kBlock), occupancy_lds(3), s, P, D, X,
                               n, g, L, kp);
        else
            hipLaunchKernelGGL(k_div_clean_mixed<false>, dim3(nblocks(n)), dim3(kBlock), occupancy_lds(3), s, P, D, X,
                               n, g, L, kp);
        return;
    }
    const dim3 grid2(nblocks(2 * n));  // two lanes per particle
    StagedLaunch sl{};
    if (staged) {
        sl = staged_launch(B, 8, 96);
        sl.block = dim3(2u * (uint32_t)B.threads);  // two lanes per particle
    }
#define SPHX_CLEAN(U, PER)                                                                                    \
    if (kp.uniform == (U ? 1 : 0) && g.periodic == PER) {                                                      \
        if (staged)                                                                                           \
            hipLaunchKernelGGL((k_div_clean_staged<U, PER>), sl.grid, sl.block, sl.dyn, s, P, D, g, L, kp, B,    \
                               cell_start, sl.cap);                                                           \
        else                                                                                                  \
            hipLaunchKernelGGL((k_div_clean<U, PER>), grid2, dim3(kBlock), occupancy_lds(4), s, P, D, n, g, L,   \
                               kp);                                                                           \
        return;                                                                                               \
    }
    SPHX_CLEAN(true, 1)
    SPHX_CLEAN(true, 0)
    SPHX_CLEAN(false, 1)
    SPHX_CLEAN(false, 0)
#undef SPHX_CLEAN
}

// ---------------------------------------------------------- pair export
__global__ void __launch_bounds__(kBlock) k_export_pairs(Particles P, uint32_t n, Grid g, NbrList L,
                                                          uint32_t* __restrict__ out,
                                                          unsigned long long* counter, uint64_t out_cap) {
    SPHX_PASS_PROLOGUE();
    const uint32_t idi = P.id[i];
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i], hi = P.h[i];
    for_neighbours(L, i, [&](uint32_t e) {
        const uint32_t j = e & kIdxMask;
        const uint32_t idj = P.id[j];
        double dx, dy, dz;
        pair_vec(g, xi, yi, zi, P.x[j], P.y[j], P.z[j], dx, dy, dz);
        if (idi < idj && exact_neighbour(dx, dy, dz, hi, P.h[j])) {
            const unsigned long long pos = atomicAdd(counter, 1ull);
            if (pos < out_cap) {
                out[2 * pos] = idi;
                out[2 * pos + 1] = idj;
            }
        }
    });
}
void launch_export_pairs(const Particles& P, uint32_t n, const Grid& g, const NbrList& L, uint32_t* out,
                         unsigned long long* counter, uint64_t out_cap, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_export_pairs, dim3(nblocks(n)), dim3(kBlock), 0, s, P, n, g, L, out, counter,
                           out_cap);
}

// ------------------------------------- broadphase candidate export (non-SPH consumers)
// broadphase_query_system (broadphase.rs:62-96): entity a queries the tree
// within 4 R_a (squared Euclidean, summed x, y, z as kiddo's SquaredEuclidean
// does) and records (a, b) when a < b.  Here one thread per sorted particle
// walks the 27 cells around its own (each distinct cell once when a periodic
// axis has fewer than three) and records (id_i, id_j) for id_i < id_j and
// d^2 < (4 R_i)^2, R in sorted order.  The host checks 4 R_max + skin <= the
// cell width, so the stencil holds every such pair of the binned particles.
__global__ void __launch_bounds__(kBlock) k_export_candidates(Particles P, uint32_t n, Grid g,
                                                               const uint32_t* __restrict__ start,
                                                               const double* __restrict__ R, double half_skin,
                                                               uint32_t* __restrict__ out,
                                                               unsigned long long* counter, uint64_t out_cap) {
    SPHX_PASS_PROLOGUE();
    const uint32_t idi = P.id[i];
    const double xs[3] = {P.x[i], P.y[i], P.z[i]};
    const double r4 = 4.0 * R[i], r2max = r4 * r4;
    // Cells hold the particles binned at the last list build, each within
    // skin/2 of its position there: every j with |x_i - x_j| < 4 R_i is binned
    // within 4 R_i + skin/2 of x_i.  Per axis, the cells overlapping that
    // interval: clamped into the grid for a closed box (the binning clamps the
    // same way), wrapped for a periodic one (each cell once).
    const double reach = r4 + half_skin;
    int lo[3], cnt[3];
    bool all[3];
    for (int a = 0; a < 3; ++a) {
        const int d = g.dims[a];
        double ua = floor((xs[a] - reach - g.lo[a]) * g.inv_w[a]), ub = floor((xs[a] + reach - g.lo[a]) * g.inv_w[a]);
        if (g.periodic) {
            all[a] = ub - ua + 1.0 >= (double)d;
            const int ia = (int)ua;
            lo[a] = all[a] ? 0 : ((ia % d) + d) % d;
            cnt[a] = all[a] ? d : (int)(ub - ua) + 1;
        } else {
            all[a] = false;
            ua = fmin(fmax(ua, 0.0), (double)(d - 1));
            ub = fmin(fmax(ub, 0.0), (double)(d - 1));
            lo[a] = (int)ua;
            cnt[a] = (int)ub - (int)ua + 1;
        }
    }
    const uint32_t nx = (uint32_t)g.dims[0], ny = (uint32_t)g.dims[1];
    for (int iz = 0; iz < cnt[2]; ++iz) {
        int cz = lo[2] + iz;
        if (cz >= g.dims[2]) cz -= g.dims[2];
        for (int iy = 0; iy < cnt[1]; ++iy) {
            int cy = lo[1] + iy;
            if (cy >= g.dims[1]) cy -= g.dims[1];
            for (int ix = 0; ix < cnt[0]; ++ix) {
                int cx = lo[0] + ix;
                if (cx >= g.dims[0]) cx -= g.dims[0];
                const uint32_t cell = ((uint32_t)cz * ny + (uint32_t)cy) * nx + (uint32_t)cx;
                for (uint32_t j = start[cell], e = start[cell + 1]; j < e; ++j) {
                    const uint32_t idj = P.id[j];
                    if (!(idi < idj)) continue;
                    double dx, dy, dz;
                    pair_vec(g, xs[0], xs[1], xs[2], P.x[j], P.y[j], P.z[j], dx, dy, dz);
                    const double d2 = (dx * dx + dy * dy) + dz * dz;
                    if (d2 < r2max) {
                        const unsigned long long pos = atomicAdd(counter, 1ull);
                        if (pos < out_cap) {
                            out[2 * pos] = idi;
                            out[2 * pos + 1] = idj;
                        }
                    }
                }
            }
        }
    }
}
void launch_export_candidates(const Particles& P, uint32_t n, const Grid& g, const uint32_t* cell_start,
                              const double* R, double half_skin, uint32_t* out, unsigned long long* counter,
                              uint64_t out_cap, hipStream_t s) {
    if (n)
        hipLaunchKernelGGL(k_export_candidates, dim3(nblocks(n)), dim3(kBlock), 0, s, P, n, g, cell_start, R,
                           half_skin, out, counter, out_cap);
}

}  // namespace sphx
