/*
 * sphx.h — C ABI of the MI355X (gfx950) SPH particle step.
 *
 * This is the drop-in boundary for the SPH step of phys-rs. Each pass
 * replaces one of the engine's ECS systems (cited per entry point), and the
 * lifecycle/data calls replace the wgpu object API of src/gpu/ (GpuContext,
 * SphPipeline::create_buffers_and_bind_group, dispatch_all, read_densities /
 * read_forces).  Everything is plain C: POD structs, host pointers and sizes;
 * the library owns device memory, the caller keeps its host arrays.
 *
 * Conventions
 *  - Every entry point returns an int status (SPHX_OK == 0) and never aborts;
 *    sphx_last_error() returns a message for the last failing call on a ctx.
 *  - Particle arrays are SoA, f64 (the reference's DVec3 / f64 components),
 *    in the caller's order ("upload order").  Results are returned in the
 *    same order.  `id` is the entity's rank under Entity's own ordering
 *    (bevy_ecs compares Entity::to_bits()), not the query iteration index:
 *    it decides pair orientation exactly like entity_a < entity_b does for
 *    CandidatePairs (src/physics/broadphase.rs:87) and is required for MHD
 *    induction, whose per-pair sign depends on it (src/physics/mhd.rs:121-131).
 *  - A ctx is not thread safe; all passes are enqueued on the ctx stream and
 *    sphx_download / sphx_export_pairs / sphx_synchronize block on it.
 *  - All fluid particles must share one Sector (src/core/coordinates.rs:51-65
 *    then reduces to local_b - local_a exactly); positions are LocalPosition.
 */
#ifndef SPHX_H
#define SPHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    SPHX_OK = 0,
    SPHX_E_NODEVICE = 1,   /* no HIP device / device index out of range      */
    SPHX_E_OOM = 2,        /* device allocation failed                        */
    SPHX_E_CAPACITY = 3,   /* n > capacity, pair buffer too small, grid too big */
    SPHX_E_INVALID = 4,    /* bad argument or parameter                        */
    SPHX_E_MULTISECTOR = 5,/* reserved: particles in different sectors        */
    SPHX_E_NONFINITE = 6,  /* NaN/Inf in inputs or state                       */
    SPHX_E_HIP = 7,        /* a HIP runtime call failed                        */
    SPHX_E_STATE = 8       /* call order violated (e.g. forces before density) */
};

/* ---- parameter enums --------------------------------------------------- */
enum SphxEos {
    SPHX_EOS_TAIT = 0,       /* P = B((rho/rho0)^gamma - 1), B = rho0 c^2/gamma, clamp >= 0 (sph.rs:157-167) */
    SPHX_EOS_POLYTROPIC = 1, /* adiabatic ideal gas at fixed entropy: P = K rho^gamma                       */
    SPHX_EOS_ISOTHERMAL = 2, /* P = rho * (k_B T / m) = rho * cs2 (thermodynamics.rs:232-238)              */
    SPHX_EOS_ADIABATIC = 3   /* ideal gas P = (gamma - 1) rho u with the specific internal energy u
                                (InternalEnergy / Mass, components/thermal.rs:9) evolved by
                                du_i/dt = sum_j m_j (P_i/rho_i^2 + Pi_ij/2) v_ij . grad_i W_ij,
                                the energy equation of the pressure and viscosity forces above */
};
enum SphxViscConvention {
    SPHX_VISC_REFERENCE = 0, /* active iff (v_i - v_j).(x_j - x_i) < 0, as sph.rs:292-295 evaluates it */
    SPHX_VISC_MONAGHAN = 1   /* active iff (v_i - v_j).(x_i - x_j) < 0 (approaching pairs)              */
};
enum SphxBoundaryFlags {
    SPHX_BND_PENALTY_CUBE = 1, /* sph_boundary_system on [-half_extent, half_extent]^3 (sph.rs:443-484) */
    SPHX_BND_REFLECT = 2,      /* specular reflection at box_lo / box_hi after the position update   */
    SPHX_BND_PERIODIC = 4      /* periodic box [box_lo, box_hi) with minimum-image pair vectors       */
};
enum SphxMhdFlags {
    SPHX_MHD_INDUCTION = 1,      /* mhd_induction_system (mhd.rs:68-152)              */
    SPHX_MHD_CLEANING = 2,       /* divergence_cleaning_system (mhd.rs:172-259)        */
    SPHX_MHD_MAXWELL_STRESS = 4, /* SPMHD Lorentz force from M = BB - |B|^2 I / 2 (mu0 = 1) */
    SPHX_MHD_LORENTZ_CHARGE = 8  /* per-particle q (E + v x B) (lorentz.rs:21-39)     */
};
enum SphxPrecision {
    SPHX_PREC_FP64 = 0,      /* plain f64 drop-in                                                  */
    SPHX_PREC_FP32_MIXED = 1 /* f32 pair math on cell-relative coordinates, f64 sums/integration   */
};

/* ---- parameters: the reference Resources, field for field -------------- */
typedef struct SphxParams {
    double dt;                 /* SimulationTime.dt (core/time.rs:18)                           */
    double sector_size;        /* UniverseConfig.sector_size (core/config.rs); informational    */
    /* TaitEquationConfig (sph.rs:125-132) */
    double rest_density;
    double speed_of_sound;     /* also the c of the viscosity term (sph.rs:280)                 */
    double gamma;
    /* other EOS kinds */
    int32_t eos;               /* enum SphxEos                                                   */
    int32_t visc_convention;   /* enum SphxViscConvention                                        */
    double poly_k;             /* SPHX_EOS_POLYTROPIC: K                                         */
    double iso_cs2;            /* SPHX_EOS_ISOTHERMAL: k_B T / m                                 */
    /* SphViscosityConfig (sph.rs:234-239) */
    double visc_alpha;
    double visc_beta;
    /* XsphConfig (sph.rs:328-331) */
    double xsph_epsilon;
    /* SphBoundaryConfig (sph.rs:415-422) */
    double half_extent;
    double activation_distance;
    double stiffness;
    int32_t boundary_flags;    /* enum SphxBoundaryFlags, OR-ed                                  */
    int32_t mhd_flags;         /* enum SphxMhdFlags, OR-ed                                       */
    double box_lo[3];          /* reflective / periodic box                                      */
    double box_hi[3];
    double gravity[3];         /* uniform body acceleration, added as F += m g                  */
    double e_field[3];         /* uniform E for SPHX_MHD_LORENTZ_CHARGE                          */
    /* MhdConfig (mhd.rs:26-39) */
    double resistivity;
    double cleaning_speed;
    double cleaning_damping;
    /* neighbour grid */
    double cell_size;          /* <= 0: 2 * max(h) * (1 + 1e-9)                                  */
    double grid_lo[3];         /* grid box; grid_hi <= grid_lo on any axis: taken from the      */
    double grid_hi[3];         /* uploaded positions (padded), or from box_lo/hi when periodic   */
    int32_t precision;         /* enum SphxPrecision                                             */
    int32_t max_neighbours;    /* per-particle neighbour-list capacity, rounded up to a multiple
                                  of 4; <= 0: 192.  Device memory: 6 bytes per particle and unit
                                  of capacity (4-byte entries, 2-byte staging slots)           */
    double neighbour_skin;     /* list radius 2*h_avg + skin, rebuilt only once some particle has
                                  moved skin/2 since the last build; < 0: 0.04 * max(h) (default);
                                  0: rebuilt every tick.  Every pass still uses exactly the pairs
                                  with r < 2*h_avg (pairs beyond contribute exactly zero). */
} SphxParams;

/* Fill *p with the reference defaults: TaitEquationConfig::default (sph.rs:134-142),
 * SphViscosityConfig::default (241-248), XsphConfig::default (333-339),
 * SphBoundaryConfig::default (424-432), MhdConfig::default (mhd.rs:41-49),
 * dt = 1/60 (time.rs:37), sector_size = 1e12, no boundary flags, FP64. */
void sphx_default_params(SphxParams* p);

/* ---- host-side particle state (the ECS components, gathered to SoA) ----- */
typedef struct SphxState {
    uint32_t n;
    double* x; double* y; double* z;      /* LocalPosition                 */
    double* vx; double* vy; double* vz;   /* Velocity                      */
    double* m;                            /* Mass                          */
    double* h;                            /* SmoothingRadius               */
    double* rho;                          /* SmoothedDensity  (output)     */
    double* p;                            /* Pressure         (output)     */
    double* ax; double* ay; double* az;   /* Acceleration = Force / Mass   */
    double* bx; double* by; double* bz;   /* MagneticField  (optional)     */
    double* psi;                          /* DivCleaningPsi (optional)     */
    double* q;                            /* Charge         (optional)     */
    uint32_t* id;                         /* rank of Entity::to_bits(); NULL -> 0..n-1 */
    double* u;                            /* InternalEnergy / Mass (J/kg); SPHX_EOS_ADIABATIC */
} SphxState;

/* ---- inputs of the reference's WGSL path (SphMode::Gpu callers) ---------- */
/* The uniform block SphUniformData (src/gpu/buffers.rs:25-38), laid out as the
 * WGSL struct SphParams (src/gpu/sph.wgsl:1-12): 40 bytes, f32, one mass and
 * one smoothing radius for every particle. */
typedef struct SphxUniformData {
    uint32_t num_particles;
    float mass;
    float h;
    float dt;
    float rest_density;
    float speed_of_sound;
    float gamma;
    float alpha_visc;
    float beta_visc;
    uint32_t _padding;
} SphxUniformData;

/* Replaces the uniform upload of create_buffers_and_bind_group
 * (src/gpu/pipeline.rs:168-181): dt, the Tait EOS (rest_density,
 * speed_of_sound, gamma) and the viscosity (alpha_visc, beta_visc) are copied
 * into *p, widened to f64; every other field of *p is kept.  SPHX_E_INVALID
 * for a NULL pointer or a non-positive dt, rest_density, speed_of_sound or
 * gamma (p is then unchanged). */
int sphx_params_from_uniform(const SphxUniformData* u, SphxParams* p);

/* Replaces the vec4 particle buffers of the same call (SphParticleVec4,
 * buffers.rs:5-22: x, y, z, w, w unused): fills the caller's SphxState arrays
 * x, y, z, vx, vy, vz from pos4 and vel4 (4 floats a particle, widened to f64),
 * m = u->mass and h = u->h for all u->num_particles particles, id = 0 .. n-1
 * (buffer order) when st->id is not NULL, and sets st->n.  SPHX_E_INVALID for
 * a NULL pointer or array or a non-positive mass or h. */
int sphx_state_from_vec4(const SphxUniformData* u, const float* pos4, const float* vel4, SphxState* st);

/* fields for sphx_download */
enum SphxFields {
    SPHX_F_POS = 1, SPHX_F_VEL = 2, SPHX_F_RHO = 4, SPHX_F_P = 8, SPHX_F_ACC = 16,
    SPHX_F_B = 32, SPHX_F_PSI = 64, SPHX_F_U = 128, SPHX_F_ALL = 255
};

typedef struct SphxCtx SphxCtx;

/* ---- lifecycle: replaces initialize_gpu_context / try_init_gpu (gpu/mod.rs:69-142)
 *      and SphPipeline::new (gpu/pipeline.rs:27) ------------------------- */
int sphx_device_count(int* count);
int sphx_create(int device, uint32_t capacity, const SphxParams* params, SphxCtx** out);
void sphx_destroy(SphxCtx* ctx);
/* New parameters for the next passes.  max_neighbours <= 0 keeps the list
 * capacity; any other value must round up (to a multiple of 4) to the
 * capacity chosen at sphx_create.  A failing call (SPHX_E_INVALID) leaves the
 * context unchanged.  Changing the box, grid, cell size or skin needs a new
 * sphx_upload. */
int sphx_set_params(SphxCtx* ctx, const SphxParams* params);
const char* sphx_last_error(const SphxCtx* ctx);

/* ---- data: replaces create_buffers_and_bind_group (gpu/pipeline.rs:168-206)
 *      and read_densities / read_forces (gpu/buffers.rs:163-213) --------- */
int sphx_upload(SphxCtx* ctx, const SphxState* host);
int sphx_download(SphxCtx* ctx, SphxState* host, uint32_t fields);
/* Residency: replaces the per-tick re-upload of buffers in the reference's
 * GPU path (create_buffers_and_bind_group, gpu/pipeline.rs:168-206, runs for
 * every dispatch) when only some components changed.  Writes the fields
 * SPHX_F_VEL, SPHX_F_B, SPHX_F_PSI and SPHX_F_U of host (upload order, host->n equal to
 * the context's n) into the resident state, keeping the sort, the grid and the
 * neighbour list; density and forces are recomputed at the next use.
 * Positions, masses and radii change through sphx_upload.  Blocks until the
 * copies are done.  SPHX_E_INVALID for other fields or a missing array,
 * SPHX_E_NONFINITE for a non-finite value, SPHX_E_STATE for B / psi / u without
 * an uploaded field. */
int sphx_update(SphxCtx* ctx, const SphxState* host, uint32_t fields);

/* ---- passes (each replaces one ECS system; async on the ctx stream) ---- */
/* build_broadphase_system + broadphase_query_system (broadphase.rs:33-96):
 * cell hash, counting sort, reorder, exact neighbour lists r < 2*h_avg.   */
int sphx_neighbours(SphxCtx* ctx);
/* sph_density_system (sph.rs:72-116) + sph_eos_system (sph.rs:149-169)
 * (the adiabatic EOS at each particle's current u)                         */
int sphx_density(SphxCtx* ctx);
/* reset_forces (forces.rs:15) + sph_pressure_force_system (sph.rs:177-228)
 * + sph_viscosity_system (sph.rs:260-322) + sph_boundary_system
 * (sph.rs:443-484) + lorentz_force_system (lorentz.rs:21-39)
 * + compute_acceleration_system (integrators.rs:110-118).
 * flags = SPHX_FORCES_VISC_ONLY reuses the pressure sum stored by the last
 * full call at the same positions (bit-identical to recomputing it); it is
 * SPHX_E_STATE with the adiabatic EOS, whose full pass also gives du/dt.   */
enum { SPHX_FORCES_FULL = 0, SPHX_FORCES_VISC_ONLY = 1 };
int sphx_forces(SphxCtx* ctx, uint32_t flags);
/* velocity_verlet_position_system (integrators.rs:63-83)                  */
int sphx_vv_position(SphxCtx* ctx);
/* velocity_verlet_velocity_system (integrators.rs:92-104); with the
 * adiabatic EOS also u += (du_n + du*) dt / 2, du_n saved by vv_position   */
int sphx_vv_velocity(SphxCtx* ctx);
/* sph_xsph_system (sph.rs:349-405)                                         */
int sphx_xsph(SphxCtx* ctx);
/* mhd_induction_system (mhd.rs:68-152)                                     */
int sphx_mhd_induction(SphxCtx* ctx);
/* divergence_cleaning_system (mhd.rs:172-259)                              */
int sphx_div_clean(SphxCtx* ctx);
/* One velocity-Verlet tick of the schedule in main.rs:103-157, repeated
 * nsteps times: forces(x_n, v_n) -> vv_position -> neighbours/density/
 * forces(x_n+1, v_n) -> vv_velocity -> xsph -> induction -> cleaning.     */
int sphx_step(SphxCtx* ctx, uint32_t nsteps);

/* ---- differentiable step (new; the reference has no adjoint) ----------- */
/* Runs nsteps ticks of sphx_step from the current state and returns
 *   L = sum_i |x_i(nsteps) - target_i|^2 / (3 n)
 * with its gradient with respect to the initial velocities (grad_v*), the
 * initial positions (grad_x*, optional: all three NULL to skip) and the
 * viscosity alpha, by a hand-written reverse pass over the same schedule
 * (main.rs:103-157; pressure sph.rs:177-228, viscosity sph.rs:260-322,
 * density sph.rs:72-116, EOS sph.rs:149-169, integrators.rs:63-104).
 * The forward keeps a checkpoint (x, v) every checkpoint_every ticks (0:
 * every nsteps) and recomputes each segment's tape during the reverse pass.
 * Arrays are in upload order.  On return the context holds the state after
 * nsteps ticks, as sphx_step would leave it.  FP64 form only; no XSPH, MHD,
 * adiabatic EOS, penalty walls or periodic boxes (SPHX_E_INVALID).  The neighbour sets and
 * the viscosity switch are those of the forward trajectory.              */
int sphx_adjoint(SphxCtx* ctx, uint32_t nsteps, uint32_t checkpoint_every,
                 const double* target_x, const double* target_y, const double* target_z,
                 double* loss, double* grad_x, double* grad_y, double* grad_z,
                 double* grad_vx, double* grad_vy, double* grad_vz, double* grad_alpha);
int sphx_synchronize(SphxCtx* ctx);

/* ---- checkpoint / resume ------------------------------------------------ */
/* The SPH state for QuickSave / QuickLoad (src/render/surface.rs:853-932),
 * whose SavedEntity (src/core/serialization.rs:7-16) keeps name, type, mass,
 * radius, position, sector, velocity and temperature but drops what the SPH
 * step needs to resume: h, PreviousAcceleration, B, psi, charge, internal
 * energy, the entity ranks and the parameters.  A checkpoint holds those, the
 * positions at the last neighbour-list build, the grid and the list's age, so
 * that a context loaded from it continues bit for bit as the saved one does.
 * The caller stores the bytes (as QuickSave writes quicksave.bin); the layout
 * is native little-endian and opens with the 8-byte magic "SPHXCKP1", then
 * u32 version, header bytes, SphxParams bytes and the particle count n.
 * sphx_checkpoint_size: bytes of a checkpoint of the current state.
 * sphx_save: writes it into buf (cap bytes); *written receives the size
 *   (also on SPHX_E_CAPACITY).  Builds the neighbour list if none is current.
 * sphx_load: replaces the parameters and state of ctx (capacity >= the saved
 *   n, the same max_neighbours: SPHX_E_CAPACITY / SPHX_E_INVALID) and keeps
 *   the caller's upload order of the saved context for download / update.
 *   SPHX_E_INVALID for a buffer that is not a checkpoint of this version. */
int sphx_checkpoint_size(SphxCtx* ctx, uint64_t* bytes);
int sphx_save(SphxCtx* ctx, void* buf, uint64_t cap, uint64_t* written);
int sphx_load(SphxCtx* ctx, const void* buf, uint64_t size);

/* ---- test and measurement -------------------------------------------- */
/* Undirected neighbour pairs (id_min, id_max) with r < 2*h_avg at the current
 * positions, sorted ascending; *count receives the total (even when > cap).
 * This is the set CandidatePairs holds after the SPH systems' r < 2h filter. */
int sphx_export_pairs(SphxCtx* ctx, uint32_t* pairs, uint64_t cap_pairs, uint64_t* count);

/* No counterpart in the reference: sphx_export_pairs for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same result as sphx_export_pairs; the
 * pairs are staged in a fixed device array of the library (2^21 pairs, 16 MiB
 * per GPU, held until the process ends) instead of a buffer it allocates, and
 * sorted in place in `pairs`.  Calls no hipMalloc / hipFree.  More than 2^21
 * pairs: SPHX_E_CAPACITY with *count set (use sphx_export_pairs).  Every call
 * on one GPU shares that array: calls that overlap on one GPU (two threads, two
 * contexts on the same device) race. */
int sphx_export_pairs_noalloc(SphxCtx* ctx, uint32_t* pairs, uint64_t cap_pairs, uint64_t* count);

/* Broadphase export for the systems beside SPH that read CandidatePairs
 * (heat_conduction_system, thermodynamics.rs:110-171; collisions): replaces
 * broadphase_query_system (src/physics/broadphase.rs:62-96).  radius[i] is
 * BoundingRadius of particle i (upload order).  Writes every (id_a, id_b) with
 * id_a < id_b and |x_a - x_b|^2 < (4 R_a)^2, R_a the radius of the smaller id
 * (the reference's query radius 4 R around the querying entity), as pairs
 * sorted ascending; *count receives the total (even when > cap_pairs).  Any
 * radius: each particle searches the grid cells within 4 R + skin/2 of it.
 * Exact equality d^2 = (4 R)^2 is not recorded.  SPHX_E_INVALID for a
 * negative or non-finite radius. */
int sphx_export_candidates(SphxCtx* ctx, const double* radius, uint32_t* pairs, uint64_t cap_pairs,
                           uint64_t* count);

typedef struct SphxTimes {       /* device milliseconds, accumulated since last reset */
    double sort_ms;              /* hash + counting sort + reorder                  */
    double neighbour_ms;         /* neighbour-list build (exact predicate)          */
    double density_ms;           /* density + EOS                                   */
    double force_ms;             /* full force pass                                 */
    double visc_only_ms;         /* viscosity-only force pass                       */
    double integrate_ms;         /* vv position + vv velocity                       */
    double mhd_ms;               /* xsph + induction + cleaning                     */
    uint64_t launches[7];        /* launches counted per slot above                 */
} SphxTimes;
/* Bracket kernels with HIP events on the ctx stream: enable = 0 off,
 * SPHX_TIMING_ALL every kernel, otherwise the slots whose SPHX_TIMING_* bits
 * are set (each event pair costs a few microseconds of the step, so a
 * throughput measurement times only the kernel it needs). */
#define SPHX_TIMING_ALL 1
#define SPHX_TIMING_SORT (1 << 1)
#define SPHX_TIMING_NEIGHBOUR (1 << 2)
#define SPHX_TIMING_DENSITY (1 << 3)
#define SPHX_TIMING_FORCE (1 << 4)
#define SPHX_TIMING_VISC_ONLY (1 << 5)
#define SPHX_TIMING_INTEGRATE (1 << 6)
#define SPHX_TIMING_MHD (1 << 7)
int sphx_enable_timing(SphxCtx* ctx, int enable);
int sphx_kernel_times(SphxCtx* ctx, SphxTimes* out, int reset);

typedef struct SphxEnergy {
    double kinetic;              /* sum 1/2 m |v|^2                      */
    double potential;            /* sum -m g.x                           */
    double internal;             /* Tait / polytropic internal energy    */
    double magnetic;             /* sum (m / rho) |B|^2 / 2  (mu0 = 1)   */
    double momentum[3];          /* sum m v                              */
    uint64_t max_neighbours;     /* largest neighbour count seen         */
    uint64_t total_neighbours;   /* directed neighbour pairs             */
} SphxEnergy;
int sphx_energy(SphxCtx* ctx, SphxEnergy* out);

#ifdef __cplusplus
}
#endif
#endif /* SPHX_H */
