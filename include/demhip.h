/*
 * demhip.h -- C ABI of the MI355X (gfx950) DEM step engine.
 *
 * The reference (lucifer1004/taichi_dem_julia_port) has no FFI: its operator API for the hot
 * path is the pair of Julia functions
 *     initialize!(g::GlobalData, threads; domain_min, cell_size, hash_table_size, dt, tolerance)
 *         src/executor.jl:155-160
 *     simulate!(g::GlobalData, threads; domain_min, cell_size, hash_table_size, dt, tolerance,
 *               gravity, global_damping)                            src/executor.jl:162-172
 * over the device-resident GlobalData (src/executor.jl:1-29) that solve() builds
 * (src/executor.jl:174-339), plus the read-backs that save_single performs
 * (src/utils.jl:30-78).  The entry points below replace exactly that seam; a DEMPort host binds
 * them with ccall.
 *
 * Conventions: C calling convention; host pointers only; every struct is POD with the layout of
 * the corresponding Julia isbits type where one exists; particle and contact indices are
 * 1-based at this boundary (the reference's convention); status 0 = OK, < 0 = error, with a
 * message from dem_last_error().  A scene handle is used by one host thread at a time.
 */
#ifndef DEMHIP_H
#define DEMHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEM_ABI_VERSION 2

enum dem_status {
    DEM_OK = 0,
    DEM_ERR_INVALID = -1,   /* bad argument or configuration */
    DEM_ERR_HIP = -2,       /* HIP runtime error (message names the call) */
    DEM_ERR_CAPACITY = -3,  /* a particle needs more than 256 contact slots */
    DEM_ERR_DUPLICATE = -4, /* duplicate contact pair in supplied state */
    DEM_ERR_NONFINITE = -5, /* non-finite particle state */
    DEM_ERR_STATE = -6      /* call not valid in the scene's current state */
};

enum dem_precision {
    DEM_PRECISION_FP64 = 0, /* every stage in FP64: the drop-in path */
    DEM_PRECISION_MIXED = 1 /* FP32 candidate geometry + compensated FP32 Hertz sums, FP64 state */
};

/* Mirrors DEMConfig (src/config.jl:3-43) field by field, without init_particles (the host reads
 * the particle file and passes its rows to dem_create). */
typedef struct dem_config {
    double domain_min[3];                       /* config.jl:4 */
    double domain_max[3];                       /* config.jl:5 */
    double particle_contact_radius_multiplier;  /* config.jl:7 */
    double neighboring_search_safety_factor;    /* config.jl:8 */
    double particle_elastic_modulus;            /* config.jl:9 */
    double particle_poisson_ratio;              /* config.jl:10 */
    double wall_normal[3];                      /* config.jl:11 */
    double wall_distance;                       /* config.jl:12 */
    double wall_density;                        /* config.jl:13 */
    double wall_elastic_modulus;                /* config.jl:14 */
    double wall_poisson_ratio;                  /* config.jl:15 */
    double bond_radius_ratio;                   /* config.jl:17 */
    double bond_elastic_modulus;                /* config.jl:18 */
    double bond_poisson_ratio;                  /* config.jl:19 */
    double bond_compressive_strength;           /* config.jl:20 */
    double bond_tensile_strength;               /* config.jl:21 */
    double bond_shear_strength;                 /* config.jl:22 */
    double pp_coefficient_friction;             /* config.jl:24 */
    double pp_coefficient_restitution;          /* config.jl:25 */
    double pp_coefficient_rolling_resistance;   /* config.jl:26 (unused, as in the reference) */
    double pw_coefficient_friction;             /* config.jl:28 */
    double pw_coefficient_restitution;          /* config.jl:29 */
    double pw_coefficient_rolling_resistance;   /* config.jl:30 (unused, as in the reference) */
    int32_t max_coordinate_number;              /* config.jl:32 */
    int32_t collision_pair_init_capacity_factor;/* config.jl:33 */
    double tolerance;                           /* config.jl:35 */
    int32_t max_particle_count;                 /* config.jl:36 */
    int32_t pad_;
    double gravity[3];                          /* config.jl:38 */
    double global_damping;                      /* config.jl:39 */
    double dt;                                  /* config.jl:40 */
    double target_time;                         /* config.jl:41 */
    double saving_interval_time;                /* config.jl:42 */
} dem_config_t;

/* Wall{Int32,Float64} (types.jl:60-64): unit normal n, offset d, material id (2 = wall). */
typedef struct dem_wall { double n[3]; double d; int32_t mid; int32_t pad_; } dem_wall_t;

/* One row of a .p4p particle file, in file order (executor.jl:219-220). */
typedef struct dem_grain_in {
    int32_t id, gid;
    double V, m;
    double k[3], v[3];
} dem_grain_in_t;

/* Grain{Int32,Float64} (types.jl:31-46), 272 bytes, byte-identical to the Julia isbits layout. */
typedef struct dem_grain {
    int32_t id, gid, mid, pad_;
    double V, m, r, r0;
    double k[3], v[3], a[3];
    double q[4];  /* w, x, y, z */
    double w[3], dw[3];
    double I[9];  /* 3x3, column-major */
} dem_grain_t;

/* Contact{Int32,Float64} (types.jl:68-78), 136 bytes.  i < j, 1-based particle indices. */
typedef struct dem_contact {
    int32_t i, j, midi, midj;
    double k[3], F[3], ti[3], tj[3], s[3];
} dem_contact_t;

/* Quantities solve() derives before the first step (executor.jl:244-280). */
typedef struct dem_derived {
    double r_max;              /* executor.jl:247-250 */
    double cell_size;          /* executor.jl:265 */
    double material_density;   /* executor.jl:244 (from the last particle) */
    int64_t hash_table_size;   /* executor.jl:266-267 (Int32-wrapped product, clamped) */
    int32_t grid_product_i32;  /* the wrapped product before clamping */
    int32_t contact_slots;     /* per-particle contact capacity of this engine */
    int64_t total_steps;       /* executor.jl:262 */
    int64_t save_per_steps;    /* executor.jl:263 */
} dem_derived_t;

typedef struct dem_options {
    int32_t precision;     /* enum dem_precision */
    int32_t contact_slots; /* contacts per particle (both roles); 0 = max(32, 2 max_coordinate_number) */
    int32_t device;        /* HIP device ordinal */
    int32_t flags;         /* DEM_FLAG_* */
} dem_options_t;

/* Broad phase owned by the lower particle index, as in the reference (kernels.jl:31), instead of
 * the lower sorted slot; the engine also falls back to it when the configuration does not
 * guarantee that touching pairs lie in both particles' windows. */
#define DEM_FLAG_INDEX_OWNED_BROAD_PHASE 1
/* Sort and search on the reference's hashed cell grid (utils.jl:1-16) every step instead of the
 * engine's dense grid of cs/2 cells (used when the domain is compact; same contact sets). */
#define DEM_FLAG_REFERENCE_GRID 2
/* Launch every step's kernels individually instead of replaying a captured two-step HIP graph. */
#define DEM_FLAG_NO_GRAPH 4
/* Sort and run the broad phase on every step.  By default (on the dense grid) a step that sorts
 * also keeps each particle's near list, the partners within r_i + r_j + 2 delta
 * (delta = 0.05 max r), and later steps test only those lists until some particle has moved
 * delta from where the lists were built; the contact sets are the same either way. */
#define DEM_FLAG_NO_NEAR_LISTS 8

/* Per-step counters (values of the last step unless marked cumulative). */
typedef struct dem_counters {
    int64_t step;              /* cumulative simulate! calls */
    int64_t n;                 /* particles */
    int64_t hash_table_size;   /* H */
    int64_t candidate_pairs;   /* P: cp_list length the reference would build (executor.jl:63);
                                  -1 unless the diagnostic passes ran (dem_set_diagnostics) */
    int64_t contacts;          /* C: active contacts after the step (save_single's count) */
    int64_t bonded;            /* C_b: active bonded contacts */
    int64_t hertz;             /* C_h = C - C_b */
    int64_t wall_contacts;     /* W: particle-wall contacts evaluated */
    int64_t created;           /* contacts appended by update_contacts! / init_bonds! */
    int64_t deactivated;       /* Hertz contacts deactivated by update_contacts! */
    int64_t bonds_broken;      /* bonds broken in resolve_collision! */
    int64_t bonds_broken_total;/* cumulative */
    int64_t fatal_gap;         /* unbonded active contacts evaluated with gap > 0 */
    int64_t overflow;          /* non-zero: contact capacity exceeded (results invalid) */
    int64_t examined;          /* candidate pairs examined by the engine's own broad phase;
                                  -1 unless the diagnostic passes ran */
    int64_t grid_cells;        /* cells of the engine's dense sort grid (0: reference grid used) */
    int64_t rebuilds;          /* cumulative steps that sorted and ran the broad phase (every step
                                  without near lists) */
    int64_t replays;           /* cumulative lean steps (merge and force kernel only) that found
                                  they had to rebuild, and were replayed as full steps */
} dem_counters_t;

typedef struct dem_scene dem_scene_t;

/* ---- lifecycle ------------------------------------------------------------------------- */
int32_t dem_abi_version(void);
/* Builds GlobalData: executor.jl:217-339 (derivations, sizing, material/surface tables).
 * walls/nwall replace the single hard-coded wall of executor.jl:313 (pass the config's wall for
 * the reference behaviour).  opt may be NULL (FP64, max_coordinate_number slots, device 0). */
int32_t dem_create(const dem_config_t* cfg, const dem_wall_t* walls, int32_t nwall,
                   const dem_grain_in_t* grains, int64_t n, const dem_options_t* opt,
                   dem_scene_t** out);
void dem_destroy(dem_scene_t* s);
const char* dem_last_error(const dem_scene_t* s); /* s may be NULL: last create error */

/* ---- the hot path -------------------------------------------------------------------- */
int32_t dem_initialize(dem_scene_t* s);              /* initialize!  executor.jl:155-160 */
/* nsteps x simulate! (executor.jl:162-172).  A step in which a particle needs more contact slots
 * than the scene has changes nothing; the rows are widened (their contents kept, up to 256 slots)
 * and the remaining steps run, with results equal to those of a scene created with the wider
 * rows.  Beyond 256 slots it returns DEM_ERR_CAPACITY with the state of the last completed step. */
int32_t dem_step(dem_scene_t* s, int64_t nsteps);

/* ---- read-back / state transfer --------------------------------------------------------- */
int32_t dem_get_derived(const dem_scene_t* s, dem_derived_t* out);
int32_t dem_get_counters(dem_scene_t* s, dem_counters_t* out);
/* Diagnostic passes on the last step of every dem_step / dem_initialize call (off by default):
 * the reference's cp_list length P (total_cps, executor.jl:62-63), counted on the reference's own
 * hashed grid, and the candidates the engine's broad phase examined.  They cost about two
 * steps' worth of extra kernels per call, so the production path leaves them off. */
int32_t dem_set_diagnostics(dem_scene_t* s, int32_t on);
/* Array(grains) (utils.jl:40): all n grains in file order. */
int32_t dem_read_grains(dem_scene_t* s, dem_grain_t* out, int64_t n);
/* Overwrite the dynamic grain state (k, v, a, q, w, dw) of all n grains. */
int32_t dem_write_grains(dem_scene_t* s, const dem_grain_t* in, int64_t n);
/* save_single's contact loop (utils.jl:48-69): active contacts sorted by (i, j), with the
 * bonded flag.  *count receives the total; at most cap records are written. */
int32_t dem_read_contacts(dem_scene_t* s, dem_contact_t* out, uint8_t* bonded, int64_t cap,
                          int64_t* count);
/* Replace the contact set (every given contact active; bonded per flag).  Used to restore a
 * saved state.  Duplicate (i, j) pairs are rejected with DEM_ERR_DUPLICATE.  A bond's tangential
 * spring s is stored as zero, the value every step of the reference leaves it at. */
int32_t dem_write_contacts(dem_scene_t* s, const dem_contact_t* in, const uint8_t* bonded,
                           int64_t n);
/* Wall shear history wall_contacts[j, i].s (kernels.jl:392,412-422) as [n][nwall][3]. */
int32_t dem_read_wall_shear(dem_scene_t* s, double* out, int64_t n);
int32_t dem_write_wall_shear(dem_scene_t* s, const double* in, int64_t n);
/* forces / moments of the last step as [n][3] (GlobalData.forces / .moments). */
int32_t dem_read_forces(dem_scene_t* s, double* F, double* T, int64_t n);

/* ---- intermediate arrays, computed by the step's own kernels from the current state --------
 * kid/kcenter/hid/neighbor_cells of contact! (executor.jl:39-43), 1-based hid and buckets. */
int32_t dem_debug_cells(dem_scene_t* s, int32_t* kid, uint32_t* hid, double* kcenter,
                        uint32_t* nbr);
/* hash_table (counts), hash_table_current after get_particle_id! (bucket start) and pid
 * (executor.jl:46-53); counts/start have hash_table_size entries, pid has n (1-based). */
int32_t dem_debug_hash_table(dem_scene_t* s, uint32_t* counts, uint32_t* start, uint32_t* pid);
/* The reference's cp_list (executor.jl:56-78) as [count][2] 1-based pairs. */
int32_t dem_debug_cp_list(dem_scene_t* s, int32_t* pairs, int64_t cap, int64_t* count);
/* Exclusive prefix sum of n uint32 values with the engine's decoupled-lookback scan (the
 * replacement of CUDA.jl accumulate!, executor.jl:50,62); *total receives the sum. */
int32_t dem_debug_scan(const uint32_t* in, uint32_t* out, int64_t n, uint64_t* total);
/* No counterpart in the reference.  The scan of dem_debug_scan for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): in, out and total (may be NULL) are DEVICE pointers, the scan is enqueued on
 * stream (a hipStream_t; NULL = the null stream) and the call returns without waiting.  Its
 * look-back scratch is a fixed array of about 1 MiB per GPU that the library holds until the
 * process ends; n is limited to 2^30 (DEM_ERR_INVALID beyond).  Every call on one GPU shares
 * that array, so calls whose scans overlap in time on one GPU (different streams, or threads)
 * race; calls ordered on one stream do not.  The first call on a device resolves the array's
 * address and should not be made under graph capture. */
int32_t dem_scan_noalloc(const uint32_t* in, uint32_t* out, int64_t n, uint64_t* total,
                         void* stream);

/* ---- per-kernel timing (HIP events on the scene's stream) ------------------------------ */
#define DEM_NUM_KERNELS 8
typedef struct dem_kernel_times {
    double ms[DEM_NUM_KERNELS];      /* summed durations */
    int64_t launches[DEM_NUM_KERNELS];
} dem_kernel_times_t;
/* Kernel order: 0 hash, 1 scan, 2 scatter, 3 broad phase, 4 merge, 5 force+integrate,
 * 6 init bonds, 7 other.  Enabling timing adds an event pair around each launch. */
int32_t dem_set_timing(dem_scene_t* s, int32_t on);
int32_t dem_get_kernel_times(dem_scene_t* s, dem_kernel_times_t* out);
const char* dem_kernel_name(int32_t k);
/* Plain device-to-device copy kernel over nbytes (rounded down to 32-byte units), run repeats
 * times after one warm-up; reports the mean duration.  Used to calibrate HBM byte counters. */
int32_t dem_debug_copy(int64_t nbytes, int32_t repeats, double* ms_per_copy);
/* Device memory of destroyed scenes kept for reuse by later ones, up to max_bytes in total
 * (0, the default: memory goes back to the runtime at once; lowering the limit frees the
 * excess).  No counterpart in the reference: for callers that create and destroy batches of
 * similar sizes in a loop (the serving re-pack), where releasing and mapping tens of gigabytes
 * costs up to hundreds of milliseconds per batch. */
int32_t dem_set_memory_cache(int64_t max_bytes);

/* ---- segmented batches (no counterpart in the reference) -----------------------------------
 * nscenes independent scenes share one set of device arrays and are advanced together by one
 * launch per kernel stage (and one graph replay per two steps).  Scene k has counts[k] grains
 * (concatenated in grains), config cfgs[k] and nwalls[k] walls (concatenated in walls).  Every
 * other entry point applies to the whole batch; particle indices are global (scene k occupies
 * offsets[k] .. offsets[k+1]-1, 0-based).  Each scene's result equals its result alone. */
int32_t dem_create_batch(int32_t nscenes, const dem_config_t* cfgs, const int32_t* nwalls,
                         const dem_wall_t* walls, const int64_t* counts,
                         const dem_grain_in_t* grains, const dem_options_t* opt, dem_scene_t** out);
int32_t dem_batch_offsets(const dem_scene_t* s, int64_t* offsets);     /* nscenes + 1 values */
/* Scenes with mask[k] == 0 get no bonds in dem_initialize (solve's no_bond). */
int32_t dem_batch_set_bond_init(dem_scene_t* s, const uint8_t* mask);
int32_t dem_batch_bonds_broken(dem_scene_t* s, int64_t* out);          /* cumulative, per scene */
/* Move the complete dynamic state of scene src_scene of src (a batch or a single scene, index 0)
 * into scene dst_scene of dst: grains, contact rows (partners re-indexed), wall history and the
 * bond-break count.  Both must hold the same particles and walls; used to re-pack live scenes
 * into a new batch when the serving loop admits or retires scenes. */
int32_t dem_batch_adopt(dem_scene_t* dst, int32_t dst_scene, dem_scene_t* src, int32_t src_scene);
/* A new batch of scenes keep[0..nkeep) of src (a batch or a single scene on the dense grid), in
 * that order, followed by nnew new scenes given as for dem_create_batch (cfgs, nwalls, walls,
 * counts, grains describe the new scenes only; they may be NULL when nnew is 0).  The kept scenes
 * arrive as dem_batch_adopt would move them, without their input rows: their derived fields,
 * order and parameters are copied from src and their state moves device to device.  Kept scenes
 * take no part in dem_initialize (mask 0); new ones keep the default (1).  src is left as it was.
 * Contact slots default to the larger of dem_create_batch's default and src's.  The serving
 * loop's re-pack at every change of the live set. */
int32_t dem_batch_repack(dem_scene_t* src, int32_t nkeep, const int32_t* keep, int32_t nnew,
                         const dem_config_t* cfgs, const int32_t* nwalls, const dem_wall_t* walls,
                         const int64_t* counts, const dem_grain_in_t* grains,
                         const dem_options_t* opt, dem_scene_t** out);

/* ---- synthetic inputs (host only; no counterpart in the reference) -------------------------
 * Loose polydisperse packing by random sequential addition: log-normal radii (median, sigma)
 * clipped to [rmin, rmax], density rho, zero velocity, placed inside [lo, hi] without overlap;
 * rows in a seeded random order with ids 1..n.  Used for the 1M-grain pour workload. */
int32_t dem_generate_pour(int64_t n, const double lo[3], const double hi[3], double median_r,
                          double sigma, double rmin, double rmax, double density, uint64_t seed,
                          dem_grain_in_t* out);

#ifdef __cplusplus
}
#endif
#endif
