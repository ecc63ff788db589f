/*
 * dem_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the reference DEM time step (lucifer1004/taichi_dem_julia_port,
 * src/kernels.jl, src/utils.jl:1-28, src/executor.jl:31-172 and the sizing/derivation part of
 * solve, src/executor.jl:217-315).  It exists to check the HIP engine and to time a CPU
 * baseline; nothing in the shipped library links, loads or calls it.
 *
 * Struct layouts deliberately equal those of include/demhip.h (and therefore the Julia isbits
 * layouts of Grain/Contact/Wall, src/types.jl:31-90) so that one ctypes definition feeds both.
 */
#ifndef DEM_ORACLE_H
#define DEM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_config {
    double domain_min[3];
    double domain_max[3];
    double particle_contact_radius_multiplier;
    double neighboring_search_safety_factor;
    double particle_elastic_modulus;
    double particle_poisson_ratio;
    double wall_normal[3];
    double wall_distance;
    double wall_density;
    double wall_elastic_modulus;
    double wall_poisson_ratio;
    double bond_radius_ratio;
    double bond_elastic_modulus;
    double bond_poisson_ratio;
    double bond_compressive_strength;
    double bond_tensile_strength;
    double bond_shear_strength;
    double pp_coefficient_friction;
    double pp_coefficient_restitution;
    double pp_coefficient_rolling_resistance;
    double pw_coefficient_friction;
    double pw_coefficient_restitution;
    double pw_coefficient_rolling_resistance;
    int32_t max_coordinate_number;
    int32_t collision_pair_init_capacity_factor;
    double tolerance;
    int32_t max_particle_count;
    int32_t pad_;
    double gravity[3];
    double global_damping;
    double dt;
    double target_time;
    double saving_interval_time;
} orc_config;

typedef struct orc_wall { double n[3]; double d; int32_t mid; int32_t pad_; } orc_wall;

typedef struct orc_grain_in { int32_t id, gid; double V, m, k[3], v[3]; } orc_grain_in;

typedef struct orc_grain {
    int32_t id, gid, mid, pad_;
    double V, m, r, r0;
    double k[3], v[3], a[3];
    double q[4];
    double w[3], dw[3];
    double I[9]; /* column-major */
} orc_grain;

typedef struct orc_contact {
    int32_t i, j, midi, midj;
    double k[3], F[3], ti[3], tj[3], s[3];
} orc_contact;

typedef struct orc_derived {
    double r_max, cell_size, material_density;
    int64_t hash_table_size;
    int32_t grid_product_i32; /* the Int32-wrapped product before clamping */
    int32_t pad_;
    int64_t total_steps, save_per_steps;
    int64_t cp_list_capacity, contact_capacity;
} orc_derived;

typedef struct orc_counters {
    int64_t step;
    int64_t total_cps;          /* cp_list length of the last contact pass */
    int64_t cp_list_capacity;   /* after doubling (executor.jl:64-71) */
    int64_t contact_ptr0;       /* contact_ptr[1] after the last pass */
    int64_t contact_ptr1;       /* contact_ptr[2] before late_clear_state! */
    int64_t active_contacts;
    int64_t bonded_contacts;
    int64_t wall_contacts;
    int64_t bonds_broken_total;
    int64_t deactivated_total;
    int64_t fatal_gap;          /* unbonded active contacts evaluated with gap > 0 */
    int64_t overflow;           /* contact array capacity exceeded */
} orc_counters;

typedef struct orc_scene orc_scene;

/* --- device-helper restatements (utils.jl:1-16) exposed for pinning --- */
uint64_t orc_morton3(int64_t x, int64_t y, int64_t z); /* interleave, x in bit 0 */
int64_t orc_cartesian3morton(const int32_t ijk[3]);    /* Morton.jl convention: M(ijk-1)+1 */
void orc_cell(const double pos[3], const double dmin[3], double cs, int32_t out[3]);
uint32_t orc_hashcode(const int32_t ijk[3], int64_t H);
void orc_center(const int32_t ijk[3], const double dmin[3], double cs, double out[3]);
void orc_neighbors(const double o[3], const int32_t ijk[3], const double c[3], int64_t H,
                   uint32_t out[8]);
int32_t orc_derive(const orc_config* cfg, const orc_grain_in* g, int64_t n, orc_derived* out);

/* --- scene --- */
orc_scene* orc_create(const orc_config* cfg, const orc_wall* walls, int32_t nwall,
                      const orc_grain_in* g, int64_t n);
void orc_destroy(orc_scene* s);
void orc_get_derived(const orc_scene* s, orc_derived* out);
int32_t orc_initialize(orc_scene* s);              /* initialize! */
int32_t orc_step(orc_scene* s, int64_t nsteps);    /* nsteps x simulate! */
int32_t orc_num_threads(void);  /* threads a step uses: 1, or OpenMP's count in the -fopenmp build */
void orc_contact_pass(orc_scene* s);               /* stages 1-3 of contact! only */
void orc_get_counters(const orc_scene* s, orc_counters* out);

/* state access (grains in file order, 0-based arrays, reference 1-based values inside) */
void orc_get_grains(const orc_scene* s, orc_grain* out);
void orc_set_grains(orc_scene* s, const orc_grain* in);
int64_t orc_num_active_contacts(const orc_scene* s);
/* active contacts sorted by (i,j); bonded flags 0/1 */
int64_t orc_get_contacts(const orc_scene* s, orc_contact* out, uint8_t* bonded, int64_t cap);
/* replace the contact set (all given contacts active; bonded per flag) */
int32_t orc_set_contacts(orc_scene* s, const orc_contact* in, const uint8_t* bonded, int64_t n);
void orc_get_wall_shear(const orc_scene* s, double* out);      /* [n][nwall][3] */
void orc_set_wall_shear(orc_scene* s, const double* in);
void orc_get_forces(const orc_scene* s, double* F, double* T); /* [n][3] of the last step */
/* Sums of |contribution| per component of the last step's forces and moments ([n][3]): the scale
 * against which a per-particle sum taken in another order is compared.  Tracked once turned on. */
void orc_set_track_scale(orc_scene* s, int32_t on);
void orc_get_force_scale(const orc_scene* s, double* F, double* T);

/* intermediate arrays of the last contact pass */
void orc_get_cells(const orc_scene* s, int32_t* kid, uint32_t* hid, double* kcenter,
                   uint32_t* nbr);
void orc_get_hash_table(const orc_scene* s, uint32_t* counts, uint32_t* current,
                        uint32_t* pid);
int64_t orc_get_cp_list(const orc_scene* s, int32_t* pairs, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
