/*
 * dem_oracle.c -- TEST INFRASTRUCTURE ONLY (parity checker and CPU-baseline port).
 *
 * A serial CPU FP64 restatement of the reference DEM step of
 * lucifer1004/taichi_dem_julia_port.  Every function cites the reference file:line it follows.
 * Build with -ffp-contract=off: the reference's Julia/LLVM code issues unfused IEEE multiplies
 * and adds, and each expression below keeps the reference's evaluation order.
 *
 * Serial execution is one valid linearisation of the reference's atomics: every atomic
 * add/sub/or/and becomes an ordinary update in loop order.  The two O(n^2) BitTables
 * (contact_active / contact_bonded, executor.jl:279-280) are replaced by one keyed set of the
 * same 64-bit pair index (i-1)*n + j with identical get/set/clear semantics, so the oracle also
 * runs at N = 1e6 where the BitTables would need 128 GiB each.
 *
 * Parity pins (tests/test_oracle_*.py, tests/oracle_pins.py): Morton known answer (README.md:23)
 * and brute-force interleave; closed forms for cell/center/neighbors and the Int32-wrapped
 * hash-table size of cfg/1-3; counting-sort invariants; O(N^2) brute-force broad phase; t=0 bond
 * counts of the reference's own particle files; Hertz static force, Coulomb cap, restitution;
 * the EBPM stiffness matrix against a Timoshenko beam element (symmetry, rigid-body modes,
 * entries) and its break thresholds against textbook stresses; frame equivariance; force and
 * moment balance; spin-up lever arms and d omega = tau / I; free flight with and without global
 * damping; the quaternion recursion; momentum.
 * Parity unpinned (no reference run is possible here): bit-level rounding of StaticArrays
 * norm/inv, Quaternionic to_rotation_matrix/normalize, Base cbrt/log, and Float64^4.
 *
 * The same source also builds with -fopenmp (liboracle_omp.so, the multi-core CPU baseline):
 * per-particle and per-pair loops run in parallel, force sums use atomic adds (so their order,
 * and the last bits of the sums, vary from run to run, as they do under the reference's CUDA
 * atomics), and update_contacts decides in parallel but applies its set changes in pair order,
 * which keeps the serial outcome.  Without -fopenmp the pragmas are ignored and the code is the
 * serial restatement.
 *
 * Planted mistakes: built with -DORC_MUTATION=k (k >= 1), the oracle carries one deliberate
 * error of the kind a restatement can plausibly make (a swapped stiffness, a flipped sign, a
 * dropped factor, a transposed rotation; the list is MUT_* below).  tests/test_oracle_mutations.py
 * builds each such variant and checks that the physics pins reject it, which shows the pins can
 * see that mistake.  The default build (ORC_MUTATION = 0) has none of them.
 */
#include "dem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int32_t orc_num_threads(void) {
#ifdef _OPENMP
    return (int32_t)omp_get_max_threads();
#else
    return 1;
#endif
}

typedef struct { double x, y, z; } v3;

#ifndef ORC_MUTATION
#define ORC_MUTATION 0
#endif
enum {
    MUT_K5K6_SWAP = 1,       /* bond moment on i takes k6 theta_i + k5 theta_j */
    MUT_K3_SIGN = 2,         /* bond shear force with -k3 (theta_i + theta_j)_z */
    MUT_GAMMA_NO_SQRT56 = 3, /* Hertz normal damping without sqrt(5/6) */
    MUT_MOMENT_ORDER = 4,    /* Hertz contact moment F x r instead of r x F */
    MUT_TORSION_SHEAR = 5,   /* bond torsional shear stress r/I instead of r/(2I) */
    MUT_DAMPING_SIGN = 6,    /* global damping adds +d|F|sign(v) */
    MUT_BOND_SCATTER_R = 7,  /* bond forces scattered with R instead of R^-1 */
    MUT_CONTACT_POINT = 8,   /* Hertz contact point at r_i + delta from i */
    MUT_VC_SIGN = 9,         /* Hertz contact velocity v_i - v_j */
    MUT_SIGMA_C_FI = 10,     /* bond normal stress from F_i instead of F_j */
    MUT_QUAT_DT = 11,        /* quaternion increment scaled by dt */
    MUT_K4_NO_NU = 12,       /* bond torsion stiffness E I / L */
    MUT_UPDATE_NEW_V = 13    /* position advanced with the new velocity */
};
#define MUT(k) (ORC_MUTATION == (k))

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vneg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline v3 vscl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }   /* a * s */
static inline v3 sclv(double s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }   /* s * a */
static inline v3 vdiv(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
/* StaticArrays dot: ((a1 b1 + a2 b2) + a3 b3) */
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* StaticArrays norm: sqrt((x^2 + y^2) + z^2) */
static inline double vnorm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
/* StaticArrays normalize: inv(norm(a)) * a */
static inline v3 vnormalize(v3 a) { double s = 1.0 / vnorm(a); return sclv(s, a); }
static inline v3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }
static inline void st3(double* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }

typedef struct { double m[3][3]; } m33; /* row-major m[row][col] */

static inline v3 mv(const m33* A, v3 u) {
    return mk(A->m[0][0] * u.x + A->m[0][1] * u.y + A->m[0][2] * u.z,
              A->m[1][0] * u.x + A->m[1][1] * u.y + A->m[1][2] * u.z,
              A->m[2][0] * u.x + A->m[2][1] * u.y + A->m[2][2] * u.z);
}
static m33 mm(const m33* A, const m33* B) {
    m33 C;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C.m[r][c] = A->m[r][0] * B->m[0][c] + A->m[r][1] * B->m[1][c] + A->m[r][2] * B->m[2][c];
    return C;
}
/* 3x3 inverse by cross products of the columns (StaticArrays _inv for Size(3,3)). */
static m33 minv(const m33* A) {
    v3 x0 = mk(A->m[0][0], A->m[1][0], A->m[2][0]);
    v3 x1 = mk(A->m[0][1], A->m[1][1], A->m[2][1]);
    v3 x2 = mk(A->m[0][2], A->m[1][2], A->m[2][2]);
    v3 y0 = vcross(x1, x2);
    double d = vdot(x0, y0);
    x0 = vdiv(x0, d);
    y0 = vdiv(y0, d);
    v3 y1 = vcross(x2, x0);
    v3 y2 = vcross(x0, x1);
    m33 R;
    R.m[0][0] = y0.x; R.m[0][1] = y0.y; R.m[0][2] = y0.z;
    R.m[1][0] = y1.x; R.m[1][1] = y1.y; R.m[1][2] = y1.z;
    R.m[2][0] = y2.x; R.m[2][1] = y2.y; R.m[2][2] = y2.z;
    return R;
}

/* Contact frame: rotation taking a onto the local x axis (kernels.jl:145-165, 332-352). */
static m33 frame(v3 a, double tol) {
    v3 b = mk(1.0, 0.0, 0.0);
    v3 v = vcross(a, b);
    double s = vnorm(v);
    double c = vdot(a, b);
    m33 R;
    if (s < tol) {
        double sg = c > 0.0 ? 1.0 : -1.0;
        R.m[0][0] = sg;  R.m[0][1] = 0.0; R.m[0][2] = 0.0;
        R.m[1][0] = 0.0; R.m[1][1] = 1.0; R.m[1][2] = 0.0;
        R.m[2][0] = 0.0; R.m[2][1] = 0.0; R.m[2][2] = sg;
    } else {
        m33 vx;
        vx.m[0][0] = 0.0;   vx.m[0][1] = -v.z; vx.m[0][2] = v.y;
        vx.m[1][0] = v.z;   vx.m[1][1] = 0.0;  vx.m[1][2] = -v.x;
        vx.m[2][0] = -v.y;  vx.m[2][1] = v.x;  vx.m[2][2] = 0.0;
        m33 vx2 = mm(&vx, &vx);
        double omc = 1.0 - c;
        double s2 = s * s;
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) {
                double id = (r == q) ? 1.0 : 0.0;
                R.m[r][q] = (id + vx.m[r][q]) + (vx2.m[r][q] * omc) / s2;
            }
    }
    return R;
}

/* ------------------------------------------------------------------------------------------
 * utils.jl:1-16 -- cell, Morton hash, centre, 8-bucket window
 * ------------------------------------------------------------------------------------------ */
static uint64_t spread3(uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (3 * b);
    return r;
}
uint64_t orc_morton3(int64_t x, int64_t y, int64_t z) {
    const uint64_t m21 = (1ull << 21) - 1;
    return spread3((uint64_t)x & m21) | (spread3((uint64_t)y & m21) << 1) |
           (spread3((uint64_t)z & m21) << 2);
}
/* Morton.jl's 1-based convention (README.md:23): cartesian3morton(c) = morton(c-1) + 1. */
int64_t orc_cartesian3morton(const int32_t ijk[3]) {
    return (int64_t)orc_morton3((int64_t)ijk[0] - 1, (int64_t)ijk[1] - 1, (int64_t)ijk[2] - 1) + 1;
}
/* utils.jl:1 cell(pos) = ceil(Int32, (pos - dmin) / cs) */
void orc_cell(const double pos[3], const double dmin[3], double cs, int32_t out[3]) {
    for (int a = 0; a < 3; ++a) out[a] = (int32_t)ceil((pos[a] - dmin[a]) / cs);
}
/* utils.jl:2 hashcode = cartesian3morton(ijk) % H + 1 (Julia rem, 1-based result) */
uint32_t orc_hashcode(const int32_t ijk[3], int64_t H) {
    int64_t m = orc_cartesian3morton(ijk);
    return (uint32_t)(m % H + 1);
}
/* utils.jl:3 center = dmin + (ijk - 0.5) * cs */
void orc_center(const int32_t ijk[3], const double dmin[3], double cs, double out[3]) {
    for (int a = 0; a < 3; ++a) out[a] = dmin[a] + ((double)ijk[a] - 0.5) * cs;
}
/* utils.jl:4-16 window order [self, x, y, z, xy, xz, yz, xyz] */
void orc_neighbors(const double o[3], const int32_t ijk[3], const double c[3], int64_t H,
                   uint32_t out[8]) {
    int32_t d[3];
    for (int a = 0; a < 3; ++a) d[a] = o[a] > c[a] ? 1 : -1;
    static const int sel[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                  {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    for (int w = 0; w < 8; ++w) {
        int32_t c3[3];
        for (int a = 0; a < 3; ++a) c3[a] = ijk[a] + (sel[w][a] ? d[a] : 0);
        out[w] = orc_hashcode(c3, H);
    }
}

/* ------------------------------------------------------------------------------------------
 * solve(): derivations and sizing (executor.jl:217-280)
 * ------------------------------------------------------------------------------------------ */
static double grain_radius(double V) { return cbrt(V / (4.0 / 3.0 * M_PI)); } /* :223 */

int32_t orc_derive(const orc_config* cfg, const orc_grain_in* g, int64_t n, orc_derived* out) {
    if (n <= 0) return -1;
    double rmax_r = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double r = grain_radius(g[i].V);
        if (i == 0 || r > rmax_r) rmax_r = r;
    }
    /* executor.jl:247-250 */
    double rmax = rmax_r * cfg->particle_contact_radius_multiplier *
                  (1.0 + cfg->bond_tensile_strength / cfg->bond_elastic_modulus) *
                  cfg->neighboring_search_safety_factor;
    double cs = 4.0 * rmax; /* :265 */
    /* :266-267 -- Int32 product with wrap-around, then clamp to [2^20, 2^22] */
    uint32_t prod = 1u;
    for (int a = 0; a < 3; ++a) {
        int32_t c = (int32_t)ceil((cfg->domain_max[a] - cfg->domain_min[a]) / cs);
        prod = prod * (uint32_t)c;
    }
    int32_t p32 = (int32_t)prod;
    int64_t H = p32;
    if (H < (1 << 20)) H = 1 << 20;
    if (H > (1 << 22)) H = 1 << 22;
    out->r_max = rmax;
    out->cell_size = cs;
    out->material_density = g[n - 1].m / g[n - 1].V; /* :244 */
    out->hash_table_size = H;
    out->grid_product_i32 = p32;
    out->pad_ = 0;
    out->total_steps = (int64_t)trunc(cfg->target_time / cfg->dt);           /* :262 */
    out->save_per_steps = (int64_t)trunc(cfg->saving_interval_time / cfg->dt); /* :263 */
    out->cp_list_capacity = n * (int64_t)cfg->collision_pair_init_capacity_factor; /* :272 */
    out->contact_capacity = n * (int64_t)cfg->max_coordinate_number;          /* :277 */
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Keyed pair state replacing the BitTables (utils.jl:18-28).  flag bit0 = active, bit1 = bonded
 * ------------------------------------------------------------------------------------------ */
typedef struct { uint64_t* keys; uint8_t* flags; uint64_t cap; uint64_t used; } pairset;
#define PS_EMPTY UINT64_MAX

static uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33; return x;
}
static void ps_init(pairset* s, uint64_t cap) {
    uint64_t c = 1024;
    while (c < cap * 2) c <<= 1;
    s->cap = c; s->used = 0;
    s->keys = (uint64_t*)malloc(c * sizeof(uint64_t));
    s->flags = (uint8_t*)calloc(c, 1);
    for (uint64_t i = 0; i < c; ++i) s->keys[i] = PS_EMPTY;
}
static void ps_free(pairset* s) { free(s->keys); free(s->flags); s->keys = 0; s->flags = 0; }
static uint64_t ps_slot(const pairset* s, uint64_t key) {
    uint64_t m = s->cap - 1, h = mix64(key) & m;
    while (s->keys[h] != PS_EMPTY && s->keys[h] != key) h = (h + 1) & m;
    return h;
}
static void ps_grow(pairset* s);
static uint8_t* ps_ref(pairset* s, uint64_t key) {
    if ((s->used + 1) * 2 > s->cap) ps_grow(s);
    uint64_t h = ps_slot(s, key);
    if (s->keys[h] == PS_EMPTY) { s->keys[h] = key; s->flags[h] = 0; s->used++; }
    return &s->flags[h];
}
static void ps_grow(pairset* s) {
    pairset t; ps_init(&t, s->cap);
    for (uint64_t i = 0; i < s->cap; ++i)
        if (s->keys[i] != PS_EMPTY && s->flags[i]) {
            uint64_t h = ps_slot(&t, s->keys[i]);
            t.keys[h] = s->keys[i]; t.flags[h] = s->flags[i]; t.used++;
        }
    ps_free(s); *s = t;
}
static uint8_t ps_get(const pairset* s, uint64_t key) {
    uint64_t h = ps_slot(s, key);
    return s->keys[h] == PS_EMPTY ? 0 : s->flags[h];
}
#define F_ACTIVE 1u
#define F_BONDED 2u
static int get_bit(const pairset* s, uint64_t ij, unsigned f) { return (ps_get(s, ij) & f) != 0; }
static void set_bit(pairset* s, uint64_t ij, unsigned f) { *ps_ref(s, ij) |= (uint8_t)f; }
static void clear_bit(pairset* s, uint64_t ij, unsigned f) {
    uint64_t h = ps_slot(s, ij);
    if (s->keys[h] != PS_EMPTY) s->flags[h] &= (uint8_t)~f;
}

/* ------------------------------------------------------------------------------------------
 * Scene (GlobalData, executor.jl:1-29)
 * ------------------------------------------------------------------------------------------ */
typedef struct { double rho, E, nu; } material;                       /* types.jl:9-13 */
typedef struct { double mu, e, mur, rho, E, nu, sc, st, ss; } surface; /* types.jl:17-27 */

struct orc_scene {
    int64_t n;
    orc_config cfg;
    orc_derived der;
    orc_grain* g;
    material mats[2];
    surface surf[2][2];
    orc_wall* walls;
    int32_t nwall;
    double dmin[3];
    double cs;
    int64_t H;
    /* per-pass arrays */
    int32_t* kid;      /* [n][3] */
    uint32_t* hid;     /* [n] 1-based */
    double* kcenter;   /* [n][3] */
    uint32_t* nbr;     /* [n][8] 1-based */
    uint32_t* ht;      /* [H] counts */
    uint32_t* htc;     /* [H] cumulative -> exclusive start */
    uint32_t* pid;     /* [n] 1-based particle index */
    uint32_t* cpr;     /* [n] */
    uint32_t* cprc;    /* [n] */
    int32_t* cpl;      /* [cap][2] 1-based */
    int64_t cpl_cap, total_cps;
    orc_contact* contacts;
    orc_contact* ctemp;
    int64_t cmax;      /* logical capacity n * max_coordinate_number */
    int64_t calloc_n;  /* allocated entries */
    uint32_t cptr[2];
    pairset st;
    orc_contact* wall_contacts; /* [n][nwall] */
    double* forces;   /* [3n] */
    double* moments;  /* [3n] */
    /* sums of |contribution| per component of forces / moments (orc_set_track_scale) */
    int track_scale;
    double* fscale;   /* [3n] */
    double* mscale;   /* [3n] */
    orc_counters cnt;
};

static void ensure_contacts(orc_scene* s, int64_t need) {
    if (need <= s->calloc_n) return;
    int64_t c = s->calloc_n ? s->calloc_n : 1024;
    while (c < need) c <<= 1;
    s->contacts = (orc_contact*)realloc(s->contacts, c * sizeof(orc_contact));
    s->ctemp = (orc_contact*)realloc(s->ctemp, c * sizeof(orc_contact));
    s->calloc_n = c;
}

orc_scene* orc_create(const orc_config* cfg, const orc_wall* walls, int32_t nwall,
                      const orc_grain_in* gi, int64_t n) {
    if (n <= 0 || nwall < 0) return NULL;
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->n = n;
    s->cfg = *cfg;
    orc_derive(cfg, gi, n, &s->der);
    s->g = (orc_grain*)calloc(n, sizeof(orc_grain));
    for (int64_t i = 0; i < n; ++i) { /* executor.jl:219-243 */
        orc_grain* g = &s->g[i];
        double r = grain_radius(gi[i].V);
        double Ii = 2.0 / 5.0 * gi[i].m * (r * r);
        g->id = gi[i].id; g->gid = gi[i].gid; g->mid = 1;
        g->V = gi[i].V; g->m = gi[i].m; g->r = r;
        g->r0 = r * cfg->particle_contact_radius_multiplier;
        for (int a = 0; a < 3; ++a) { g->k[a] = gi[i].k[a]; g->v[a] = gi[i].v[a]; }
        g->q[0] = 1.0;
        memset(g->I, 0, sizeof(g->I));
        g->I[0] = Ii; g->I[4] = Ii; g->I[8] = Ii;
    }
    /* executor.jl:283-310 */
    s->mats[0].rho = s->der.material_density;
    s->mats[0].E = cfg->particle_elastic_modulus;
    s->mats[0].nu = cfg->particle_poisson_ratio;
    s->mats[1].rho = cfg->wall_density;
    s->mats[1].E = cfg->wall_elastic_modulus;
    s->mats[1].nu = cfg->wall_poisson_ratio;
    surface pp = {cfg->pp_coefficient_friction, cfg->pp_coefficient_restitution,
                  cfg->pp_coefficient_rolling_resistance, cfg->bond_radius_ratio,
                  cfg->bond_elastic_modulus, cfg->bond_poisson_ratio,
                  cfg->bond_compressive_strength, cfg->bond_tensile_strength,
                  cfg->bond_shear_strength};
    surface pw = {cfg->pw_coefficient_friction, cfg->pw_coefficient_restitution,
                  cfg->pw_coefficient_rolling_resistance, 0, 0, 0, 0, 0, 0};
    s->surf[0][0] = pp; s->surf[0][1] = pw; s->surf[1][0] = pw; s->surf[1][1] = pw;
    s->nwall = nwall;
    s->walls = (orc_wall*)calloc(nwall > 0 ? nwall : 1, sizeof(orc_wall));
    if (nwall) memcpy(s->walls, walls, nwall * sizeof(orc_wall));
    for (int a = 0; a < 3; ++a) s->dmin[a] = cfg->domain_min[a];
    s->cs = s->der.cell_size;
    s->H = s->der.hash_table_size;
    s->kid = (int32_t*)calloc(3 * n, sizeof(int32_t));
    s->hid = (uint32_t*)calloc(n, sizeof(uint32_t));
    s->kcenter = (double*)calloc(3 * n, sizeof(double));
    s->nbr = (uint32_t*)calloc(8 * n, sizeof(uint32_t));
    s->ht = (uint32_t*)calloc(s->H, sizeof(uint32_t));
    s->htc = (uint32_t*)calloc(s->H, sizeof(uint32_t));
    s->pid = (uint32_t*)calloc(n, sizeof(uint32_t));
    s->cpr = (uint32_t*)calloc(n, sizeof(uint32_t));
    s->cprc = (uint32_t*)calloc(n, sizeof(uint32_t));
    s->cpl_cap = s->der.cp_list_capacity > 0 ? s->der.cp_list_capacity : 1;
    s->cpl = (int32_t*)malloc(2 * s->cpl_cap * sizeof(int32_t));
    s->cmax = s->der.contact_capacity;
    ensure_contacts(s, n < 1024 ? 1024 : n);
    ps_init(&s->st, 4 * n);
    s->wall_contacts = (orc_contact*)calloc(n * (nwall > 0 ? nwall : 1), sizeof(orc_contact));
    s->forces = (double*)calloc(3 * n, sizeof(double));
    s->moments = (double*)calloc(3 * n, sizeof(double));
    return s;
}

void orc_destroy(orc_scene* s) {
    if (!s) return;
    free(s->g); free(s->walls); free(s->kid); free(s->hid); free(s->kcenter); free(s->nbr);
    free(s->ht); free(s->htc); free(s->pid); free(s->cpr); free(s->cprc); free(s->cpl);
    free(s->contacts); free(s->ctemp); ps_free(&s->st); free(s->wall_contacts);
    free(s->forces); free(s->moments); free(s->fscale); free(s->mscale); free(s);
}

void orc_get_derived(const orc_scene* s, orc_derived* out) { *out = s->der; }

static inline uint64_t pairkey(const orc_scene* s, int32_t i, int32_t j) {
    return (uint64_t)(i - 1) * (uint64_t)s->n + (uint64_t)j; /* kernels.jl:73,99 */
}

/* ------------------------------------------------------------------------------------------
 * contact! stages 1-3 (executor.jl:38-78)
 * ------------------------------------------------------------------------------------------ */
static void contact_pass(orc_scene* s) {
    const int64_t n = s->n;
    /* executor.jl:39-43 */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        int32_t* kid = &s->kid[3 * i];
        orc_cell(s->g[i].k, s->dmin, s->cs, kid);
        s->hid[i] = orc_hashcode(kid, s->H);
        orc_center(kid, s->dmin, s->cs, &s->kcenter[3 * i]);
        orc_neighbors(s->g[i].k, kid, &s->kcenter[3 * i], s->H, &s->nbr[8 * i]);
    }
    /* executor.jl:46-53, kernels.jl:7-22 */
    memset(s->ht, 0, s->H * sizeof(uint32_t));
    for (int64_t i = 0; i < n; ++i) s->ht[s->hid[i] - 1] += 1u;
    uint32_t acc = 0;
    for (int64_t b = 0; b < s->H; ++b) { acc += s->ht[b]; s->htc[b] = acc; } /* inclusive */
    for (int64_t i = 0; i < n; ++i) {
        uint32_t id = s->htc[s->hid[i] - 1];     /* atomic_sub returns the old value */
        s->htc[s->hid[i] - 1] = id - 1u;
        s->pid[id - 1] = (uint32_t)(i + 1);
    }
    /* executor.jl:56-62, kernels.jl:24-37 */
    memset(s->cpr, 0, n * sizeof(uint32_t));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
        for (int w = 0; w < 8; ++w) {
            uint32_t b = s->nbr[8 * i + w] - 1;
            for (uint32_t k = s->htc[b] + 1; k <= s->htc[b] + s->ht[b]; ++k) {
                uint32_t j = s->pid[k - 1];
                if ((uint32_t)(i + 1) < j) s->cpr[i] += 1u;
            }
        }
    }
    acc = 0;
    for (int64_t i = 0; i < n; ++i) { acc += s->cpr[i]; s->cprc[i] = acc; }
    s->total_cps = acc; /* :63 */
    int64_t cap = s->cpl_cap;
    while (cap < s->total_cps) cap <<= 1; /* :64-71 */
    if (cap > s->cpl_cap) {
        s->cpl = (int32_t*)realloc(s->cpl, 2 * cap * sizeof(int32_t));
        s->cpl_cap = cap;
    }
    /* kernels.jl:39-58 */
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
        for (int w = 0; w < 8; ++w) {
            uint32_t b = s->nbr[8 * i + w] - 1;
            for (uint32_t k = s->htc[b] + 1; k <= s->htc[b] + s->ht[b]; ++k) {
                uint32_t j = s->pid[k - 1];
                if ((uint32_t)(i + 1) < j) {
                    uint32_t cur = s->cprc[i];
                    s->cprc[i] = cur - 1u;
                    s->cpl[2 * (cur - 1)] = (int32_t)(i + 1);
                    s->cpl[2 * (cur - 1) + 1] = (int32_t)j;
                }
            }
        }
    }
    s->cnt.total_cps = s->total_cps;
    s->cnt.cp_list_capacity = s->cpl_cap;
}

void orc_contact_pass(orc_scene* s) { contact_pass(s); }

static int append_contact(orc_scene* s, int32_t i, int32_t j) {
    int64_t off = s->cptr[0];
    if (off >= s->cmax) { s->cnt.overflow += 1; return -1; }
    ensure_contacts(s, off + 1);
    s->cptr[0] += 1u;
    orc_contact* c = &s->contacts[off];
    memset(c, 0, sizeof(*c));
    c->i = i; c->j = j;
    c->midi = s->g[i - 1].mid; c->midj = s->g[j - 1].mid;
    return 0;
}

/* kernels.jl:60-85 */
static void init_bonds(orc_scene* s) {
    for (int64_t idx = 0; idx < s->total_cps; ++idx) {
        int32_t i = s->cpl[2 * idx], j = s->cpl[2 * idx + 1];
        const orc_grain* gi = &s->g[i - 1];
        const orc_grain* gj = &s->g[j - 1];
        if (vnorm(vsub(ld3(gi->k), ld3(gj->k))) < gi->r0 + gj->r0) {
            uint64_t ij = pairkey(s, i, j);
            /* serial linearisation: a duplicate (i,j) sees the earlier set_bit! */
            if (get_bit(&s->st, ij, F_ACTIVE)) continue;
            if (append_contact(s, i, j)) continue;
            set_bit(&s->st, ij, F_ACTIVE);
            set_bit(&s->st, ij, F_BONDED);
        }
    }
}

/* kernels.jl:87-120 for one pair, given whether the pair is valid (touching) and the current
 * flags; applies the set change (serial linearisation of the atomics). */
static void apply_update(orc_scene* s, int64_t idx, int valid) {
    int32_t i = s->cpl[2 * idx], j = s->cpl[2 * idx + 1];
    uint64_t ij = pairkey(s, i, j);
    if (get_bit(&s->st, ij, F_BONDED)) return;
    int exist = get_bit(&s->st, ij, F_ACTIVE);
    if (exist && !valid) {
        s->cptr[1] += 1u;
        clear_bit(&s->st, ij, F_ACTIVE);
        s->cnt.deactivated_total += 1;
    } else if (!exist && valid) {
        if (append_contact(s, i, j)) return;
        set_bit(&s->st, ij, F_ACTIVE);
    }
}

/* kernels.jl:87-120.  The pairs that may change the set (valid and not yet active, or active and
 * no longer valid) are found in parallel; apply_update then runs on them in pair order, which
 * re-reads the flags and so gives the serial loop's outcome (a duplicate pair sees the earlier
 * copy's change). */
static void update_contacts(orc_scene* s) {
    const int64_t P = s->total_cps;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    int64_t** lists = (int64_t**)calloc((size_t)nt, sizeof(int64_t*));
    int64_t* lens = (int64_t*)calloc((size_t)nt, sizeof(int64_t));
    int used = 1;
#pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#pragma omp single
        used = omp_get_num_threads();
#endif
        int64_t cap = 1024, len = 0;
        int64_t* l = (int64_t*)malloc((size_t)cap * sizeof(int64_t));
        /* static schedule: thread t takes the t-th contiguous block of pairs */
#pragma omp for schedule(static)
        for (int64_t idx = 0; idx < P; ++idx) {
            int32_t i = s->cpl[2 * idx], j = s->cpl[2 * idx + 1];
            uint8_t f = ps_get(&s->st, pairkey(s, i, j));
            if (f & F_BONDED) continue;
            const orc_grain* gi = &s->g[i - 1];
            const orc_grain* gj = &s->g[j - 1];
            int valid = vnorm(vsub(ld3(gi->k), ld3(gj->k))) < gi->r + gj->r;
            int exist = (f & F_ACTIVE) != 0;
            if (exist == valid) continue;
            if (len == cap) { cap *= 2; l = (int64_t*)realloc(l, (size_t)cap * sizeof(int64_t)); }
            l[len++] = valid ? idx : -1 - idx;
        }
        lists[t] = l;
        lens[t] = len;
    }
    for (int t = 0; t < used; ++t) {
        for (int64_t k = 0; k < lens[t]; ++k) {
            int64_t v = lists[t][k];
            if (v >= 0) apply_update(s, v, 1);
            else apply_update(s, -1 - v, 0);
        }
        free(lists[t]);
    }
    free(lists);
    free(lens);
}

static void atomic_add_vec3(double* arr, int64_t i0, v3 v) { /* kernels.jl:1-5 */
#pragma omp atomic
    arr[i0] += v.x;
#pragma omp atomic
    arr[i0 + 1] += v.y;
#pragma omp atomic
    arr[i0 + 2] += v.z;
}
/* the same sum where only one thread touches the particle */
static void add_vec3(double* arr, int64_t i0, v3 v) {
    arr[i0] += v.x; arr[i0 + 1] += v.y; arr[i0 + 2] += v.z;
}
/* |v| per component into a scale array (sc may be NULL: not tracked) */
static void atomic_add_abs3(double* sc, int64_t i0, v3 v) {
    if (!sc) return;
    atomic_add_vec3(sc, i0, mk(fabs(v.x), fabs(v.y), fabs(v.z)));
}
static void add_abs3(double* sc, int64_t i0, v3 v) {
    if (!sc) return;
    add_vec3(sc, i0, mk(fabs(v.x), fabs(v.y), fabs(v.z)));
}

/* Hertz-Mindlin tangential/normal law shared by particle and wall contacts
 * (kernels.jl:261-303 and 369-406); returns local force, updates s. */
static v3 hertz_mindlin(double gap, double dn, v3 vc, double nui, double Ei, double nuj,
                        double Ej, double Rs, double ms, double e, double mu, double dt,
                        v3* s_io, int64_t* fatal) {
    double Ys = 1.0 / ((1.0 - nui * nui) / Ei + (1.0 - nuj * nuj) / Ej);
    double Gs = 1.0 / (2.0 * (2.0 - nui) * (1.0 + nui) / Ei + 2.0 * (2.0 - nuj) * (1.0 + nuj) / Ej);
    double le = log(e);
    double beta = le / sqrt(le * le + M_PI * M_PI);
    double srd = sqrt(Rs * dn);
    double Sn = 2.0 * Ys * srd;
    double St = 8.0 * Gs * srd;
    double kn = 4.0 / 3.0 * Ys * srd;
    double gn = MUT(MUT_GAMMA_NO_SQRT56) ? -2.0 * beta * sqrt(Sn * ms)
                                         : -2.0 * beta * sqrt(5.0 / 6.0 * Sn * ms);
    if (gn < 0) *fatal += 1;
    double kt = 8.0 * Gs * srd;
    double gt = -2.0 * beta * sqrt(5.0 / 6.0 * St * ms);
    if (gt < 0) *fatal += 1;
    v3 ds = mk(vc.x * 0.0, vc.y * dt, vc.z * dt);
    v3 sv = vadd(*s_io, ds);
    double F1 = -kn * gap - gn * vc.x;
    v3 Ft = sclv(-kt, sv);
    double F2, F3;
    double nFt = vnorm(Ft);
    if (nFt >= mu * F1) { /* sliding */
        double ratio = mu * F1 / nFt;
        F2 = ratio * Ft.y;
        F3 = ratio * Ft.z;
        sv = mk(sv.x, F2 / kt, F3 / kt);
    } else {
        F2 = Ft.y - gt * vc.y;
        F3 = Ft.z - gt * vc.z;
    }
    *s_io = sv;
    return mk(F1, F2, F3);
}

/* kernels.jl:122-315 */
static void resolve_collision(orc_scene* s) {
    const double dt = s->cfg.dt, tol = s->cfg.tolerance;
    int64_t total = s->cptr[0];
    int64_t nbroken = 0, nfatal = 0;
    double* fsc = s->track_scale ? s->fscale : NULL;
    double* msc = s->track_scale ? s->mscale : NULL;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nbroken, nfatal)
    for (int64_t idx = 0; idx < total; ++idx) {
        orc_contact* c = &s->contacts[idx];
        int32_t i = c->i, j = c->j;
        uint64_t ij = pairkey(s, i, j);
        if (!get_bit(&s->st, ij, F_ACTIVE)) continue;
        const orc_grain* gi = &s->g[i - 1];
        const orc_grain* gj = &s->g[j - 1];
        v3 ki = ld3(gi->k), kj = ld3(gj->k);
        v3 a = vnormalize(vsub(kj, ki));
        m33 R = frame(a, tol);
        double Li = vnorm(vsub(kj, ki));
        if (get_bit(&s->st, ij, F_BONDED)) {
            /* EBPM (kernels.jl:170-239) */
            v3 k = vscl(vadd(ki, kj), 0.5);
            v3 di = mv(&R, vscl(ld3(gi->v), dt));   /* R * v * dt == R * (v * dt) */
            v3 dj = mv(&R, vscl(ld3(gj->v), dt));
            v3 thi = mv(&R, vscl(ld3(gi->w), dt));
            v3 thj = mv(&R, vscl(ld3(gj->w), dt));
            int32_t midi = gi->mid, midj = gj->mid;
            const surface* sf = &s->surf[midi - 1][midj - 1];
            double rmin = gi->r < gj->r ? gi->r : gj->r;
            double rj = sf->rho * rmin;
            double Lj = Li;
            double Ej = sf->E, nu = sf->nu;
            double rj2 = rj * rj;
            double Ij = rj2 * rj2 * M_PI / 4.0;
            double phi = 20.0 / 3.0 * rj2 / (Lj * Lj) * (1.0 + nu);
            double Aj = rj2 * M_PI;
            double k1 = Ej * Aj / Lj;
            double k2 = 12.0 * Ej * Ij / (Lj * Lj * Lj) / (1.0 + phi);
            double k3 = 6.0 * Ej * Ij / (Lj * Lj) / (1.0 + phi);
            double k4 = MUT(MUT_K4_NO_NU) ? Ej * Ij / Lj : Ej * Ij / Lj / (1.0 + nu);
            double k5 = Ej * Ij * (4.0 + phi) / Lj / (1.0 + phi);
            double k6 = Ej * Ij * (2.0 - phi) / Lj / (1.0 + phi);
            v3 dF = mk(k1 * (di.x - dj.x),
                       MUT(MUT_K3_SIGN) ? k2 * (di.y - dj.y) - k3 * (thi.z + thj.z)
                                        : k2 * (di.y - dj.y) + k3 * (thi.z + thj.z),
                       k2 * (di.z - dj.z) - k3 * (thi.y + thj.y));
            v3 dTi = MUT(MUT_K5K6_SWAP)
                         ? mk(k4 * (thi.x - thj.x), k3 * (dj.z - di.z) + k6 * thi.y + k5 * thj.y,
                              k3 * (di.y - dj.y) + k6 * thi.z + k5 * thj.z)
                         : mk(k4 * (thi.x - thj.x),
                              k3 * (dj.z - di.z) + k5 * thi.y + k6 * thj.y,
                              k3 * (di.y - dj.y) + k5 * thi.z + k6 * thj.z);
            v3 dTj = mk(k4 * (thj.x - thi.x),
                        k3 * (dj.z - di.z) + k6 * thi.y + k5 * thj.y,
                        k3 * (di.y - dj.y) + k6 * thi.z + k5 * thj.z);
            v3 Fi = vadd(ld3(c->F), dF);
            v3 Fj = vneg(Fi);
            v3 Ti = vadd(ld3(c->ti), dTi);
            v3 Tj = vadd(ld3(c->tj), dTj);
            double Fn = MUT(MUT_SIGMA_C_FI) ? Fi.x : Fj.x;
            double sci = Fn / Aj - rj / Ij * sqrt(Ti.y * Ti.y + Ti.z * Ti.z);
            double scj = Fn / Aj - rj / Ij * sqrt(Tj.y * Tj.y + Tj.z * Tj.z);
            double sc = -(sci < scj ? sci : scj);
            double st = sci > scj ? sci : scj;
            double ss = (MUT(MUT_TORSION_SHEAR) ? fabs(Ti.x) * rj / Ij : fabs(Ti.x) * rj / 2.0 / Ij) +
                        4.0 / 3.0 / Aj * sqrt(Fi.y * Fi.y + Fi.z * Fi.z);
            if (sc >= sf->sc || st >= sf->st || ss >= sf->ss) {
                /* flags of distinct pairs are distinct bytes; the set is not resized here */
                clear_bit(&s->st, ij, F_ACTIVE);
                clear_bit(&s->st, ij, F_BONDED);
                nbroken += 1;
            } else {
                m33 Ri = MUT(MUT_BOND_SCATTER_R) ? R : minv(&R);
                v3 fi = mv(&Ri, vneg(Fi)), fj = mv(&Ri, vneg(Fj));
                v3 ti = mv(&Ri, vneg(Ti)), tj = mv(&Ri, vneg(Tj));
                atomic_add_vec3(s->forces, 3 * (i - 1), fi);
                atomic_add_vec3(s->forces, 3 * (j - 1), fj);
                atomic_add_vec3(s->moments, 3 * (i - 1), ti);
                atomic_add_vec3(s->moments, 3 * (j - 1), tj);
                atomic_add_abs3(fsc, 3 * (i - 1), fi);
                atomic_add_abs3(fsc, 3 * (j - 1), fj);
                atomic_add_abs3(msc, 3 * (i - 1), ti);
                atomic_add_abs3(msc, 3 * (j - 1), tj);
                st3(c->k, k); st3(c->F, Fi); st3(c->ti, Ti); st3(c->tj, Tj);
                st3(c->s, mk(0.0, 0.0, 0.0));
            }
        } else {
            /* Hertz-Mindlin (kernels.jl:240-313) */
            double gap = Li - gi->r - gj->r;
            if (gap > 0) nfatal += 1;
            double dn = fabs(gap);
            v3 k = vadd(ki, vscl(vnormalize(vsub(kj, ki)), MUT(MUT_CONTACT_POINT) ? gi->r + dn : gi->r - dn));
            v3 kip = vsub(k, ki), kjp = vsub(k, kj);
            v3 vci = vadd(vcross(ld3(gi->w), kip), ld3(gi->v));
            v3 vcj = vadd(vcross(ld3(gj->w), kjp), ld3(gj->v));
            v3 vc = MUT(MUT_VC_SIGN) ? mv(&R, vsub(vci, vcj)) : mv(&R, vsub(vcj, vci));
            int32_t midi = gi->mid, midj = gj->mid;
            const surface* sf = &s->surf[midi - 1][midj - 1];
            double Rs = 1.0 / (1.0 / gi->r + 1.0 / gj->r);
            double ms = 1.0 / (1.0 / gi->m + 1.0 / gj->m);
            v3 sv = ld3(c->s);
            v3 F = hertz_mindlin(gap, dn, vc, s->mats[midi - 1].nu, s->mats[midi - 1].E,
                                 s->mats[midj - 1].nu, s->mats[midj - 1].E, Rs, ms, sf->e,
                                 sf->mu, dt, &sv, &nfatal);
            m33 Ri = minv(&R);
            v3 Fg = mv(&Ri, vneg(F));
            v3 mi = MUT(MUT_MOMENT_ORDER) ? vcross(Fg, kip) : vcross(kip, Fg);
            v3 mj = MUT(MUT_MOMENT_ORDER) ? vcross(vneg(Fg), kjp) : vcross(kjp, vneg(Fg));
            atomic_add_vec3(s->forces, 3 * (i - 1), Fg);
            atomic_add_vec3(s->forces, 3 * (j - 1), vneg(Fg));
            atomic_add_vec3(s->moments, 3 * (i - 1), mi);
            atomic_add_vec3(s->moments, 3 * (j - 1), mj);
            atomic_add_abs3(fsc, 3 * (i - 1), Fg);
            atomic_add_abs3(fsc, 3 * (j - 1), Fg);
            atomic_add_abs3(msc, 3 * (i - 1), mi);
            atomic_add_abs3(msc, 3 * (j - 1), mj);
            st3(c->k, k); st3(c->F, F); st3(c->s, sv);
        }
    }
    s->cptr[1] += (uint32_t)nbroken;
    s->cnt.bonds_broken_total += nbroken;
    s->cnt.fatal_gap += nfatal;
}

/* kernels.jl:317-426 */
static void resolve_wall(orc_scene* s) {
    const double dt = s->cfg.dt, tol = s->cfg.tolerance;
    int64_t wc = 0, nfatal = 0;
#pragma omp parallel for schedule(static) reduction(+ : wc, nfatal)
    for (int64_t i = 0; i < s->n; ++i) {
        const orc_grain* gi = &s->g[i];
        for (int32_t j = 0; j < s->nwall; ++j) {
            const orc_wall* w = &s->walls[j];
            orc_contact* c = &s->wall_contacts[i * s->nwall + j];
            v3 nrm = ld3(w->n);
            v3 ki = ld3(gi->k);
            if (fabs(vdot(ki, nrm) - w->d) < gi->r) {
                wc++;
                m33 R = frame(nrm, tol);
                double L = vdot(ki, nrm) - w->d;
                double gap = fabs(L) - gi->r;
                double dn = fabs(gap);
                v3 kip = vscl(vdiv(sclv(-L, nrm), fabs(L)), fabs(L) + dn / 2.0);
                v3 k = vadd(ki, kip);
                v3 vci = vadd(vcross(ld3(gi->w), kip), ld3(gi->v));
                v3 vc = mv(&R, vneg(vci));
                int32_t midi = gi->mid, midj = w->mid;
                const surface* sf = &s->surf[midi - 1][midj - 1];
                v3 sv = ld3(c->s);
                v3 F = hertz_mindlin(gap, dn, vc, s->mats[midi - 1].nu, s->mats[midi - 1].E,
                                     s->mats[midj - 1].nu, s->mats[midj - 1].E, gi->r, gi->m,
                                     sf->e, sf->mu, dt, &sv, &nfatal);
                m33 Ri = minv(&R);
                v3 Fg = mv(&Ri, vneg(F));
                add_vec3(s->forces, 3 * i, Fg);
                add_vec3(s->moments, 3 * i, vcross(kip, Fg));
                if (s->track_scale) {
                    add_abs3(s->fscale, 3 * i, Fg);
                    add_abs3(s->mscale, 3 * i, vcross(kip, Fg));
                }
                c->i = (int32_t)(i + 1); c->j = 0; c->midi = midi; c->midj = midj;
                st3(c->k, k); st3(c->F, F);
                st3(c->ti, mk(0, 0, 0)); st3(c->tj, mk(0, 0, 0)); st3(c->s, sv);
            } else {
                memset(c, 0, sizeof(*c));
            }
        }
    }
    s->cnt.wall_contacts = wc;
    s->cnt.fatal_gap += nfatal;
}

static double jsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }

/* kernels.jl:428-443 */
static void apply_body_force(orc_scene* s) {
    const double* gr = s->cfg.gravity;
    const double dmp = s->cfg.global_damping;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < s->n; ++i) {
        const orc_grain* g = &s->g[i];
        const double nd = MUT(MUT_DAMPING_SIGN) ? dmp : -dmp;
        add_vec3(s->forces, 3 * i, vscl(ld3(gr), g->m));
        v3 F = ld3(&s->forces[3 * i]);
        v3 T = ld3(&s->moments[3 * i]);
        v3 dF = mk(nd * fabs(F.x) * jsign(g->v[0]), nd * fabs(F.y) * jsign(g->v[1]),
                   nd * fabs(F.z) * jsign(g->v[2]));
        v3 dT = mk(nd * fabs(T.x) * jsign(g->w[0]), nd * fabs(T.y) * jsign(g->w[1]),
                   nd * fabs(T.z) * jsign(g->w[2]));
        add_vec3(s->forces, 3 * i, dF);
        add_vec3(s->moments, 3 * i, dT);
        if (s->track_scale) {
            add_abs3(s->fscale, 3 * i, vscl(ld3(gr), g->m));
            add_abs3(s->fscale, 3 * i, dF);
            add_abs3(s->mscale, 3 * i, dT);
        }
    }
}

/* Quaternion -> rotation matrix (normalising form, unit-q equivalent of dem.py:317-344). */
static m33 quat_matrix(const double q[4]) {
    double w = q[0], x = q[1], y = q[2], z = q[3];
    double nn = 1.0 / (w * w + x * x + y * y + z * z);
    m33 M;
    M.m[0][0] = 1.0 - 2.0 * (y * y + z * z) * nn;
    M.m[0][1] = 2.0 * (x * y - z * w) * nn;
    M.m[0][2] = 2.0 * (x * z + y * w) * nn;
    M.m[1][0] = 2.0 * (x * y + z * w) * nn;
    M.m[1][1] = 1.0 - 2.0 * (x * x + z * z) * nn;
    M.m[1][2] = 2.0 * (y * z - x * w) * nn;
    M.m[2][0] = 2.0 * (x * z - y * w) * nn;
    M.m[2][1] = 2.0 * (y * z + x * w) * nn;
    M.m[2][2] = 1.0 - 2.0 * (x * x + y * y) * nn;
    return M;
}

/* kernels.jl:445-485 */
static void update(orc_scene* s) {
    const double dt = s->cfg.dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < s->n; ++i) {
        orc_grain* g = &s->g[i];
        v3 F = ld3(&s->forces[3 * i]);
        v3 a = vdiv(F, g->m);
        v3 v = vadd(ld3(g->v), vscl(a, dt));
        v3 k = vadd(vadd(ld3(g->k), vscl(MUT(MUT_UPDATE_NEW_V) ? v : ld3(g->v), dt)),
                    vscl(sclv(0.5, a), dt * dt));
        double q[4] = {g->q[0], g->q[1], g->q[2], g->q[3]};
        v3 w = ld3(g->w);
        m33 R = quat_matrix(q);
        v3 T = ld3(&s->moments[3 * i]);
        v3 Tl = mv(&R, T);
        v3 wl = mv(&R, w);
        m33 I;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) I.m[r][c] = g->I[c * 3 + r];
        m33 Iinv = minv(&I);
        v3 dwl = mv(&Iinv, vsub(Tl, vcross(wl, mv(&I, wl))));
        m33 Rinv = minv(&R);
        v3 dw = mv(&Rinv, dwl);
        double dq0 = -0.5 * (q[1] * w.x + q[2] * w.y + q[3] * w.z);
        double dq1 = 0.5 * (q[0] * w.x + q[2] * w.z - q[3] * w.y);
        double dq2 = 0.5 * (q[0] * w.y + q[1] * w.z + q[3] * w.x);
        double dq3 = 0.5 * (q[0] * w.z + q[1] * w.y - q[2] * w.x);
        const double qs = MUT(MUT_QUAT_DT) ? dt : 1.0;
        double nq[4] = {q[0] + qs * dq0, q[1] + qs * dq1, q[2] + qs * dq2, q[3] + qs * dq3};
        double qn = sqrt(nq[0] * nq[0] + nq[1] * nq[1] + nq[2] * nq[2] + nq[3] * nq[3]);
        for (int c = 0; c < 4; ++c) g->q[c] = nq[c] / qn;
        w = vadd(w, vscl(dw, dt));
        st3(g->k, k); st3(g->v, v); st3(g->a, a); st3(g->w, w); st3(g->dw, dw);
    }
}

/* executor.jl:140-153, kernels.jl:487-506 (serial cursor keeps survivor order) */
static void late_clear_state(orc_scene* s) {
    uint32_t total = s->cptr[0], inactive = s->cptr[1];
    s->cnt.contact_ptr1 = inactive;
    if (inactive > 0) {
        s->cptr[0] = 0; s->cptr[1] = 0;
        for (uint32_t idx = 0; idx < total; ++idx) {
            const orc_contact* c = &s->contacts[idx];
            if (get_bit(&s->st, pairkey(s, c->i, c->j), F_ACTIVE)) {
                s->ctemp[s->cptr[0]] = *c;
                s->cptr[0] += 1u;
            }
        }
        orc_contact* t = s->contacts; s->contacts = s->ctemp; s->ctemp = t;
        /* keep the key set small: rebuild it from the survivors (no flag changes) */
        pairset nps; ps_init(&nps, 2 * (uint64_t)s->cptr[0] + 1024);
        for (uint32_t idx = 0; idx < s->cptr[0]; ++idx) {
            const orc_contact* c = &s->contacts[idx];
            uint64_t key = pairkey(s, c->i, c->j);
            *ps_ref(&nps, key) = ps_get(&s->st, key);
        }
        ps_free(&s->st); s->st = nps;
    }
}

static void refresh_counts(orc_scene* s) {
    int64_t a = 0, b = 0;
    for (uint32_t idx = 0; idx < s->cptr[0]; ++idx) {
        uint8_t f = ps_get(&s->st, pairkey(s, s->contacts[idx].i, s->contacts[idx].j));
        if (f & F_ACTIVE) a++;
        if (f & F_BONDED) b++;
    }
    s->cnt.active_contacts = a;
    s->cnt.bonded_contacts = b;
    s->cnt.contact_ptr0 = s->cptr[0];
}

int32_t orc_initialize(orc_scene* s) { /* executor.jl:155-160 */
    contact_pass(s);
    init_bonds(s);
    refresh_counts(s);
    return s->cnt.overflow ? -3 : 0;
}

int32_t orc_step(orc_scene* s, int64_t nsteps) { /* executor.jl:162-172 */
    for (int64_t t = 0; t < nsteps; ++t) {
        memset(s->forces, 0, 3 * s->n * sizeof(double)); /* clear_state! */
        memset(s->moments, 0, 3 * s->n * sizeof(double));
        if (s->track_scale) {
            memset(s->fscale, 0, 3 * s->n * sizeof(double));
            memset(s->mscale, 0, 3 * s->n * sizeof(double));
        }
        contact_pass(s);
        update_contacts(s);
        uint32_t total = s->cptr[0], inactive = s->cptr[1];
        if ((uint32_t)(total - inactive) > 0u) resolve_collision(s); /* executor.jl:99 */
        resolve_wall(s);
        apply_body_force(s);
        update(s);
        late_clear_state(s);
        s->cnt.step += 1;
    }
    refresh_counts(s);
    return s->cnt.overflow ? -3 : 0;
}

void orc_get_counters(const orc_scene* s, orc_counters* out) { *out = s->cnt; }

void orc_get_grains(const orc_scene* s, orc_grain* out) {
    memcpy(out, s->g, s->n * sizeof(orc_grain));
}
void orc_set_grains(orc_scene* s, const orc_grain* in) {
    memcpy(s->g, in, s->n * sizeof(orc_grain));
}

int64_t orc_num_active_contacts(const orc_scene* s) {
    int64_t a = 0;
    for (uint32_t idx = 0; idx < s->cptr[0]; ++idx)
        if (get_bit(&s->st, pairkey(s, s->contacts[idx].i, s->contacts[idx].j), F_ACTIVE)) a++;
    return a;
}

static int cmp_contact(const void* pa, const void* pb) {
    const orc_contact* a = (const orc_contact*)pa;
    const orc_contact* b = (const orc_contact*)pb;
    if (a->i != b->i) return a->i < b->i ? -1 : 1;
    if (a->j != b->j) return a->j < b->j ? -1 : 1;
    return 0;
}

int64_t orc_get_contacts(const orc_scene* s, orc_contact* out, uint8_t* bonded, int64_t cap) {
    int64_t m = 0;
    orc_contact* tmp = (orc_contact*)malloc((s->cptr[0] + 1) * sizeof(orc_contact));
    for (uint32_t idx = 0; idx < s->cptr[0]; ++idx)
        if (get_bit(&s->st, pairkey(s, s->contacts[idx].i, s->contacts[idx].j), F_ACTIVE))
            tmp[m++] = s->contacts[idx];
    qsort(tmp, m, sizeof(orc_contact), cmp_contact);
    int64_t w = m < cap ? m : cap;
    for (int64_t k = 0; k < w; ++k) {
        out[k] = tmp[k];
        if (bonded) bonded[k] = get_bit(&s->st, pairkey(s, tmp[k].i, tmp[k].j), F_BONDED) ? 1 : 0;
    }
    free(tmp);
    return m;
}

int32_t orc_set_contacts(orc_scene* s, const orc_contact* in, const uint8_t* bonded, int64_t n) {
    if (n > s->cmax) return -3;
    ensure_contacts(s, n);
    ps_free(&s->st);
    ps_init(&s->st, 2 * n + 1024);
    for (int64_t k = 0; k < n; ++k) {
        s->contacts[k] = in[k];
        uint64_t key = pairkey(s, in[k].i, in[k].j);
        set_bit(&s->st, key, F_ACTIVE);
        if (bonded && bonded[k]) set_bit(&s->st, key, F_BONDED);
    }
    s->cptr[0] = (uint32_t)n;
    s->cptr[1] = 0;
    refresh_counts(s);
    return 0;
}

void orc_get_wall_shear(const orc_scene* s, double* out) {
    for (int64_t k = 0; k < s->n * s->nwall; ++k)
        for (int a = 0; a < 3; ++a) out[3 * k + a] = s->wall_contacts[k].s[a];
}
void orc_set_wall_shear(orc_scene* s, const double* in) {
    for (int64_t k = 0; k < s->n * s->nwall; ++k)
        for (int a = 0; a < 3; ++a) s->wall_contacts[k].s[a] = in[3 * k + a];
}
void orc_get_forces(const orc_scene* s, double* F, double* T) {
    memcpy(F, s->forces, 3 * s->n * sizeof(double));
    memcpy(T, s->moments, 3 * s->n * sizeof(double));
}

void orc_set_track_scale(orc_scene* s, int32_t on) {
    if (on && !s->fscale) {
        s->fscale = (double*)calloc(3 * s->n, sizeof(double));
        s->mscale = (double*)calloc(3 * s->n, sizeof(double));
    }
    s->track_scale = on != 0;
}

void orc_get_force_scale(const orc_scene* s, double* F, double* T) {
    if (!s->fscale) {
        memset(F, 0, 3 * s->n * sizeof(double));
        memset(T, 0, 3 * s->n * sizeof(double));
        return;
    }
    memcpy(F, s->fscale, 3 * s->n * sizeof(double));
    memcpy(T, s->mscale, 3 * s->n * sizeof(double));
}

void orc_get_cells(const orc_scene* s, int32_t* kid, uint32_t* hid, double* kcenter,
                   uint32_t* nbr) {
    if (kid) memcpy(kid, s->kid, 3 * s->n * sizeof(int32_t));
    if (hid) memcpy(hid, s->hid, s->n * sizeof(uint32_t));
    if (kcenter) memcpy(kcenter, s->kcenter, 3 * s->n * sizeof(double));
    if (nbr) memcpy(nbr, s->nbr, 8 * s->n * sizeof(uint32_t));
}
void orc_get_hash_table(const orc_scene* s, uint32_t* counts, uint32_t* current,
                        uint32_t* pid) {
    if (counts) memcpy(counts, s->ht, s->H * sizeof(uint32_t));
    if (current) memcpy(current, s->htc, s->H * sizeof(uint32_t));
    if (pid) memcpy(pid, s->pid, s->n * sizeof(uint32_t));
}
int64_t orc_get_cp_list(const orc_scene* s, int32_t* pairs, int64_t cap) {
    int64_t w = s->total_cps < cap ? s->total_cps : cap;
    if (pairs) memcpy(pairs, s->cpl, 2 * w * sizeof(int32_t));
    return s->total_cps;
}
