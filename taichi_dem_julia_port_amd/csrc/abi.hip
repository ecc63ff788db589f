// abi.hip -- the extern "C" boundary of libdemhip (declared in include/demhip.h).
//
// dem_create performs the setup half of the reference's solve() (src/executor.jl:217-339):
// per-grain derivations, sizing, the material/surface tables and the device allocation.
// dem_initialize / dem_step replace initialize! / simulate! (src/executor.jl:155-172).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "dem_internal.hpp"

using namespace dem;

namespace {

thread_local std::string g_create_err;

int32_t fail(dem_scene* s, int32_t code, const std::string& msg) {
    if (s) s->err = msg;
    else g_create_err = msg;
    return code;
}

int32_t hip_fail(dem_scene* s, const char* what, hipError_t e) {
    return fail(s, DEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(s, call)                                      \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(s, #call, e_); \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e == hipSuccess) e = hipMemset(*p, 0, count * sizeof(T));
    return e;
}

// Sub-arrays of one device allocation, each 256-byte aligned.  Laid out once with a null base to
// size the allocation, then again over it.
struct Carve {
    char* base;
    size_t off = 0;
    explicit Carve(void* b) : base((char*)b) {}
    template <typename T>
    void take(T** p, int64_t count) {
        *p = (T*)(base + off);
        off += ((size_t)std::max<int64_t>(count, 1) * sizeof(T) + 255) & ~(size_t)255;
    }
};

// Scene allocations released by dem_destroy, kept for reuse when a cache is enabled
// (dem_set_memory_cache): a serving loop releases and creates batches of similar sizes at every
// change of its live set, and handing tens of gigabytes back to the driver and mapping them again
// costs up to hundreds of milliseconds per batch.  A request takes the smallest cached block of
// its device that holds it and wastes at most half of it.
struct CachedBlock {
    int device;
    void* ptr;
    size_t bytes;
};
std::mutex g_cache_mu;
std::vector<CachedBlock> g_cache;  // oldest first
size_t g_cache_cap = 0, g_cache_bytes = 0;

// Frees cached blocks, oldest first, until at most keep bytes remain.  Needs g_cache_mu.
void cache_trim(size_t keep) {
    while (g_cache_bytes > keep && !g_cache.empty()) {
        const CachedBlock b = g_cache.front();
        g_cache.erase(g_cache.begin());
        g_cache_bytes -= b.bytes;
        int cur = 0;
        hipGetDevice(&cur);
        hipSetDevice(b.device);
        hipFree(b.ptr);
        hipSetDevice(cur);
    }
}

// A device allocation of at least bytes on the current device (*got: its size).
hipError_t block_alloc(int device, void** p, size_t bytes, size_t* got) {
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        int best = -1;
        for (int i = 0; i < (int)g_cache.size(); ++i) {
            const CachedBlock& b = g_cache[i];
            if (b.device == device && b.bytes >= bytes && b.bytes <= 2 * bytes &&
                (best < 0 || b.bytes < g_cache[best].bytes))
                best = i;
        }
        if (best >= 0) {
            *p = g_cache[best].ptr;
            *got = g_cache[best].bytes;
            g_cache_bytes -= *got;
            g_cache.erase(g_cache.begin() + best);
            return hipSuccess;
        }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {  // cached blocks may be what fills the device
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(g_cache_mu);
            cache_trim(0);
        }
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) *p = nullptr;
    *got = bytes;
    return e;
}

void block_free(int device, void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    if (bytes > g_cache_cap) {
        hipFree(p);
        return;
    }
    g_cache.push_back({device, p, bytes});
    g_cache_bytes += bytes;
    cache_trim(g_cache_cap);
}

// The contact rows of K slots per particle, zeroed.
hipError_t alloc_rows(dem_scene* s, int32_t K) {
    DevState& d = s->d;
    const int64_t KN = (int64_t)K * s->n;
    auto layout = [&](void* base) {
        Carve c(base);
        c.take(&d.touch, KN);
        for (int b = 0; b < 2; ++b) { c.take(&d.pq[b], KN); c.take(&d.st[b], (int64_t)S_NCOMP * KN); }
        return c.off;
    };
    const size_t bytes = layout(nullptr);
    hipError_t e = block_alloc(s->device, &d.rows_arena, bytes, &d.rows_arena_bytes);
    if (e != hipSuccess) return e;
    layout(d.rows_arena);
    return hipMemset(d.rows_arena, 0, bytes);
}

// Run independent host tasks on threads of their own (serially where a thread cannot start).
void run_parallel(std::vector<std::function<void()>>& tasks) {
    std::vector<std::thread> th;
    size_t started = 0;
    try {
        for (; started < tasks.size(); ++started) th.emplace_back(tasks[started]);
    } catch (const std::system_error&) {
        // fewer threads than tasks: the rest run here
    }
    for (size_t k = started; k < tasks.size(); ++k) tasks[k]();
    for (auto& t : th) t.join();
}

// The near lists of KN slots per particle (nothing when the scene does not use them).  Entries
// are written by a rebuilding step's broad phase before any step reads them, so no fill.
hipError_t alloc_near(dem_scene* s, int32_t KN) {
    DevState& d = s->d;
    d.nb = nullptr;
    if (!s->P.skin) return hipSuccess;
    hipError_t e = block_alloc(s->device, &d.nb_arena, (size_t)KN * s->n * sizeof(uint32_t),
                               &d.nb_arena_bytes);
    if (e != hipSuccess) return e;
    d.nb = (uint32_t*)d.nb_arena;
    return hipSuccess;
}

// Near-list slots per particle at scene creation: the contact slots, at least 16.
int32_t near_slots(int32_t K) { return std::max(16, K); }

// Copies between host memory and a scene's device arrays, on the scene's stream and complete on
// return: ordered after every launch on that stream, and finished before the next one.
hipError_t to_host(const dem_scene* s, void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream);
    return e == hipSuccess ? hipStreamSynchronize(s->stream) : e;
}
hipError_t to_device(const dem_scene* s, void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream);
    return e == hipSuccess ? hipStreamSynchronize(s->stream) : e;
}

// executor.jl:223 r = cbrt(V / (4/3 pi))
double grain_radius(double V) { return cbrt(V / (4.0 / 3.0 * M_PI)); }

void drop_graphs(dem_scene* s) {
    for (int l = 0; l < 2; ++l)
        for (int g = 0; g < 2; ++g)
            for (int b = 0; b < 2; ++b) {
                hipGraphExec_t& ge = s->graph_exec[l][g][b];
                hipGraph_t& gr = s->graph[l][g][b];
                if (ge) { hipGraphExecDestroy(ge); ge = nullptr; }
                if (gr) { hipGraphDestroy(gr); gr = nullptr; }
                s->graph_tpb[l][g][b] = 0;
            }
}

void free_scene(dem_scene* s) {
    if (!s) return;
    DevState& d = s->d;
    // nothing queued may still use memory that goes back to the cache
    if (s->stream) hipStreamSynchronize(s->stream);
    block_free(s->device, d.arena, d.arena_bytes);
    block_free(s->device, d.rows_arena, d.rows_arena_bytes);
    block_free(s->device, d.nb_arena, d.nb_arena_bytes);
    if (s->h_ctr) hipHostFree(s->h_ctr);
    if (s->h_ctr_raw) hipHostFree(s->h_ctr_raw);
    for (hipEvent_t e : s->ev)
        if (e) hipEventDestroy(e);
    drop_graphs(s);
    if (s->stream) hipStreamDestroy(s->stream);
    delete[] s->h_id; delete[] s->h_gid; delete[] s->h_mid; delete[] s->h_V; delete[] s->h_m;
    delete[] s->h_r; delete[] s->h_r0; delete[] s->h_I;
    delete s;
}

int32_t sync_counters(dem_scene* s) {
    HIPCHK(s, hipMemcpyAsync(s->h_ctr_raw, s->d.ctr, sizeof(unsigned long long) * CTR_WORDS,
                             hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    for (int k = 0; k < CT_NCOUNTERS; ++k) {
        unsigned long long v = 0;
        for (int c = 0; c < CTR_COPIES; ++c) v += s->h_ctr_raw[c * CTR_STRIDE + k];
        s->h_ctr[k] = v;
    }
    return DEM_OK;
}

// The force kernel's block size for the next steps, from the row entries counted on the last
// counted step (sync_counters must have run), and whether a single scene runs contact-free.
void choose_force_blocks(dem_scene* s) {
    s->force_tpb = 4 * s->h_ctr[CT_ENTRIES] >= (unsigned long long)s->n ? 64 : 256;
    s->force_free = !s->batched && s->h_ctr[CT_ENTRIES] == 0;
}

// What a captured graph's force and merge launches depend on besides the buffer parity.
int force_key(const dem_scene* s) { return s->force_tpb + (s->force_free ? 1 : 0); }

}  // namespace

void dem::set_last_error_global(const std::string& msg) { g_create_err = msg; }

int32_t dem::timing_flush(dem_scene* s) {
    if (s->nev == 0) return DEM_OK;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    for (int k = 0; k < s->nev; ++k) {
        float ms = 0.f;
        HIPCHK(s, hipEventElapsedTime(&ms, s->ev[2 * k], s->ev[2 * k + 1]));
        s->kms[s->ev_kernel[k]] += ms;
        s->klaunch[s->ev_kernel[k]] += 1;
    }
    s->nev = 0;
    return DEM_OK;
}

namespace {

int32_t collect_timing(dem_scene* s) { return s->timing ? timing_flush(s) : DEM_OK; }

// Widen the contact rows to k2 slots per particle, keeping every committed row and its state:
// slot m of particle p stays at m * n + p in the partner words, and each state component moves
// to its new stride.  On an allocation failure the scene keeps its rows.
int32_t widen_rows(dem_scene* s, int32_t k2) {
    DevState& d = s->d;
    const int64_t n = s->n, K = s->K;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    uint32_t* o_touch = d.touch;
    uint32_t* o_pq[2] = {d.pq[0], d.pq[1]};
    double* o_st[2] = {d.st[0], d.st[1]};
    void* o_arena = d.rows_arena;
    const size_t o_bytes = d.rows_arena_bytes;
    hipError_t e = alloc_rows(s, k2);
    // the zero fill runs on the null stream, which the scene's non-blocking stream does not wait for
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        if (d.rows_arena && d.rows_arena != o_arena) block_free(s->device, d.rows_arena, d.rows_arena_bytes);
        d.touch = o_touch;
        for (int b = 0; b < 2; ++b) { d.pq[b] = o_pq[b]; d.st[b] = o_st[b]; }
        d.rows_arena = o_arena;
        d.rows_arena_bytes = o_bytes;
        return hip_fail(s, "widening the contact rows", e);
    }
    for (int b = 0; b < 2 && e == hipSuccess; ++b) {
        e = hipMemcpyAsync(d.pq[b], o_pq[b], (size_t)K * n * sizeof(uint32_t),
                           hipMemcpyDeviceToDevice, s->stream);
        for (int c = 0; c < S_NCOMP && e == hipSuccess; ++c)
            e = hipMemcpyAsync(d.st[b] + (int64_t)c * k2 * n, o_st[b] + (int64_t)c * K * n,
                               (size_t)K * n * sizeof(double), hipMemcpyDeviceToDevice, s->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return hip_fail(s, "copying the contact rows", e);
    block_free(s->device, o_arena, o_bytes);
    s->K = k2;
    s->P.K = k2;
    s->der.contact_slots = k2;
    drop_graphs(s);
    return DEM_OK;
}

// Widen the near lists to hold `need` entries per particle (at least twice as many slots, at most
// MAX_NEAR_SLOTS), after a rebuilding step found a longer list than they hold.  Their contents
// are not kept: the next step rebuilds them.  Past MAX_NEAR_SLOTS every step keeps rebuilding
// (a rebuilding step finds the contacts by itself, so results do not change).
constexpr int32_t MAX_NEAR_SLOTS = 512;
int32_t widen_near(dem_scene* s, unsigned long long need) {
    DevState& d = s->d;
    const int64_t want = std::min<int64_t>(
        std::max<int64_t>(2 * (int64_t)s->P.KN, ((int64_t)need + 7) & ~(int64_t)7), MAX_NEAR_SLOTS);
    if (want <= s->P.KN) return DEM_OK;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    block_free(s->device, d.nb_arena, d.nb_arena_bytes);
    d.nb_arena = nullptr;
    d.nb_arena_bytes = 0;
    s->P.KN = (int32_t)want;
    hipError_t e = alloc_near(s, s->P.KN);
    if (e != hipSuccess) return hip_fail(s, "widening the near lists", e);
    drop_graphs(s);
    s->skin_dirty = true;
    return DEM_OK;
}

// The next row width after an overflow that needed `need` slots: at least twice the current one.
int32_t wider_slots(const dem_scene* s, unsigned long long need) {
    const int64_t want = std::max<int64_t>(2 * (int64_t)s->K, ((int64_t)need + 7) & ~(int64_t)7);
    return (int32_t)std::min<int64_t>(want, MAX_SLOTS);
}

}  // namespace

extern "C" {

int32_t dem_abi_version(void) { return DEM_ABI_VERSION; }

const char* dem_last_error(const dem_scene_t* s) {
    return s ? s->err.c_str() : g_create_err.c_str();
}

const char* dem_kernel_name(int32_t k) {
    static const char* names[DEM_NUM_KERNELS] = {"hash", "scan", "scatter", "broad_phase",
                                                 "merge", "force_integrate", "init_bonds", "other"};
    return (k >= 0 && k < DEM_NUM_KERNELS) ? names[k] : "";
}

}  // extern "C"

namespace {

int32_t check_scene(const dem_config_t* cfg, const dem_wall_t* walls, int32_t nwall,
                    const dem_grain_in_t* g, int64_t n) {
    if (!cfg || !g) return fail(nullptr, DEM_ERR_INVALID, "cfg and grains must be non-NULL");
    if (n <= 0 || n > (int64_t)IDX_MASK) return fail(nullptr, DEM_ERR_INVALID, "n out of range");
    if (nwall < 0 || nwall > MAX_WALLS || (nwall > 0 && !walls))
        return fail(nullptr, DEM_ERR_INVALID, "nwall must be in [0, 16] with a wall array");
    if (!(cfg->dt > 0.0)) return fail(nullptr, DEM_ERR_INVALID, "dt must be positive");
    for (int64_t i = 0; i < n; ++i)
        if (!(g[i].V > 0.0) || !(g[i].m > 0.0))
            return fail(nullptr, DEM_ERR_INVALID, "grain volume and mass must be positive");
    for (int j = 0; j < nwall; ++j)
        if (walls[j].mid < 1 || walls[j].mid > 2)
            return fail(nullptr, DEM_ERR_INVALID, "wall material id must be 1 or 2");
    return DEM_OK;
}

// Near-list skin delta as a fraction of the largest radius.
constexpr double NEAR_SKIN = 0.05;

// The setup half of solve() for one scene (executor.jl:219-315): per-grain derivations into the
// host arrays at [off, off + n), sizing, and the kernel parameters.  want_dense selects the
// engine's dense sort grid; force_dense widens its cells instead of falling back to the
// reference's grid when the domain is large.
int32_t derive_scene(dem_scene* s, int64_t off, const dem_config_t* cfg, const dem_wall_t* walls,
                     int32_t nwall, const dem_grain_in_t* g, int64_t n, bool want_dense,
                     bool force_dense, int32_t flags, SceneP& S, dem_derived_t& der,
                     bool& slot_owned) {
    double rmax_r = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double r = grain_radius(g[i].V);
        const int64_t k = off + i;
        s->h_id[k] = g[i].id; s->h_gid[k] = g[i].gid; s->h_mid[k] = 1;
        s->h_V[k] = g[i].V; s->h_m[k] = g[i].m; s->h_r[k] = r;
        s->h_r0[k] = r * cfg->particle_contact_radius_multiplier;
        s->h_I[k] = 2.0 / 5.0 * g[i].m * (r * r);
        if (i == 0 || r > rmax_r) rmax_r = r;
    }
    der.r_max = rmax_r * cfg->particle_contact_radius_multiplier *
                (1.0 + cfg->bond_tensile_strength / cfg->bond_elastic_modulus) *
                cfg->neighboring_search_safety_factor;
    der.cell_size = 4.0 * der.r_max;
    der.material_density = g[n - 1].m / g[n - 1].V;
    if (!(der.cell_size > 0.0) || !std::isfinite(der.cell_size))
        return fail(nullptr, DEM_ERR_INVALID, "derived cell size is not positive and finite");
    uint32_t prod = 1u;
    for (int a = 0; a < 3; ++a)
        prod *= (uint32_t)(int32_t)ceil((cfg->domain_max[a] - cfg->domain_min[a]) / der.cell_size);
    der.grid_product_i32 = (int32_t)prod;
    int64_t H = der.grid_product_i32;
    H = std::min<int64_t>(std::max<int64_t>(H, 1 << 20), 1 << 22);
    der.hash_table_size = H;
    der.total_steps = (int64_t)trunc(cfg->target_time / cfg->dt);
    der.save_per_steps = (int64_t)trunc(cfg->saving_interval_time / cfg->dt);
    // largest distance at which a pair can be in contact or bonded
    const double reach = 2.0 * rmax_r * std::max(1.0, cfg->particle_contact_radius_multiplier);
    slot_owned = reach * (1.0 + 1e-9) < 0.5 * der.cell_size &&
                 !(flags & DEM_FLAG_INDEX_OWNED_BROAD_PHASE);

    memset(&S, 0, sizeof(S));
    // near-list skin delta: the lists hold the partners within r_i + r_j + 2 delta
    S.skin = (flags & DEM_FLAG_NO_NEAR_LISTS) ? 0.0 : NEAR_SKIN * rmax_r;
    // A touching (or bond-distance) pair lies in both particles' reference windows when the
    // contact reach is below cs/2 with margin; otherwise the broad phase checks the rule per pair.
    S.ref_check = reach * (1.0 + 1e-6) < 0.5 * der.cell_size ? 0 : 1;
    for (int a = 0; a < 3; ++a) S.grid.dmin[a] = cfg->domain_min[a];
    S.grid.cs = der.cell_size;
    S.grid.H = (uint64_t)H;
    S.grid.Hmask = ((uint64_t)H & ((uint64_t)H - 1)) == 0 ? (uint64_t)H - 1 : 0;
    S.nwall = nwall;
    S.bond_init = 1;
    S.dt = cfg->dt;
    S.tol = cfg->tolerance;
    S.damping = cfg->global_damping;
    for (int a = 0; a < 3; ++a) S.g[a] = cfg->gravity[a];
    S.matE[0] = cfg->particle_elastic_modulus; S.matNu[0] = cfg->particle_poisson_ratio;
    S.matE[1] = cfg->wall_elastic_modulus;     S.matNu[1] = cfg->wall_poisson_ratio;
    for (int m = 0; m < 2; ++m) {  // the same FP64 operations, in the same order, as per contact
        const double nui = S.matNu[0], Ei = S.matE[0], nuj = S.matNu[m], Ej = S.matE[m];
        S.ysm[m] = 1.0 / ((1.0 - nui * nui) / Ei + (1.0 - nuj * nuj) / Ej);
        S.gsm[m] = 1.0 / (2.0 * (2.0 - nui) * (1.0 + nui) / Ei + 2.0 * (2.0 - nuj) * (1.0 + nuj) / Ej);
    }
    auto beta_of = [](double e) { const double le = log(e); return le / sqrt(le * le + M_PI * M_PI); };
    SurfaceP pp{cfg->pp_coefficient_friction, cfg->pp_coefficient_restitution,
                beta_of(cfg->pp_coefficient_restitution), cfg->bond_radius_ratio,
                cfg->bond_elastic_modulus, cfg->bond_poisson_ratio, cfg->bond_compressive_strength,
                cfg->bond_tensile_strength, cfg->bond_shear_strength};
    SurfaceP pw{cfg->pw_coefficient_friction, cfg->pw_coefficient_restitution,
                beta_of(cfg->pw_coefficient_restitution), 0, 0, 0, 0, 0, 0};
    S.surf[0][0] = pp; S.surf[0][1] = pw; S.surf[1][0] = pw; S.surf[1][1] = pw;
    for (int j = 0; j < nwall; ++j) {
        for (int a = 0; a < 3; ++a) S.walls[j].n[a] = walls[j].n[a];
        S.walls[j].d = walls[j].d;
        S.walls[j].mid = walls[j].mid;
    }
    S.bond_mult = cfg->particle_contact_radius_multiplier;
    {
        const double E = cfg->particle_elastic_modulus, nu = cfg->particle_poisson_ratio;
        S.ys_pp = (float)(1.0 / ((1.0 - nu * nu) / E + (1.0 - nu * nu) / E));
        S.gs_pp = (float)(1.0 / (2.0 * (2.0 - nu) * (1.0 + nu) / E + 2.0 * (2.0 - nu) * (1.0 + nu) / E));
        S.beta_pp = (float)pp.beta;
        S.mu_pp = (float)pp.mu;
    }
    // Dense sort/broad-phase grids covering the domain with at most max(4 n, 2^20) cells
    // (batches widen the cells to at most 8 n + 4096): the per-step grid has cells of at least
    // the near-list reach 2 max r + 2 delta, the bond pass's of at least the bond reach 2 max r0.
    S.dense.cells = 0;
    S.dense_init.cells = 0;
    auto make_grid = [&](double cell, DenseGrid& G) -> bool {
        const double budget = force_dense ? 8.0 * (double)n + 4096.0
                                          : std::max(4.0 * (double)n, (double)(1 << 20));
        for (int it = 0; it < 64; ++it) {
            double cells = 1.0;
            int32_t dims[3];
            for (int a = 0; a < 3; ++a) {
                const double ext = cfg->domain_max[a] - cfg->domain_min[a];
                const double dd = std::max(1.0, ceil(ext / cell));
                dims[a] = (int32_t)std::min(dd, 2.0e9);
                cells *= dd;
            }
            if (cells <= budget && cells < 2.0e9) {
                for (int a = 0; a < 3; ++a) { G.origin[a] = cfg->domain_min[a]; G.dims[a] = dims[a]; }
                G.inv_cell = 1.0 / cell;
                G.cell = 1.0 / G.inv_cell;
                G.inv_dx = 1.0 / (double)dims[0];
                G.inv_dxy = 1.0 / ((double)dims[0] * (double)dims[1]);
                G.cells = (uint32_t)cells;
                return true;
            }
            if (!force_dense) return false;
            cell *= std::max(1.05, cbrt(cells / budget));
        }
        return false;
    };
    if (want_dense) {
        const double touch = (2.0 * rmax_r + 2.0 * S.skin) * (1.0 + 1e-6);
        const bool ok = make_grid(touch, S.dense) && make_grid(reach * (1.0 + 1e-6), S.dense_init);
        if (!ok) {
            S.dense.cells = 0;
            S.dense_init.cells = 0;
            if (force_dense) return fail(nullptr, DEM_ERR_INVALID, "scene domain cannot be gridded");
        } else if (!make_grid(2.0 * rmax_r * (1.0 + 1e-6), S.order)) {
            S.order = S.dense;
        }
    }
    return DEM_OK;
}

// The internal order of the particles: within each scene, by the cell of the initial position in
// the scene's per-step dense grid (or in the reference grid when the dense grid is unused), x
// fastest, ties by file index.  Neighbours in space become neighbours in memory, so the sort,
// the broad phase and the partner gathers of the force kernel read nearby addresses.
// Orders scenes k0.. only, whose input rows (file index gbase..) grows holds.
void spatial_order(dem_scene* s, const dem_grain_in_t* grows, int64_t gbase, int k0) {
    const int64_t n = s->n;
    s->perm.resize(n);
    s->iperm.resize(n);
    const int nseg = s->batched ? s->nseg : 1;
    std::vector<uint64_t> key;
    std::vector<uint32_t> cnt;
    for (int k = k0; k < nseg; ++k) {
        const int64_t a = s->batched ? s->seg_off[k] : 0, b = s->batched ? s->seg_off[k + 1] : n;
        const SceneP& S = s->batched ? s->seg_host[k] : s->P.sc;
        if (S.dense.cells > 0) {
            // counting sort by dense cell (stable: file order within a cell)
            const DenseGrid& G = S.order;
            key.assign(b - a, 0);
            cnt.assign((size_t)G.cells + 1, 0u);
            for (int64_t i = a; i < b; ++i) {
                uint64_t c[3];
                for (int ax = 0; ax < 3; ++ax) {
                    const double v = (grows[i - gbase].k[ax] - G.origin[ax]) * G.inv_cell;
                    int64_t q = std::isfinite(v) ? (int64_t)std::floor(std::max(-1.0, std::min(v, 4.0e9))) : 0;
                    q = q < 0 ? 0 : (q >= G.dims[ax] ? G.dims[ax] - 1 : q);
                    c[ax] = (uint64_t)q;
                }
                key[i - a] = c[0] + (uint64_t)G.dims[0] * (c[1] + (uint64_t)G.dims[1] * c[2]);
                ++cnt[key[i - a] + 1];
            }
            for (size_t c = 1; c < cnt.size(); ++c) cnt[c] += cnt[c - 1];
            for (int64_t i = a; i < b; ++i) s->perm[a + cnt[key[i - a]]++] = (uint32_t)i;
        } else {
            // reference grid: sort by cell coordinates, x fastest
            key.assign(b - a, 0);
            std::vector<uint32_t> idx(b - a);
            for (int64_t i = a; i < b; ++i) {
                uint64_t c[3];
                for (int ax = 0; ax < 3; ++ax) {
                    const double v = (grows[i - gbase].k[ax] - S.grid.dmin[ax]) / S.grid.cs;
                    int64_t q = std::isfinite(v) ? (int64_t)std::floor(std::max(-1.0, std::min(v, 4.0e6))) : 0;
                    q = q < 0 ? 0 : (q > (1 << 21) - 1 ? (1 << 21) - 1 : q);
                    c[ax] = (uint64_t)q;
                }
                key[i - a] = (c[2] << 42) | (c[1] << 21) | c[0];
                idx[i - a] = (uint32_t)i;
            }
            std::stable_sort(idx.begin(), idx.end(),
                             [&](uint32_t x, uint32_t y) { return key[x - a] < key[y - a]; });
            std::copy(idx.begin(), idx.end(), s->perm.begin() + a);
        }
    }
    for (int64_t u = gbase; u < n; ++u) s->iperm[s->perm[u]] = (uint32_t)u;
}

// Device allocation and initial state (executor.jl:229-242: q = 1, w = a = dw = 0) of the grains
// already derived into the host arrays.  grows holds the input rows of file index gbase.. (scenes
// k0..); particles below gbase are left for the caller to move in (dem_batch_repack).
int32_t allocate_scene(dem_scene* s, const dem_grain_in_t* grows, int64_t gbase, int k0,
                       int32_t nwall_max, int64_t nkeys) {
    DevState& d = s->d;
    const int64_t n = s->n;
    const int64_t KN = (int64_t)s->K * n;
    hipError_t e = hipSuccess;
    const int64_t max_tiles = scan_tiles_max(std::max<int64_t>(nkeys + 1, n + 1));
    auto layout = [&](void* base) {
        Carve c(base);
        for (int b = 0; b < 2; ++b) {
            c.take(&d.pos[b], n); c.take(&d.vel[b], n); c.take(&d.omg[b], n); c.take(&d.quat[b], n);
            c.take(&d.rcnt[b], n);
        }
        c.take(&d.acc, n); c.take(&d.dw, n); c.take(&d.ftot, n); c.take(&d.ttot, n); c.take(&d.r0, n);
        c.take(&d.wact, n); c.take(&d.wshear, (int64_t)(nwall_max > 0 ? nwall_max : 1) * n);
        c.take(&d.hid, n); c.take(&d.rank, n); c.take(&d.x32, n); c.take(&d.bcnt, nkeys + 1);
        c.take(&d.bstart, nkeys + 1); c.take(&d.sidxb[0], n); c.take(&d.sidxb[1], n); c.take(&d.skey, n);
        c.take(&d.spos, n); c.take(&d.spos32, n); c.take(&d.eid, n); c.take(&d.sfid, n);
        c.take(&d.tcnt, n); c.take(&d.sel, n);
        c.take(&d.nbc, n); c.take(&d.xref, n); c.take(&d.gmin, n); c.take(&d.gskip, (n + 63) / 64);
        c.take(&d.ctr, CTR_WORDS);
        c.take(&d.scan.status, max_tiles); c.take(&d.scan.counters, 2);
        c.take(&d.scan.tagg, max_tiles * TAGG_COPIES);
        c.take(&d.scratch_u32, n + 1); c.take(&d.scratch2_u32, n + 1); c.take(&d.total, 1);
        if (s->batched) { c.take(&d.seg, s->nseg); c.take(&d.sid, n); c.take(&d.seg_broken, s->nseg); }
        return c.off;
    };
    const size_t bytes = layout(nullptr);
    e = block_alloc(s->device, &d.arena, bytes, &d.arena_bytes);
    if (e != hipSuccess)
        return fail(nullptr, DEM_ERR_HIP, std::string("device allocation failed: ") + hipGetErrorString(e));
    layout(d.arena);
    d.sidx = d.sidxb[0];
    d.scan.max_tiles = max_tiles;
    e = hipMemset(d.arena, 0, bytes);
    if (e == hipSuccess) e = alloc_rows(s, s->K);
    if (e == hipSuccess) e = alloc_near(s, s->P.KN);
    if (e != hipSuccess)
        return fail(nullptr, DEM_ERR_HIP, std::string("device allocation failed: ") + hipGetErrorString(e));
    if (hipHostMalloc((void**)&s->h_ctr, sizeof(unsigned long long) * CT_NCOUNTERS) != hipSuccess)
        return fail(nullptr, DEM_ERR_HIP, "hipHostMalloc failed");
    memset(s->h_ctr, 0, sizeof(unsigned long long) * CT_NCOUNTERS);
    if (hipHostMalloc((void**)&s->h_ctr_raw, sizeof(unsigned long long) * CTR_WORDS) != hipSuccess)
        return fail(nullptr, DEM_ERR_HIP, "hipHostMalloc failed");
    memset(s->h_ctr_raw, 0, sizeof(unsigned long long) * CTR_WORDS);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(nullptr, DEM_ERR_HIP, "hipStreamCreate failed");
    spatial_order(s, grows, gbase, k0);
    // one staging allocation: input rows, radii and inertias of the new grains, r0 of all grains
    // (file order) and the scene offsets; the gather into internal order runs on the device
    const int64_t nn = n - gbase;
    const size_t b_rows = (size_t)nn * sizeof(dem_grain_in_t), b_rI = (size_t)nn * sizeof(double);
    const size_t b_r0 = (size_t)n * sizeof(double), b_off = (size_t)(s->nseg + 1) * sizeof(int64_t);
    char* tmp = nullptr;
    e = hipMalloc((void**)&tmp, b_rows + 2 * b_rI + b_r0 + b_off);
    if (e != hipSuccess) return fail(nullptr, DEM_ERR_HIP, std::string("staging allocation failed: ") + hipGetErrorString(e));
    dem_grain_in_t* t_rows = (dem_grain_in_t*)tmp;
    double* t_r = (double*)(tmp + b_rows);
    double* t_I = t_r + nn;
    double* t_r0 = t_I + nn;
    int64_t* t_off = (int64_t*)(t_r0 + n);
    if (nn > 0) {
        e = hipMemcpy(t_rows, grows, b_rows, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(t_r, s->h_r + gbase, b_rI, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(t_I, s->h_I + gbase, b_rI, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(t_r0, s->h_r0, b_r0, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t_off, s->seg_off.data(), b_off, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.eid, s->perm.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && s->batched)
        e = hipMemcpy(d.seg, s->seg_host.data(), s->nseg * sizeof(SceneP), hipMemcpyHostToDevice);
    // the zero fills and copies above run on the null stream, which the scene's non-blocking
    // stream does not wait for
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = run_init_state(s, t_rows, t_r, t_I, t_r0, t_off, gbase);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    hipFree(tmp);
    if (e != hipSuccess) return fail(nullptr, DEM_ERR_HIP, std::string("initial upload failed: ") + hipGetErrorString(e));
    s->P.seg = d.seg;
    s->P.sid = d.sid;
    s->P.seg_broken = d.seg_broken;
    return DEM_OK;
}

void alloc_host_fields(dem_scene* s, int64_t n) {
    s->h_id = new int32_t[n]; s->h_gid = new int32_t[n]; s->h_mid = new int32_t[n];
    s->h_V = new double[n]; s->h_m = new double[n]; s->h_r = new double[n];
    s->h_r0 = new double[n]; s->h_I = new double[n];
}

int32_t options_of(const dem_options_t* opt, dem_options_t& o) {
    o = dem_options_t{DEM_PRECISION_FP64, 0, 0, 0};
    if (opt) o = *opt;
    if (o.precision != DEM_PRECISION_FP64 && o.precision != DEM_PRECISION_MIXED)
        return fail(nullptr, DEM_ERR_INVALID, "unknown precision");
    return DEM_OK;
}

// A row holds a particle's contacts in both roles (i and j), so the default is twice the
// reference's per-particle average budget (contacts = n * max_coordinate_number,
// executor.jl:277), at least 32.
int32_t slots_of(const dem_options_t& o, int32_t max_coord) {
    return o.contact_slots > 0 ? o.contact_slots : std::min(128, std::max(32, 2 * max_coord));
}

}  // namespace

extern "C" {

int32_t dem_create(const dem_config_t* cfg, const dem_wall_t* walls, int32_t nwall,
                   const dem_grain_in_t* g, int64_t n, const dem_options_t* opt,
                   dem_scene_t** out) {
    g_create_err.clear();
    if (!out) return fail(nullptr, DEM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int32_t rc = check_scene(cfg, walls, nwall, g, n);
    if (rc) return rc;
    dem_options_t o;
    if ((rc = options_of(opt, o))) return rc;
    const int32_t K = slots_of(o, cfg->max_coordinate_number);
    if (K < 1 || K > MAX_SLOTS) return fail(nullptr, DEM_ERR_INVALID, "contact_slots must be in [1, 256]");
    if (hipSetDevice(o.device) != hipSuccess) return fail(nullptr, DEM_ERR_HIP, "hipSetDevice failed");

    dem_scene* s = new dem_scene();
    s->n = n;
    s->nwall = nwall;
    s->precision = o.precision;
    s->K = K;
    s->device = o.device;
    s->cfg = *cfg;
    s->use_graph = !(o.flags & DEM_FLAG_NO_GRAPH);
    s->seg_off = {0, n};
    alloc_host_fields(s, n);
    const bool want_dense = !(o.flags & (DEM_FLAG_REFERENCE_GRID | DEM_FLAG_INDEX_OWNED_BROAD_PHASE));
    if ((rc = derive_scene(s, 0, cfg, walls, nwall, g, n, want_dense, false, o.flags, s->P.sc,
                           s->der, s->slot_owned))) {
        free_scene(s);
        return rc;
    }
    s->der.contact_slots = K;
    s->seg_cfg = {*cfg};
    s->seg_der = {s->der};
    Params& P = s->P;
    P.n = n;
    P.K = K;
    P.nseg = 1;
    s->dense = P.sc.dense.cells > 0;
    P.skin = s->dense && !(o.flags & DEM_FLAG_NO_NEAR_LISTS) ? 1 : 0;
    P.KN = near_slots(K);
    s->skin_min = P.sc.skin;
    P.sort_keys = s->dense ? P.sc.dense.cells : (uint32_t)P.sc.grid.H;
    P.sort_keys_init = s->dense ? P.sc.dense_init.cells : (uint32_t)P.sc.grid.H;
    const int64_t nkeys = std::max<int64_t>(
        (int64_t)P.sc.grid.H, std::max<int64_t>((int64_t)P.sc.dense.cells, (int64_t)P.sc.dense_init.cells));
    if ((rc = allocate_scene(s, g, 0, 0, nwall, nkeys))) {
        free_scene(s);
        return rc;
    }
    *out = s;
    return DEM_OK;
}

int32_t dem_create_batch(int32_t nscenes, const dem_config_t* cfgs, const int32_t* nwalls,
                         const dem_wall_t* walls, const int64_t* counts,
                         const dem_grain_in_t* grains, const dem_options_t* opt,
                         dem_scene_t** out) {
    g_create_err.clear();
    if (!out) return fail(nullptr, DEM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (nscenes < 1 || nscenes > 65535 || !cfgs || !nwalls || !counts || !grains)
        return fail(nullptr, DEM_ERR_INVALID, "need 1..65535 scenes with configs, wall counts and grains");
    dem_options_t o;
    int32_t rc = options_of(opt, o);
    if (rc) return rc;
    if (o.flags & (DEM_FLAG_REFERENCE_GRID | DEM_FLAG_INDEX_OWNED_BROAD_PHASE))
        return fail(nullptr, DEM_ERR_INVALID, "batches always use the dense sort grid");
    int64_t total = 0, woff = 0;
    int32_t nwall_max = 0, maxcoord = 0;
    for (int k = 0; k < nscenes; ++k) {
        const int64_t nk = counts[k];
        if ((rc = check_scene(&cfgs[k], walls ? walls + woff : nullptr, nwalls[k], grains + total, nk)))
            return rc;
        total += nk;
        woff += nwalls[k];
        nwall_max = std::max(nwall_max, nwalls[k]);
        maxcoord = std::max(maxcoord, cfgs[k].max_coordinate_number);
    }
    if (total > (int64_t)IDX_MASK) return fail(nullptr, DEM_ERR_INVALID, "too many particles");
    const int32_t K = slots_of(o, maxcoord);
    if (K < 1 || K > MAX_SLOTS) return fail(nullptr, DEM_ERR_INVALID, "contact_slots must be in [1, 256]");
    if (hipSetDevice(o.device) != hipSuccess) return fail(nullptr, DEM_ERR_HIP, "hipSetDevice failed");

    dem_scene* s = new dem_scene();
    s->n = total;
    s->nwall = nwall_max;
    s->precision = o.precision;
    s->K = K;
    s->device = o.device;
    s->cfg = cfgs[0];
    s->use_graph = !(o.flags & DEM_FLAG_NO_GRAPH);
    s->batched = true;
    s->nseg = nscenes;
    s->seg_off.assign(nscenes + 1, 0);
    s->seg_host.resize(nscenes);
    s->seg_cfg.resize(nscenes);
    s->seg_der.resize(nscenes);
    alloc_host_fields(s, total);
    int64_t off = 0;
    uint64_t keys = 0, keys_init = 0;
    woff = 0;
    for (int k = 0; k < nscenes; ++k) {
        dem_derived_t der{};
        bool slot_owned = false;
        SceneP& S = s->seg_host[k];
        if ((rc = derive_scene(s, off, &cfgs[k], walls ? walls + woff : nullptr, nwalls[k], grains + off,
                               counts[k], true, true, o.flags, S, der, slot_owned))) {
            free_scene(s);
            return rc;
        }
        S.cell_base = (uint32_t)keys;
        keys += S.dense.cells;
        S.cell_base_init = (uint32_t)keys_init;
        keys_init += S.dense_init.cells;
        if (k == 0) s->der = der;
        s->seg_cfg[k] = cfgs[k];
        s->seg_der[k] = der;
        s->seg_off[k] = off;
        off += counts[k];
        woff += nwalls[k];
    }
    s->seg_off[nscenes] = off;
    if (keys >= (1ull << 31) || keys_init >= (1ull << 31)) {
        free_scene(s);
        return fail(nullptr, DEM_ERR_INVALID, "batch sort grid too large");
    }
    s->der.contact_slots = K;
    Params& P = s->P;
    P.n = total;
    P.K = K;
    P.nseg = nscenes;
    P.sc = s->seg_host[0];
    s->dense = true;
    P.skin = (o.flags & DEM_FLAG_NO_NEAR_LISTS) ? 0 : 1;
    P.KN = near_slots(K);
    s->skin_min = s->seg_host[0].skin;
    for (const SceneP& Sk : s->seg_host) s->skin_min = std::min(s->skin_min, Sk.skin);
    P.sort_keys = (uint32_t)keys;
    P.sort_keys_init = (uint32_t)keys_init;
    if ((rc = allocate_scene(s, grains, 0, 0, nwall_max, (int64_t)std::max(keys, keys_init)))) {
        free_scene(s);
        return rc;
    }
    *out = s;
    return DEM_OK;
}

int32_t dem_batch_offsets(const dem_scene_t* s, int64_t* offsets) {
    if (!s || !offsets) return DEM_ERR_INVALID;
    for (int k = 0; k <= s->nseg; ++k) offsets[k] = s->seg_off[k];
    return DEM_OK;
}

int32_t dem_batch_set_bond_init(dem_scene_t* s, const uint8_t* mask) {
    if (!s || !mask) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    if (!s->batched) {
        s->P.sc.bond_init = mask[0] ? 1 : 0;
        return DEM_OK;
    }
    for (int k = 0; k < s->nseg; ++k) s->seg_host[k].bond_init = mask[k] ? 1 : 0;
    s->P.sc = s->seg_host[0];
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, to_device(s, s->d.seg, s->seg_host.data(), s->nseg * sizeof(SceneP)));
    return DEM_OK;
}

int32_t dem_batch_adopt(dem_scene_t* dst, int32_t dst_scene, dem_scene_t* src, int32_t src_scene) {
    if (!dst || !src || dst == src) return DEM_ERR_INVALID;
    if (dst_scene < 0 || dst_scene >= dst->nseg || src_scene < 0 || src_scene >= src->nseg)
        return fail(dst, DEM_ERR_INVALID, "scene index out of range");
    if (dst->device != src->device) return fail(dst, DEM_ERR_INVALID, "scenes on different devices");
    const int64_t n = src->seg_off[src_scene + 1] - src->seg_off[src_scene];
    if (dst->seg_off[dst_scene + 1] - dst->seg_off[dst_scene] != n)
        return fail(dst, DEM_ERR_INVALID, "scenes differ in particle count");
    const SceneP& Ss = src->batched ? src->seg_host[src_scene] : src->P.sc;
    const SceneP& Sd = dst->batched ? dst->seg_host[dst_scene] : dst->P.sc;
    if (Ss.nwall != Sd.nwall) return fail(dst, DEM_ERR_INVALID, "scenes differ in wall count");
    {  // state moves slot for slot, so both sides must store the scene's particles in one order
        const int64_t so = src->seg_off[src_scene], dof = dst->seg_off[dst_scene];
        for (int64_t k = 0; k < n; ++k)
            if ((int64_t)src->perm[so + k] - so != (int64_t)dst->perm[dof + k] - dof)
                return fail(dst, DEM_ERR_INVALID,
                            "scenes store their particles in different orders (different initial grains)");
    }
    HIPCHK(dst, hipSetDevice(dst->device));
    HIPCHK(dst, hipStreamSynchronize(src->stream));
    HIPCHK(dst, hipStreamSynchronize(dst->stream));
    uint32_t* ovf = nullptr;
    HIPCHK(dst, hipMalloc(&ovf, sizeof(uint32_t)));
    // ordered on the stream k_adopt runs on (a null-stream fill would not be waited for)
    hipError_t e = hipMemsetAsync(ovf, 0, sizeof(uint32_t), dst->stream);
    if (e == hipSuccess)
        e = run_adopt(dst, dst->seg_off[dst_scene], src, src->seg_off[src_scene], n, Ss.nwall, ovf);
    uint32_t o = 0;
    if (e == hipSuccess) e = hipStreamSynchronize(dst->stream);
    if (e == hipSuccess) e = to_host(dst, &o, ovf, sizeof(uint32_t));
    hipFree(ovf);
    if (e != hipSuccess) return hip_fail(dst, "dem_batch_adopt", e);
    dst->skin_dirty = true;  // the destination scene's particles were replaced
    dst->force_free = false;  // and may have rows now
    if (o) return fail(dst, DEM_ERR_CAPACITY, "destination has fewer contact slots than needed");
    // cumulative bond breaks of the scene travel with it
    unsigned long long nb = 0;
    if (src->batched) {
        HIPCHK(dst, to_host(src, &nb, src->d.seg_broken + src_scene, sizeof(nb)));
    } else {
        int32_t rc = sync_counters(src);
        if (rc) return fail(dst, rc, "dem_batch_adopt: reading the source counters failed");
        nb = src->h_ctr[CT_BROKEN_TOTAL];
    }
    if (dst->batched) {
        HIPCHK(dst, to_device(dst, dst->d.seg_broken + dst_scene, &nb, sizeof(nb)));
    } else {
        // the total goes into copy 0, the other copies restart from zero
        std::vector<unsigned long long> raw(CTR_WORDS);
        HIPCHK(dst, to_host(dst, raw.data(), dst->d.ctr, CTR_WORDS * sizeof(unsigned long long)));
        for (int c = 0; c < CTR_COPIES; ++c) raw[c * CTR_STRIDE + CT_BROKEN_TOTAL] = c == 0 ? nb : 0ull;
        HIPCHK(dst, to_device(dst, dst->d.ctr, raw.data(), CTR_WORDS * sizeof(unsigned long long)));
    }
    return DEM_OK;
}

int32_t dem_batch_repack(dem_scene_t* src, int32_t nkeep, const int32_t* keep, int32_t nnew,
                         const dem_config_t* cfgs, const int32_t* nwalls, const dem_wall_t* walls,
                         const int64_t* counts, const dem_grain_in_t* grains,
                         const dem_options_t* opt, dem_scene_t** out) {
    g_create_err.clear();
    if (!out) return fail(nullptr, DEM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!src) return fail(nullptr, DEM_ERR_INVALID, "src is NULL");
    if (nkeep < 0 || nnew < 0 || nkeep > src->nseg || (nkeep > 0 && !keep) || nkeep + nnew < 1 ||
        nkeep + nnew > 65535)
        return fail(nullptr, DEM_ERR_INVALID, "need 1..65535 scenes: kept scenes of src plus new ones");
    if (nnew > 0 && (!cfgs || !nwalls || !counts || !grains))
        return fail(nullptr, DEM_ERR_INVALID, "new scenes need configs, wall counts and grains");
    dem_options_t o;
    int32_t rc = options_of(opt, o);
    if (rc) return rc;
    if (o.flags & (DEM_FLAG_REFERENCE_GRID | DEM_FLAG_INDEX_OWNED_BROAD_PHASE))
        return fail(nullptr, DEM_ERR_INVALID, "batches always use the dense sort grid");
    if (o.device != src->device) return fail(nullptr, DEM_ERR_INVALID, "src is on another device");
    const int32_t nscenes = nkeep + nnew;
    std::vector<char> seen(src->nseg, 0);
    int64_t total = 0, woff = 0;
    int32_t nwall_max = 0, maxcoord = 0;
    for (int j = 0; j < nkeep; ++j) {
        const int32_t k = keep[j];
        if (k < 0 || k >= src->nseg || seen[k])
            return fail(nullptr, DEM_ERR_INVALID, "keep lists a scene index out of range or twice");
        seen[k] = 1;
        const SceneP& Sk = src->batched ? src->seg_host[k] : src->P.sc;
        if (Sk.dense.cells == 0 || Sk.dense_init.cells == 0)
            return fail(nullptr, DEM_ERR_INVALID, "a kept scene does not use the dense sort grid");
        total += src->seg_off[k + 1] - src->seg_off[k];
        nwall_max = std::max(nwall_max, Sk.nwall);
        maxcoord = std::max(maxcoord, src->seg_cfg[k].max_coordinate_number);
    }
    const int64_t kept_total = total;
    int64_t goff = 0;
    for (int t = 0; t < nnew; ++t) {
        if ((rc = check_scene(&cfgs[t], walls ? walls + woff : nullptr, nwalls[t], grains + goff, counts[t])))
            return rc;
        goff += counts[t];
        woff += nwalls[t];
        nwall_max = std::max(nwall_max, nwalls[t]);
        maxcoord = std::max(maxcoord, cfgs[t].max_coordinate_number);
    }
    total += goff;
    if (total > (int64_t)IDX_MASK) return fail(nullptr, DEM_ERR_INVALID, "too many particles");
    int32_t K = slots_of(o, maxcoord);
    // rows that dem_initialize widened stay wide: the kept scenes' bonds must fit
    if (o.contact_slots <= 0 && nkeep > 0) K = std::max(K, src->K);
    if (K < 1 || K > MAX_SLOTS) return fail(nullptr, DEM_ERR_INVALID, "contact_slots must be in [1, 256]");
    if (hipSetDevice(o.device) != hipSuccess) return fail(nullptr, DEM_ERR_HIP, "hipSetDevice failed");

    dem_scene* s = new dem_scene();
    s->n = total;
    s->nwall = nwall_max;
    s->precision = o.precision;
    s->K = K;
    s->device = o.device;
    s->use_graph = !(o.flags & DEM_FLAG_NO_GRAPH);
    s->batched = true;
    s->nseg = nscenes;
    s->seg_off.assign(nscenes + 1, 0);
    s->seg_host.resize(nscenes);
    s->seg_cfg.resize(nscenes);
    s->seg_der.resize(nscenes);
    s->perm.resize(total);
    s->iperm.resize(total);
    alloc_host_fields(s, total);
    int64_t off = 0;
    uint64_t keys = 0, keys_init = 0;
    auto place = [&](int j, int64_t m) {
        SceneP& S = s->seg_host[j];
        S.cell_base = (uint32_t)keys;
        keys += S.dense.cells;
        S.cell_base_init = (uint32_t)keys_init;
        keys_init += S.dense_init.cells;
        s->seg_off[j] = off;
        off += m;
    };
    // kept scenes: their derived host fields, order and parameters move over unchanged
    std::vector<int64_t> kept_at(nkeep);  // first particle of kept scene j in the new batch
    for (int j = 0; j < nkeep; ++j) {
        const int32_t k = keep[j];
        s->seg_host[j] = src->batched ? src->seg_host[k] : src->P.sc;
        s->seg_host[j].bond_init = 0;
        s->seg_cfg[j] = src->seg_cfg[k];
        s->seg_der[j] = src->seg_der[k];
        kept_at[j] = off;
        place(j, src->seg_off[k + 1] - src->seg_off[k]);
    }
    {
        // the per-grain host fields, one task per field: the copies are independent, and the
        // first touch of the new arrays' pages makes a serial copy of a large batch slow
        auto each_kept = [&](auto fn) {
            for (int j = 0; j < nkeep; ++j) {
                const int64_t a = src->seg_off[keep[j]];
                fn(a, src->seg_off[keep[j] + 1] - a, kept_at[j]);
            }
        };
        auto field = [&](auto* dst, const auto* from) {
            return [=, &each_kept] {
                each_kept([&](int64_t a, int64_t m, int64_t at) { std::copy(from + a, from + a + m, dst + at); });
            };
        };
        std::vector<std::function<void()>> tasks = {
            field(s->h_id, src->h_id), field(s->h_gid, src->h_gid), field(s->h_mid, src->h_mid),
            field(s->h_V, src->h_V), field(s->h_m, src->h_m), field(s->h_r, src->h_r),
            field(s->h_r0, src->h_r0), field(s->h_I, src->h_I),
            [&] {
                each_kept([&](int64_t a, int64_t m, int64_t at) {
                    for (int64_t u = 0; u < m; ++u) {
                        s->perm[at + u] = (uint32_t)(src->perm[a + u] - a + at);
                        s->iperm[s->perm[at + u]] = (uint32_t)(at + u);
                    }
                });
            }};
        run_parallel(tasks);
    }
    // new scenes: derived from their input rows as dem_create_batch does
    woff = 0;
    goff = 0;
    for (int t = 0; t < nnew; ++t) {
        const int j = nkeep + t;
        bool slot_owned = false;
        if ((rc = derive_scene(s, off, &cfgs[t], walls ? walls + woff : nullptr, nwalls[t], grains + goff,
                               counts[t], true, true, o.flags, s->seg_host[j], s->seg_der[j], slot_owned))) {
            free_scene(s);
            return rc;
        }
        s->seg_cfg[j] = cfgs[t];
        place(j, counts[t]);
        goff += counts[t];
        woff += nwalls[t];
    }
    s->seg_off[nscenes] = off;
    if (keys >= (1ull << 31) || keys_init >= (1ull << 31)) {
        free_scene(s);
        return fail(nullptr, DEM_ERR_INVALID, "batch sort grid too large");
    }
    s->cfg = s->seg_cfg[0];
    s->der = s->seg_der[0];
    s->der.contact_slots = K;
    // the kept scenes bring their contacts: start with the source's force-kernel block size
    if (nkeep > 0) s->force_tpb = src->force_tpb;
    Params& P = s->P;
    P.n = total;
    P.K = K;
    P.nseg = nscenes;
    P.sc = s->seg_host[0];
    s->dense = true;
    P.skin = (o.flags & DEM_FLAG_NO_NEAR_LISTS) ? 0 : 1;
    P.KN = near_slots(K);
    s->skin_min = s->seg_host[0].skin;
    for (const SceneP& Sk : s->seg_host) s->skin_min = std::min(s->skin_min, Sk.skin);
    P.sort_keys = (uint32_t)keys;
    P.sort_keys_init = (uint32_t)keys_init;
    if ((rc = allocate_scene(s, grains, kept_total, nkeep, nwall_max, (int64_t)std::max(keys, keys_init)))) {
        free_scene(s);
        return rc;
    }
    // the kept scenes' state, contact rows and bond-break totals, all on the new batch's stream
    std::vector<unsigned long long> nb(nscenes, 0ull);
    hipError_t e = hipStreamSynchronize(src->stream);
    if (e == hipSuccess && nkeep > 0) {
        if (src->batched) {
            std::vector<unsigned long long> v(src->nseg);
            e = to_host(src, v.data(), src->d.seg_broken, src->nseg * sizeof(unsigned long long));
            for (int j = 0; j < nkeep; ++j) nb[j] = v[keep[j]];
        } else {
            rc = sync_counters(src);
            if (rc) {
                free_scene(s);
                return fail(nullptr, rc, "dem_batch_repack: reading the source counters failed");
            }
            nb[0] = src->h_ctr[CT_BROKEN_TOTAL];
        }
    }
    uint32_t* ovf = nullptr;
    if (e == hipSuccess) e = hipMalloc(&ovf, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(ovf, 0, sizeof(uint32_t), s->stream);
    for (int j = 0; j < nkeep && e == hipSuccess; ++j) {
        const int32_t k = keep[j];
        const SceneP& Sk = s->seg_host[j];
        e = run_adopt(s, s->seg_off[j], src, src->seg_off[k], src->seg_off[k + 1] - src->seg_off[k],
                      Sk.nwall, ovf);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d.seg_broken, nb.data(), nscenes * sizeof(unsigned long long),
                           hipMemcpyHostToDevice, s->stream);
    uint32_t o_flag = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&o_flag, ovf, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (ovf) hipFree(ovf);
    if (e != hipSuccess) {
        free_scene(s);
        return fail(nullptr, DEM_ERR_HIP, std::string("dem_batch_repack: ") + hipGetErrorString(e));
    }
    if (o_flag) {
        free_scene(s);
        return fail(nullptr, DEM_ERR_CAPACITY, "dem_batch_repack: a kept scene has more contacts than the new rows hold");
    }
    *out = s;
    return DEM_OK;
}

int32_t dem_batch_bonds_broken(dem_scene_t* s, int64_t* out) {
    if (!s || !out) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    if (!s->batched) {
        int32_t rc = sync_counters(s);
        if (rc) return rc;
        out[0] = (int64_t)s->h_ctr[CT_BROKEN_TOTAL];
        return DEM_OK;
    }
    std::vector<unsigned long long> v(s->nseg);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, to_host(s, v.data(), s->d.seg_broken, s->nseg * sizeof(unsigned long long)));
    for (int k = 0; k < s->nseg; ++k) out[k] = (int64_t)v[k];
    return DEM_OK;
}

void dem_destroy(dem_scene_t* s) {
    if (!s) return;
    hipSetDevice(s->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    free_scene(s);
}

int32_t dem_get_derived(const dem_scene_t* s, dem_derived_t* out) {
    if (!s || !out) return DEM_ERR_INVALID;
    *out = s->der;
    return DEM_OK;
}

int32_t dem_initialize(dem_scene_t* s) {
    if (!s) return DEM_ERR_INVALID;
    if (s->initialized || s->steps > 0)
        return fail(s, DEM_ERR_STATE, "dem_initialize must be called once, before the first step");
    HIPCHK(s, hipSetDevice(s->device));
    while (true) {
        HIPCHK(s, run_contact_detection(s, true, true));
        int32_t rc = sync_counters(s);
        if (rc) return rc;
        if ((rc = collect_timing(s))) return rc;
        if (!s->h_ctr[CT_FAIL]) {
            HIPCHK(s, run_rows_finish(s, false));  // commit the rows the bond pass built
            HIPCHK(s, hipStreamSynchronize(s->stream));
            break;
        }
        // a particle has more bonds than its row holds: drop the pass's rows (rows that existed
        // before it, such as those of a re-packed batch's kept scenes, stay), widen, redo it
        HIPCHK(s, run_rows_finish(s, true));
        if (s->K >= MAX_SLOTS)
            return fail(s, DEM_ERR_CAPACITY, "more than 256 contacts on one particle during bond init");
        if ((rc = widen_rows(s, wider_slots(s, s->h_ctr[CT_MAXROW])))) return rc;
    }
    s->initialized = true;
    choose_force_blocks(s);
    return DEM_OK;
}

// Steps per long graph.  Between two graph launches the GPU idles about 8 us (1M pour, lean
// steps of 73 us), so long call bodies replay graphs of this many steps, the rest two-step ones.
constexpr int64_t GRAPH_LONG_STEPS = 16;

// Capture the steps of one graph starting from the current buffer parity as that parity's graph:
// plain steps (no output writes, no diagnostics) or lean steps, GRAPH_LONG_STEPS or two of them.
static int32_t capture_graph(dem_scene* s, bool lean, bool lng) {
    const int start = s->cur;
    hipGraph_t& gr = s->graph[lean][lng][start];
    hipGraphExec_t& ge = s->graph_exec[lean][lng][start];
    if (ge) { hipGraphExecDestroy(ge); ge = nullptr; }
    if (gr) { hipGraphDestroy(gr); gr = nullptr; }
    HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    const int64_t nst = lng ? GRAPH_LONG_STEPS : 2;
    for (int64_t k = 0; k < nst && e == hipSuccess; ++k) {
        if (lean) {
            e = run_lean_step(s);
        } else {
            e = run_contact_detection(s, false, false);
            if (e == hipSuccess) e = run_force_integrate(s, false);
        }
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s->stream, &g);
    s->cur = start;  // capture only recorded the launches
    if (e != hipSuccess) { if (g) hipGraphDestroy(g); return hip_fail(s, "graph capture", e); }
    if (e2 != hipSuccess) { if (g) hipGraphDestroy(g); return hip_fail(s, "hipStreamEndCapture", e2); }
    gr = g;
    HIPCHK(s, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    s->graph_tpb[lean][lng][start] = force_key(s);
    return DEM_OK;
}

// Replay `count` graphs of plain or lean steps, long or two-step ones, from the current buffer
// parity (where they also end), capturing that parity's graph first if it is missing or was
// captured with another force block size.
static int32_t replay_graphs(dem_scene* s, bool lean, bool lng, int64_t count) {
    if (count <= 0) return DEM_OK;
    const int b = s->cur;
    if (!s->graph_exec[lean][lng][b] || s->graph_tpb[lean][lng][b] != force_key(s)) {
        int32_t rc = capture_graph(s, lean, lng);
        if (rc) return rc;
    }
    for (int64_t k = 0; k < count; ++k) {
        HIPCHK(s, hipGraphLaunch(s->graph_exec[lean][lng][b], s->stream));
        s->steps += lng ? GRAPH_LONG_STEPS : 2;
    }
    return DEM_OK;
}

// Make the next step (buffer parity cur) rebuild the sort, the broad phase and the near lists.
static hipError_t request_rebuild(dem_scene* s) {
    return hipMemsetAsync(s->d.ctr + CT_SKIN + s->cur, 0xff, sizeof(unsigned long long), s->stream);
}

// One full step, eager; `last`: the last step of the call, with its outputs and counters.
static hipError_t full_step(dem_scene* s, bool last) {
    // the diagnostic passes read this step's sort
    if (s->P.skin && s->diagnostics && last) {
        hipError_t e = request_rebuild(s);
        if (e != hipSuccess) return e;
    }
    hipError_t e = run_contact_detection(s, false, last);
    if (e == hipSuccess) e = run_force_integrate(s, last);
    if (e == hipSuccess) s->steps++;
    return e;
}

// Enqueue n steps; call_end: the last of them ends the dem_step call.  With lean steps: a full
// first step when it has to rebuild (the counters read last hold its flag word), lean steps
// (replayed from the long, then the two-step lean graphs of the current buffer parity, an odd
// one eager), and when it ends the call an eager last step that counts and writes the outputs:
// lean, or full when it is also the rebuilding first step or runs the diagnostic passes.
// Otherwise full steps replayed the same way, the call's last step eager.  Graphs are not used
// while timing individual kernels.
static int32_t enqueue_steps(dem_scene* s, int64_t nsteps, bool call_end, bool lean) {
    int64_t t = 0;
    bool rebuild = false;
    if (s->P.skin && s->skin_dirty && nsteps > 0) {
        HIPCHK(s, request_rebuild(s));
        s->skin_dirty = false;
        rebuild = true;
    } else if (s->P.skin) {
        rebuild = s->h_ctr[CT_SKIN + s->cur] != 0ull;
    }
    const bool graphs = s->use_graph && !s->timing;
    const int64_t body_end = call_end ? nsteps - 1 : nsteps;  // the call's last step is full
    if (lean) {
        if (rebuild && t < body_end) {
            HIPCHK(s, full_step(s, false));
            ++t;
        }
        // the body's lean steps of a single scene in the contact-free mode are one launch each,
        // whose run starts by clearing the next step's words
        if (s->force_free && !s->batched && t < body_end) HIPCHK(s, run_lean_free_prologue(s));
        if (graphs) {
            const int64_t nlong = (body_end - t) / GRAPH_LONG_STEPS;
            int32_t rc = replay_graphs(s, true, true, nlong);
            if (rc) return rc;
            t += nlong * GRAPH_LONG_STEPS;
            const int64_t npair = (body_end - t) / 2;
            if ((rc = replay_graphs(s, true, false, npair))) return rc;
            t += 2 * npair;
        }
        for (; t < body_end; ++t) {
            HIPCHK(s, run_lean_step(s));
            s->steps++;
        }
    } else if (graphs) {
        // the call's last step stays eager (it writes the outputs)
        const int64_t nlong = (nsteps - 1) / GRAPH_LONG_STEPS;
        int32_t rc = replay_graphs(s, false, true, nlong);
        if (rc) return rc;
        t += nlong * GRAPH_LONG_STEPS;
        const int64_t npair = (nsteps - 1 - t) / 2;
        if ((rc = replay_graphs(s, false, false, npair))) return rc;
        t += 2 * npair;
    }
    for (; t < nsteps; ++t) {
        const bool last = call_end && t == nsteps - 1;
        if (lean && last && !s->diagnostics && !(rebuild && t == 0)) {
            // a lean last step: if it has to rebuild it fails, and the host replays it as a full one
            HIPCHK(s, run_lean_step(s, true));
            s->steps++;
        } else {
            HIPCHK(s, full_step(s, last));
        }
    }
    return DEM_OK;
}

// Lean-step chunks: between LEAN_CHUNK_MIN and LEAN_CHUNK_MAX steps.  Rebuilds closer together
// than LEAN_CHUNK_MIN steps switch the scene to full steps; a call of full steps that rebuilt at
// most once per 2 LEAN_CHUNK_MIN steps switches it back.
constexpr int64_t LEAN_CHUNK_MIN = 16, LEAN_CHUNK_MAX = 1024;

// The next lean chunk after one that completed: twice as long, unless the displacement bound grew
// since the previous read with no rebuild between them; then half the steps it needs at that rate
// to reach the skin (so a rebuild falls on the first step of a chunk, not inside one), at least 4.
static void next_lean_chunk(dem_scene* s) {
    const int cur = s->cur;
    float D = 0.0f;
    for (int c = 0; c < CTR_COPIES; ++c) {
        const uint32_t bits = (uint32_t)s->h_ctr_raw[c * CTR_STRIDE + CT_DISP + cur];
        float v;
        memcpy(&v, &bits, sizeof(v));
        D = std::max(D, v);
    }
    int64_t chunk = std::min(2 * s->lean_chunk, LEAN_CHUNK_MAX);
    const unsigned long long rb = s->h_ctr[CT_REBUILDS];
    if (rb == s->pred_rebuilds && s->steps > s->pred_step && (double)D > s->pred_disp) {
        const double rate = ((double)D - s->pred_disp) / (double)(s->steps - s->pred_step);
        const double left = 0.5 * (0.99 * s->skin_min - (double)D) / rate;
        chunk = std::max<int64_t>(4, std::min<int64_t>(chunk, (int64_t)std::max(0.0, left)));
    }
    s->lean_chunk = chunk;
    s->pred_disp = (double)D;
    s->pred_step = s->steps;
    s->pred_rebuilds = rb;
}

int32_t dem_step(dem_scene_t* s, int64_t nsteps) {
    if (!s || nsteps < 0) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    int64_t left = nsteps;
    while (left > 0) {
        const int64_t steps0 = s->steps;
        const int cur0 = s->cur;
        const unsigned long long done0 = s->done_count, rebuilds0 = s->h_ctr[CT_REBUILDS];
        const bool lean = s->P.skin && s->lean_chunk > 0;
        const int64_t n = lean ? std::min(left, s->lean_chunk) : left;
        // per-kernel timing: this chunk's launches are kept only if none of its steps failed
        int32_t rc = collect_timing(s);
        if (rc) return rc;
        if ((rc = enqueue_steps(s, n, n == left, lean))) return rc;
        if ((rc = sync_counters(s))) return rc;
        s->done_count = s->h_ctr[CT_DONE];
        if (s->h_ctr[CT_FAIL]) s->nev = 0;
        if (!s->h_ctr[CT_FAIL]) {
            left -= n;
            if (lean) {
                next_lean_chunk(s);
            } else if (s->P.skin && (s->h_ctr[CT_REBUILDS] - rebuilds0) * 2 * LEAN_CHUNK_MIN <=
                                        (unsigned long long)n) {
                s->lean_chunk = LEAN_CHUNK_MIN;
            }
            continue;
        }
        // A step failed: it and the later steps changed nothing.  Resume from the state after the
        // steps that completed.
        const int64_t done = (int64_t)(s->done_count - done0);
        s->steps = steps0 + done;
        s->cur = (done & 1) ? 1 - cur0 : cur0;
        left -= done;
        s->skin_dirty = true;  // the failed steps' broad phases may have rewritten the near lists
        const unsigned long long aborted = s->h_ctr[CT_MAXROW] ? 0ull : s->h_ctr[CT_ABORT];
        HIPCHK(s, run_rows_finish(s, true));
        if (aborted & 2ull) s->force_free = false;  // contacts appeared: the full force kernel
        if (aborted & 1ull) {
            ++s->replays;
            // a lean step had to rebuild: the next chunk starts with a full step there, and is
            // as long as the interval since the previous such rebuild
            const int64_t gap = s->steps - s->last_rebuild_step;
            s->last_rebuild_step = s->steps;
            s->lean_chunk = gap < LEAN_CHUNK_MIN ? 0 : std::min(gap, LEAN_CHUNK_MAX);
        }
        if (aborted) continue;
        // a row overflowed: widen the rows
        if (s->K >= MAX_SLOTS) {
            HIPCHK(s, hipStreamSynchronize(s->stream));
            return fail(s, DEM_ERR_CAPACITY,
                        "more than 256 contacts on one particle (the state is that of the last "
                        "completed step)");
        }
        if ((rc = widen_rows(s, wider_slots(s, s->h_ctr[CT_MAXROW])))) return rc;
    }
    if (s->P.skin && s->h_ctr[CT_NBNEED]) {
        // a near list overflowed (the steps after it rebuilt): widen them for the next call
        const unsigned long long need = s->h_ctr[CT_NBNEED];
        HIPCHK(s, hipMemsetAsync(s->d.ctr + CT_NBNEED, 0, sizeof(unsigned long long), s->stream));
        s->h_ctr[CT_NBNEED] = 0;
        int32_t rc = widen_near(s, need);
        if (rc) return rc;
    }
    if (nsteps > 0) choose_force_blocks(s);
    return DEM_OK;
}

int32_t dem_set_diagnostics(dem_scene_t* s, int32_t on) {
    if (!s) return DEM_ERR_INVALID;
    s->diagnostics = on != 0;
    return DEM_OK;
}

int32_t dem_get_counters(dem_scene_t* s, dem_counters_t* out) {
    if (!s || !out) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    int32_t rc = sync_counters(s);
    if (rc) return rc;
    const unsigned long long* c = s->h_ctr;
    memset(out, 0, sizeof(*out));
    out->step = s->steps;
    out->n = s->n;
    out->hash_table_size = s->der.hash_table_size;
    // counted only by the diagnostic passes (dem_set_diagnostics); -1: not counted
    out->candidate_pairs = s->diag_valid ? (int64_t)c[CT_CAND] : -1;
    if (s->steps == 0) {  // after initialize only: every created contact is a bond
        out->contacts = (int64_t)c[CT_CREATED];
        out->bonded = (int64_t)c[CT_CREATED];
    } else {
        out->contacts = (int64_t)c[CT_ACTIVE];
        out->bonded = (int64_t)c[CT_ACTIVE_BOND];
    }
    out->hertz = out->contacts - out->bonded;
    out->wall_contacts = (int64_t)c[CT_WALL];
    out->created = (int64_t)c[CT_CREATED];
    out->deactivated = (int64_t)c[CT_DEACT];
    out->bonds_broken = (int64_t)c[CT_BROKEN];
    out->bonds_broken_total = (int64_t)c[CT_BROKEN_TOTAL];
    out->fatal_gap = (int64_t)c[CT_FATAL];
    out->overflow = (int64_t)c[CT_OVERFLOW];
    out->examined = s->diag_valid ? (int64_t)c[CT_EXAMINED] : -1;
    out->grid_cells = s->dense ? (int64_t)s->P.sort_keys : 0;
    out->rebuilds = s->P.skin ? (int64_t)c[CT_REBUILDS] : s->steps;
    out->replays = s->replays;
    return DEM_OK;
}

int32_t dem_read_grains(dem_scene_t* s, dem_grain_t* out, int64_t n) {
    if (!s || !out || n != s->n) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    std::vector<double4> p(n), v(n), w(n), q(n), a(n), dw(n);
    const DevState& d = s->d;
    const int c = s->cur;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, to_host(s, p.data(), d.pos[c], n * sizeof(double4)));
    HIPCHK(s, to_host(s, v.data(), d.vel[c], n * sizeof(double4)));
    HIPCHK(s, to_host(s, w.data(), d.omg[c], n * sizeof(double4)));
    HIPCHK(s, to_host(s, q.data(), d.quat[c], n * sizeof(double4)));
    HIPCHK(s, to_host(s, a.data(), d.acc, n * sizeof(double4)));
    HIPCHK(s, to_host(s, dw.data(), d.dw, n * sizeof(double4)));
    for (int64_t i = 0; i < n; ++i) {
        const int64_t u = s->iperm[i];
        dem_grain_t& o = out[i];
        memset(&o, 0, sizeof(o));
        o.id = s->h_id[i]; o.gid = s->h_gid[i]; o.mid = s->h_mid[i];
        o.V = s->h_V[i]; o.m = s->h_m[i]; o.r = s->h_r[i]; o.r0 = s->h_r0[i];
        o.k[0] = p[u].x; o.k[1] = p[u].y; o.k[2] = p[u].z;
        o.v[0] = v[u].x; o.v[1] = v[u].y; o.v[2] = v[u].z;
        o.a[0] = a[u].x; o.a[1] = a[u].y; o.a[2] = a[u].z;
        o.q[0] = q[u].x; o.q[1] = q[u].y; o.q[2] = q[u].z; o.q[3] = q[u].w;
        o.w[0] = w[u].x; o.w[1] = w[u].y; o.w[2] = w[u].z;
        o.dw[0] = dw[u].x; o.dw[1] = dw[u].y; o.dw[2] = dw[u].z;
        o.I[0] = s->h_I[i]; o.I[4] = s->h_I[i]; o.I[8] = s->h_I[i];
    }
    return DEM_OK;
}

int32_t dem_write_grains(dem_scene_t* s, const dem_grain_t* in, int64_t n) {
    if (!s || !in || n != s->n) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    std::vector<double4> p(n), v(n), w(n), q(n), a(n), dw(n);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t u = s->iperm[i];
        const dem_grain_t& g = in[i];
        p[u] = make_double4(g.k[0], g.k[1], g.k[2], s->h_r[i]);
        v[u] = make_double4(g.v[0], g.v[1], g.v[2], s->h_m[i]);
        w[u] = make_double4(g.w[0], g.w[1], g.w[2], s->h_I[i]);
        q[u] = make_double4(g.q[0], g.q[1], g.q[2], g.q[3]);
        a[u] = make_double4(g.a[0], g.a[1], g.a[2], 0.0);
        dw[u] = make_double4(g.dw[0], g.dw[1], g.dw[2], 0.0);
    }
    DevState& d = s->d;
    const int c = s->cur;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, to_device(s, d.pos[c], p.data(), n * sizeof(double4)));
    HIPCHK(s, to_device(s, d.vel[c], v.data(), n * sizeof(double4)));
    HIPCHK(s, to_device(s, d.omg[c], w.data(), n * sizeof(double4)));
    HIPCHK(s, to_device(s, d.quat[c], q.data(), n * sizeof(double4)));
    HIPCHK(s, to_device(s, d.acc, a.data(), n * sizeof(double4)));
    HIPCHK(s, to_device(s, d.dw, dw.data(), n * sizeof(double4)));
    s->skin_dirty = true;  // the positions moved outside a step
    return DEM_OK;
}

int32_t dem_read_contacts(dem_scene_t* s, dem_contact_t* out, uint8_t* bonded, int64_t cap,
                          int64_t* count) {
    if (!s || !count || cap < 0 || (cap > 0 && !out)) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    dem_contact_t* dout = nullptr;
    uint8_t* dbd = nullptr;
    int64_t total = 0;
    HIPCHK(s, run_read_contacts(s, &dout, &dbd, &total));
    HIPCHK(s, hipStreamSynchronize(s->stream));  // the records are written on the scene's stream
    const int64_t w = std::min(cap, total);
    hipError_t e = hipSuccess;
    if (w > 0) {
        // records come out in internal row order: sort by (i, j) in file indices
        std::vector<dem_contact_t> all(total);
        std::vector<uint8_t> bd(total);
        e = to_host(s, all.data(), dout, total * sizeof(dem_contact_t));
        if (e == hipSuccess) e = to_host(s, bd.data(), dbd, total);
        if (e == hipSuccess) {
            std::vector<int64_t> ord(total);
            for (int64_t k = 0; k < total; ++k) ord[k] = k;
            std::sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) {
                return all[x].i != all[y].i ? all[x].i < all[y].i : all[x].j < all[y].j;
            });
            for (int64_t k = 0; k < w; ++k) {
                out[k] = all[ord[k]];
                if (bonded) bonded[k] = bd[ord[k]];
            }
        }
    }
    hipFree(dout);
    hipFree(dbd);
    if (e != hipSuccess) return hip_fail(s, "hipMemcpy(contacts)", e);
    *count = total;
    return DEM_OK;
}

int32_t dem_write_contacts(dem_scene_t* s, const dem_contact_t* in, const uint8_t* bonded,
                           int64_t m) {
    if (!s || m < 0 || (m > 0 && !in)) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    const int64_t n = s->n;
    const int K = s->K;
    struct Ent { uint32_t row, partner; int64_t src; bool is_i; };
    std::vector<Ent> ents;
    ents.reserve(2 * m);
    for (int64_t c = 0; c < m; ++c) {
        const int64_t i = in[c].i, j = in[c].j;
        if (i < 1 || j < 1 || i > n || j > n || i >= j)
            return fail(s, DEM_ERR_INVALID, "contacts need 1 <= i < j <= n");
        const uint32_t ui = s->iperm[i - 1], uj = s->iperm[j - 1];
        ents.push_back({ui, uj, c, true});
        ents.push_back({uj, ui, c, false});
    }
    std::sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
        return a.row != b.row ? a.row < b.row : a.partner < b.partner;
    });
    std::vector<uint32_t> rc(n, 0);
    std::vector<uint64_t> dst(ents.size());
    std::vector<uint32_t> word(ents.size());
    std::vector<double> state(ents.size() * S_NCOMP);
    for (size_t k = 0; k < ents.size(); ++k) {
        const Ent& en = ents[k];
        if (k > 0 && ents[k - 1].row == en.row && ents[k - 1].partner == en.partner)
            return fail(s, DEM_ERR_DUPLICATE, "duplicate contact pair");
        const uint32_t slot = rc[en.row]++;
        if ((int)slot >= K) return fail(s, DEM_ERR_CAPACITY, "contact slots per particle exceeded");
        dst[k] = (uint64_t)slot * n + en.row;
        const bool bd = bonded && bonded[en.src];
        word[k] = en.partner | (bd ? BONDED : 0u) | (en.is_i ? IS_I : 0u);
        const dem_contact_t& c = in[en.src];
        double* sv = &state[k * S_NCOMP];
        for (int a = 0; a < 3; ++a) {
            sv[S_KX + a] = c.k[a]; sv[S_FX + a] = c.F[a]; sv[S_TIX + a] = c.ti[a];
            sv[S_TJX + a] = c.tj[a]; sv[S_SX + a] = c.s[a];
        }
    }
    DevState& d = s->d;
    uint64_t* ddst = nullptr;
    uint32_t* dword = nullptr;
    double* dstate = nullptr;
    const size_t me = std::max<size_t>(ents.size(), 1);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMalloc(&ddst, me * sizeof(uint64_t)));
    HIPCHK(s, hipMalloc(&dword, me * sizeof(uint32_t)));
    HIPCHK(s, hipMalloc(&dstate, me * S_NCOMP * sizeof(double)));
    hipError_t e = to_device(s, ddst, dst.data(), ents.size() * sizeof(uint64_t));
    if (e == hipSuccess) e = to_device(s, dword, word.data(), ents.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = to_device(s, dstate, state.data(), ents.size() * S_NCOMP * sizeof(double));
    if (e == hipSuccess) e = to_device(s, d.rcnt[0], rc.data(), n * sizeof(uint32_t));
    // on the scene's stream, ahead of k_write_rows and the next merge
    if (e == hipSuccess) e = hipMemsetAsync(d.rcnt[1], 0, n * sizeof(uint32_t), s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d.sel, 0, n, s->stream);
    // the merges' per-group values and the contact-free mode assumed the old rows: no group is left
    // out until its value is renewed, and the force kernel reads rows again
    if (e == hipSuccess && d.gskip) e = hipMemsetAsync(d.gskip, 0, ((n + 63) / 64) * sizeof(uint32_t), s->stream);
    s->force_free = false;
    s->skin_dirty = true;  // the near lists' bond marks are those of the old rows: rebuild them
    if (e == hipSuccess) e = run_write_contacts(s, ddst, dword, dstate, (int64_t)ents.size());
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    hipFree(ddst); hipFree(dword); hipFree(dstate);
    if (e != hipSuccess) return hip_fail(s, "dem_write_contacts", e);
    return DEM_OK;
}

int32_t dem_read_wall_shear(dem_scene_t* s, double* out, int64_t n) {
    if (!s || !out || n != s->n) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    const int nw = s->nwall;
    if (nw == 0) return DEM_OK;
    std::vector<double4> ws((size_t)nw * n);
    std::vector<uint32_t> act(n);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, to_host(s, ws.data(), s->d.wshear, ws.size() * sizeof(double4)));
    HIPCHK(s, to_host(s, act.data(), s->d.wact, n * sizeof(uint32_t)));
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < nw; ++j) {
            const int64_t u = s->iperm[i];
            const bool on = (act[u] >> j) & 1u;
            const double4 v = ws[(size_t)j * n + u];
            double* o = &out[(i * nw + j) * 3];
            o[0] = on ? v.x : 0.0; o[1] = on ? v.y : 0.0; o[2] = on ? v.z : 0.0;
        }
    return DEM_OK;
}

int32_t dem_write_wall_shear(dem_scene_t* s, const double* in, int64_t n) {
    if (!s || !in || n != s->n) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    const int nw = s->nwall;
    if (nw == 0) return DEM_OK;
    std::vector<double4> ws((size_t)nw * n);
    std::vector<uint32_t> act(n, 0u);
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < nw; ++j) {
            const int64_t u = s->iperm[i];
            const double* v = &in[(i * nw + j) * 3];
            ws[(size_t)j * n + u] = make_double4(v[0], v[1], v[2], 0.0);
            if (v[0] != 0.0 || v[1] != 0.0 || v[2] != 0.0) act[u] |= 1u << j;
        }
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, to_device(s, s->d.wshear, ws.data(), ws.size() * sizeof(double4)));
    HIPCHK(s, to_device(s, s->d.wact, act.data(), n * sizeof(uint32_t)));
    return DEM_OK;
}

int32_t dem_read_forces(dem_scene_t* s, double* F, double* T, int64_t n) {
    if (!s || n != s->n) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    std::vector<double4> f(n), t(n);
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, to_host(s, f.data(), s->d.ftot, n * sizeof(double4)));
    HIPCHK(s, to_host(s, t.data(), s->d.ttot, n * sizeof(double4)));
    for (int64_t i = 0; i < n; ++i) {
        const int64_t u = s->iperm[i];
        if (F) { F[3 * i] = f[u].x; F[3 * i + 1] = f[u].y; F[3 * i + 2] = f[u].z; }
        if (T) { T[3 * i] = t[u].x; T[3 * i + 1] = t[u].y; T[3 * i + 2] = t[u].z; }
    }
    return DEM_OK;
}

int32_t dem_debug_cells(dem_scene_t* s, int32_t* kid, uint32_t* hid, double* kcenter,
                        uint32_t* nbr) {
    if (!s) return DEM_ERR_INVALID;
    if (s->batched) return fail(s, DEM_ERR_STATE, "not available for batches");
    HIPCHK(s, hipSetDevice(s->device));
    const int64_t n = s->n;
    int32_t* dk = nullptr; uint32_t* dh = nullptr; double* dc = nullptr; uint32_t* dn = nullptr;
    HIPCHK(s, hipMalloc(&dk, 3 * n * sizeof(int32_t)));
    HIPCHK(s, hipMalloc(&dh, n * sizeof(uint32_t)));
    HIPCHK(s, hipMalloc(&dc, 3 * n * sizeof(double)));
    HIPCHK(s, hipMalloc(&dn, 8 * n * sizeof(uint32_t)));
    hipError_t e = run_debug_cells(s, dk, dh, dc, dn);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    std::vector<int32_t> k3(3 * n);
    std::vector<uint32_t> h1(n), n8(8 * n);
    std::vector<double> c3(3 * n);
    if (e == hipSuccess) e = to_host(s, k3.data(), dk, 3 * n * sizeof(int32_t));
    if (e == hipSuccess) e = to_host(s, h1.data(), dh, n * sizeof(uint32_t));
    if (e == hipSuccess) e = to_host(s, c3.data(), dc, 3 * n * sizeof(double));
    if (e == hipSuccess) e = to_host(s, n8.data(), dn, 8 * n * sizeof(uint32_t));
    hipFree(dk); hipFree(dh); hipFree(dc); hipFree(dn);
    if (e == hipSuccess)
        for (int64_t i = 0; i < n; ++i) {  // back to file order
            const int64_t u = s->iperm[i];
            for (int a = 0; a < 3; ++a) {
                if (kid) kid[3 * i + a] = k3[3 * u + a];
                if (kcenter) kcenter[3 * i + a] = c3[3 * u + a];
            }
            if (hid) hid[i] = h1[u];
            if (nbr) for (int a = 0; a < 8; ++a) nbr[8 * i + a] = n8[8 * u + a];
        }
    if (e != hipSuccess) return hip_fail(s, "dem_debug_cells", e);
    return DEM_OK;
}

int32_t dem_debug_hash_table(dem_scene_t* s, uint32_t* counts, uint32_t* start, uint32_t* pid) {
    if (!s) return DEM_ERR_INVALID;
    if (s->batched) return fail(s, DEM_ERR_STATE, "not available for batches");
    HIPCHK(s, hipSetDevice(s->device));
    HIPCHK(s, run_debug_sort(s));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    const int64_t H = s->der.hash_table_size, n = s->n;
    std::vector<uint32_t> bs(H + 1), si(n);
    HIPCHK(s, to_host(s, bs.data(), s->d.bstart, (H + 1) * sizeof(uint32_t)));
    HIPCHK(s, to_host(s, si.data(), s->d.sidx, n * sizeof(uint32_t)));
    for (int64_t b = 0; b < H; ++b) {
        if (counts) counts[b] = bs[b + 1] - bs[b];
        if (start) start[b] = bs[b];
    }
    if (pid)
        for (int64_t k = 0; k < n; ++k) pid[k] = s->perm[si[k]] + 1u;
    return DEM_OK;
}

int32_t dem_debug_cp_list(dem_scene_t* s, int32_t* pairs, int64_t cap, int64_t* count) {
    if (!s || !count || cap < 0 || (cap > 0 && !pairs)) return DEM_ERR_INVALID;
    if (s->batched) return fail(s, DEM_ERR_STATE, "not available for batches");
    HIPCHK(s, hipSetDevice(s->device));
    int32_t* dp = nullptr;
    int64_t total = 0;
    HIPCHK(s, run_debug_cp_list(s, &dp, &total));
    const int64_t w = std::min(cap, total);
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess && w > 0) e = to_host(s, pairs, dp, 2 * w * sizeof(int32_t));
    hipFree(dp);
    if (e != hipSuccess) return hip_fail(s, "dem_debug_cp_list", e);
    *count = total;
    return DEM_OK;
}

int32_t dem_debug_scan(const uint32_t* in, uint32_t* out, int64_t n, uint64_t* total) {
    if (n < 0 || (n > 0 && (!in || !out))) return DEM_ERR_INVALID;
    if (n == 0) { if (total) *total = 0; return DEM_OK; }
    ScanWorkspace ws;
    ws.max_tiles = scan_tiles(n);
    uint32_t *din = nullptr, *dout = nullptr;
    uint64_t* dtot = nullptr;
    hipError_t e = dalloc(&ws.status, ws.max_tiles);
    if (e == hipSuccess) e = dalloc(&ws.counters, 2);
    if (e == hipSuccess) e = dalloc(&din, n);
    if (e == hipSuccess) e = dalloc(&dout, n);
    if (e == hipSuccess) e = dalloc(&dtot, 1);
    if (e == hipSuccess) e = hipMemcpy(din, in, n * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_scan(ws, din, dout, n, false, dtot, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, dout, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
    uint64_t t = 0;
    if (e == hipSuccess) e = hipMemcpy(&t, dtot, sizeof(uint64_t), hipMemcpyDeviceToHost);
    hipFree(ws.status); hipFree(ws.counters); hipFree(din); hipFree(dout); hipFree(dtot);
    if (e != hipSuccess) { g_create_err = std::string("dem_debug_scan: ") + hipGetErrorString(e); return DEM_ERR_HIP; }
    if (total) *total = t;
    return DEM_OK;
}

int32_t dem_set_memory_cache(int64_t max_bytes) {
    if (max_bytes < 0) return DEM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache_cap = (size_t)max_bytes;
    cache_trim(g_cache_cap);
    return DEM_OK;
}

int32_t dem_debug_copy(int64_t nbytes, int32_t repeats, double* ms_per_copy) {
    if (nbytes <= 0 || repeats <= 0) return DEM_ERR_INVALID;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, nbytes);
    if (e == hipSuccess) e = hipMalloc(&b, nbytes);
    if (e == hipSuccess) e = hipMemset(a, 1, nbytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = run_copy(a, b, nbytes, nullptr);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int r = 0; r < repeats && e == hipSuccess; ++r) e = run_copy(a, b, nbytes, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (a) hipFree(a);
    if (b) hipFree(b);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (e != hipSuccess) { g_create_err = std::string("dem_debug_copy: ") + hipGetErrorString(e); return DEM_ERR_HIP; }
    if (ms_per_copy) *ms_per_copy = ms / repeats;
    return DEM_OK;
}

int32_t dem_set_timing(dem_scene_t* s, int32_t on) {
    if (!s) return DEM_ERR_INVALID;
    HIPCHK(s, hipSetDevice(s->device));
    if (on && s->ev.empty()) {
        const int cap = 4096;
        s->ev.assign(2 * cap, nullptr);
        s->ev_kernel.assign(cap, 0);
        for (auto& e : s->ev) HIPCHK(s, hipEventCreate(&e));
    }
    if (!on && s->nev) { int32_t rc = timing_flush(s); if (rc) return rc; }
    s->timing = on != 0;
    s->nev = 0;
    for (int k = 0; k < DEM_NUM_KERNELS; ++k) { s->kms[k] = 0.0; s->klaunch[k] = 0; }
    return DEM_OK;
}

int32_t dem_get_kernel_times(dem_scene_t* s, dem_kernel_times_t* out) {
    if (!s || !out) return DEM_ERR_INVALID;
    int32_t rc = collect_timing(s);
    if (rc) return rc;
    for (int k = 0; k < DEM_NUM_KERNELS; ++k) { out->ms[k] = s->kms[k]; out->launches[k] = s->klaunch[k]; }
    return DEM_OK;
}

}  // extern "C"
