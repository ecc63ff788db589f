// dem_internal.hpp -- scene state and kernel launchers shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/demhip.h"
#include "dem_device.hpp"

namespace dem {

constexpr int MAX_WALLS = 16;
constexpr int MAX_SLOTS = 256;  // contact slots per particle the rows may widen to
constexpr uint32_t IDX_MASK = 0x1fffffffu;  // partner index bits of a row entry
constexpr uint32_t IS_I = 1u << 29;         // the row's particle is the contact's i (the lower
                                            // file index): the canonical orientation
constexpr uint32_t BONDED = 1u << 30;       // entry is an EBPM bond
constexpr uint32_t BROKEN = 1u << 31;       // bond broke this step; dropped by the next merge
// A near-list entry is partner | IS_I, with NB_BOND when the partner was the particle's intact
// bond at the rebuild: the rebuilding merge sets it, the force kernel clears it when the bond
// breaks, and lean merges leave such entries untested (a bond persists whether or not its grains
// touch, so testing it changes no row).
constexpr uint32_t NB_BOND = 1u << 30;
// Set in a particle's near-list count (nbc) by the force kernel when one of its bonds breaks: the
// next lean merge clears the particle's NB_BOND marks (its lists are then all tested until the
// next rebuild marks them again), since a broken bond's pair may touch as a Hertz contact.
constexpr uint32_t NB_BROKE = 1u << 31;

// Row-entry state components (each an array of contact_slots * n doubles per buffer).
enum StateComp { S_KX = 0, S_KY, S_KZ, S_FX, S_FY, S_FZ, S_TIX, S_TIY, S_TIZ,
                 S_TJX, S_TJY, S_TJZ, S_SX, S_SY, S_SZ, S_NCOMP };

enum CounterIdx {
    CT_CAND = 0,     // candidate pairs i<j over the 8-bucket windows (cp_list length)
    CT_TOUCH,        // touching pairs found by the broad phase
    CT_ENTRIES,      // row entries after merge (both roles)
    CT_CREATED,      // contacts created
    CT_DEACT,        // Hertz contacts deactivated
    CT_BROKEN,       // bonds broken
    CT_WALL,         // wall contacts
    CT_FATAL,        // gap > 0 or negative damping coefficient evaluations
    CT_ACTIVE,       // active contacts after the step (counted once per pair)
    CT_ACTIVE_BOND,  // active bonds after the step
    CT_EXAMINED,     // candidate pairs examined by the engine's broad phase
    CT_NSTEP,        // counters below are reset at the start of every step
    CT_BROKEN_TOTAL = 12,  // cumulative
    CT_OVERFLOW = 13,      // a row overflowed in the failed step (cleared by the recovery)
    CT_DONE = 14,          // cumulative: steps whose force kernel ran (copy 0 only)
    CT_FAIL = 15,          // a row overflowed: every later step of the call does nothing (copy 0)
    CT_MAXROW = 16,        // the longest row an overflowing particle needed (copy 0, maximum)
    CT_REBUILDS = 17,      // cumulative: completed steps that rebuilt the sort and near lists (copy 0)
    CT_NBNEED = 18,        // the longest near list a particle needed when one overflowed (copy 0)
    CT_ABORT = 19,         // a lean step found that it had to rebuild: it and the later steps of the
                           // call did nothing (copy 0; CT_FAIL is set with it)
    CT_SKIN = 20,          // CT_SKIN + c: the step of buffer parity c rebuilds (copy 0; two words)
    CT_NCOUNTERS = 22,     // the host sums copies of the counters below this index
    CT_DISP = 22,          // CT_DISP + c: FP32 bits of a bound on every particle's displacement since
                           // the last rebuild, for the step of parity c (the maximum over all copies)
    CT_TICKET = 24,        // a one-kernel lean step: its blocks done so far, per class of blocks
                           // (copy c: blocks b % CTR_COPIES == c), and the classes done (copy 0,
                           // CT_TICKET + 2)
    CT_MDONE = 25          // the same for its merge blocks (classes done: copy 0, CT_MDONE + 2)
};

// Counters are kept in CTR_COPIES copies, 256 B apart, and a block adds to copy blockIdx % CTR_COPIES:
// device-scope atomics on one address serialise, and one copy per wave of a million-particle
// launch would cost more than the launch.  The host sums the copies.
constexpr int CTR_COPIES = 64;
constexpr int CTR_STRIDE = 32;
constexpr int CTR_WORDS = CTR_COPIES * CTR_STRIDE;

__device__ __forceinline__ unsigned long long* ctr_at(unsigned long long* ctr, int k) {
    return ctr + (size_t)(blockIdx.x % CTR_COPIES) * CTR_STRIDE + k;
}

// A step is a transaction: the merge writes changed rows into the other buffer and marks them
// pending (ROW_PENDING | committed buffer) in the row selector, and the force kernel commits them.
// A row that does not fit fails the step: the force kernel of that step and of every later step
// of the call then changes nothing (the merge only ever writes a particle's uncommitted buffer),
// and the host drops the pending rows, widens the rows and replays the steps that did not run.
constexpr uint8_t ROW_PENDING = 2;

// The failure flag as the force kernel reads it: written by an earlier kernel of the stream, so a
// plain load sees it (a kernel starts with the caches invalidated).
__device__ __forceinline__ bool step_failed(const unsigned long long* ctr) {
    return ctr[CT_FAIL] != 0ull;
}
__device__ __forceinline__ void fail_step(unsigned long long* ctr, uint32_t need) {
    atomicAdd(ctr_at(ctr, CT_OVERFLOW), 1ull);
    atomicMax(&ctr[CT_MAXROW], (unsigned long long)need);
    atomicOr(&ctr[CT_FAIL], 1ull);
}

// Near lists (a Verlet skin).  A step that rebuilds (sorts, runs the broad phase) also records,
// for every particle, the partners within r_i + r_j + 2 delta of it, and the force kernel keeps
// each particle's position at that step.  While no particle has moved delta or more from it, a
// pair that touches is in both partners' near lists (|x_i - x_j| can only have shrunk by less
// than 2 delta), so the following steps skip the sort and the broad phase and test only the
// near lists, with the broad phase's exact test: the touching set is the same.  The flag word
// CT_SKIN + c says whether the step of buffer parity c rebuilds; the step of parity c clears
// the other parity's word in its first kernel, and its broad phase (when a near list
// overflows) and force kernel (when a particle has moved too far) set it.
//
// A lean step is the merge and the force kernel alone, the merge testing the near lists; its
// merge clears the other parity's words itself.  A lean step whose flag word asks for a rebuild
// fails the step instead (CT_ABORT with CT_FAIL): it and every later step of the call change
// nothing, and the host replays from it with a full step.
__device__ __forceinline__ bool skin_rebuilds(const unsigned long long* ctr, int cur) {
    return ctr[CT_SKIN + cur] != 0ull;
}
__device__ __forceinline__ void skin_request(unsigned long long* ctr, int cur) {
    __hip_atomic_store(&ctr[CT_SKIN + (1 - cur)], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// A particle whose nearest near-list partner was a gap g >= 0 away at the rebuild (gmin) touches
// none of them while g > 2 D, D bounding every particle's displacement since then: the gap of a
// pair shrinks by at most the sum of its grains' displacements.  The force kernel keeps D for the
// next step as the maximum of per-wave maxima over CTR_COPIES copies (non-negative FP32 bits
// order like unsigned integers).

struct SurfaceP {
    double mu, e, beta;          // friction, restitution, log(e)/sqrt(log(e)^2 + pi^2)
    double rho, E, nu;           // bond radius ratio, modulus, Poisson ratio
    double sc, st, ss;           // bond strengths
};

struct WallP { double n[3]; double d; int32_t mid; int32_t pad; };

// Everything that may differ between the scenes of a batch.
struct SceneP {
    HashGrid grid;      // the reference's hashed cell grid (utils.jl:1-16)
    DenseGrid dense;    // dense grid of the per-step sort/broad phase, cells >= 2 max r
                        // (dense.cells == 0: unused)
    DenseGrid dense_init;    // dense grid of the bond pass, cells >= the bond reach
    DenseGrid order;    // host only: the grid of cells >= 2 max r whose cells order the particles
                        // in memory at creation (the same with and without near lists)
    uint32_t cell_base;      // first global sort key of this scene's dense cells
    uint32_t cell_base_init; // the same for dense_init
    int32_t nwall;
    int32_t bond_init;  // scene takes part in dem_initialize
    int32_t ref_check;  // touching pairs must be checked against the reference window rule
    double skin;        // near-list skin delta (the per-step dense grid's cells are >= 2 max r + 2 delta)
    double bond_mult;   // particle_contact_radius_multiplier (r0 = r * bond_mult)
    double dt, tol, damping;
    double g[3];
    double matE[2], matNu[2];
    // effective Young's and shear moduli Y*, G* of a particle against material m (kernels.jl:262-263,
    // 371-372): per-scene constants the reference evaluates per contact, evaluated once here
    double ysm[2], gsm[2];
    SurfaceP surf[2][2];
    WallP walls[MAX_WALLS];
    // particle-particle Hertz-Mindlin constants in FP32 (mixed precision)
    float ys_pp, gs_pp, beta_pp, mu_pp;
};

// Everything a kernel needs by value.  Single-scene kernels read scene 0 from the kernel
// arguments (sc); batched kernels look a particle's scene up in the device table seg[sid[p]].
struct Params {
    int64_t n;
    int32_t K;            // contact slots per particle
    int32_t nseg;         // scenes (1 for a single scene)
    int32_t skin;         // near lists in use (dense grid only)
    int32_t KN;           // near-list slots per particle
    uint32_t sort_keys;   // keys of the per-step sort (dense cells of all scenes, or H)
    uint32_t sort_keys_init;  // keys of the bond-pass sort
    SceneP sc;
    const SceneP* seg;
    const uint16_t* sid;
    unsigned long long* seg_broken;  // cumulative bond breaks per scene
};

template <bool SEG>
__device__ __forceinline__ const SceneP& scene_of(const Params& P, uint32_t p) {
    if constexpr (SEG) return P.seg[P.sid[p]];
    else return P.sc;
}

// Copies of each tile total of a counting sort's histogram: block b of the histogram pass adds
// into copy b % TAGG_COPIES, which on a part that deals blocks round-robin to its 8 XCDs keeps
// blocks of different XCDs off one address.
constexpr int TAGG_COPIES = 8;

struct ScanWorkspace {
    uint64_t* status = nullptr;
    uint32_t* counters = nullptr;
    uint32_t* tagg = nullptr;  // [tile][TAGG_COPIES] histogram tile totals (zero between sorts)
    int64_t max_tiles = 0;
};

int64_t scan_tiles(int64_t n);
int64_t scan_tiles_max(int64_t n);  // the most tiles any scan of at most n values takes
int scan_tile_shift(int64_t n);     // log2 of the tile length a scan of n values uses
hipError_t launch_scan(ScanWorkspace& ws, uint32_t* in, uint32_t* out, int64_t n, bool zero_input,
                       uint64_t* total, hipStream_t stream);
// Exclusive scan of n counts whose per-tile totals are in ws.tagg; clears the counts.
// gate: when not null, the scan does nothing unless *gate is nonzero (a step that reuses the
// near lists skips its sort).
hipError_t launch_tile_scan(ScanWorkspace& ws, uint32_t* in, uint32_t* out, int64_t n,
                            hipStream_t stream, const unsigned long long* gate = nullptr);

// Device arrays of one scene.
struct DevState {
    // particle state, double-buffered (index cur / 1-cur)
    double4* pos[2] = {nullptr, nullptr};   // x, y, z, r
    double4* vel[2] = {nullptr, nullptr};   // vx, vy, vz, m
    double4* omg[2] = {nullptr, nullptr};   // wx, wy, wz, I (isotropic inertia)
    double4* quat[2] = {nullptr, nullptr};  // w, x, y, z
    double4* acc = nullptr;                 // a (output of update!)
    double4* dw = nullptr;                  // dω (output of update!)
    double4* ftot = nullptr;                // force of the last step
    double4* ttot = nullptr;                // moment of the last step
    double* r0 = nullptr;
    uint32_t* wact = nullptr;               // bit w: particle touched wall w last step
    double4* wshear = nullptr;              // [nwall][n] shear displacement
    // cell hashing / counting sort.  The sort walks the particles in the previous step's sorted
    // order (sidxb[1 - cur]) and writes this step's order to sidxb[cur]; hid, rank and x32 are
    // indexed by the previous sorted slot.
    uint32_t* hid = nullptr;     // 0-based bucket
    uint32_t* rank = nullptr;    // slot within the bucket
    float4* x32 = nullptr;       // FP32 position and radius
    uint32_t* bcnt = nullptr;    // [H + 1] bucket counts (kept zero between steps)
    uint32_t* bstart = nullptr;  // [H + 1] exclusive bucket starts
    uint32_t* sidxb[2] = {nullptr, nullptr};
    uint32_t* sidx = nullptr;    // sorted slot -> particle (sidxb[cur] after a sort)
    uint32_t* skey = nullptr;    // sorted slot -> bucket
    uint32_t* eid = nullptr;     // internal index -> file index
    uint32_t* sfid = nullptr;    // sorted slot -> file index (reference-grid sorts only)
    double4* spos = nullptr;     // sorted copy of pos (reference-grid broad phases only)
    float4* spos32 = nullptr;    // sorted FP32 copy (broad-phase pre-test)
    // broad phase output
    uint32_t* tcnt = nullptr;    // [n]
    uint32_t* touch = nullptr;   // [K][n]
    // near lists: partner | IS_I words of every pair within r_i + r_j + 2 delta at the last
    // rebuild, and each particle's position then (FP32)
    uint32_t* nbc = nullptr;     // [n] entries (may exceed KN after an overflow)
    uint32_t* nb = nullptr;      // [KN][n]
    float4* xref = nullptr;      // [n]
    uint32_t* gmin = nullptr;    // [n] FP32 bits: the smallest near-list gap at the rebuild (>= 0)
    // [ceil(n / 64)] FP32 bits per group of 64 particles (a wave's): while none of them has a row
    // entry or a touching partner, the smallest of their gmin, else 0; written by every merge
    // wave that runs, so a lean merge leaves out a group whose value exceeds 2 D
    uint32_t* gskip = nullptr;
    // contact rows, double-buffered per particle (sel[p])
    uint8_t* sel = nullptr;
    uint32_t* rcnt[2] = {nullptr, nullptr};
    uint32_t* pq[2] = {nullptr, nullptr};   // [K][n] partner | flags
    double* st[2] = {nullptr, nullptr};     // [S_NCOMP][K][n]
    // counters
    unsigned long long* ctr = nullptr;
    // scan workspace and scratch
    ScanWorkspace scan;
    uint32_t* scratch_u32 = nullptr;  // [n + 1]
    uint32_t* scratch2_u32 = nullptr; // [n + 1]
    uint64_t* total = nullptr;
    // batched scenes
    SceneP* seg = nullptr;
    uint16_t* sid = nullptr;
    unsigned long long* seg_broken = nullptr;
    // every array above lives in one of two allocations: the contact rows (touch, pq, st), which
    // change size when rows are widened, and everything else
    void* arena = nullptr;
    void* rows_arena = nullptr;
    void* nb_arena = nullptr;  // the near lists nb, which change size when they widen
    size_t arena_bytes = 0, rows_arena_bytes = 0, nb_arena_bytes = 0;
};

}  // namespace dem

struct dem_scene {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0;
    int32_t nwall = 0;
    int32_t precision = 0;
    int32_t K = 0;
    int cur = 0;
    bool initialized = false;
    bool slot_owned = true;  // broad phase owned by sorted slot (needs touching => mutual windows)
    bool dense = false;      // per-step sort on the dense cs/2 grid (else the reference's grid)
    bool batched = false;    // several scenes in one set of arrays (per-particle scene ids)
    int32_t nseg = 1;
    // Particles are stored in a spatial order (by cell of the initial position, within each
    // scene): internal index u holds the particle of file index perm[u]; iperm inverts it.
    // Every array on the device uses internal indices; the read, write and debug entry points
    // translate, and orientation (i < j) always compares file indices.
    std::vector<uint32_t> perm, iperm;
    std::vector<int64_t> seg_off;   // [nseg + 1] first particle of each scene
    std::vector<dem::SceneP> seg_host;
    std::vector<dem_config_t> seg_cfg;   // [nseg] each scene's configuration
    std::vector<dem_derived_t> seg_der;  // [nseg] and its derived quantities
    int64_t steps = 0;
    uint64_t broken_total_host = 0;
    dem_config_t cfg{};
    dem_derived_t der{};
    dem::Params P{};
    dem::DevState d{};
    // host copies of static grain fields
    int32_t* h_id = nullptr;
    int32_t* h_gid = nullptr;
    int32_t* h_mid = nullptr;
    double* h_V = nullptr;
    double* h_m = nullptr;
    double* h_r = nullptr;
    double* h_r0 = nullptr;
    double* h_I = nullptr;
    unsigned long long* h_ctr = nullptr;      // pinned, summed over the copies
    unsigned long long* h_ctr_raw = nullptr;  // pinned, CTR_WORDS
    std::string err;
    // steps captured as HIP graphs and replayed for the body of every dem_step call: the
    // particle buffers alternate, so an even number of steps restores the buffer parity.  One
    // graph for each buffer parity cur it starts from, of plain or lean steps, GRAPH_LONG_STEPS
    // long (long bodies) or two steps long (the rest), indexed [lean][long][parity]; each holds
    // the force kernel's block size it was captured with.
    hipGraph_t graph[2][2][2] = {};
    hipGraphExec_t graph_exec[2][2][2] = {};
    int graph_tpb[2][2][2] = {};
    // lean steps run in chunks of at most lean_chunk steps between host checks: twice as long
    // after a chunk without a rebuild, the last interval between rebuilds after one
    int64_t lean_chunk = 64;
    int64_t last_rebuild_step = 0;
    int64_t replays = 0;  // lean steps that had to rebuild (replayed as full steps)
    // the displacement bound D at the last counter read, its step and the rebuild count then: two
    // reads with no rebuild between them give D's growth per step, from which the next chunk
    // ends before D reaches the skin
    double pred_disp = 0.0;
    int64_t pred_step = 0;
    unsigned long long pred_rebuilds = 0;
    double skin_min = 0.0;  // the smallest skin delta of the scene's (batch's) scenes
    // threads per block of the force kernel: 64 when the last counted step had at least one row
    // entry per four particles, else 256; each graph holds the value it was captured with
    int force_tpb = 256;
    // contact-free mode (single scenes): the last counted step left no row entry, so the force
    // kernel runs a variant that reads and walks no rows; a merge that would make a row entry
    // fails its step instead (CT_ABORT bit 2) and the host replays it with this mode off.  The
    // graphs hold the mode they were captured in with the block size (force_key).
    bool force_free = false;
    // diagnostic passes on the last step of a call (candidate pairs of the reference's cp_list,
    // candidates examined by the broad phase): off unless dem_set_diagnostics turns them on
    bool diagnostics = false;
    bool diag_valid = false;  // the last counted step ran them
    // steps the device has completed (its CT_DONE counter) as of the last counter read
    unsigned long long done_count = 0;
    // the near lists may not hold the current positions' partners (scene creation, a host write,
    // a replayed step): the next step rebuilds
    bool skin_dirty = true;
    bool use_graph = true;
    // timing: an event pair around every launch, collected when the pool fills or on request
    bool timing = false;
    std::vector<hipEvent_t> ev;   // 2 * capacity
    std::vector<int> ev_kernel;   // capacity
    int nev = 0;
    double kms[DEM_NUM_KERNELS] = {};
    int64_t klaunch[DEM_NUM_KERNELS] = {};
};

namespace dem {
// kernel launchers (kernels.hip)
enum KernelId { K_HASH = 0, K_SCAN, K_SCATTER, K_BROAD, K_MERGE, K_FORCE, K_INIT, K_OTHER };
// last: the last step of a call, which counts the per-step counters (and runs the diagnostic
// passes when the scene has them on)
hipError_t run_contact_detection(dem_scene* s, bool init, bool last);
hipError_t run_force_integrate(dem_scene* s, bool write_forces);
// one lean step: the merge over the near lists (or the abort) and the force kernel; last: the
// last step of a call, which counts the per-step counters and writes the outputs
hipError_t run_lean_step(dem_scene* s, bool last = false);
hipError_t run_lean_free_prologue(dem_scene* s);
// after a bond pass or a failed step: commit the pending rows (rollback = false), or drop them,
// clear the touching counts and the failure flags (rollback = true)
hipError_t run_rows_finish(dem_scene* s, bool rollback);
hipError_t run_debug_cells(dem_scene* s, int32_t* kid, uint32_t* hid, double* kc, uint32_t* nbr);
hipError_t run_debug_sort(dem_scene* s);
hipError_t run_debug_cp_list(dem_scene* s, int32_t** pairs_dev, int64_t* count);
hipError_t run_read_contacts(dem_scene* s, dem_contact_t** out_dev, uint8_t** bonded_dev,
                             int64_t* count);
hipError_t run_write_contacts(dem_scene* s, const uint64_t* dst_dev, const uint32_t* word_dev,
                              const double* state_dev, int64_t m);
// the message dem_last_error(NULL) returns (entry points without a scene handle)
void set_last_error_global(const std::string& msg);
void timing_begin(dem_scene* s, int kernel);
void timing_end(dem_scene* s);
int32_t timing_flush(dem_scene* s);  // abi.hip: collect elapsed times of recorded pairs
hipError_t run_copy(const void* src, void* dst, int64_t nbytes, hipStream_t stream);
hipError_t run_adopt(dem_scene* dst, int64_t dst_off, const dem_scene* src, int64_t src_off,
                     int64_t n, int32_t nwall, uint32_t* overflow_dev);
// initial state of the particles of file index >= gbase from device copies of their input rows,
// radii and inertias; r0, sort order and scene ids of all particles
hipError_t run_init_state(dem_scene* s, const dem_grain_in_t* rows, const double* r,
                          const double* I, const double* r0f, const int64_t* seg_off, int64_t gbase);
}  // namespace dem
