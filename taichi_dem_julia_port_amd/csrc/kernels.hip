// kernels.hip -- the DEM step as gfx950 kernels.
//
// One step of the reference (simulate!, src/executor.jl:162-172) becomes six launches with no
// host synchronisation:
//   k_hash     cell + Morton bucket per particle, bucket histogram (kernels.jl:7-13, utils.jl:1-2)
//   scan       exclusive bucket starts (accumulate!, executor.jl:50)
//   k_scatter  counting-sort scatter into a sorted mirror of positions (kernels.jl:15-22)
//   k_broad    per-particle walk over the 8 window buckets (kernels.jl:24-58) that keeps only
//              touching pairs i<j (the test of kernels.jl:103 / :72) and appends each to the
//              rows of both particles
//   k_merge    per-particle merge of the new touching set with the particle's persistent
//              contact row -- the BitTable state machine of update_contacts! (kernels.jl:87-120)
//   k_force    per-particle loop over its row: Hertz-Mindlin / EBPM (kernels.jl:122-315), walls
//              (kernels.jl:317-426), body force (kernels.jl:428-443) and update! (445-485)
// Between near-list rebuilds a step is k_merge (testing the near lists) and k_force; while a
// single scene has no contact, one k_lean_free launch does both.
//
// Contact state lives in per-particle rows (column-major [slot][particle] arrays, sorted by
// partner index).  Every contact appears in the rows of both its particles and both copies are
// advanced with bit-identical arithmetic in the canonical (i<j) orientation, so each particle
// sums its own forces without atomics and the sum order is deterministic.
#include "dem_internal.hpp"

#include <float.h>

namespace dem {

namespace {

constexpr int TPB = 256;
constexpr int MERGE_TPB = 128;

// Blocks are dispatched to the 8 XCDs round-robin (block b runs on XCD b % 8), and each XCD has its
// own L2.  Renumbering so that XCD x runs the x-th contiguous eighth of the logical blocks keeps
// the neighbour cells that adjacent blocks share inside one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    constexpr uint32_t NXCD = 8;
    const uint32_t x = b % NXCD, local = b / NXCD;
    const uint32_t q = nb / NXCD, r = nb % NXCD;
    return x < r ? x * (q + 1) + local : r * (q + 1) + (x - r) * q + local;
}

// The same for batches, whose scenes differ in cost per particle: XCD x runs chunks of
// XCD_CHUNK contiguous logical blocks, every 8th chunk, so each XCD gets a share of every part of
// the batch.  Blocks past the last whole round of 8 chunks keep their own number.
constexpr uint32_t XCD_CHUNK = 32;
__device__ __forceinline__ uint32_t xcd_block_chunked(uint32_t b, uint32_t nb) {
    constexpr uint32_t NXCD = 8, ROUND = NXCD * XCD_CHUNK;
    const uint32_t nfull = nb / ROUND * ROUND;
    if (b >= nfull) return b;
    const uint32_t x = b % NXCD, local = b / NXCD;
    return local / XCD_CHUNK * ROUND + x * XCD_CHUNK + local % XCD_CHUNK;
}

template <bool SEG>
__device__ __forceinline__ uint32_t xcd_map(uint32_t b, uint32_t nb) {
    if constexpr (SEG) return xcd_block_chunked(b, nb);
    else return xcd_block(b, nb);
}

__device__ __forceinline__ double jsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }

// ------------------------------------------------------------------------------------------
// hashing and counting sort
// ------------------------------------------------------------------------------------------
// Counting sort, pass 1.  Thread t takes the particle in slot t of the previous step's sorted
// order, so the lanes of a wave mostly share cells: each run of equal buckets among adjacent lanes
// claims its slots with one atomic.  Any permutation is a valid starting order; a coherent one
// only makes the pass cheaper.
template <bool DENSE, bool SEG>
__global__ void __launch_bounds__(TPB)
k_hash(const Params P, const double4* __restrict__ pos, const uint32_t* __restrict__ sidx_prev,
       uint32_t* __restrict__ hid, uint32_t* __restrict__ rank, float4* __restrict__ x32,
       uint32_t* __restrict__ bcnt, uint32_t* __restrict__ tagg, int tile_shift,
       unsigned long long* ctr, int reset_counters, int init, uint32_t* __restrict__ nbc,
       uint32_t* __restrict__ gmin, int skin_cur) {
    if (reset_counters && blockIdx.x == 0) {
        for (int i = threadIdx.x; i < CTR_COPIES * CT_NSTEP; i += TPB)
            ctr[(i / CT_NSTEP) * CTR_STRIDE + i % CT_NSTEP] = 0ull;
        if (skin_cur >= 0) {
            // the next step's rebuild flag and displacement bound, set by this step's broad
            // phase and force kernel
            if (threadIdx.x == 0) ctr[CT_SKIN + 1 - skin_cur] = 0ull;
            for (int c = threadIdx.x; c < CTR_COPIES; c += TPB)
                ctr[c * CTR_STRIDE + CT_DISP + 1 - skin_cur] = 0ull;
        }
    }
    // near lists in use (skin_cur = the buffer parity): a step that reuses them does not sort
    if (skin_cur >= 0 && !skin_rebuilds(ctr, skin_cur)) return;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const bool live = t < P.n;
    uint32_t b = 0xffffffffu;
    double4 x = make_double4(0.0, 0.0, 0.0, 0.0);
    if (live) {
        const uint32_t p = sidx_prev[t];
        x = pos[p];
        if (skin_cur >= 0) {  // this step's broad phase refills the near lists
            nbc[p] = 0u;
            gmin[p] = 0x7f800000u;  // +inf
        }
        const SceneP& S = scene_of<SEG>(P, p);
        int32_t c[3];
        b = DENSE ? (init ? S.cell_base_init + dense_cell(S.dense_init, mk3(x.x, x.y, x.z), c)
                          : S.cell_base + dense_cell(S.dense, mk3(x.x, x.y, x.z), c))
                  : ref_bucket(S.grid, mk3(x.x, x.y, x.z));
    }
    // runs of equal buckets among adjacent live lanes (live lanes are a prefix of the wave)
    const int lane = threadIdx.x & 63;
    const uint32_t bprev = __shfl_up(b, 1, 64);
    const bool head = live && (lane == 0 || bprev != b);
    const uint64_t heads = __ballot(head);
    const uint64_t nlive = (uint64_t)__popcll(__ballot(live));
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int h = 63 - __clzll(heads & upto);  // this lane's run head (valid for live lanes)
    const uint64_t above = heads & ~upto;
    const uint32_t end = above ? (uint32_t)__ffsll((unsigned long long)above) - 1u : (uint32_t)nlive;
    uint32_t base = 0u;
    if (head) base = atomicAdd(&bcnt[b], end - (uint32_t)lane);
    base = __shfl(base, live ? h : lane, 64);
    if (live) {
        hid[t] = b;
        rank[t] = base + (uint32_t)(lane - h);
        x32[t] = make_float4((float)x.x, (float)x.y, (float)x.z, (float)x.w);
    }
    // The same run aggregation over scan tiles, so the scan of bcnt needs no look-back.  A wave's
    // first tile run is combined with the other waves' in LDS (a block's grains are adjacent in
    // the previous order, so they nearly always share one tile); later runs add directly.  The
    // reference's hashed buckets scatter a block over many tiles, so that grid keeps the
    // look-back scan instead.
    if constexpr (!DENSE) return;
    __shared__ uint32_t s_tile[TPB / 64], s_tcnt[TPB / 64];
    const int wave = threadIdx.x >> 6;
    uint32_t* const tcopy = tagg + (blockIdx.x % TAGG_COPIES);
    const uint32_t tl = b >> tile_shift;
    const uint32_t tlprev = __shfl_up(tl, 1, 64);
    const bool thead = live && (lane == 0 || tlprev != tl);
    const uint64_t tabove = __ballot(thead) & ~upto;
    const uint32_t tend = tabove ? (uint32_t)__ffsll((unsigned long long)tabove) - 1u : (uint32_t)nlive;
    if (lane == 0) {
        s_tile[wave] = thead ? tl : 0xffffffffu;
        s_tcnt[wave] = thead ? tend : 0u;
    } else if (thead) {
        atomicAdd(&tcopy[(size_t)tl * TAGG_COPIES], tend - (uint32_t)lane);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t0 = s_tile[0], c0 = s_tcnt[0];
#pragma unroll
        for (int w = 1; w < TPB / 64; ++w) {
            if (s_tile[w] == t0) {
                c0 += s_tcnt[w];
            } else {
                if (c0) atomicAdd(&tcopy[(size_t)t0 * TAGG_COPIES], c0);
                t0 = s_tile[w];
                c0 = s_tcnt[w];
            }
        }
        if (c0) atomicAdd(&tcopy[(size_t)t0 * TAGG_COPIES], c0);
    }
}

// Counting sort, pass 2: scatter slot t of the previous order to its new slot.  The order moves
// little from step to step, so the writes are close to sequential.
__global__ void __launch_bounds__(TPB)
k_scatter(const Params P, const double4* __restrict__ pos, const uint32_t* __restrict__ sidx_prev,
          const uint32_t* __restrict__ hid, const uint32_t* __restrict__ rank,
          const float4* __restrict__ x32, const uint32_t* __restrict__ bstart,
          uint32_t* __restrict__ sidx, uint32_t* __restrict__ skey, double4* __restrict__ spos,
          float4* __restrict__ spos32, const uint32_t* __restrict__ eid, uint32_t* __restrict__ sfid,
          uint32_t* __restrict__ tagg, int ntiles, const unsigned long long* __restrict__ gate) {
    if (gate && *gate == 0ull) return;  // the step reuses the near lists: no sort
    // the scan that read the histogram's tile totals has finished: clear them for the next sort
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < ntiles * TAGG_COPIES; i += TPB) tagg[i] = 0u;
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= P.n) return;
    const uint32_t b = hid[t];
    const uint32_t slot = bstart[b] + rank[t];
    const uint32_t p = sidx_prev[t];
    sidx[slot] = p;
    skey[slot] = b;
    spos32[slot] = x32[t];
    if (spos) spos[slot] = pos[p];  // the reference-grid broad phases read FP64 sorted positions
    if (sfid) sfid[slot] = eid[p];  // file index by slot, for the candidate count
}

// ------------------------------------------------------------------------------------------
// broad phase: touching pairs of the 8-bucket window
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void row_append(uint32_t* __restrict__ tcnt, uint32_t* __restrict__ touch,
                                           int64_t n, int K, uint32_t row, uint32_t partner,
                                           uint32_t& overflow) {
    const uint32_t m = atomicAdd(&tcnt[row], 1u);
    if ((int)m < K) touch[(int64_t)m * n + row] = partner;
    else overflow = 1;
}

template <bool INIT, bool MIXED>
__global__ void __launch_bounds__(TPB)
k_broad(const Params P, const uint32_t* __restrict__ sidx, const double4* __restrict__ spos,
        const float4* __restrict__ spos32, const uint32_t* __restrict__ bstart,
        const double* __restrict__ r0, const uint32_t* __restrict__ eid, uint32_t* __restrict__ tcnt,
        uint32_t* __restrict__ touch, unsigned long long* ctr) {
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    uint64_t cand = 0, found = 0;
    uint32_t overflow = 0;
    if (t < P.n) {
        const uint32_t p = sidx[t];
        const uint32_t ep = eid[p];
        const double4 me = spos[t];
        const d3 xp = mk3(me.x, me.y, me.z);
        const double rp = INIT ? r0[p] : me.w;
        float4 me32 = make_float4(0.f, 0.f, 0.f, 0.f);
        float slack = 0.f;
        if (MIXED) {
            me32 = spos32[t];
            slack = 8.0f * FLT_EPSILON * (fabsf(me32.x) + fabsf(me32.y) + fabsf(me32.z));
        }
        uint32_t w[8];
        window_of(P.sc.grid, xp, w);
#pragma unroll 1
        for (int wi = 0; wi < 8; ++wi) {
            const uint32_t b = w[wi];
            bool dup = false;
            for (int wj = 0; wj < wi; ++wj) dup |= (w[wj] == b);
            const uint32_t s0 = bstart[b], s1 = bstart[b + 1];
            for (uint32_t k = s0; k < s1; ++k) {
                const uint32_t q = sidx[k];
                if (eid[q] <= ep) continue;  // the pair (i, j), i < j in file order
                ++cand;
                if (dup) continue;  // an aliased window bucket repeats the same candidates
                if (MIXED) {
                    const float4 o = spos32[k];
                    const float dx = me32.x - o.x, dy = me32.y - o.y, dz = me32.z - o.z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    const float rs = me32.w + o.w;
                    const float thr = rs * (1.0f + 64.0f * FLT_EPSILON) + slack;
                    if (d2 > thr * thr) continue;
                }
                const double4 o = spos[k];
                const double rq = INIT ? r0[q] : o.w;
                // kernels.jl:103 (norm(k_i - k_j) < r_i + r_j), kernels.jl:72 with r0 at t = 0
                if (norm(xp - mk3(o.x, o.y, o.z)) < rp + rq) {
                    ++found;
                    row_append(tcnt, touch, P.n, P.K, p, q | IS_I, overflow);
                    row_append(tcnt, touch, P.n, P.K, q, p, overflow);
                }
            }
        }
    }
    wave_add(ctr_at(ctr, CT_CAND), cand);
    wave_add(ctr_at(ctr, CT_TOUCH), found);
    if (overflow) fail_step(ctr, (uint32_t)P.K + 1u);
}

// Slot-owned broad phase.  A touching pair lies inside both particles' windows (the window
// reaches at least cs/2 = 2 r_max past a particle on every axis, and r_i + r_j <= 2 max r <
// 2 r_max), so it can be owned by the particle with the lower sorted slot: each thread scans
// only slots k > t of its window buckets and reads the partner's index only for touching pairs.
// The reference's pair rule -- (lo, hi) is a candidate iff hi's bucket is one of lo's window
// buckets -- is then checked exactly: when the owner is hi, lo's window is recomputed.
template <bool INIT, bool MIXED>
__global__ void __launch_bounds__(TPB)
k_broad_slot(const Params P, const uint32_t* __restrict__ sidx, const double4* __restrict__ spos,
             const float4* __restrict__ spos32, const uint32_t* __restrict__ bstart,
             const double* __restrict__ r0, const uint32_t* __restrict__ eid,
             uint32_t* __restrict__ tcnt, uint32_t* __restrict__ touch, unsigned long long* ctr) {
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    uint64_t found = 0;
    uint32_t overflow = 0;
    if (t < P.n) {
        const double4 me = spos[t];
        const d3 xp = mk3(me.x, me.y, me.z);
        uint32_t p = 0xffffffffu;  // own index, loaded on the first touching pair
        float4 me32 = make_float4(0.f, 0.f, 0.f, 0.f);
        float slack = 0.f;
        if (MIXED) {
            me32 = spos32[t];
            slack = 8.0f * FLT_EPSILON * (fabsf(me32.x) + fabsf(me32.y) + fabsf(me32.z));
        }
        uint32_t w[8];
        window_of(P.sc.grid, xp, w);
        const uint32_t own = w[0];
#pragma unroll 1
        for (int wi = 0; wi < 8; ++wi) {
            const uint32_t b = w[wi];
            bool dup = false;
            for (int wj = 0; wj < wi; ++wj) dup |= (w[wj] == b);
            if (dup) continue;
            uint32_t k = bstart[b];
            const uint32_t s1 = bstart[b + 1];
            if (k <= (uint32_t)t) k = (uint32_t)t + 1;
            for (; k < s1; ++k) {
                if (MIXED) {
                    const float4 o = spos32[k];
                    const float dx = me32.x - o.x, dy = me32.y - o.y, dz = me32.z - o.z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    const float rs = me32.w + o.w;
                    const float thr = rs * (1.0f + 64.0f * FLT_EPSILON) + slack;
                    if (!INIT && d2 > thr * thr) continue;
                }
                const double4 o = spos[k];
                if (p == 0xffffffffu) p = sidx[t];
                const uint32_t q = sidx[k];
                const bool pi = eid[p] < eid[q];  // p is the pair's i (file order)
                const uint32_t lo = pi ? p : q, hi = pi ? q : p;
                const double4 xl = pi ? me : o, xh = pi ? o : me;
                const double rl = INIT ? r0[lo] : xl.w, rh = INIT ? r0[hi] : xh.w;
                // kernels.jl:103 (norm(k_i - k_j) < r_i + r_j), kernels.jl:72 with r0 at t = 0
                if (!(norm(mk3(xl.x, xl.y, xl.z) - mk3(xh.x, xh.y, xh.z)) < rl + rh)) continue;
                if (!pi) {  // owner is hi: is hi's bucket in lo's window?
                    uint32_t wl[8];
                    window_of(P.sc.grid, mk3(o.x, o.y, o.z), wl);
                    bool in = false;
#pragma unroll
                    for (int m = 0; m < 8; ++m) in |= (wl[m] == own);
                    if (!in) continue;
                }
                ++found;
                row_append(tcnt, touch, P.n, P.K, p, q | (pi ? IS_I : 0u), overflow);
                row_append(tcnt, touch, P.n, P.K, q, p | (pi ? 0u : IS_I), overflow);
            }
        }
    }
    wave_add(ctr_at(ctr, CT_TOUCH), found);
    if (overflow) fail_step(ctr, (uint32_t)P.K + 1u);
}

// Dense-grid broad phase.  Cells have side >= the largest contact distance, so every touching
// pair lies in the 27-cell stencil; slots are sorted by cell index (x fastest), so the three
// x-neighbours of a cell row form one contiguous slot range.  Pairs are owned by the lower sorted
// slot, and the four stencil rows below the particle's own (dz = -1, or dz = 0 and dy = -1) hold
// only lower slots, so each thread walks five slot ranges, (dy, dz) = (0,0), (1,0), (-1,1),
// (0,1), (1,1).
//
// A block's 256 consecutive slots lie in cells [cf, cl], so the j-th range of each of its threads
// lies inside the slots of cells [cf + off_j - 1, cl + off_j + 2) (off_j the linear offset of the
// row).  The block stages the FP32 copies of those five unions in LDS once and every thread then
// pre-tests its candidates from LDS, four per batch: |d|^2 against (r_i + r_j)^2 with a margin
// that covers the FP32 rounding of coordinates and radii, so no touching pair is rejected.
// Survivors are queued per wave and, once the scan is done, confirmed by all 64 lanes together
// (a survivor is rare per candidate but present in nearly every batch of a dense packing, so
// confirming in place would run the FP64 path at a fraction of the lanes).  They get the FP64
// squared-distance test, the reference's exact test
// (norm(k_i - k_j) < r_i + r_j, kernels.jl:103; r0 at t = 0, kernels.jl:72) and the reference's
// candidate rule (hi's bucket in lo's window, kernels.jl:24-58), so the contact set is the
// reference's in either precision.  Blocks whose unions exceed the stage, or that straddle two
// scenes of a batch, read the FP32 copies from global memory instead.
// The slot ranges [lo_j, lo_j + len_j) (all above slot t) of the five stencil rows a particle in
// cell c scans for partners: the row itself from x - 1 to x + 1 and the rows (y+1, z), (y-1, z+1),
// (y, z+1), (y+1, z+1).  A pair is owned by its lower slot, which covers the other four rows.
// A partner can be at most R = r_i * mult + cell / 2 away (the cell edge is at least the largest
// contact reach), so a neighbour cell whose nearest face is farther than R from this particle
// holds no partner: every particle of cell x - 1 lies below this cell's lower x face, and so on.
// Per row the perpendicular gap is combined with the x gap.  The gaps are FP32 with a margin over
// their rounding; a particle clamped into a boundary cell only has larger gaps to the cells it is
// clamped away from.  `extra` widens the reach by the near lists' 2 delta (the cells are then at
// least 2 max r + 2 delta).
__device__ __forceinline__ void broad_ranges(const DenseGrid& G, uint32_t cbase, const int32_t c[3],
                                             float4 me32, double mult, float extra, int64_t t,
                                             bool work, const uint32_t* __restrict__ bstart,
                                             uint32_t lo[5], uint32_t len[5]) {
    const float cell = (float)G.cell;
    const float R = fmaf(me32.w, (float)mult, 0.5f * cell + extra) * (1.0f + 1e-5f) +
                    16.0f * FLT_EPSILON * (fabsf(me32.x) + fabsf(me32.y) + fabsf(me32.z) + 3.0f * cell);
    const float R2 = R * R;
    float gl[3], gh[3];  // distances to the cell's lower and upper faces (0 if outside)
    {
        const float pc[3] = {me32.x, me32.y, me32.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float flo = (float)(G.origin[a] + (double)c[a] * G.cell);
            gl[a] = fmaxf(pc[a] - flo, 0.0f);
            gh[a] = fmaxf(flo + cell - pc[a], 0.0f);
        }
    }
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const int dy = j == 0 ? 0 : (j == 1 ? 1 : j - 3);
        const int dz = j < 2 ? 0 : 1;
        const int32_t y = c[1] + dy, z = c[2] + dz;
        const float gy = dy > 0 ? gh[1] : (dy < 0 ? gl[1] : 0.0f);
        const float gz = dz > 0 ? gh[2] : 0.0f;
        const float g2 = gy * gy + gz * gz;  // squared gap to the row
        const int32_t x0 = (c[0] > 0 && gl[0] * gl[0] + g2 <= R2) ? c[0] - 1 : c[0];
        const int32_t x1 = (c[0] + 1 < G.dims[0] && gh[0] * gh[0] + g2 <= R2) ? c[0] + 1 : c[0];
        lo[j] = 0u; len[j] = 0u;
        if (work && y >= 0 && y < G.dims[1] && z < G.dims[2] && g2 <= R2) {
            const uint32_t row = cbase +
                                 (uint32_t)G.dims[0] * ((uint32_t)y + (uint32_t)G.dims[1] * (uint32_t)z);
            uint32_t a = bstart[row + (uint32_t)x0];
            const uint32_t e = bstart[row + (uint32_t)x1 + 1];
            if (a <= (uint32_t)t) a = (uint32_t)t + 1;
            lo[j] = a;
            len[j] = e > a ? e - a : 0u;
        }
    }
}

constexpr int BROAD_BATCH = 4;
constexpr int STAGE_CAP = 1344;  // float4 per block (21 KiB): six blocks per CU fit in the LDS
constexpr int HIT_QUEUE = 256;   // pre-test survivors queued per wave

// Append a near-list entry; `need` keeps the longest list that did not fit.
__device__ __forceinline__ void near_append(uint32_t* __restrict__ nbc, uint32_t* __restrict__ nb,
                                            int64_t n, int KN, uint32_t row, uint32_t word,
                                            uint32_t& need) {
    const uint32_t m = atomicAdd(&nbc[row], 1u);
    if ((int)m < KN) nb[(int64_t)m * n + row] = word;
    else need = need > m + 1u ? need : m + 1u;
}

// A thread waits on a chain of dependent loads (key, bucket bounds, union bounds, stage), so the
// kernel is held to the register count of six waves per SIMD (80 VGPRs) for single scenes and
// five (96) for batches; neither per-step variant spills.
//
// With near lists (NEAR; skin_cur is the buffer parity) the broad phase runs only on steps that
// rebuild; its FP32 pre-test then reaches r_i + r_j + 2 delta, every pair that passes it goes to
// both particles' near lists, and the touching ones to the touching rows as before.  That variant
// needs more registers and runs at five waves per SIMD (four for batches).
template <bool INIT, bool SEG, bool NEAR>
__global__ void __launch_bounds__(TPB)
__attribute__((amdgpu_waves_per_eu(SEG && NEAR ? 4 : (SEG || NEAR ? 5 : 6))))
k_broad_dense(const Params P, const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ skey,
              const double4* __restrict__ pos, const float4* __restrict__ spos32,
              const uint32_t* __restrict__ bstart, const double* __restrict__ r0,
              const uint32_t* __restrict__ eid, uint32_t* __restrict__ tcnt,
              uint32_t* __restrict__ touch, unsigned long long* ctr, int diag,
              uint32_t* __restrict__ nbc, uint32_t* __restrict__ nb, uint32_t* __restrict__ gmin,
              int skin_cur) {
    static_assert(!(INIT && NEAR), "the bond pass keeps no near lists");
    if (NEAR && !skin_rebuilds(ctr, skin_cur)) return;
    constexpr bool near_lists = NEAR;
    // BROAD_BATCH - 1 entries of padding: a batch's reads past the end of the last range stay
    // inside the array (and are masked)
    __shared__ float4 stage[STAGE_CAP + BROAD_BATCH];
    __shared__ uint32_t u_lo[5], u_off[6];  // union slot ranges and their LDS offsets
    __shared__ uint32_t b_cell[2], b_scene[2];
    // per-wave queue of candidates that passed the FP32 pre-test: slot k and the owner's lane
    __shared__ uint32_t hq_k[TPB / 64][HIT_QUEUE];
    __shared__ uint8_t hq_l[TPB / 64][HIT_QUEUE];
    __shared__ uint32_t hq_n[TPB / 64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) hq_n[wv] = 0u;
    const int64_t n = P.n;
    const int64_t t0 = (int64_t)xcd_map<SEG>(blockIdx.x, gridDim.x) * TPB;
    const int64_t t = t0 + threadIdx.x;
    const int64_t tl = (t0 + TPB < n ? t0 + TPB : n) - 1;
    const bool live = t < n;
    // this thread's particle and cell; threads t0 and tl also give the block's first and last
    // cells.  The FP64 position is gathered only for candidates that pass the FP32 pre-test.
    float4 me32 = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t p = 0u, key = 0u;
    if (live) {
        me32 = spos32[t];
        p = sidx[t];
        key = skey[t];
    }
    const SceneP& S = scene_of<SEG>(P, p);
    const DenseGrid& G = INIT ? S.dense_init : S.dense;
    const uint32_t cbase = INIT ? S.cell_base_init : S.cell_base;
    const uint32_t mycell = live ? key - cbase : 0u;
    int32_t c[3];
    dense_coords(G, mycell, c);
    if (threadIdx.x == 0) { b_cell[0] = mycell; b_scene[0] = SEG ? (uint32_t)P.sid[p] : 0u; }
    if (t == tl) { b_cell[1] = mycell; b_scene[1] = SEG ? (uint32_t)P.sid[p] : 0u; }
    // this thread's five slot ranges [lo_j, lo_j + len_j), all above t
    const bool work = live && (!INIT || S.bond_init);
    // the near lists' reach past r_i + r_j (rounded up to FP32)
    const float ext2 = near_lists ? (float)(2.0 * S.skin) * (1.0f + FLT_EPSILON) : 0.0f;
    uint32_t lo[5], len[5];
    broad_ranges(G, cbase, c, me32, INIT ? S.bond_mult * (1.0 + 1e-9) : 1.0, ext2, t, work, bstart,
                 lo, len);
    __syncthreads();
    // union ranges (threads 0-4), then their LDS offsets (thread 0)
    const bool one_scene = !SEG || b_scene[0] == b_scene[1];
    if (threadIdx.x < 5) {
        const int j = threadIdx.x;
        uint32_t ulo = 0u, uhi = 0u;
        if (one_scene) {
            const SceneP& Sb = SEG ? P.seg[b_scene[0]] : P.sc;
            const DenseGrid& Gb = INIT ? Sb.dense_init : Sb.dense;
            const uint32_t bbase = INIT ? Sb.cell_base_init : Sb.cell_base;
            const int dy = j == 0 ? 0 : (j == 1 ? 1 : j - 3);
            const int dz = j < 2 ? 0 : 1;
            const int64_t off = (int64_t)Gb.dims[0] * ((int64_t)dy + (int64_t)Gb.dims[1] * dz);
            int64_t a = (int64_t)b_cell[0] + off - 1, b = (int64_t)b_cell[1] + off + 2;
            const int64_t cells = (int64_t)Gb.cells;
            a = a < 0 ? 0 : (a > cells ? cells : a);
            b = b < 0 ? 0 : (b > cells ? cells : b);
            if (b > a) {
                ulo = bstart[bbase + (uint32_t)a];
                uhi = bstart[bbase + (uint32_t)b];
                if (j == 0 && ulo <= (uint32_t)t0) ulo = (uint32_t)t0 + 1;
                if (uhi < ulo) uhi = ulo;
            }
        }
        u_lo[j] = ulo;
        u_off[j + 1] = uhi - ulo;  // length for now
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0u;
        u_off[0] = 0u;
        for (int j = 0; j < 5; ++j) { acc += u_off[j + 1]; u_off[j + 1] = acc; }
        if (!one_scene || acc > (uint32_t)STAGE_CAP) u_off[5] = 0xffffffffu;  // not staged
    }
    __syncthreads();
    const bool staged = u_off[5] != 0xffffffffu;
    if (staged) {
        for (uint32_t i = threadIdx.x; i < u_off[5]; i += TPB) {
            const int j = (i >= u_off[1]) + (i >= u_off[2]) + (i >= u_off[3]) + (i >= u_off[4]);
            stage[i] = spos32[u_lo[j] + (i - u_off[j])];
        }
    }
    __syncthreads();

    uint32_t found = 0u;  // touching pairs this thread owns (few)
    uint32_t overflow = 0, nb_need = 0;
    const double mult = INIT ? S.bond_mult * (1.0 + 1e-9) : 1.0;
    // FP64 tests of a pre-test survivor (slot k) of the particle in slot to (index po), and the
    // row appends of a touching pair; with near lists every survivor is a near pair
    auto confirm = [&](const SceneP& Sc, uint32_t po, uint32_t k) {
        const uint32_t q = sidx[k];
        if (near_lists) {
            const bool pn = eid[po] < eid[q];
            near_append(nbc, nb, n, P.KN, po, q | (pn ? IS_I : 0u), nb_need);
            near_append(nbc, nb, n, P.KN, q, po | (pn ? 0u : IS_I), nb_need);
        }
        const double4 mo = pos[po], ov = pos[q];
        const double ex = mo.x - ov.x, ey = mo.y - ov.y, ez = mo.z - ov.z;
        if (near_lists) {
            // the pair's gap, rounded down (two roundings of sqrt and difference, then FP32)
            const double dd = sqrt(ex * ex + ey * ey + ez * ez), rr = mo.w + ov.w;
            const double gap = (dd - rr) - 4e-16 * (dd + rr);
            const uint32_t gb = __float_as_uint(gap > 0.0 ? __double2float_rd(gap) : 0.0f);
            atomicMin(&gmin[po], gb);
            atomicMin(&gmin[q], gb);
        }
        const double rs = (mo.w + ov.w) * (INIT ? Sc.bond_mult * (1.0 + 1e-9) : 1.0);
        if (ex * ex + ey * ey + ez * ez > rs * rs * (1.0 + 1e-12)) return;  // sqrt(d2) >= rs
        const bool pi = eid[po] < eid[q];  // po is the pair's i (file order)
        const uint32_t il = pi ? po : q, ih = pi ? q : po;
        const double4 xl = pi ? mo : ov, xh = pi ? ov : mo;
        const double rl = INIT ? r0[il] : xl.w, rh = INIT ? r0[ih] : xh.w;
        if (!(norm(mk3(xl.x, xl.y, xl.z) - mk3(xh.x, xh.y, xh.z)) < rl + rh)) return;
        // when the contact reach is below cs/2 every touching pair passes this test
        if (Sc.ref_check && !ref_candidate(Sc.grid, mk3(xl.x, xl.y, xl.z), mk3(xh.x, xh.y, xh.z)))
            return;
        ++found;
        row_append(tcnt, touch, n, P.K, po, q | (pi ? IS_I : 0u), overflow);
        row_append(tcnt, touch, n, P.K, q, po | (pi ? 0u : IS_I), overflow);
    };
    if (work) {
        const float slack = 8.0f * FLT_EPSILON * (fabsf(me32.x) + fabsf(me32.y) + fabsf(me32.z));
        const float fmult = (float)mult * (1.0f + 64.0f * FLT_EPSILON);
        // each range lies inside its union by construction; checked so a stage read can never
        // stray.  Range j starts at stage slot u_off[j] + (lo[j] - u_lo[j]), recomputed from LDS
        // in the loop rather than held in registers.
        bool in_stage = staged;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            if (staged && len[j]) {
                const uint32_t ulen = u_off[j + 1] - u_off[j];
                if (lo[j] < u_lo[j] || lo[j] - u_lo[j] + len[j] > ulen) in_stage = false;
            }
        }
        // FP32 pre-test |d|^2 <= thr^2, thr = (r_i + r_j + ext2) * fmult + slack, evaluated with
        // fused multiply-adds; the 64-ulp margin in fmult covers their rounding
        const float c0 = fmaf(me32.w + ext2, fmult, slack);
        auto near = [&](const float4 o) -> bool {
            const float ex = me32.x - o.x, ey = me32.y - o.y, ez = me32.z - o.z;
            const float d2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
            const float thr = fmaf(o.w, fmult, c0);
            return fmaf(-thr, thr, d2) <= 0.0f;  // false for a non-finite position
        };
        // one batch of candidates o[] at slots rlo + i + u of a range of rlen slots: survivors go
        // to the wave's queue, confirmed below by all lanes together; a survivor that finds the
        // queue full is confirmed at once
        auto batch = [&](const float4 (&o)[BROAD_BATCH], uint32_t rlo, uint32_t rlen, uint32_t i) {
            bool hit[BROAD_BATCH];
            bool any = false;
#pragma unroll
            for (int u = 0; u < BROAD_BATCH; ++u) {
                hit[u] = near(o[u]) && (u == 0 || i + u < rlen);
                any |= hit[u];
            }
            if (!any) return;
#pragma unroll
            for (int u = 0; u < BROAD_BATCH; ++u) {
                const uint64_t m = __ballot(hit[u]);
                if (!m) continue;
                // the queue count is shared by the wave's lanes: relaxed atomics keep the compiler
                // from reusing a value another lane has since replaced
                const uint32_t base =
                    __hip_atomic_load(&hq_n[wv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                const uint32_t cnt = (uint32_t)__popcll(m);
                const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                const bool fits = base + cnt <= (uint32_t)HIT_QUEUE;
                if (hit[u]) {
                    if (fits) {
                        hq_k[wv][base + rank] = rlo + i + (uint32_t)u;
                        hq_l[wv][base + rank] = (uint8_t)lane;
                    } else {
                        confirm(S, p, rlo + i + (uint32_t)u);
                    }
                }
                if (fits && rank == 0 && hit[u])
                    __hip_atomic_store(&hq_n[wv], base + cnt, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
        };
        // range j without indexing the arrays by a runtime value (which would spill them)
#define BROAD_RANGE(j)                                                                           \
    const uint32_t rlo = j == 0 ? lo[0] : j == 1 ? lo[1] : j == 2 ? lo[2] : j == 3 ? lo[3] : lo[4]; \
    const uint32_t rlen = j == 0 ? len[0] : j == 1 ? len[1] : j == 2 ? len[2] : j == 3 ? len[3] : len[4]
        if (in_stage) {
#pragma unroll 1
            for (int j = 0; j < 5; ++j) {
                BROAD_RANGE(j);
                const uint32_t rlb = rlen ? u_off[j] + (rlo - u_lo[j]) : 0u;
                const float4* sp = stage + rlb;  // reads past the range stay in the padded stage
#pragma unroll 1
                for (uint32_t i = 0; i < rlen; i += BROAD_BATCH) {
                    float4 o[BROAD_BATCH];
#pragma unroll
                    for (int u = 0; u < BROAD_BATCH; ++u) o[u] = sp[i + u];
                    batch(o, rlo, rlen, i);
                }
            }
        } else {
#pragma unroll 1
            for (int j = 0; j < 5; ++j) {
                BROAD_RANGE(j);
#pragma unroll 1
                for (uint32_t i = 0; i < rlen; i += BROAD_BATCH) {
                    float4 o[BROAD_BATCH];
#pragma unroll
                    for (int u = 0; u < BROAD_BATCH; ++u)
                        o[u] = i + u < rlen ? spos32[rlo + i + u] : make_float4(0.f, 0.f, 0.f, 0.f);
                    batch(o, rlo, rlen, i);
                }
            }
        }
#undef BROAD_RANGE
    }
    // every lane of the wave confirms queued survivors (the queue holds only live owners)
    {
        const uint32_t qn =
            __hip_atomic_load(&hq_n[wv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        for (uint32_t e = (uint32_t)lane; e < qn; e += 64u) {
            const uint32_t k = hq_k[wv][e];
            const int64_t to = t0 + (int64_t)wv * 64 + hq_l[wv][e];
            const uint32_t po = sidx[to];
            confirm(scene_of<SEG>(P, po), po, k);
        }
    }
    if (diag) wave_add(ctr_at(ctr, CT_TOUCH), found);
    if (overflow) fail_step(ctr, (uint32_t)P.K + 1u);
    if (nb_need) {  // a near list is incomplete: the next step rebuilds (and the host widens them)
        atomicMax(&ctr[CT_NBNEED], (unsigned long long)nb_need);
        skin_request(ctr, skin_cur);
    }
}

// Candidate pairs the dense broad phase examines (its five trimmed slot ranges): a diagnostic,
// evaluated on the last step of each dem_step call.  (Counting inside the broad-phase
// loop itself costs more than the loop.)
template <bool SEG>
__global__ void __launch_bounds__(TPB)
k_examined_count(const Params P, const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ skey,
                 const float4* __restrict__ spos32, const uint32_t* __restrict__ bstart,
                 unsigned long long* ctr) {
    // the ranges the broad phase walks, with the near lists' reach when they are in use
    const int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x;
    uint64_t c = 0;
    if (t < P.n) {
        const SceneP& S = scene_of<SEG>(P, SEG ? sidx[t] : 0u);
        const DenseGrid& G = S.dense;
        int32_t cc[3];
        dense_coords(G, skey[t] - S.cell_base, cc);
        uint32_t lo[5], len[5];
        const float ext2 = P.skin ? (float)(2.0 * S.skin) * (1.0f + FLT_EPSILON) : 0.0f;
        broad_ranges(G, S.cell_base, cc, spos32[t], 1.0, ext2, t, true, bstart, lo, len);
#pragma unroll
        for (int j = 0; j < 5; ++j) c += len[j];
    }
    wave_add(ctr_at(ctr, CT_EXAMINED), c);
}

// Candidate-pair count of the reference's cp_list (kernels.jl:24-37): a diagnostic, evaluated on
// the last step of each dem_step call.
__global__ void __launch_bounds__(TPB)
k_cand_count(const Params P, const double4* __restrict__ pos, const uint32_t* __restrict__ sfid,
             const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ eid,
             unsigned long long* ctr) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    uint64_t c = 0;
    if (i < P.n) {
        const double4 x = pos[i];
        const uint32_t ei = eid[i];
        uint32_t w[8];
        window_of(P.sc.grid, mk3(x.x, x.y, x.z), w);
        uint32_t lo[8], hi[8];
#pragma unroll
        for (int wi = 0; wi < 8; ++wi) { lo[wi] = bstart[w[wi]]; hi[wi] = bstart[w[wi] + 1]; }
        for (int wi = 0; wi < 8; ++wi) {
            uint32_t k = lo[wi];
            for (; k + 3 < hi[wi]; k += 4) {  // four independent loads in flight
                const uint32_t a = sfid[k], b = sfid[k + 1], d = sfid[k + 2], e = sfid[k + 3];
                c += (uint64_t)(ei < a) + (ei < b) + (ei < d) + (ei < e);
            }
            for (; k < hi[wi]; ++k) c += ei < sfid[k];
        }
    }
    wave_add(ctr_at(ctr, CT_CAND), c);
}

// ------------------------------------------------------------------------------------------
// merge: persistent contact rows vs the new touching set (update_contacts! / init_bonds!)
// ------------------------------------------------------------------------------------------

// Is hi's bucket one of lo's 8 window buckets (the pair is in the reference's cp_list)?
__device__ __forceinline__ bool in_window(const SceneP& S, const double4* __restrict__ pos,
                                          uint32_t lo, uint32_t hi) {
    const double4 xl = pos[lo], xh = pos[hi];
    return ref_candidate(S.grid, mk3(xl.x, xl.y, xl.z), mk3(xh.x, xh.y, xh.z));
}

// State an entry carries from step to step: a bond's force and moments (F, ti, tj), a Hertz
// contact's tangential spring s.  The contact point, a Hertz contact's local force and the zero
// moments of a Hertz contact are outputs, rewritten on every output step, so a row rebuilt by the
// merge moves only the state (the bond pass, which runs before any step, moves everything).
template <bool INIT>
__device__ __forceinline__ void move_state(double* __restrict__ sn, const double* __restrict__ so,
                                           int64_t cs, int64_t dst, int64_t src, uint32_t word) {
    const int c0 = INIT ? 0 : ((word & BONDED) ? S_FX : S_SX);
    const int c1 = INIT ? S_NCOMP : ((word & BONDED) ? S_SX : S_NCOMP);
    for (int c = c0; c < c1; ++c) sn[c * cs + dst] = so[c * cs + src];
}
// A new entry's state is zero (outside the bond pass only a Hertz contact is ever created).
template <bool INIT>
__device__ __forceinline__ void zero_state(double* __restrict__ sn, int64_t cs, int64_t dst) {
#pragma unroll
    for (int c = INIT ? 0 : S_SX; c < S_NCOMP; ++c) sn[c * cs + dst] = 0.0;
}

// Per-particle merge of the sorted old row with the sorted new touching set.  `old_at(m)` and
// `touch_at(m)` read the m-th old-row word and the m-th touching partner.
template <bool INIT, typename OldAt, typename TouchAt>
__device__ __forceinline__ void merge_rows(const SceneP& S, const double4* __restrict__ pos, uint32_t p,
                                           int64_t n, int K, uint32_t nold, uint32_t nu, int b,
                                           OldAt old_at, TouchAt touch_at, uint8_t* __restrict__ sel,
                                           uint32_t* __restrict__ rcnt0, uint32_t* __restrict__ rcnt1,
                                           uint32_t* __restrict__ pq0, uint32_t* __restrict__ pq1,
                                           double* __restrict__ st0, double* __restrict__ st1,
                                           uint64_t& created, uint64_t& deact, uint64_t& entries,
                                           uint32_t& overflow) {
    const uint32_t INF = 0xffffffffu;
    // pass 1: decide the new row
    uint32_t io = 0, in = 0, nout = 0;
    bool changed = false;
    while (io < nold || in < nu) {
        const uint32_t eo = io < nold ? old_at(io) : INF;
        const uint32_t qo = io < nold ? (eo & IDX_MASK) : INF;
        const uint32_t tw = in < nu ? touch_at(in) : INF;
        const uint32_t qn = in < nu ? (tw & IDX_MASK) : INF;
        if (qo < qn) {  // only in the old row
            ++io;
            if (eo & BROKEN) { changed = true; continue; }
            if ((eo & BONDED) || INIT) { ++nout; continue; }
            const uint32_t lo = (eo & IS_I) ? p : qo, hi = (eo & IS_I) ? qo : p;
            if (in_window(S, pos, lo, hi)) {  // in cp_list and not touching: deactivate
                changed = true;
                if (eo & IS_I) ++deact;
            } else {
                ++nout;  // not a candidate this step: untouched (kernels.jl:97-119)
            }
        } else if (qn < qo) {  // only in the new set: append a zeroed contact
            ++in; ++nout; changed = true;
            if (tw & IS_I) ++created;
        } else {  // in both
            ++io; ++in; ++nout;
            if (eo & BROKEN) {  // bond removed at the end of last step, re-created as Hertz
                changed = true;
                if (tw & IS_I) ++created;
            } else if (INIT && !(eo & BONDED)) {
                changed = true;
            }
        }
    }
    if ((int)nout > K) overflow = 1;
    entries = nout;
    if (!changed || (int)nout > K) return;
    // pass 2: write the new row into the other buffer
    const int nb = 1 - b;
    uint32_t* pqn = nb ? pq1 : pq0;
    const double* so = b ? st1 : st0;
    double* sn = nb ? st1 : st0;
    const int64_t cs = (int64_t)K * n;  // component stride
    io = 0; in = 0;
    uint32_t o = 0;
    while (io < nold || in < nu) {
        const uint32_t eo = io < nold ? old_at(io) : INF;
        const uint32_t qo = io < nold ? (eo & IDX_MASK) : INF;
        const uint32_t tw = in < nu ? touch_at(in) : INF;
        const uint32_t qn = in < nu ? (tw & IDX_MASK) : INF;
        const int64_t dst = (int64_t)o * n + p;
        if (qo < qn) {
            const int64_t src = (int64_t)io * n + p;
            ++io;
            if (eo & BROKEN) continue;
            if (!(eo & BONDED) && !INIT) {
                const uint32_t lo = (eo & IS_I) ? p : qo, hi = (eo & IS_I) ? qo : p;
                if (in_window(S, pos, lo, hi)) continue;
            }
            pqn[dst] = eo;
            move_state<INIT>(sn, so, cs, dst, src, eo);
            ++o;
        } else if (qn < qo) {
            ++in;
            pqn[dst] = (tw & (IDX_MASK | IS_I)) | (INIT ? BONDED : 0u);
            zero_state<INIT>(sn, cs, dst);
            ++o;
        } else {
            const int64_t src = (int64_t)io * n + p;
            ++io; ++in;
            if (eo & BROKEN) {
                pqn[dst] = tw & (IDX_MASK | IS_I);
                zero_state<INIT>(sn, cs, dst);
            } else if (INIT && !(eo & BONDED)) {
                pqn[dst] = eo | BONDED;
                zero_state<INIT>(sn, cs, dst);
            } else {
                pqn[dst] = eo;
                move_state<INIT>(sn, so, cs, dst, src, eo);
            }
            ++o;
        }
    }
    (nb ? rcnt1 : rcnt0)[p] = o;
    sel[p] = (uint8_t)(ROW_PENDING | b);  // committed by the force kernel (or the bond pass)
}

// Sort the touch words v[0..nt) by partner index (insertion sort: rows are short) and drop
// repeats; returns the count.
template <typename Get, typename Set>
__device__ __forceinline__ uint32_t sort_unique(uint32_t nt, Get get, Set set) {
    for (uint32_t a = 1; a < nt; ++a) {
        const uint32_t v = get(a);
        int b = (int)a - 1;
        while (b >= 0 && (get(b) & IDX_MASK) > (v & IDX_MASK)) { set(b + 1, get(b)); --b; }
        set(b + 1, v);
    }
    uint32_t nu = 0;
    for (uint32_t a = 0; a < nt; ++a)
        if (nu == 0 || (get(a) & IDX_MASK) != (get(nu - 1) & IDX_MASK)) { set(nu, get(a)); ++nu; }
    return nu;
}

// The broad phase's test of a touching pair (kernels.jl:103, norm(k_i - k_j) < r_i + r_j, and the
// reference's window rule when the contact reach exceeds half its cell), for a particle at mo and
// its partner at ov; pi: the particle is the pair's i.  The same operations as the broad phase's
// confirmation, so a near-list step finds the touching set a rebuilding step would.
__device__ __forceinline__ bool touching_pair(const SceneP& S, const double4 mo, const double4 ov,
                                              bool pi) {
    const double ex = mo.x - ov.x, ey = mo.y - ov.y, ez = mo.z - ov.z;
    const double rs = mo.w + ov.w;
    if (ex * ex + ey * ey + ez * ez > rs * rs * (1.0 + 1e-12)) return false;  // sqrt(d2) >= rs
    const double4 xl = pi ? mo : ov, xh = pi ? ov : mo;
    if (!(norm(mk3(xl.x, xl.y, xl.z) - mk3(xh.x, xh.y, xh.z)) < xl.w + xh.w)) return false;
    return !S.ref_check || ref_candidate(S.grid, mk3(xl.x, xl.y, xl.z), mk3(xh.x, xh.y, xh.z));
}

constexpr int MERGE_LDS = 16;

// The arrays the merge reads and writes.
struct MergeArgs {
    const double4* __restrict__ pos;  // positions of the step (the current buffer)
    uint32_t* __restrict__ nbc; uint32_t* __restrict__ nb; const uint32_t* __restrict__ gmin;
    uint32_t* __restrict__ gmin_w;  // the same array, written by k_mark_bonds
    uint32_t* __restrict__ tcnt; uint32_t* __restrict__ touch;
    uint8_t* __restrict__ sel;
    uint32_t* __restrict__ rcnt0; uint32_t* __restrict__ rcnt1;
    uint32_t* __restrict__ pq0; uint32_t* __restrict__ pq1;
    double* __restrict__ st0; double* __restrict__ st1;
    uint32_t* __restrict__ gskip;
};

// What a wave of the merge counts.
struct MergeTally {
    uint64_t created = 0, deact = 0, entries = 0, found = 0;
    uint32_t overflow = 0, need = 0;
    uint32_t contact = 0;  // contact-free mode: a particle would get a row entry
};

// The merge of particle p: its new touching set (from the broad phase, or on a step that reuses
// the near lists the near-list entries that pass the broad phase's test) merged into its row.
// lt, lw: this thread's column of the block's LDS stage (stride BS).  twoD: twice the
// displacement bound D (rounded up); a particle whose near-list gaps at the rebuild all exceed it
// touches none of those partners now.
// A particle's merge inputs, loaded first (independent loads): the touching (or near-list) count,
// the smallest near-list gap, the row selector and both row counts.
// `near`: near lists in use, so the step may reuse them (then the near-list count and smallest gap
// are the inputs) or rebuild them (then the touching count is).
struct MergeIn {
    uint32_t tc, nc, c0, c1;  // nc: the near-list count, with NB_BROKE
    float gm;
    uint8_t sl;
    __device__ __forceinline__ uint32_t cnt_in(bool reuse) const { return reuse ? (nc & ~NB_BROKE) : tc; }
};
__device__ __forceinline__ MergeIn load_merge_in(const MergeArgs& M, uint32_t p, bool near, bool lean) {
    MergeIn m;
    m.tc = lean ? 0u : M.tcnt[p];
    m.nc = near ? M.nbc[p] : 0u;
    m.gm = near ? __uint_as_float(M.gmin[p]) : 0.0f;
    m.sl = M.sel[p];
    m.c0 = M.rcnt0[p];
    m.c1 = M.rcnt1[p];
    return m;
}

template <bool INIT, bool SEG, int BS, bool BONDS = false>
__device__ __forceinline__ bool merge_particle(const Params& P, const MergeArgs& M, uint32_t p,
                                               const MergeIn& in, uint32_t* lt, uint32_t* lw,
                                               bool reuse, float twoD, MergeTally& t, bool free = false) {
    const int64_t n = P.n;
    const int K = P.K;
    // The committed buffer: after a failed step of the call a row may still be marked
    // pending; this merge then overwrites the uncommitted buffer again, which the host drops.
    const uint32_t cnt_in = in.cnt_in(reuse);
    const float gm = reuse ? in.gm : 0.0f;
    const int b = in.sl & 1;
    const uint32_t nold = b ? in.c1 : in.c0;
    uint32_t nt = cnt_in;
    bool t_lds = false;  // the touching partners are in lt already
    // On a lean step the old row (when it fits the LDS stage) is staged before the near lists are
    // walked.  Entries marked as intact bonds are left untested, unless a bond of this particle
    // has broken since the rebuild (a BROKEN entry now, or NB_BROKE set by an earlier merge that
    // saw one: then its lists are tested in full until the next rebuild) or its row is long.
    const uint32_t* pqo = b ? M.pq1 : M.pq0;
    bool lw_staged = false, broke = (in.nc & NB_BROKE) != 0u;
    if (BONDS && reuse && nold && nold <= (uint32_t)MERGE_LDS) {
        uint32_t any = 0u;
        for (uint32_t m = 0; m < nold; m += 8) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = m + u < nold ? pqo[(int64_t)(m + u) * n + p] : 0u;
#pragma unroll
            for (int u = 0; u < 8; ++u) if (m + u < nold) { lw[(m + u) * BS] = v[u]; any |= v[u]; }
        }
        lw_staged = true;
        if ((any & BROKEN) && !broke) {
            broke = true;
            M.nbc[p] = in.nc | NB_BROKE;
        }
    } else if (BONDS && reuse && nold > (uint32_t)MERGE_LDS && !broke) {
        // a long row is not searched for BROKEN entries: its lists are tested in full until the
        // next rebuild
        broke = true;
        M.nbc[p] = in.nc | NB_BROKE;
    }
    const uint32_t bond_mask = (!BONDS || broke || nold > (uint32_t)MERGE_LDS) ? 0u : NB_BOND;
    if (reuse) {
        // the first MERGE_LDS touching partners go to lt, later ones (when there are more) to
        // the touching row, which then also takes a copy of the first ones
        nt = 0;
        // gmin skips the walk; with a broken bond (bond_mask 0 on a bonded scene's lean step) it
        // may be the smallest gap of the unmarked entries only
        const uint32_t nn = (gm > twoD && !(BONDS && broke)) ? 0u : min(cnt_in, (uint32_t)P.KN);
        auto take = [&](uint32_t w) {  // a touching near-list entry, in list order
            if (w & IS_I) ++t.found;
            if (nt < (uint32_t)MERGE_LDS) {
                lt[nt * BS] = w;
            } else {
                if (nt == (uint32_t)MERGE_LDS)
                    for (uint32_t a = 0; a < (uint32_t)MERGE_LDS && (int)a < K; ++a)
                        M.touch[(int64_t)a * n + p] = lt[a * BS];
                if ((int)nt < K) M.touch[(int64_t)nt * n + p] = w;
            }
            ++nt;
        };
        if constexpr (BS == 64 && !SEG) {
          if (nn) {
            // contact-dense single scenes (one-wave blocks): long lists, four entries at a time --
            // four word loads, then four partner gathers, then four tests (packed 1M pour merge
            // 97 -> 90 us; batches keep the one-deep walk, whose registers allow five waves per
            // SIMD instead of three: 256-cube batch merge 101 against 112 us)
            const SceneP& S = scene_of<SEG>(P, p);
            const double4 mo = M.pos[p];
            for (uint32_t m0 = 0; m0 < nn; m0 += 4) {
                uint32_t w4[4];
                double4 o4[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) w4[u] = m0 + u < nn ? M.nb[(int64_t)(m0 + u) * n + p] : 0u;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    o4[u] = (m0 + u < nn && !(w4[u] & bond_mask)) ? M.pos[w4[u] & IDX_MASK] : mo;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (m0 + u < nn && !(w4[u] & bond_mask) && touching_pair(S, mo, o4[u], (w4[u] & IS_I) != 0u))
                        take(w4[u] & ~NB_BOND);
            }
          }
        } else if (nn) {
            const SceneP& S = scene_of<SEG>(P, p);
            const double4 mo = M.pos[p];
            // entry m + 1's word and partner position are loaded while entry m is tested (no
            // position for an intact bond's entry, which is not tested)
            uint32_t w = M.nb[p];
            double4 o = (w & bond_mask) ? mo : M.pos[w & IDX_MASK];
            for (uint32_t m = 0; m < nn; ++m) {
                uint32_t wn = 0u;
                double4 on = mo;
                if (m + 1 < nn) {
                    wn = M.nb[(int64_t)(m + 1) * n + p];
                    if (!(wn & bond_mask)) on = M.pos[wn & IDX_MASK];
                }
                if (!(w & bond_mask) && touching_pair(S, mo, o, (w & IS_I) != 0u)) {
                    w &= ~NB_BOND;
                    if (w & IS_I) ++t.found;
                    if (nt < (uint32_t)MERGE_LDS) {
                        lt[nt * BS] = w;
                    } else {
                        if (nt == (uint32_t)MERGE_LDS)
                            for (uint32_t a = 0; a < (uint32_t)MERGE_LDS && (int)a < K; ++a)
                                M.touch[(int64_t)a * n + p] = lt[a * BS];
                        if ((int)nt < K) M.touch[(int64_t)nt * n + p] = w;
                    }
                    ++nt;
                }
                w = wn;
                o = on;
            }
        }
        t_lds = nt <= (uint32_t)MERGE_LDS;
    }
    t.need = max(t.need, nt);
    if (nt | nold) {
        const SceneP& S = scene_of<SEG>(P, p);
        if (nt && !reuse) M.tcnt[p] = 0u;
        if (free) {  // contact-free mode: the step fails, to be replayed with the full force kernel
            t.contact = 1;
            return false;
        }
        uint32_t overflow = 0;
        if ((int)nt > K) { overflow = 1; nt = K; }
        uint64_t entries = 0;
        if (nt <= (uint32_t)MERGE_LDS && nold <= (uint32_t)MERGE_LDS) {
            for (uint32_t m = 0; m < (t_lds ? 0u : nt); m += 8) {
                uint32_t v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = m + u < nt ? M.touch[(int64_t)(m + u) * n + p] : 0u;
#pragma unroll
                for (int u = 0; u < 8; ++u) if (m + u < nt) lt[(m + u) * BS] = v[u];
            }
            for (uint32_t m = 0; m < (lw_staged ? 0u : nold); m += 8) {
                uint32_t v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = m + u < nold ? pqo[(int64_t)(m + u) * n + p] : 0u;
#pragma unroll
                for (int u = 0; u < 8; ++u) if (m + u < nold) lw[(m + u) * BS] = v[u];
            }
            const uint32_t nu = sort_unique(
                nt, [&](uint32_t m) { return lt[m * BS]; },
                [&](uint32_t m, uint32_t v) { lt[m * BS] = v; });
            merge_rows<INIT>(S, M.pos, p, n, K, nold, nu, b,
                             [&](uint32_t m) { return lw[m * BS]; },
                             [&](uint32_t m) { return lt[m * BS]; }, M.sel, M.rcnt0,
                             M.rcnt1, M.pq0, M.pq1, M.st0, M.st1, t.created, t.deact, entries, overflow);
        } else {
            if (t_lds)
                for (uint32_t a = 0; a < nt; ++a) M.touch[(int64_t)a * n + p] = lt[a * BS];
            // long rows: the touching list is sorted in place (it is rebuilt every step)
            const uint32_t nu = sort_unique(
                nt, [&](uint32_t m) { return M.touch[(int64_t)m * n + p]; },
                [&](uint32_t m, uint32_t v) { M.touch[(int64_t)m * n + p] = v; });
            merge_rows<INIT>(S, M.pos, p, n, K, nold, nu, b,
                             [&](uint32_t m) { return pqo[(int64_t)m * n + p]; },
                             [&](uint32_t m) { return M.touch[(int64_t)m * n + p]; }, M.sel, M.rcnt0,
                             M.rcnt1, M.pq0, M.pq1, M.st0, M.st1, t.created, t.deact, entries, overflow);
        }
        t.entries += entries;
        if (overflow) {
            t.overflow = 1;
            t.need = max(t.need, (uint32_t)entries);
        }
    }
    // no row and no touching partner: nothing changed (only the variant that keeps group values
    // asks; computing it costs the others registers)
    if constexpr (!INIT && !SEG && BS >= MERGE_TPB) return (nt | nold) == 0u;  // sparse single-scene merges
    else return false;
}

// The merge's view of the step, decided for the whole block (so a block leaves or stays as a
// whole), after the block's particle loads are issued: on a lean step nothing after a failed step
// of the call, a step that has to rebuild fails the call instead, otherwise the step runs and
// block 0 clears the next step's words; whether the step reuses the near lists; and twice the
// displacement bound D of the step (rounded up, the maximum over the counter copies).
// clear_next false (the one-kernel lean step): the next step's words are left alone, since force
// blocks of the same launch write them.
struct MergeStep { bool go, reuse; float twoD; };
__device__ __forceinline__ MergeStep merge_step_begins(unsigned long long* ctr, int skin_cur,
                                                       bool near, bool lean, int bs,
                                                       bool clear_next = true) {
    __shared__ int s_go, s_reuse;
    __shared__ float s_disp;
    if (threadIdx.x < 64) {
        float v = 0.0f;
        if (near)
            for (int c = threadIdx.x; c < CTR_COPIES; c += 64)
                v = fmaxf(v, __uint_as_float((uint32_t)ctr[c * CTR_STRIDE + CT_DISP + skin_cur]));
        for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
        if (threadIdx.x == 0) {
            int go = 1;
            if (lean) {
                if (step_failed(ctr)) {
                    go = 0;
                } else if (skin_rebuilds(ctr, skin_cur)) {
                    atomicOr(&ctr[CT_ABORT], 1ull);
                    atomicOr(&ctr[CT_FAIL], 1ull);
                    go = 0;
                } else if (blockIdx.x == 0 && clear_next) {
                    ctr[CT_SKIN + 1 - skin_cur] = 0ull;
                }
            }
            s_go = go;
            s_reuse = near && (lean || !skin_rebuilds(ctr, skin_cur));
            s_disp = v;
        }
    }
    __syncthreads();
    MergeStep m;
    m.go = s_go != 0;
    m.reuse = s_reuse != 0;
    m.twoD = 2.0f * s_disp * (1.0f + 1e-6f);
    if (m.go && lean && blockIdx.x == 0 && clear_next)
        for (int c = threadIdx.x; c < CTR_COPIES; c += bs)
            ctr[c * CTR_STRIDE + CT_DISP + 1 - skin_cur] = 0ull;
    return m;
}

__device__ __forceinline__ void merge_counters(unsigned long long* ctr, bool diag, bool reuse,
                                               const MergeTally& t, bool free = false) {
    if (free && __ballot(t.contact != 0u) != 0ull && (threadIdx.x & 63) == 0) {
        atomicOr(&ctr[CT_ABORT], 2ull);
        atomicOr(&ctr[CT_FAIL], 1ull);
    }
    if (diag) {
        wave_add(ctr_at(ctr, CT_CREATED), t.created);
        wave_add(ctr_at(ctr, CT_DEACT), t.deact);
        wave_add(ctr_at(ctr, CT_ENTRIES), t.entries);
        if (reuse) wave_add(ctr_at(ctr, CT_TOUCH), t.found);
    }
    // the row this particle needs: its merged length, or at least its touching count
    if (t.overflow) fail_step(ctr, t.need);
}

// The merge kernel.  Rows of up to MERGE_LDS touching partners and old entries are staged in LDS
// with batched independent loads; longer rows are merged in place in global memory.  On a step
// that reuses the near lists the touching partners are the near-list entries that pass the
// broad phase's test, found here (the broad phase did not run).  On a lean step of a sparse
// single scene (blocks of MERGE_TPB threads) a wave whose group of 64 particles has no row entry
// and no near-list partner within 2 D (gskip) leaves before loading anything of its particles:
// their merges would change nothing.  The other single-scene variants set their groups' values
// to 0 (never left out), which keeps them valid when the block size changes.  free: the scene is
// in contact-free mode (the sparse single-scene variant only): no particle has a row, the force
// kernel evaluates none, and a particle that would get one fails the step instead (CT_ABORT bit
// 2), which the host replays with the full force kernel.
// BONDS: the scene's bond pass ran, so lean steps stage the old row first and leave the
// near-list entries of intact bonds untested.  Those variants are held to five waves per SIMD
// (four for the single-scene one-wave blocks, whose walk loads four near-list entries at a
// time): the stage and the marks need a few registers more, and the merges are bound by their
// chains of dependent loads.
template <bool INIT, bool SEG, int BS, bool BONDS = false>
__global__ void __launch_bounds__(BS)
__attribute__((amdgpu_waves_per_eu(!BONDS ? 1 : (SEG || BS != 64 ? 5 : 4))))
k_merge(const Params P, const MergeArgs M, unsigned long long* ctr, int diag, int skin_cur, int lean,
        int free) {
    __shared__ uint32_t lt[MERGE_LDS * BS];  // touching partners
    __shared__ uint32_t lw[MERGE_LDS * BS];  // old-row words
    const int64_t p64 = (int64_t)xcd_map<SEG>(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    const bool in = p64 < P.n;
    const uint32_t p = in ? (uint32_t)p64 : 0u;
    const bool near = !INIT && skin_cur >= 0;
    constexpr bool GROUPS = !INIT && !SEG && BS == MERGE_TPB;
    if constexpr (!GROUPS) {
        // the particle's loads first: the block's look at the step's words waits for them together
        const MergeIn mi = load_merge_in(M, p, near, lean != 0);
        const MergeStep st = merge_step_begins(ctr, skin_cur, near, lean != 0, BS);
        if (!st.go) return;
        MergeTally t;
        if (in) merge_particle<INIT, SEG, BS, BONDS>(P, M, p, mi, lt + threadIdx.x, lw + threadIdx.x, st.reuse, st.twoD, t);
        merge_counters(ctr, diag != 0, st.reuse, t);
        if constexpr (!SEG)
            if ((threadIdx.x & 63) == 0 && p64 < P.n) M.gskip[p64 >> 6] = 0u;
    } else {
        const int64_t g = p64 >> 6;       // the wave's group
        const bool gin = (g << 6) < P.n;  // it holds particles
        MergeIn mi;
        MergeStep st;
        if (lean) {
            // the group's value first, the particle's loads only if the wave stays
            const float gs = gin ? __uint_as_float(M.gskip[g]) : __builtin_inff();
            st = merge_step_begins(ctr, skin_cur, near, true, BS);
            if (!st.go || gs > st.twoD) return;
            mi = load_merge_in(M, p, near, true);
        } else {
            mi = load_merge_in(M, p, near, false);
            st = merge_step_begins(ctr, skin_cur, near, false, BS);
            if (!st.go) return;
        }
        MergeTally t;
        bool idle = true;
        if (in) idle = merge_particle<INIT, SEG, BS, BONDS>(P, M, p, mi, lt + threadIdx.x, lw + threadIdx.x, st.reuse,
                                                     st.twoD, t, free != 0);
        merge_counters(ctr, diag != 0, st.reuse, t, free != 0);
        // the group's value for the next lean steps (gmin stays until the next rebuild)
        float v = !in ? __builtin_inff() : (idle && near && mi.gm > 0.0f) ? mi.gm : 0.0f;
        for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0 && gin) M.gskip[g] = __float_as_uint(v);
    }
}

// A rebuilding step's near lists, after its merge: mark each entry whose partner is an intact
// bond in the particle's new row (NB_BOND), so that lean merges leave it untested, and make the
// particle's smallest near-list gap (gmin) that of its unmarked entries, with the broad phase's
// rounding (a lean merge whose particle has a broken bond walks its list whatever gmin says).
// Rows of more than MERGE_LDS entries are not marked (all their entries are tested).  Launched
// on every full step with near lists; it leaves at once unless the step rebuilt them.
template <bool SEG>
__global__ void __launch_bounds__(TPB)
k_mark_bonds(const Params P, const MergeArgs M, const unsigned long long* ctr, int skin_cur) {
    __shared__ uint32_t rw[MERGE_LDS * TPB];  // the new row, one column per thread
    if (!skin_rebuilds(ctr, skin_cur)) return;
    const int64_t p64 = (int64_t)xcd_map<SEG>(blockIdx.x, gridDim.x) * TPB + threadIdx.x;
    if (p64 >= P.n) return;
    const uint32_t p = (uint32_t)p64;
    const int64_t n = P.n;
    const uint32_t nc = min(M.nbc[p] & ~NB_BROKE, (uint32_t)P.KN);
    const uint8_t sl = M.sel[p];
    const int b = (sl & ROW_PENDING) ? ((sl & 1) ^ 1) : (sl & 1);  // the row the force kernel commits
    const uint32_t cnt = b ? M.rcnt1[p] : M.rcnt0[p];
    if (!nc || !cnt || cnt > (uint32_t)MERGE_LDS) return;
    const uint32_t* pq = b ? M.pq1 : M.pq0;
    uint32_t* r = rw + threadIdx.x;
    for (uint32_t m = 0; m < cnt; ++m) r[m * TPB] = pq[(int64_t)m * n + p];
    uint32_t marked = 0u, gb = 0x7f800000u;  // +inf
    const double4 mo = M.pos[p];
    for (uint32_t m = 0; m < nc; ++m) {
        const int64_t e = (int64_t)m * n + p;
        const uint32_t w = M.nb[e];
        const uint32_t q = w & IDX_MASK;
        bool bond = false;
        for (uint32_t k = 0; k < cnt; ++k) {
            const uint32_t eo = r[k * TPB];
            if ((eo & IDX_MASK) == q) {
                bond = (eo & BONDED) && !(eo & BROKEN);
                break;
            }
        }
        if (bond) {
            M.nb[e] = w | NB_BOND;
            ++marked;
        } else {  // the pair's gap, rounded down, as the broad phase takes it
            const double4 ov = M.pos[q];
            const double ex = mo.x - ov.x, ey = mo.y - ov.y, ez = mo.z - ov.z;
            const double dd = sqrt(ex * ex + ey * ey + ez * ez), rr = mo.w + ov.w;
            const double gap = (dd - rr) - 4e-16 * (dd + rr);
            gb = min(gb, __float_as_uint(gap > 0.0 ? __double2float_rd(gap) : 0.0f));
        }
    }
    if (marked) M.gmin_w[p] = gb;
}

// ------------------------------------------------------------------------------------------
// forces + integration
// ------------------------------------------------------------------------------------------
struct HMOut { d3 F; d3 s; bool fatal; };

typedef double dvec4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_nt(double4* p, double4 v) {
    const dvec4 t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<dvec4*>(p));
}

// Hertz-Mindlin law (kernels.jl:261-303, 369-406).  Ys and Gs are the effective moduli
// (kernels.jl:262-263), per-scene constants evaluated once on the host with the same operations.
__device__ __forceinline__ HMOut hertz_mindlin(double gap, double dn, d3 vc, double Ys, double Gs,
                                               double Rs, double ms, const SurfaceP& sf, double dt,
                                               d3 s) {
    HMOut o;
    const double beta = sf.beta;
    const double srd = sqrt(Rs * dn);
    const double Sn = 2.0 * Ys * srd;
    const double St = 8.0 * Gs * srd;
    const double kn = 4.0 / 3.0 * Ys * srd;
    const double gn = -2.0 * beta * sqrt(5.0 / 6.0 * Sn * ms);
    const double kt = 8.0 * Gs * srd;
    const double gt = -2.0 * beta * sqrt(5.0 / 6.0 * St * ms);
    o.fatal = (gn < 0) || (gt < 0);
    const d3 sv = s + mk3(vc.x * 0.0, vc.y * dt, vc.z * dt);
    const double F1 = -kn * gap - gn * vc.x;
    const d3 Ft = (-kt) * sv;
    const double nFt = norm(Ft);
    double F2, F3;
    if (nFt >= sf.mu * F1) {
        const double ratio = sf.mu * F1 / nFt;
        F2 = ratio * Ft.y;
        F3 = ratio * Ft.z;
        o.s = mk3(sv.x, F2 / kt, F3 / kt);
    } else {
        F2 = Ft.y - gt * vc.y;
        F3 = Ft.z - gt * vc.z;
        o.s = sv;
    }
    o.F = mk3(F1, F2, F3);
    return o;
}

// ---- mixed precision: FP32 contact frame and Hertz-Mindlin law, FP64 spring state ----------
struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mkf(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 tof(d3 a) { return mkf((float)a.x, (float)a.y, (float)a.z); }
__device__ __forceinline__ d3 tod(f3 a) { return mk3((double)a.x, (double)a.y, (double)a.z); }
__device__ __forceinline__ f3 fsub(f3 a, f3 b) { return mkf(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 fadd(f3 a, f3 b) { return mkf(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 fscl(f3 a, float s) { return mkf(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 fcross(f3 a, f3 b) {
    return mkf(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Reciprocal and square root from the hardware approximations (about 1 ulp): the mixed path's
// accuracy is set by FP32 rounding anyway, and an IEEE-exact FP32 division or square root is a
// sequence of about ten instructions.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fnorm(f3 a) { return fsqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ f3 fneg(f3 a) { return mkf(-a.x, -a.y, -a.z); }

struct fm3 { float a[3][3]; };
__device__ __forceinline__ f3 fmul(const fm3& R, f3 u) {
    return mkf(R.a[0][0] * u.x + R.a[0][1] * u.y + R.a[0][2] * u.z,
               R.a[1][0] * u.x + R.a[1][1] * u.y + R.a[1][2] * u.z,
               R.a[2][0] * u.x + R.a[2][1] * u.y + R.a[2][2] * u.z);
}
// R^T u: R is a rotation, so its transpose is its inverse to FP32 precision
__device__ __forceinline__ f3 fmul_t(const fm3& R, f3 u) {
    return mkf(R.a[0][0] * u.x + R.a[1][0] * u.y + R.a[2][0] * u.z,
               R.a[0][1] * u.x + R.a[1][1] * u.y + R.a[2][1] * u.z,
               R.a[0][2] * u.x + R.a[1][2] * u.y + R.a[2][2] * u.z);
}
// The same frame as contact_frame (normal -> local x), in FP32; near-degenerate normals
// (|a x ex| < 1e-6) take the reference's axis-aligned branch.
__device__ __forceinline__ fm3 contact_frame_f32(f3 a, float tol) {
    const f3 v = mkf(0.0f, a.z, -a.y);  // a x (1, 0, 0)
    const float s = fnorm(v);
    const float c = a.x;
    fm3 R;
    if (s < tol) {
        const float sg = c > 0.0f ? 1.0f : -1.0f;
        R.a[0][0] = sg;   R.a[0][1] = 0.0f; R.a[0][2] = 0.0f;
        R.a[1][0] = 0.0f; R.a[1][1] = 1.0f; R.a[1][2] = 0.0f;
        R.a[2][0] = 0.0f; R.a[2][1] = 0.0f; R.a[2][2] = sg;
        return R;
    }
    float K[3][3] = {{0.0f, -v.z, v.y}, {v.z, 0.0f, -v.x}, {-v.y, v.x, 0.0f}};
    const float f = (1.0f - c) * frcp(s * s);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float k2 = K[r][0] * K[0][q] + K[r][1] * K[1][q] + K[r][2] * K[2][q];
            R.a[r][q] = ((r == q) ? 1.0f : 0.0f) + K[r][q] + k2 * f;
        }
    return R;
}

// Neumaier-compensated FP32 accumulator (mixed precision force sums).
struct CSum3 {
    float s[3] = {0.f, 0.f, 0.f};
    float c[3] = {0.f, 0.f, 0.f};
    __device__ __forceinline__ void add(int k, float v) {
        const float t = s[k] + v;
        if (fabsf(s[k]) >= fabsf(v)) c[k] += (s[k] - t) + v;
        else c[k] += (v - t) + s[k];
        s[k] = t;
    }
    __device__ __forceinline__ void add(d3 v) { add(0, (float)v.x); add(1, (float)v.y); add(2, (float)v.z); }
    __device__ __forceinline__ void add(f3 v) { add(0, v.x); add(1, v.y); add(2, v.z); }
    __device__ __forceinline__ d3 value() const {
        return mk3((double)s[0] + (double)c[0], (double)s[1] + (double)c[1], (double)s[2] + (double)c[2]);
    }
};

// What one contact adds to its two grains, in its canonical orientation (i the lower file index,
// j the higher): the force on i (j receives its negation, which is what the reference computes for
// j bit for bit), the moments on i and j, and whether the FP64 sums or the compensated FP32 sums of
// mixed precision take them.  A bond that breaks adds nothing.
constexpr uint8_t CK_NONE = 0, CK_FP64 = 1, CK_FP32 = 2;
struct Contrib {
    d3 F, Ti, Tj;
    uint8_t kind = CK_NONE;
    bool broke = false, bonded = false, fatal = false;
};

// The contact of row entry `word` (kernels.jl:122-315): EBPM for a bond (kernels.jl:170-239),
// Hertz-Mindlin otherwise (kernels.jl:240-313).  The contact's state (a bond's F, ti, tj; a Hertz
// contact's spring s) is read from and written to the row entry at st[c * cs + e]; the contact
// point and a Hertz contact's local force are written only when write_forces is set.  plo: the
// row's particle is i; inv_r, inv_m are that particle's reciprocal radius and mass.
template <bool MIXED>
__device__ __forceinline__ Contrib eval_contact(const SceneP& S, double dt, uint32_t word, bool plo,
                                                d3 ki, d3 kj, d3 vi, d3 vj, d3 wi, d3 wj, double ri,
                                                double rj, double mi, double mj, double inv_r,
                                                double inv_m, double* __restrict__ st, int64_t cs,
                                                int64_t e, int write_forces) {
    Contrib o;
    const double Li = norm(kj - ki);
    if (word & BONDED) {
      if (MIXED) {
        // EBPM with the frame, stiffnesses, increments and stresses in FP32, added to the
        // FP64 force and moment state.  Relative motion comes from FP64 velocity
        // differences (bonded grains share most of their motion).  A stress within FP32
        // rounding of its strength, or a normal near -x, takes the FP64 law below.
        const f3 d32 = tof(kj - ki);
        const f3 a32 = fscl(d32, frcp((float)Li));
        if (a32.x >= -0.99f) {
            const fm3 R32 = contact_frame_f32(a32, 1e-6f);
            const float dtf = (float)dt;
            const f3 dd = fmul(R32, fscl(tof(vi - vj), dtf));    // d_i - d_j
            const f3 dth = fmul(R32, fscl(tof(wi - wj), dtf));   // theta_i - theta_j
            const f3 thi = fmul(R32, fscl(tof(wi), dtf)), thj = fmul(R32, fscl(tof(wj), dtf));
            const SurfaceP& sf = S.surf[0][0];
            const float rb = (float)sf.rho * (float)(ri < rj ? ri : rj);
            const float iL = frcp((float)Li);
            const float E = (float)sf.E, nu = (float)sf.nu;
            const float rb2 = rb * rb;
            const float Ib = rb2 * rb2 * (float)(M_PI / 4.0);
            const float phi = (20.0f / 3.0f) * rb2 * iL * iL * (1.0f + nu);
            const float Ab = rb2 * (float)M_PI;
            const float i1p = frcp(1.0f + phi);
            const float EI = E * Ib;
            const float k1 = E * Ab * iL;
            const float k2 = 12.0f * EI * iL * iL * iL * i1p;
            const float k3 = 6.0f * EI * iL * iL * i1p;
            const float k4 = EI * iL * frcp(1.0f + nu);
            const float k5 = EI * (4.0f + phi) * iL * i1p;
            const float k6 = EI * (2.0f - phi) * iL * i1p;
            const f3 dF = mkf(k1 * dd.x, k2 * dd.y + k3 * (thi.z + thj.z), k2 * dd.z - k3 * (thi.y + thj.y));
            const f3 dTi = mkf(k4 * dth.x, -k3 * dd.z + k5 * thi.y + k6 * thj.y, k3 * dd.y + k5 * thi.z + k6 * thj.z);
            const f3 dTj = mkf(-k4 * dth.x, -k3 * dd.z + k6 * thi.y + k5 * thj.y, k3 * dd.y + k6 * thi.z + k5 * thj.z);
            const d3 Fi = mk3(st[S_FX * cs + e], st[S_FY * cs + e], st[S_FZ * cs + e]) + tod(dF);
            const d3 Ti = mk3(st[S_TIX * cs + e], st[S_TIY * cs + e], st[S_TIZ * cs + e]) + tod(dTi);
            const d3 Tj = mk3(st[S_TJX * cs + e], st[S_TJY * cs + e], st[S_TJZ * cs + e]) + tod(dTj);
            const f3 F32 = tof(Fi), Ti32 = tof(Ti), Tj32 = tof(Tj);
            const float iAb = frcp(Ab), rbIb = rb * frcp(Ib);
            const float sci = -F32.x * iAb - rbIb * fsqrt(Ti32.y * Ti32.y + Ti32.z * Ti32.z);
            const float scj = -F32.x * iAb - rbIb * fsqrt(Tj32.y * Tj32.y + Tj32.z * Tj32.z);
            const float sc = -fminf(sci, scj), stt = fmaxf(sci, scj);
            const float ss = fabsf(Ti32.x) * rbIb * 0.5f + (4.0f / 3.0f) * iAb * fsqrt(F32.y * F32.y + F32.z * F32.z);
            const float SC = (float)sf.sc, ST = (float)sf.st, SS = (float)sf.ss;
            auto close = [](float a, float b) { return fabsf(a - b) <= 1e-4f * fmaxf(fabsf(a), fabsf(b)); };
            if (!(close(sc, SC) || close(stt, ST) || close(ss, SS))) {
                if (sc >= SC || stt >= ST || ss >= SS) {
                    o.broke = true;  // both copies take the same decision
                    return o;
                }
                o.kind = CK_FP32;
                o.F = tod(fmul_t(R32, fneg(F32)));
                o.Ti = tod(fmul_t(R32, fneg(Ti32)));
                o.Tj = tod(fmul_t(R32, fneg(Tj32)));
                const d3 kc = (ki + kj) * 0.5;
                if (write_forces) {  // the contact point is an output, not state
                    st[S_KX * cs + e] = kc.x; st[S_KY * cs + e] = kc.y; st[S_KZ * cs + e] = kc.z;
                }
                st[S_FX * cs + e] = Fi.x; st[S_FY * cs + e] = Fi.y; st[S_FZ * cs + e] = Fi.z;
                st[S_TIX * cs + e] = Ti.x; st[S_TIY * cs + e] = Ti.y; st[S_TIZ * cs + e] = Ti.z;
                st[S_TJX * cs + e] = Tj.x; st[S_TJY * cs + e] = Tj.y; st[S_TJZ * cs + e] = Tj.z;
                o.bonded = true;
                return o;
            }
        }
      }
        // EBPM bond (kernels.jl:170-239)
        const m3 R = contact_frame(normalize(kj - ki), S.tol);
        const d3 kc = (ki + kj) * 0.5;
        const d3 di = mul(R, vi * dt), dj = mul(R, vj * dt);
        const d3 thi = mul(R, wi * dt), thj = mul(R, wj * dt);
        const SurfaceP& sf = S.surf[0][0];
        const double rmin = ri < rj ? ri : rj;
        const double rb = sf.rho * rmin;
        const double L = Li;
        const double E = sf.E, nu = sf.nu;
        const double rb2 = rb * rb;
        const double Ib = rb2 * rb2 * M_PI / 4.0;
        const double phi = 20.0 / 3.0 * rb2 / (L * L) * (1.0 + nu);
        const double Ab = rb2 * M_PI;
        const double k1 = E * Ab / L;
        const double k2 = 12.0 * E * Ib / (L * L * L) / (1.0 + phi);
        const double k3 = 6.0 * E * Ib / (L * L) / (1.0 + phi);
        const double k4 = E * Ib / L / (1.0 + nu);
        const double k5 = E * Ib * (4.0 + phi) / L / (1.0 + phi);
        const double k6 = E * Ib * (2.0 - phi) / L / (1.0 + phi);
        const d3 dF = mk3(k1 * (di.x - dj.x), k2 * (di.y - dj.y) + k3 * (thi.z + thj.z),
                          k2 * (di.z - dj.z) - k3 * (thi.y + thj.y));
        const d3 dTi = mk3(k4 * (thi.x - thj.x), k3 * (dj.z - di.z) + k5 * thi.y + k6 * thj.y,
                           k3 * (di.y - dj.y) + k5 * thi.z + k6 * thj.z);
        const d3 dTj = mk3(k4 * (thj.x - thi.x), k3 * (dj.z - di.z) + k6 * thi.y + k5 * thj.y,
                           k3 * (di.y - dj.y) + k6 * thi.z + k5 * thj.z);
        const d3 Fi = mk3(st[S_FX * cs + e], st[S_FY * cs + e], st[S_FZ * cs + e]) + dF;
        const d3 Fj = -Fi;
        const d3 Ti = mk3(st[S_TIX * cs + e], st[S_TIY * cs + e], st[S_TIZ * cs + e]) + dTi;
        const d3 Tj = mk3(st[S_TJX * cs + e], st[S_TJY * cs + e], st[S_TJZ * cs + e]) + dTj;
        const double sci = Fj.x / Ab - rb / Ib * sqrt(Ti.y * Ti.y + Ti.z * Ti.z);
        const double scj = Fj.x / Ab - rb / Ib * sqrt(Tj.y * Tj.y + Tj.z * Tj.z);
        const double sc = -(sci < scj ? sci : scj);
        const double stt = sci > scj ? sci : scj;
        const double ss = fabs(Ti.x) * rb / 2.0 / Ib + 4.0 / 3.0 / Ab * sqrt(Fi.y * Fi.y + Fi.z * Fi.z);
        if (sc >= sf.sc || stt >= sf.st || ss >= sf.ss) {
            o.broke = true;  // both copies take the same decision
            return o;
        }
        const m3 Ri = inverse(R);
        // the force on j, mul(Ri, -Fj) with Fj = -Fi, is -o.F bit for bit
        o.kind = CK_FP64;
        o.F = mul(Ri, -Fi);
        o.Ti = mul(Ri, -Ti);
        o.Tj = mul(Ri, -Tj);
        if (write_forces) {  // the contact point is an output, not state
            st[S_KX * cs + e] = kc.x; st[S_KY * cs + e] = kc.y; st[S_KZ * cs + e] = kc.z;
        }
        st[S_FX * cs + e] = Fi.x; st[S_FY * cs + e] = Fi.y; st[S_FZ * cs + e] = Fi.z;
        st[S_TIX * cs + e] = Ti.x; st[S_TIY * cs + e] = Ti.y; st[S_TIZ * cs + e] = Ti.z;
        st[S_TJX * cs + e] = Tj.x; st[S_TJY * cs + e] = Tj.y; st[S_TJZ * cs + e] = Tj.z;
        // s of a bond stays at the zero it was created (or restored) with; kernels.jl:230-238 stores zero
        o.bonded = true;
        return o;
    } else {
      if (MIXED) {
        // Hertz-Mindlin with FP32 geometry and law; overlap from the FP64 distance.
        // Contacts it cannot decide like FP64 fall through to the FP64 law below.
        const double gap = Li - ri - rj;
        const double dn = fabs(gap);
        const f3 d32 = tof(kj - ki);
        const f3 a32 = fscl(d32, frcp((float)Li));
        // near a = -x the reference's frame is ill-conditioned (its rotation about the
        // normal follows the normal's tiny transverse components): FP64 there
        bool fp64 = a32.x < -0.99f;
        const fm3 R32 = contact_frame_f32(a32, 1e-6f);
        const float lever = (float)(ri - dn);
        const f3 kip = fscl(a32, lever);
        const f3 kjp = fsub(kip, d32);
        // (w_j x k_j' + v_j) - (w_i x k_i' + v_i) with v_j - v_i taken in FP64: grains in
        // contact share most of their velocity, and subtracting FP32 copies of two fast
        // grains' velocities would lose the small relative part
        const f3 dv = tof(vj - vi);
        const f3 vc = fmul(R32, fadd(fsub(fcross(tof(wj), kjp), fcross(tof(wi), kip)), dv));
        const float Rs = frcp(frcp((float)ri) + frcp((float)rj));
        const float ms = frcp(frcp((float)mi) + frcp((float)mj));
        const float srd = fsqrt(Rs * (float)dn);
        const float kn = (4.0f / 3.0f) * S.ys_pp * srd;
        const float kt = 8.0f * S.gs_pp * srd;
        const float gn = -2.0f * S.beta_pp * fsqrt((5.0f / 6.0f) * (2.0f * S.ys_pp * srd) * ms);
        const float gt = -2.0f * S.beta_pp * fsqrt((5.0f / 6.0f) * kt * ms);
        // the tangential spring accumulates in FP64: v dt increments vanish in FP32
        d3 sv = mk3(st[S_SX * cs + e], st[S_SY * cs + e], st[S_SZ * cs + e]);
        sv = sv + mk3((double)vc.x * 0.0, (double)vc.y * dt, (double)vc.z * dt);
        const float F1 = -kn * (float)gap - gn * vc.x;
        const f3 Ft = fscl(tof(sv), -kt);
        const float nFt = fnorm(Ft);
        // a slip decision within FP32 rounding of the threshold is taken in FP64: the
        // reference's slip rule reverses the stored spring, so a flipped decision is an
        // O(1) change of the tangential force
        const float thr = S.mu_pp * F1;
        fp64 = fp64 || fabsf(nFt - thr) <= 1e-4f * fmaxf(fabsf(nFt), fabsf(thr));
        float F2, F3;
        if (fp64) {
            F2 = F3 = 0.0f;
        } else if (nFt >= thr) {
            const float ratio = S.mu_pp * F1 * frcp(nFt);
            F2 = ratio * Ft.y;
            F3 = ratio * Ft.z;
            const float ikt = frcp(kt);
            sv = mk3(sv.x, (double)(F2 * ikt), (double)(F3 * ikt));
        } else {
            F2 = Ft.y - gt * vc.y;
            F3 = Ft.z - gt * vc.z;
        }
        if (!fp64) {
        const f3 Fl = mkf(F1, F2, F3);
        const f3 Fg = fmul_t(R32, mkf(-F1, -F2, -F3));
        o.kind = CK_FP32;
        o.F = tod(Fg);
        o.Ti = tod(fcross(kip, Fg));
        o.Tj = tod(fcross(kjp, mkf(-Fg.x, -Fg.y, -Fg.z)));
        const d3 kc = ki + tod(kip);
        if (write_forces) {  // contact point and local force are outputs, not state
            st[S_KX * cs + e] = kc.x; st[S_KY * cs + e] = kc.y; st[S_KZ * cs + e] = kc.z;
            st[S_FX * cs + e] = (double)Fl.x; st[S_FY * cs + e] = (double)Fl.y; st[S_FZ * cs + e] = (double)Fl.z;
        }
        st[S_SX * cs + e] = sv.x; st[S_SY * cs + e] = sv.y; st[S_SZ * cs + e] = sv.z;
        o.fatal = gap > 0 || gn < 0.0f || gt < 0.0f;
        return o;
        }
      }
      {
        // Hertz-Mindlin (kernels.jl:240-313)
        const m3 R = contact_frame(normalize(kj - ki), S.tol);
        const double gap = Li - ri - rj;
        const double dn = fabs(gap);
        const d3 kc = ki + normalize(kj - ki) * (ri - dn);
        const d3 kip = kc - ki, kjp = kc - kj;
        const d3 vci = cross(wi, kip) + vi;
        const d3 vcj = cross(wj, kjp) + vj;
        const d3 vc = mul(R, vcj - vci);
        const SurfaceP& sf = S.surf[0][0];
        // 1/r_i + 1/r_j with this particle's reciprocal taken once per row (IEEE addition
        // commutes, so the sum is the canonical one bit for bit)
        const double Rs = 1.0 / (inv_r + 1.0 / (plo ? rj : ri));
        const double ms = 1.0 / (inv_m + 1.0 / (plo ? mj : mi));
        const d3 s0 = mk3(st[S_SX * cs + e], st[S_SY * cs + e], st[S_SZ * cs + e]);
        const HMOut h = hertz_mindlin(gap, dn, vc, S.ysm[0], S.gsm[0], Rs, ms, sf, dt, s0);
        const m3 Ri = inverse(R);
        const d3 Fg = mul(Ri, -h.F);
        o.kind = CK_FP64;
        o.F = Fg;
        o.Ti = cross(kip, Fg);
        o.Tj = cross(kjp, -Fg);
        if (write_forces) {  // contact point and local force are outputs, not state
            st[S_KX * cs + e] = kc.x; st[S_KY * cs + e] = kc.y; st[S_KZ * cs + e] = kc.z;
            st[S_FX * cs + e] = h.F.x; st[S_FY * cs + e] = h.F.y; st[S_FZ * cs + e] = h.F.z;
        }
        st[S_SX * cs + e] = h.s.x; st[S_SY * cs + e] = h.s.y; st[S_SZ * cs + e] = h.s.z;
        o.fatal = gap > 0 || h.fatal;
        return o;
      }
    }
}

// The walls particle at kp with radius r touches (the test of kernels.jl:331, |k . n - d| < r),
// bit j for wall j.  The planes are read four at a time, so their loads wait once per group
// rather than once per wall (a wait per wall cost the 1M pour's force kernel 7.5 us of 68).
__device__ __forceinline__ uint32_t walls_touched(const SceneP& S, d3 kp, double r) {
    uint32_t m = 0u;
    const int nw = S.nwall;
    for (int j0 = 0; j0 < nw; j0 += 8) {
        WallP w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j0 + u < nw) w[u] = S.walls[j0 + u];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j0 + u < nw && fabs(dot(kp, mk3(w[u].n[0], w[u].n[1], w[u].n[2])) - w[u].d) < r)
                m |= 1u << (j0 + u);
    }
    return m;
}

// What a wave of a force kernel counts for the step's counters and the near-list check.
struct ForceTally {
    uint64_t broken = 0, nwallc = 0, fatal = 0, active = 0, active_b = 0;
    bool moved = false;  // a particle's new position is delta or more from its rebuild position
    float disp = 0.0f;   // and its distance from there (an upper bound), the largest of the lane's
};

// The state arrays a force kernel reads and writes.
struct ForceArgs {
    const double4* __restrict__ pos; const double4* __restrict__ vel;
    const double4* __restrict__ omg; const double4* __restrict__ quat;
    double4* __restrict__ pos_n; double4* __restrict__ vel_n;
    double4* __restrict__ omg_n; double4* __restrict__ quat_n;
    double4* __restrict__ acc; double4* __restrict__ dwo;
    double4* __restrict__ ftot; double4* __restrict__ ttot;
    uint8_t* __restrict__ sel;
    const uint32_t* __restrict__ rcnt0; const uint32_t* __restrict__ rcnt1;
    uint32_t* __restrict__ pq0; uint32_t* __restrict__ pq1;
    double* __restrict__ st0; double* __restrict__ st1;
    uint32_t* __restrict__ wact; double4* __restrict__ wshear;
    float4* __restrict__ xref;
};

// apply_body_force! (kernels.jl:428-443) and update! (kernels.jl:445-485) of particle p from the
// summed contact and wall force F and moment T; the near-list displacement check; the next state.
// Q and xr: the orientation and rebuild position, already loaded, or loaded here (LATE).
template <bool LATE>
__device__ __forceinline__ void integrate_particle(const SceneP& S, const ForceArgs& A, uint32_t p,
                                                   const double4& X, const double4& Vv,
                                                   const double4& W, double4 Q, float4 xr, d3 F,
                                                   d3 T, int write_forces, int skin_cur,
                                                   bool rebuilt, ForceTally& t) {
    const double dt = S.dt;
    const d3 kp = mk3(X.x, X.y, X.z), vp = mk3(Vv.x, Vv.y, Vv.z), wp = mk3(W.x, W.y, W.z);
    F = F + mk3(S.g[0] * Vv.w, S.g[1] * Vv.w, S.g[2] * Vv.w);
    const double nd = -S.damping;
    F = F + mk3(nd * fabs(F.x) * jsign(vp.x), nd * fabs(F.y) * jsign(vp.y), nd * fabs(F.z) * jsign(vp.z));
    T = T + mk3(nd * fabs(T.x) * jsign(wp.x), nd * fabs(T.y) * jsign(wp.y), nd * fabs(T.z) * jsign(wp.z));
    if (write_forces) {
        A.ftot[p] = make_double4(F.x, F.y, F.z, 0.0);
        A.ttot[p] = make_double4(T.x, T.y, T.z, 0.0);
    }
    // update! (kernels.jl:445-485)
    if constexpr (LATE) Q = A.quat[p];
    const d3 av = F / Vv.w;
    const d3 knew = (kp + vp * dt) + (0.5 * av) * (dt * dt);
    const d3 vnew = vp + av * dt;
    const m3 Rq = quat_matrix(Q);
    const d3 Tl = mul(Rq, T);
    const d3 wl = mul(Rq, wp);
    m3 I;
    I.a[0][0] = W.w; I.a[0][1] = 0.0; I.a[0][2] = 0.0;
    I.a[1][0] = 0.0; I.a[1][1] = W.w; I.a[1][2] = 0.0;
    I.a[2][0] = 0.0; I.a[2][1] = 0.0; I.a[2][2] = W.w;
    const m3 Iinv = inverse(I);
    const d3 dwl = mul(Iinv, Tl - cross(wl, mul(I, wl)));
    const m3 Rqi = inverse(Rq);
    const d3 dw = mul(Rqi, dwl);
    const double q0 = Q.x, q1 = Q.y, q2 = Q.z, q3 = Q.w;
    const double dq0 = -0.5 * (q1 * wp.x + q2 * wp.y + q3 * wp.z);
    const double dq1 = 0.5 * (q0 * wp.x + q2 * wp.z - q3 * wp.y);
    const double dq2 = 0.5 * (q0 * wp.y + q1 * wp.z + q3 * wp.x);
    const double dq3 = 0.5 * (q0 * wp.z + q1 * wp.y - q2 * wp.x);
    const double n0 = q0 + dq0, n1 = q1 + dq1, n2 = q2 + dq2, n3 = q3 + dq3;
    const double qn = sqrt(n0 * n0 + n1 * n1 + n2 * n2 + n3 * n3);
    const d3 wnew = wp + dw * dt;
    if (skin_cur >= 0) {
        // |x_new - x_rebuild| < delta, with x_rebuild rounded to FP32: the bound is reduced by
        // that rounding (at most 2^-24 |x| per axis) and by 1e-6 relative for the FP64 sums
        if (rebuilt) {
            xr = make_float4((float)X.x, (float)X.y, (float)X.z, 0.f);
            A.xref[p] = xr;
        } else if (LATE) {
            xr = A.xref[p];
        }
        const double ex = knew.x - (double)xr.x, ey = knew.y - (double)xr.y,
                     ez = knew.z - (double)xr.z;
        const double er = 0x1p-22 * (fabs((double)xr.x) + fabs((double)xr.y) + fabs((double)xr.z));
        const double chk = S.skin * (1.0 - 1e-6) - er;
        const double d2 = ex * ex + ey * ey + ez * ez;
        // A grain whose position is not finite (the reference's exactly normal impacts make
        // NaN grains) touches nothing in any step, so it never needs a rebuild.
        const bool finite = isfinite(knew.x) && isfinite(knew.y) && isfinite(knew.z);
        t.moved |= finite && (!(chk > 0.0) || !(d2 < chk * chk));
        // the displacement bound for the next step, at least sqrt(d2) + er: FP32 from values
        // rounded up, each later rounding (at most an ulp, 2^-23) covered by a factor 1 + 2^-20
        if (finite) {
            const float up = 1.0f + 0x1p-20f;
            const float b = (sqrtf(__double2float_ru(d2)) * up + __double2float_ru(er)) * up;
            t.disp = fmaxf(t.disp, b);
        }
    }
    // streaming stores: the next state is not read again before this launch ends
    store_nt(A.pos_n + p, make_double4(knew.x, knew.y, knew.z, X.w));
    store_nt(A.vel_n + p, make_double4(vnew.x, vnew.y, vnew.z, Vv.w));
    store_nt(A.omg_n + p, make_double4(wnew.x, wnew.y, wnew.z, W.w));
    store_nt(A.quat_n + p, make_double4(n0 / qn, n1 / qn, n2 / qn, n3 / qn));
    if (write_forces) {  // a and dw are outputs of update!; the step never reads them back
        A.acc[p] = make_double4(av.x, av.y, av.z, 0.0);
        A.dwo[p] = make_double4(dw.x, dw.y, dw.z, 0.0);
    }
}

// A particle's own state as the force kernels load it first: position, velocity, spin, and unless
// LATE also the orientation, the rebuild position and the wall flags (one memory round trip less
// per wave).  LATE (mixed precision on contact-dense states) loads those after the row, where
// they hold no registers during the contact loop: the mixed kernel is at its register limit, and
// with them live across the loop it spills (packed 1M pour: 249 -> 267 us).
struct Own {
    double4 X, Vv, W, Q;
    float4 xr;    // the position at the last rebuild
    uint32_t wa;  // the walls this particle touched last step
};
template <bool LATE>
__device__ __forceinline__ Own load_own(const SceneP& S, const ForceArgs& A, uint32_t p,
                                        int skin_cur, bool rebuilt) {
    Own o;
    o.X = A.pos[p]; o.Vv = A.vel[p]; o.W = A.omg[p];
    o.Q = make_double4(0.0, 0.0, 0.0, 0.0);
    o.xr = make_float4(0.f, 0.f, 0.f, 0.f);
    o.wa = 0u;
    if constexpr (!LATE) {
        o.Q = A.quat[p];
        // loaded whether or not the step rebuilt: no state load waits for the step's flag word
        if (skin_cur >= 0) o.xr = A.xref[p];
        if (S.nwall) o.wa = A.wact[p];
    }
    return o;
}

// The one-kernel lean step (k_lean_free): every merge block of the step is done and the step has
// not failed, so a force wave may write wall history.  The wait is bounded: merge blocks precede
// the force blocks in the grid and finish without waiting on anything, and on time-out the step
// fails (CT_ABORT bit 2), which the host replays with separate launches.
// The flags and counts are device-scope atomics throughout (a merge wave's flag atomics have
// completed before its block is counted done), so no fence is needed; an agent-scope fence here
// writes back and invalidates the XCD's L2, which cost the step 18 times its time.
constexpr int GATE_SPIN = 4096;
__device__ __forceinline__ bool merge_gate(unsigned long long* ctr, uint32_t nmb) {
    const uint32_t ncls = nmb < (uint32_t)CTR_COPIES ? nmb : (uint32_t)CTR_COPIES;
    for (int i = 0; i < GATE_SPIN; ++i) {
        if (__hip_atomic_load(&ctr[CT_MDONE + 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= ncls)
            return __hip_atomic_load(&ctr[CT_FAIL], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull;
        __builtin_amdgcn_s_sleep(4);
    }
    atomicOr(&ctr[CT_ABORT], 2ull);
    atomicOr(&ctr[CT_FAIL], 1ull);
    return false;
}

// Block b of nblocks is done: counted in class b % CTR_COPIES (that copy's word idx), since atomics
// on one address serialise (one count per block of a million-particle launch cost the step four
// times its time); the block that completes its class resets the class's word and counts the
// class in copy 0's word idx + 2.  True for the block that completes the last class.
__device__ __forceinline__ bool count_done(unsigned long long* ctr, int idx, uint32_t b, uint32_t nblocks) {
    const uint32_t c = b % (uint32_t)CTR_COPIES;
    const uint32_t ncls = nblocks < (uint32_t)CTR_COPIES ? nblocks : (uint32_t)CTR_COPIES;
    const uint32_t in_cls = (nblocks - c + CTR_COPIES - 1) / CTR_COPIES;
    unsigned long long* w = ctr + (size_t)c * CTR_STRIDE + idx;
    if (__hip_atomic_fetch_add(w, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != in_cls - 1) return false;
    __hip_atomic_store(w, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __hip_atomic_fetch_add(&ctr[idx + 2], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == ncls - 1;
}

// Contacts (kernels.jl:122-315), walls (kernels.jl:317-426), body force and update! of particle p,
// whose own state `own` is loaded.  failed: a row of this step or an earlier one overflowed, so
// nothing changes (tested before the first store, so the flag's load hides behind the state
// loads); rebuilt: near lists in use (skin_cur >= 0) and this step rebuilt them.  FREE: the
// contact-free mode, in which no particle has a row (the merge fails a step that would give one a
// row), so the row is neither read nor walked and the kernel needs few registers.  GATE: the force
// blocks of a one-kernel lean step, whose wave writes wall history only once the step's gate_n
// merge blocks are done and the step has not failed.
template <bool MIXED, bool SEG, bool LATE, bool WALLS_LATE, bool FREE = false, bool GATE = false>
__device__ __forceinline__ void force_particle(const Params& P, const ForceArgs& A, uint32_t p,
                                               const Own& own, int write_forces, int skin_cur,
                                               bool failed, bool rebuilt,
                                               unsigned long long* ctr, ForceTally& t,
                                               uint32_t gate_n = 0u) {
    const int64_t n = P.n;
    const SceneP& S = scene_of<SEG>(P, p);
    const double dt = S.dt;
    const double4 X = own.X, Vv = own.Vv, W = own.W;
    const double4 Q = own.Q;
    const float4 xr = own.xr;
    uint32_t wa = own.wa;
    const d3 kp = mk3(X.x, X.y, X.z), vp = mk3(Vv.x, Vv.y, Vv.z), wp = mk3(W.x, W.y, W.z);
    // the walls touched now: on sparse states tested before the row so the planes' loads overlap
    // the row's; on contact-dense states (WALLS_LATE) after it, holding no register across the
    // contact loop (mixed packed 1M pour: 248 against 222 us)
    const uint32_t wtouch_early = WALLS_LATE ? 0u : walls_touched(S, kp, X.w);
    d3 F = mk3(0.0, 0.0, 0.0), T = mk3(0.0, 0.0, 0.0);
    CSum3 FH, TH;  // mixed precision: compensated FP32 sums of the Hertz contributions
    // commit a row the merge rebuilt this step (it sits in the other buffer)
    // both row counts with the selector (independent loads), the committed one chosen after
    const uint8_t sl = FREE ? 0 : A.sel[p];
    const uint32_t c0 = FREE ? 0u : A.rcnt0[p], c1 = FREE ? 0u : A.rcnt1[p];
    const int b = (sl & ROW_PENDING) ? ((sl & 1) ^ 1) : sl;
    if (failed) return;
    if (sl & ROW_PENDING) A.sel[p] = (uint8_t)b;
    uint32_t* pq = b ? A.pq1 : A.pq0;
    double* st = b ? A.st1 : A.st0;
    const uint32_t cnt = FREE ? 0u : (b ? c1 : c0);
    const int64_t cs = (int64_t)P.K * n;
    const double inv_r = cnt ? 1.0 / X.w : 0.0, inv_m = cnt ? 1.0 / Vv.w : 0.0;
    // software pipeline over the row: entry m+1's partner state and entry m+2's word are
    // loaded while entry m is evaluated
    uint32_t w_cur = 0u, w_nxt = 0u;
    double4 Xq_c = make_double4(0.0, 0.0, 0.0, 0.0), Vq_c = Xq_c, Wq_c = Xq_c;
    if (cnt > 0) {
        w_cur = pq[p];
        if (cnt > 1) w_nxt = pq[n + p];
        const uint32_t q0 = w_cur & IDX_MASK;
        Xq_c = A.pos[q0]; Vq_c = A.vel[q0]; Wq_c = A.omg[q0];
    }
    for (uint32_t m = 0; m < cnt; ++m) {
        const int64_t e = (int64_t)m * n + p;
        const uint32_t word = w_cur;
        const double4 Xq = Xq_c, Vq = Vq_c, Wq = Wq_c;
        if (m + 1 < cnt) {
            const uint32_t qn = w_nxt & IDX_MASK;
            Xq_c = A.pos[qn]; Vq_c = A.vel[qn]; Wq_c = A.omg[qn];
            w_cur = w_nxt;
            if (m + 2 < cnt) w_nxt = pq[(int64_t)(m + 2) * n + p];
        }
        const bool plo = (word & IS_I) != 0u;  // p is the contact's i (the lower file index)
        // canonical orientation: i is the lower file index, j the higher
        const d3 ki = plo ? kp : mk3(Xq.x, Xq.y, Xq.z);
        const d3 kj = plo ? mk3(Xq.x, Xq.y, Xq.z) : kp;
        const d3 vi = plo ? vp : mk3(Vq.x, Vq.y, Vq.z);
        const d3 vj = plo ? mk3(Vq.x, Vq.y, Vq.z) : vp;
        const d3 wi = plo ? wp : mk3(Wq.x, Wq.y, Wq.z);
        const d3 wj = plo ? mk3(Wq.x, Wq.y, Wq.z) : wp;
        const double ri = plo ? X.w : Xq.w, rj = plo ? Xq.w : X.w;
        const double mi = plo ? Vv.w : Vq.w, mj = plo ? Vq.w : Vv.w;
        const Contrib o = eval_contact<MIXED>(S, dt, word, plo, ki, kj, vi, vj, wi, wj, ri, rj, mi,
                                              mj, inv_r, inv_m, st, cs, e, write_forces);
        if (o.broke) {
            pq[e] = word | BROKEN;  // both copies take the same decision
            if (plo) {
                ++t.broken;
                if (SEG) atomicAdd(&P.seg_broken[P.sid[p]], 1ull);
            }
            continue;
        }
        if (o.kind == CK_FP32) {
            if (plo) { FH.add(tof(o.F)); TH.add(tof(o.Ti)); }
            else { FH.add(tof(-o.F)); TH.add(tof(o.Tj)); }
        } else {
            if (plo) { F = F + o.F; T = T + o.Ti; }
            else { F = F + (-o.F); T = T + o.Tj; }
        }
        if (plo) {
            ++t.active;
            if (o.bonded) ++t.active_b;
            if (o.fatal) ++t.fatal;
        }
    }
    if (MIXED) {
        F = F + FH.value();
        T = T + TH.value();
    }
    // walls (kernels.jl:317-426): two-sided planes, shear history reset when not touching
    if (LATE && S.nwall) wa = A.wact[p];
    const uint32_t wtouch = WALLS_LATE ? walls_touched(S, kp, X.w) : wtouch_early;
    if constexpr (GATE)  // wall history is written in place (a failed step must leave it)
        if (__ballot(wtouch != 0u || wa != 0u) != 0ull && !merge_gate(ctr, gate_n)) return;
    for (uint32_t wm = wtouch; wm; wm &= wm - 1u) {  // in wall order
        const int j = __ffs(wm) - 1;
        const WallP& wl = S.walls[j];
        const d3 nrm = mk3(wl.n[0], wl.n[1], wl.n[2]);
        {
            ++t.nwallc;
            const m3 R = contact_frame(nrm, S.tol);
            const double L = dot(kp, nrm) - wl.d;
            const double gap = fabs(L) - X.w;
            const double dn = fabs(gap);
            const d3 kip = (((-L) * nrm) / fabs(L)) * (fabs(L) + dn / 2.0);
            const d3 vci = cross(wp, kip) + vp;
            const d3 vc = mul(R, -vci);
            const int mj = wl.mid - 1;
            const SurfaceP& sf = S.surf[0][mj];
            d3 s0 = mk3(0.0, 0.0, 0.0);
            if (wa & (1u << j)) {
                const double4 ws = A.wshear[(int64_t)j * n + p];
                s0 = mk3(ws.x, ws.y, ws.z);
            }
            const HMOut h = hertz_mindlin(gap, dn, vc, S.ysm[mj], S.gsm[mj], X.w, Vv.w, sf, dt, s0);
            if (h.fatal) ++t.fatal;
            const m3 Ri = inverse(R);
            const d3 Fg = mul(Ri, -h.F);
            F = F + Fg;
            T = T + cross(kip, Fg);
            A.wshear[(int64_t)j * n + p] = make_double4(h.s.x, h.s.y, h.s.z, 0.0);
        }
    }
    if (S.nwall && wtouch != wa) A.wact[p] = wtouch;
    integrate_particle<LATE>(S, A, p, X, Vv, W, Q, xr, F, T, write_forces, skin_cur, rebuilt, t);
}

// A force kernel's counters, per wave (every lane calls it).  done: this launch ends the step
// (counts it, and its rebuild).
__device__ __forceinline__ void force_counters(unsigned long long* ctr, int skin_cur, bool rebuilt,
                                               int write_forces, bool done, ForceTally& t) {
    if (done && blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&ctr[CT_DONE], 1ull);
        if (rebuilt) atomicAdd(&ctr[CT_REBUILDS], 1ull);
    }
    if (skin_cur >= 0) {
        if (__ballot(t.moved) != 0ull && (threadIdx.x & 63) == 0) skin_request(ctr, skin_cur);
        float disp = t.disp;
        for (int o = 32; o > 0; o >>= 1) disp = fmaxf(disp, __shfl_xor(disp, o, 64));
        if ((threadIdx.x & 63) == 0 && disp > 0.0f)
            atomicMax(ctr_at(ctr, CT_DISP + 1 - skin_cur), (unsigned long long)__float_as_uint(disp));
    }
    wave_add(ctr_at(ctr, CT_BROKEN), t.broken);
    wave_add(ctr_at(ctr, CT_BROKEN_TOTAL), t.broken);
    if (write_forces) {
        wave_add(ctr_at(ctr, CT_WALL), t.nwallc);
        wave_add(ctr_at(ctr, CT_FATAL), t.fatal);
        wave_add(ctr_at(ctr, CT_ACTIVE), t.active);
        wave_add(ctr_at(ctr, CT_ACTIVE_BOND), t.active_b);
    }
}

// The mixed kernel carries both contact laws in two precisions; left alone the compiler gives the
// single-scene variants and the batched sparse-state variant more than 256 registers and one wave
// per SIMD, so they are held to two waves per SIMD.  The other variants fit in 256 and keep the
// schedule the compiler picks without the bound.
// BS threads a block.  With contacts, rows differ in length and a one-wave block frees its slot as
// soon as its wave is done (packed 1M pour: 269 us at 256 threads, 247 at 64); without, 256-thread
// blocks stream the state 1 % faster.  The host picks by the last counted row entries per particle.
// FREE (single scenes, 256 threads): the contact-free mode, no row read or walked.
template <bool MIXED, bool SEG, int BS, bool FREE = false>
__global__ void __launch_bounds__(BS)
__attribute__((amdgpu_waves_per_eu(FREE ? 3 : (MIXED && (!SEG || BS == 256) ? 2 : 1))))
k_force(const Params P, const ForceArgs A, unsigned long long* ctr, int write_forces, int skin_cur) {
    static_assert(!FREE || (!SEG && BS == 256), "contact-free mode: sparse single scenes");
    const bool failed = step_failed(ctr);
    // near lists in use (skin_cur >= 0): did this step rebuild them?
    const bool rebuilt = skin_cur >= 0 && skin_rebuilds(ctr, skin_cur);
    // XCD-contiguous blocks: partners are spatial neighbours, so their state is in this XCD's L2
    const int64_t p64 = (int64_t)xcd_map<SEG>(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    ForceTally t;
    if (p64 < P.n) {
        const uint32_t p = (uint32_t)p64;
        // mixed precision on contact-dense states (one-wave blocks) loads part of the state late
        constexpr bool LATE = MIXED && BS == 64;
        const Own own = load_own<LATE>(scene_of<SEG>(P, p), A, p, skin_cur, rebuilt);
        force_particle<MIXED, SEG, LATE, BS == 64, FREE>(P, A, p, own, write_forces, skin_cur, failed, rebuilt, ctr,
                                                         t);
    }
    if (failed) return;
    force_counters(ctr, skin_cur, rebuilt, write_forces, true, t);
}

// A lean step of a single scene in the contact-free mode, as one launch.  The contact-free force
// kernel reads nothing the merge writes (no row), so the first nmb blocks test the near lists as
// the lean merge does (a wave per 64-grain group, groups left out by their gskip value) while the
// other blocks run the force kernel: the merge's launch and chains of dependent loads overlap the
// state streaming instead of preceding it.  What the order of two launches gave is kept:
// * a merge that finds a contact fails the step (CT_ABORT bit 2 with CT_FAIL) after force blocks
//   may have written the next state; that buffer is not the step's input, so the host's replay
//   from the step starts from its state unchanged.  Wall history is written in place, so waves
//   with a wall contact pass merge_gate first;
// * the step is counted (CT_DONE) by the block that finishes last, if the step did not fail;
// * merge blocks read this step's rebuild flag and displacement bound while force blocks write
//   the next step's, so no block clears the next step's words: the last block clears this
//   step's (those of the step after next; a failed step keeps them for its replay), and a run of
//   these steps starts with k_lean_free_prologue.
// The grid is nmb merge blocks, then force blocks (both counts multiples of 8, so force block f
// runs on XCD f % 8 as xcd_block assumes).  The merge blocks are few (at most 128): wave w of them
// takes the groups g = w (mod its count), reads 64 of their gskip values at once and merges the
// groups that stay one after another, so the merge holds few of the slots the force blocks stream
// the state with.
template <bool MIXED>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(3)))
k_lean_free(const Params P, const MergeArgs M, const ForceArgs A, unsigned long long* ctr, int skin_cur,
            uint32_t nmb) {
    __shared__ uint32_t lt[MERGE_LDS * TPB];
    __shared__ uint32_t lw[MERGE_LDS * TPB];
    __shared__ int s_last;
    if (blockIdx.x < nmb) {
        const MergeStep st = merge_step_begins(ctr, skin_cur, true, true, TPB, false);
        if (st.go) {
            const int lane = threadIdx.x & 63;
            const uint32_t nw = nmb * (TPB / 64), w = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);
            const uint32_t ng = (uint32_t)((P.n + 63) / 64);
            for (uint32_t g0 = w; g0 < ng; g0 += 64u * nw) {
                const uint64_t gl = (uint64_t)g0 + (uint64_t)lane * nw;
                const float gs = gl < ng ? __uint_as_float(M.gskip[gl]) : __builtin_inff();
                for (uint64_t stay = __ballot(gl < ng && gs <= st.twoD); stay; stay &= stay - 1ull) {
                    const uint32_t g = g0 + (uint32_t)(__ffsll((unsigned long long)stay) - 1) * nw;
                    const int64_t p64 = ((int64_t)g << 6) + lane;
                    const bool in = p64 < P.n;
                    const uint32_t p = in ? (uint32_t)p64 : 0u;
                    const MergeIn mi = load_merge_in(M, p, true, true);
                    MergeTally t;
                    bool idle = true;
                    if (in)
                        idle = merge_particle<false, false, TPB, false>(P, M, p, mi, lt + threadIdx.x,
                                                                        lw + threadIdx.x, true, st.twoD, t, true);
                    merge_counters(ctr, false, true, t, true);
                    float v = !in ? __builtin_inff() : (idle && mi.gm > 0.0f) ? mi.gm : 0.0f;
                    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
                    if (lane == 0) M.gskip[g] = __float_as_uint(v);
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0);  // this wave's flag atomics are done
        __syncthreads();
        if (threadIdx.x == 0) count_done(ctr, CT_MDONE, blockIdx.x, nmb);
    } else {
        const uint32_t f = blockIdx.x - nmb;
        const int64_t p64 = (int64_t)xcd_block(f, gridDim.x - nmb) * TPB + threadIdx.x;
        // a step that has to rebuild fails in its merge blocks
        const bool failed = step_failed(ctr) || skin_rebuilds(ctr, skin_cur);
        ForceTally t;
        if (p64 < P.n) {
            const uint32_t p = (uint32_t)p64;
            const Own own = load_own<false>(scene_of<false>(P, p), A, p, skin_cur, false);
            force_particle<MIXED, false, false, false, true, true>(P, A, p, own, 0, skin_cur, failed, false, ctr,
                                                                   t, nmb);
        }
        if (!failed) force_counters(ctr, skin_cur, false, 0, false, t);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) s_last = count_done(ctr, CT_TICKET, blockIdx.x, gridDim.x);
    __syncthreads();
    if (s_last) {
        if (__hip_atomic_load(&ctr[CT_FAIL], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull) {
            for (int c = threadIdx.x; c < CTR_COPIES; c += TPB) ctr[c * CTR_STRIDE + CT_DISP + skin_cur] = 0ull;
            if (threadIdx.x == 0) {
                ctr[CT_SKIN + skin_cur] = 0ull;
                atomicAdd(&ctr[CT_DONE], 1ull);
            }
        }
        if (threadIdx.x == 0) {
            ctr[CT_MDONE + 2] = 0ull;
            ctr[CT_TICKET + 2] = 0ull;
        }
    }
}

// Before a run of one-kernel lean steps: the next step's words (a lean step of two launches
// leaves them for its successor to clear) and the block counts.
__global__ void __launch_bounds__(64) k_lean_free_prologue(unsigned long long* ctr, int skin_cur) {
    for (int c = threadIdx.x; c < CTR_COPIES; c += 64) {
        ctr[c * CTR_STRIDE + CT_DISP + 1 - skin_cur] = 0ull;
        ctr[c * CTR_STRIDE + CT_TICKET] = 0ull;
        ctr[c * CTR_STRIDE + CT_MDONE] = 0ull;
    }
    if (threadIdx.x == 0) {
        ctr[CT_SKIN + 1 - skin_cur] = 0ull;
        ctr[CT_TICKET + 2] = 0ull;
        ctr[CT_MDONE + 2] = 0ull;
    }
}

// After a bond pass (rollback = 0): commit the rows it rebuilt.  After a failed step
// (rollback = 1): keep the committed rows, clear the touching counts the skipped merges left, and
// clear the failure flags (copy 0 of the flags, the lean-step abort; every copy of the overflow
// count).
__global__ void __launch_bounds__(TPB)
k_rows_finish(int64_t n, uint8_t* __restrict__ sel, uint32_t* __restrict__ tcnt,
              unsigned long long* __restrict__ ctr, int rollback) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p < n) {
        const uint8_t sl = sel[p];
        if (sl & ROW_PENDING) sel[p] = rollback ? (uint8_t)(sl & 1) : (uint8_t)((sl & 1) ^ 1);
        if (rollback) tcnt[p] = 0u;
    }
    if (rollback && blockIdx.x == 0) {
        for (int c = threadIdx.x; c < CTR_COPIES; c += TPB) ctr[c * CTR_STRIDE + CT_OVERFLOW] = 0ull;
        if (threadIdx.x == 0) { ctr[CT_FAIL] = 0ull; ctr[CT_MAXROW] = 0ull; ctr[CT_ABORT] = 0ull; }
    }
}

// ------------------------------------------------------------------------------------------
// read-back, state restore and intermediate arrays
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(TPB)
k_debug_cells(const Params P, const double4* __restrict__ pos, int32_t* kid, uint32_t* hid,
              double* kc, uint32_t* nbr) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= P.n) return;
    const double4 x = pos[p];
    uint32_t w[8];
    int32_t c[3];
    double ctr[3];
    window_of(P.sc.grid, mk3(x.x, x.y, x.z), w, c, ctr);
    for (int a = 0; a < 3; ++a) { kid[3 * p + a] = c[a]; kc[3 * p + a] = ctr[a]; }
    hid[p] = w[0] + 1u;
    for (int k = 0; k < 8; ++k) nbr[8 * p + k] = w[k] + 1u;
}

// cp_list of the reference (kernels.jl:24-58): every window slot, aliases included.
template <bool EMIT>
__global__ void __launch_bounds__(TPB)
k_cp_list(const Params P, const double4* __restrict__ pos, const uint32_t* __restrict__ sidx,
          const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ eid,
          uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off, int32_t* __restrict__ pairs) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= P.n) return;
    const double4 x = pos[i];
    const uint32_t ei = eid[i];
    uint32_t w[8];
    window_of(P.sc.grid, mk3(x.x, x.y, x.z), w);
    uint32_t c = 0;
    const uint32_t base = EMIT ? off[i] : 0u;
    for (int wi = 0; wi < 8; ++wi) {
        const uint32_t b = w[wi];
        for (uint32_t k = bstart[b]; k < bstart[b + 1]; ++k) {
            const uint32_t ej = eid[sidx[k]];
            if (ei < ej) {  // 1-based file indices, as the reference's cp_list holds them
                if (EMIT) {
                    pairs[2 * ((int64_t)base + c)] = (int32_t)ei + 1;
                    pairs[2 * ((int64_t)base + c) + 1] = (int32_t)ej + 1;
                }
                ++c;
            }
        }
    }
    if (!EMIT) cnt[i] = c;
}

// Active contacts, one record per pair (from the row of its i), with 1-based file indices.
template <bool EMIT>
__global__ void __launch_bounds__(TPB)
k_lo_rows(const Params P, const uint8_t* __restrict__ sel, const uint32_t* __restrict__ rcnt0,
          const uint32_t* __restrict__ rcnt1, const uint32_t* __restrict__ pq0,
          const uint32_t* __restrict__ pq1, const double* __restrict__ st0,
          const double* __restrict__ st1, const uint32_t* __restrict__ eid, uint32_t* __restrict__ cnt,
          const uint32_t* __restrict__ off, dem_contact_t* __restrict__ out,
          uint8_t* __restrict__ bonded) {
    const int64_t p64 = (int64_t)blockIdx.x * TPB + threadIdx.x;
    const int64_t n = P.n;
    if (p64 >= n) return;
    const uint32_t p = (uint32_t)p64;
    const int b = sel[p];
    const uint32_t* pq = b ? pq1 : pq0;
    const double* st = b ? st1 : st0;
    const uint32_t rc = (b ? rcnt1 : rcnt0)[p];
    const int64_t cs = (int64_t)P.K * n;
    uint32_t c = 0;
    const uint32_t base = EMIT ? off[p] : 0u;
    for (uint32_t m = 0; m < rc; ++m) {
        const int64_t e = (int64_t)m * n + p;
        const uint32_t word = pq[e];
        const uint32_t q = word & IDX_MASK;
        if (!(word & IS_I) || (word & BROKEN)) continue;
        if (EMIT) {
            dem_contact_t& o = out[base + c];
            const bool bd = (word & BONDED) != 0;
            o.i = (int32_t)eid[p] + 1; o.j = (int32_t)eid[q] + 1; o.midi = 1; o.midj = 1;
            o.k[0] = st[S_KX * cs + e]; o.k[1] = st[S_KY * cs + e]; o.k[2] = st[S_KZ * cs + e];
            o.F[0] = st[S_FX * cs + e]; o.F[1] = st[S_FY * cs + e]; o.F[2] = st[S_FZ * cs + e];
            o.ti[0] = bd ? st[S_TIX * cs + e] : 0.0; o.ti[1] = bd ? st[S_TIY * cs + e] : 0.0;
            o.ti[2] = bd ? st[S_TIZ * cs + e] : 0.0;
            o.tj[0] = bd ? st[S_TJX * cs + e] : 0.0; o.tj[1] = bd ? st[S_TJY * cs + e] : 0.0;
            o.tj[2] = bd ? st[S_TJZ * cs + e] : 0.0;
            // a bond's tangential spring is zero in every reachable state (kernels.jl:230-238)
            o.s[0] = bd ? 0.0 : st[S_SX * cs + e]; o.s[1] = bd ? 0.0 : st[S_SY * cs + e];
            o.s[2] = bd ? 0.0 : st[S_SZ * cs + e];
            bonded[base + c] = bd ? 1 : 0;
        }
        ++c;
    }
    if (!EMIT) cnt[p] = c;
}

__global__ void __launch_bounds__(TPB)
k_write_rows(const Params P, int64_t m, const uint64_t* __restrict__ dst,
             const uint32_t* __restrict__ word, const double* __restrict__ state,
             uint32_t* __restrict__ pq, double* __restrict__ st) {
    const int64_t e = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (e >= m) return;
    const int64_t d = (int64_t)dst[e];
    const int64_t cs = (int64_t)P.K * P.n;
    pq[d] = word[e];
    // a bond's tangential spring is zero in every reachable state (kernels.jl:230-238)
    const bool bond = (word[e] & BONDED) != 0u;
    for (int c = 0; c < S_NCOMP; ++c)
        st[c * cs + d] = (bond && c >= S_SX) ? 0.0 : state[e * S_NCOMP + c];
}

__global__ void __launch_bounds__(TPB)
k_copy(const double4* __restrict__ src, double4* __restrict__ dst, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * TPB + threadIdx.x; k < n; k += (int64_t)gridDim.x * TPB)
        dst[k] = src[k];
}

// Copy one scene's complete dynamic state between two scenes/batches on the same device:
// particle buffers, wall history, and contact rows with partner indices moved from src_off to
// dst_off (rows land in buffer 0 of the destination).
struct AdoptArgs {
    int64_t n, src_off, dst_off, Ns, Nd;
    int32_t Ks, Kd, nwall, src_cur, dst_cur;
    const double4 *spos, *svel, *somg, *squat, *sacc, *sdw, *swshear;
    double4 *dpos, *dvel, *domg, *dquat, *dacc, *ddw, *dwshear;
    const uint32_t* swact;
    uint32_t* dwact;
    const uint8_t* ssel;
    uint8_t* dsel;
    const uint32_t *srcnt0, *srcnt1, *spq0, *spq1;
    uint32_t *drcnt0, *drcnt1, *dpq0;
    const double *sst0, *sst1;
    double* dst0;
    uint32_t* overflow;
};

__global__ void __launch_bounds__(TPB) k_adopt(const AdoptArgs a) {
    const int64_t p = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (p >= a.n) return;
    const int64_t ps = a.src_off + p, pd = a.dst_off + p;
    a.dpos[pd] = a.spos[ps];
    a.dvel[pd] = a.svel[ps];
    a.domg[pd] = a.somg[ps];
    a.dquat[pd] = a.squat[ps];
    a.dacc[pd] = a.sacc[ps];
    a.ddw[pd] = a.sdw[ps];
    a.dwact[pd] = a.swact[ps];
    for (int j = 0; j < a.nwall; ++j) a.dwshear[(int64_t)j * a.Nd + pd] = a.swshear[(int64_t)j * a.Ns + ps];
    const int b = a.ssel[ps];
    const uint32_t* pq = b ? a.spq1 : a.spq0;
    const double* st = b ? a.sst1 : a.sst0;
    uint32_t cnt = (b ? a.srcnt1 : a.srcnt0)[ps];
    if ((int)cnt > a.Kd) { *a.overflow = 1u; cnt = (uint32_t)a.Kd; }
    const int64_t css = (int64_t)a.Ks * a.Ns, csd = (int64_t)a.Kd * a.Nd;
    for (uint32_t m = 0; m < cnt; ++m) {
        const uint32_t w = pq[(int64_t)m * a.Ns + ps];
        const uint32_t q = (uint32_t)((int64_t)(w & IDX_MASK) - a.src_off + a.dst_off);
        a.dpq0[(int64_t)m * a.Nd + pd] = (w & ~IDX_MASK) | q;
        for (int c = 0; c < S_NCOMP; ++c)
            a.dst0[c * csd + (int64_t)m * a.Nd + pd] = st[c * css + (int64_t)m * a.Ns + ps];
    }
    a.drcnt0[pd] = cnt;
    a.drcnt1[pd] = 0u;
    a.dsel[pd] = 0;
}

// Initial particle state (executor.jl:229-242: q = 1, w = a = dw = 0) from the file-order input
// rows: internal index u holds file index eid[u].  The rows, radii and inertias cover file indices
// gbase..n-1; particles below gbase already hold state (moved in by k_adopt) and get only their
// r0, their starting sort order and their scene id.
struct InitArgs {
    int64_t n, gbase;
    int32_t nseg;
    const dem_grain_in_t* rows;
    const double *r, *I;   // [n - gbase]
    const double* r0f;     // [n], file order
    const int64_t* seg_off;  // [nseg + 1] (batches), else null
    const uint32_t* eid;
    double4 *pos, *vel, *omg, *quat;
    double* r0;
    uint32_t *sidx0, *sidx1;
    uint16_t* sid;
};

__global__ void __launch_bounds__(TPB) k_init_state(const InitArgs a) {
    const int64_t u = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (u >= a.n) return;
    const int64_t i = a.eid[u];
    a.r0[u] = a.r0f[i];
    a.sidx0[u] = (uint32_t)u;  // the first sort starts from the internal (spatial) order
    a.sidx1[u] = (uint32_t)u;
    if (a.seg_off) {
        // the last scene whose first particle is at or before u (scenes are never empty)
        int lo = 0, hi = a.nseg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.seg_off[mid] <= u) lo = mid;
            else hi = mid - 1;
        }
        a.sid[u] = (uint16_t)lo;
    }
    if (i < a.gbase) return;
    const int64_t k = i - a.gbase;
    const dem_grain_in_t& g = a.rows[k];
    a.pos[u] = make_double4(g.k[0], g.k[1], g.k[2], a.r[k]);
    a.vel[u] = make_double4(g.v[0], g.v[1], g.v[2], g.m);
    a.omg[u] = make_double4(0.0, 0.0, 0.0, a.I[k]);
    a.quat[u] = make_double4(1.0, 0.0, 0.0, 0.0);
}

inline dim3 grid_for(int64_t n, int tpb) { return dim3((uint32_t)((n + tpb - 1) / tpb)); }

}  // namespace

// ------------------------------------------------------------------------------------------
// timing helpers
// ------------------------------------------------------------------------------------------
void timing_begin(dem_scene* s, int kernel) {
    if (!s->timing) return;
    if (s->nev >= (int)s->ev_kernel.size()) timing_flush(s);
    hipEventRecord(s->ev[2 * s->nev], s->stream);
    s->ev_kernel[s->nev] = kernel;
}
void timing_end(dem_scene* s) {
    if (!s->timing) return;
    hipEventRecord(s->ev[2 * s->nev + 1], s->stream);
    s->nev++;
}

#define LAUNCH(s, kid, ...)                 \
    do {                                    \
        timing_begin(s, kid);               \
        __VA_ARGS__;                        \
        timing_end(s);                      \
        hipError_t e_ = hipGetLastError();  \
        if (e_ != hipSuccess) return e_;    \
    } while (0)

// skin_cur >= 0: the per-step sort of buffer parity skin_cur with near lists in use, which does
// nothing on a step that reuses them
template <bool SEG>
static hipError_t sort_impl(dem_scene* s, bool reset_counters, bool dense, bool init = false,
                            int skin_cur = -1) {
    DevState& d = s->d;
    const Params& P = s->P;
    const int c = s->cur;
    // this step's order goes to sidxb[cur], starting from the previous step's in sidxb[1 - cur]
    const uint32_t* prev = d.sidxb[1 - c];
    d.sidx = d.sidxb[c];
    const int64_t nkeys = dense ? (int64_t)(init ? P.sort_keys_init : P.sort_keys) : (int64_t)P.sc.grid.H;
    const int64_t ntiles = scan_tiles(nkeys + 1);
    if (ntiles > d.scan.max_tiles) return hipErrorInvalidValue;
    const int shift = scan_tile_shift(nkeys + 1);
    if (dense) {
        LAUNCH(s, K_HASH, hipLaunchKernelGGL((k_hash<true, SEG>), grid_for(P.n, TPB), dim3(TPB), 0,
                                             s->stream, P, d.pos[c], prev, d.hid, d.rank, d.x32,
                                             d.bcnt, d.scan.tagg, shift, d.ctr, reset_counters ? 1 : 0,
                                             init ? 1 : 0, d.nbc, d.gmin, skin_cur));
    } else {  // the reference's hashed grid: single scenes only
        LAUNCH(s, K_HASH, hipLaunchKernelGGL((k_hash<false, false>), grid_for(P.n, TPB), dim3(TPB), 0,
                                             s->stream, P, d.pos[c], prev, d.hid, d.rank, d.x32,
                                             d.bcnt, d.scan.tagg, shift, d.ctr, reset_counters ? 1 : 0,
                                             0, d.nbc, d.gmin, -1));
    }
    const unsigned long long* gate = skin_cur >= 0 ? d.ctr + CT_SKIN + skin_cur : nullptr;
    timing_begin(s, K_SCAN);
    hipError_t e = dense ? launch_tile_scan(d.scan, d.bcnt, d.bstart, nkeys + 1, s->stream, gate)
                         : launch_scan(d.scan, d.bcnt, d.bstart, nkeys + 1, true, nullptr, s->stream);
    timing_end(s);
    if (e != hipSuccess) return e;
    LAUNCH(s, K_SCATTER, hipLaunchKernelGGL(k_scatter, grid_for(P.n, TPB), dim3(TPB), 0, s->stream,
                                            P, d.pos[c], prev, d.hid, d.rank, d.x32, d.bstart, d.sidx,
                                            d.skey, dense ? nullptr : d.spos, d.spos32, d.eid,
                                            dense ? nullptr : d.sfid, d.scan.tagg,
                                            dense ? (int)ntiles : 0, gate));
    return hipSuccess;
}

static hipError_t sort_particles(dem_scene* s, bool reset_counters, bool dense) {
    return s->batched ? sort_impl<true>(s, reset_counters, dense)
                      : sort_impl<false>(s, reset_counters, dense);
}

static MergeArgs merge_args(dem_scene* s) {
    DevState& d = s->d;
    MergeArgs M;
    M.pos = d.pos[s->cur]; M.nbc = d.nbc; M.nb = d.nb; M.gmin = d.gmin; M.gmin_w = d.gmin; M.gskip = d.gskip;
    M.tcnt = d.tcnt;
    M.touch = d.touch; M.sel = d.sel; M.rcnt0 = d.rcnt[0]; M.rcnt1 = d.rcnt[1];
    M.pq0 = d.pq[0]; M.pq1 = d.pq[1]; M.st0 = d.st[0]; M.st1 = d.st[1];
    return M;
}

static ForceArgs force_args(dem_scene* s) {
    DevState& d = s->d;
    const int c = s->cur, nx = 1 - s->cur;
    ForceArgs A;
    A.pos = d.pos[c]; A.vel = d.vel[c]; A.omg = d.omg[c]; A.quat = d.quat[c];
    A.pos_n = d.pos[nx]; A.vel_n = d.vel[nx]; A.omg_n = d.omg[nx]; A.quat_n = d.quat[nx];
    A.acc = d.acc; A.dwo = d.dw; A.ftot = d.ftot; A.ttot = d.ttot; A.sel = d.sel;
    A.rcnt0 = d.rcnt[0]; A.rcnt1 = d.rcnt[1]; A.pq0 = d.pq[0]; A.pq1 = d.pq[1];
    A.st0 = d.st[0]; A.st1 = d.st[1]; A.wact = d.wact; A.wshear = d.wshear; A.xref = d.xref;
    return A;
}

template <bool SEG>
static hipError_t detect_impl(dem_scene* s, bool init, bool last) {
    // per-step counters are only read after the last step of a call; the diagnostic passes (the
    // reference's cp_list length, candidates examined) run there only when asked for
    const int diag = last ? 1 : 0;
    const bool count_candidates = last && s->diagnostics;
    DevState& d = s->d;
    const Params& P = s->P;
    const int c = s->cur;
    const bool dense = s->dense;
    if (SEG && !dense) return hipErrorInvalidValue;
    // near lists: the buffer parity, whose flag word says whether this step rebuilds
    const int skin_cur = (P.skin && dense && !init) ? c : -1;
    hipError_t e = sort_impl<SEG>(s, true, dense, init, skin_cur);
    if (e != hipSuccess) return e;
    const bool mixed = s->precision == DEM_PRECISION_MIXED && !init;
    const dim3 g = grid_for(P.n, TPB);
    if (count_candidates && !dense)
        LAUNCH(s, K_OTHER, hipLaunchKernelGGL(k_cand_count, g, dim3(TPB), 0, s->stream, P, d.pos[c],
                                              d.sfid, d.bstart, d.eid, d.ctr));
    if (dense) {
        if (init) {
            LAUNCH(s, K_INIT, hipLaunchKernelGGL((k_broad_dense<true, SEG, false>), g, dim3(TPB), 0,
                                                 s->stream, P, d.sidx, d.skey, d.pos[c], d.spos32,
                                                 d.bstart, d.r0, d.eid, d.tcnt, d.touch, d.ctr, diag,
                                                 d.nbc, d.nb, d.gmin, -1));
        } else if (skin_cur >= 0) {
            LAUNCH(s, K_BROAD, hipLaunchKernelGGL((k_broad_dense<false, SEG, true>), g, dim3(TPB), 0,
                                                  s->stream, P, d.sidx, d.skey, d.pos[c], d.spos32,
                                                  d.bstart, d.r0, d.eid, d.tcnt, d.touch, d.ctr, diag,
                                                  d.nbc, d.nb, d.gmin, skin_cur));
        } else {
            LAUNCH(s, K_BROAD, hipLaunchKernelGGL((k_broad_dense<false, SEG, false>), g, dim3(TPB), 0,
                                                  s->stream, P, d.sidx, d.skey, d.pos[c], d.spos32,
                                                  d.bstart, d.r0, d.eid, d.tcnt, d.touch, d.ctr, diag,
                                                  d.nbc, d.nb, d.gmin, -1));
        }
    } else if (s->slot_owned) {
        if (init) {
            LAUNCH(s, K_INIT, hipLaunchKernelGGL((k_broad_slot<true, false>), g, dim3(TPB), 0, s->stream,
                                                 P, d.sidx, d.spos, d.spos32, d.bstart, d.r0, d.eid,
                                                 d.tcnt, d.touch, d.ctr));
        } else if (mixed) {
            LAUNCH(s, K_BROAD, hipLaunchKernelGGL((k_broad_slot<false, true>), g, dim3(TPB), 0, s->stream,
                                                  P, d.sidx, d.spos, d.spos32, d.bstart, d.r0, d.eid,
                                                  d.tcnt, d.touch, d.ctr));
        } else {
            LAUNCH(s, K_BROAD, hipLaunchKernelGGL((k_broad_slot<false, false>), g, dim3(TPB), 0, s->stream,
                                                  P, d.sidx, d.spos, d.spos32, d.bstart, d.r0, d.eid,
                                                  d.tcnt, d.touch, d.ctr));
        }
    } else if (init) {
        LAUNCH(s, K_INIT, hipLaunchKernelGGL((k_broad<true, false>), g, dim3(TPB), 0, s->stream, P,
                                             d.sidx, d.spos, d.spos32, d.bstart, d.r0, d.eid,
                                             d.tcnt, d.touch, d.ctr));
    } else if (mixed) {
        LAUNCH(s, K_BROAD, hipLaunchKernelGGL((k_broad<false, true>), g, dim3(TPB), 0, s->stream, P,
                                              d.sidx, d.spos, d.spos32, d.bstart, d.r0, d.eid,
                                              d.tcnt, d.touch, d.ctr));
    } else {
        LAUNCH(s, K_BROAD, hipLaunchKernelGGL((k_broad<false, false>), g, dim3(TPB), 0, s->stream, P,
                                              d.sidx, d.spos, d.spos32, d.bstart, d.r0, d.eid,
                                              d.tcnt, d.touch, d.ctr));
    }
    const size_t lds = 0;
    const dim3 gm = grid_for(P.n, MERGE_TPB);
    if (init) {
        LAUNCH(s, K_MERGE, hipLaunchKernelGGL((k_merge<true, SEG, MERGE_TPB>), gm, dim3(MERGE_TPB), lds,
                                              s->stream, P, merge_args(s), d.ctr, diag, skin_cur, 0, 0));
    } else if (s->force_tpb == 64 && s->initialized) {
        // contact-dense, as the force kernel's block choice: one-wave blocks for uneven rows; a
        // scene whose bond pass ran keeps its near-list bond marks current (a step that does not
        // rebuild but walks the lists sees bonds that broke)
        LAUNCH(s, K_MERGE, hipLaunchKernelGGL((k_merge<false, SEG, 64, true>), grid_for(P.n, 64), dim3(64), lds,
                                              s->stream, P, merge_args(s), d.ctr, diag, skin_cur, 0, 0));
    } else if (s->force_tpb == 64) {
        LAUNCH(s, K_MERGE, hipLaunchKernelGGL((k_merge<false, SEG, 64>), grid_for(P.n, 64), dim3(64), lds,
                                              s->stream, P, merge_args(s), d.ctr, diag, skin_cur, 0, 0));
    } else if (s->initialized) {
        LAUNCH(s, K_MERGE, hipLaunchKernelGGL((k_merge<false, SEG, MERGE_TPB, true>), gm, dim3(MERGE_TPB), lds,
                                              s->stream, P, merge_args(s), d.ctr, diag, skin_cur, 0,
                                              s->force_free ? 1 : 0));
    } else {
        LAUNCH(s, K_MERGE, hipLaunchKernelGGL((k_merge<false, SEG, MERGE_TPB>), gm, dim3(MERGE_TPB), lds,
                                              s->stream, P, merge_args(s), d.ctr, diag, skin_cur, 0,
                                              s->force_free ? 1 : 0));
    }
    if (skin_cur >= 0 && s->initialized)  // a rebuilding step marks the near lists' intact bonds
        LAUNCH(s, K_OTHER, hipLaunchKernelGGL((k_mark_bonds<SEG>), g, dim3(TPB), 0, s->stream, P, merge_args(s),
                                              d.ctr, skin_cur));
    if (count_candidates && dense)
        LAUNCH(s, K_OTHER, hipLaunchKernelGGL((k_examined_count<SEG>), g, dim3(TPB), 0, s->stream, P,
                                              d.sidx, d.skey, d.spos32, d.bstart, d.ctr));
    if (count_candidates && dense && !SEG) {
        // the cp_list length needs the reference's bucket sort: rebuild it after the merge
        if ((e = sort_impl<false>(s, false, false)) != hipSuccess) return e;
        LAUNCH(s, K_OTHER, hipLaunchKernelGGL(k_cand_count, g, dim3(TPB), 0, s->stream, P, d.pos[c],
                                              d.sfid, d.bstart, d.eid, d.ctr));
    }
    return hipSuccess;
}

hipError_t run_contact_detection(dem_scene* s, bool init, bool last) {
    if (last) s->diag_valid = s->diagnostics;
    return s->batched ? detect_impl<true>(s, init, last) : detect_impl<false>(s, init, last);
}

template <bool SEG, int BS, bool FREE = false>
static hipError_t launch_force(dem_scene* s, bool write_forces) {
    const Params& P = s->P;
    const int skin = (P.skin && s->dense) ? s->cur : -1;
    const dim3 g = grid_for(P.n, BS);
    const ForceArgs A = force_args(s);
    if (s->precision == DEM_PRECISION_MIXED) {
        LAUNCH(s, K_FORCE, hipLaunchKernelGGL((k_force<true, SEG, BS, FREE>), g, dim3(BS), 0, s->stream, P,
                                              A, s->d.ctr, write_forces ? 1 : 0, skin));
    } else {
        LAUNCH(s, K_FORCE, hipLaunchKernelGGL((k_force<false, SEG, BS, FREE>), g, dim3(BS), 0, s->stream, P,
                                              A, s->d.ctr, write_forces ? 1 : 0, skin));
    }
    return hipSuccess;
}

template <bool SEG>
static hipError_t force_impl(dem_scene* s, bool write_forces) {
    hipError_t e;
    if constexpr (!SEG) {
        if (s->force_free) e = launch_force<false, 256, true>(s, write_forces);
        else e = s->force_tpb == 64 ? launch_force<false, 64>(s, write_forces)
                                    : launch_force<false, 256>(s, write_forces);
    } else {
        e = s->force_tpb == 64 ? launch_force<true, 64>(s, write_forces) : launch_force<true, 256>(s, write_forces);
    }
    if (e != hipSuccess) return e;
    s->cur = 1 - s->cur;
    return hipSuccess;
}

hipError_t run_force_integrate(dem_scene* s, bool write_forces) {
    return s->batched ? force_impl<true>(s, write_forces) : force_impl<false>(s, write_forces);
}

// The per-step counters (those below CT_NSTEP), cleared before a lean step that ends a dem_step
// call as the first kernel of a full step clears them: the counters read after the call are
// that step's.
__global__ void __launch_bounds__(TPB) k_clear_step_counters(unsigned long long* ctr) {
    for (int i = threadIdx.x; i < CTR_COPIES * CT_NSTEP; i += TPB)
        ctr[(i / CT_NSTEP) * CTR_STRIDE + i % CT_NSTEP] = 0ull;
}

// One lean step of a single scene in the contact-free mode (k_lean_free): a merge block per four
// 64-grain groups up to 128 blocks (1M pour: 57.8 us a step; 256 blocks 58.2, 64 blocks 66.9), then
// the force blocks, each count rounded up to a multiple of 8.
static hipError_t launch_lean_free(dem_scene* s) {
    const Params& P = s->P;
    const uint32_t nb = (uint32_t)((P.n + TPB - 1) / TPB);
    const uint32_t nfb = (nb + 7u) / 8u * 8u;
    const uint32_t nmb = ((nb < 128u ? nb : 128u) + 7u) & ~7u;
    const MergeArgs M = merge_args(s);
    const ForceArgs A = force_args(s);
    if (s->precision == DEM_PRECISION_MIXED) {
        LAUNCH(s, K_FORCE, hipLaunchKernelGGL((k_lean_free<true>), dim3(nmb + nfb), dim3(TPB), 0, s->stream, P, M, A,
                                              s->d.ctr, s->cur, nmb));
    } else {
        LAUNCH(s, K_FORCE, hipLaunchKernelGGL((k_lean_free<false>), dim3(nmb + nfb), dim3(TPB), 0, s->stream, P, M,
                                              A, s->d.ctr, s->cur, nmb));
    }
    s->cur = 1 - s->cur;
    return hipSuccess;
}

hipError_t run_lean_free_prologue(dem_scene* s) {
    LAUNCH(s, K_OTHER, hipLaunchKernelGGL(k_lean_free_prologue, dim3(1), dim3(64), 0, s->stream, s->d.ctr, s->cur));
    return hipSuccess;
}

template <bool SEG>
static hipError_t lean_impl(dem_scene* s, bool last) {
    DevState& d = s->d;
    const Params& P = s->P;
    const int c = s->cur;
    if (!P.skin || !s->dense) return hipErrorInvalidValue;
    const int diag = last ? 1 : 0;
    if (last)
        LAUNCH(s, K_OTHER, hipLaunchKernelGGL(k_clear_step_counters, dim3(1), dim3(TPB), 0, s->stream, d.ctr));
    if constexpr (!SEG)
        if (s->force_free && !last) return launch_lean_free(s);
    // a scene whose bond pass ran leaves the near-list entries of intact bonds untested
    const bool bonds = s->initialized;
    const int fr = s->force_free ? 1 : 0;
    if (s->force_tpb == 64 && bonds) {
        LAUNCH(s, K_MERGE, hipLaunchKernelGGL((k_merge<false, SEG, 64, true>), grid_for(P.n, 64), dim3(64), 0,
                                              s->stream, P, merge_args(s), d.ctr, diag, c, 1, 0));
    } else if (s->force_tpb == 64) {
        LAUNCH(s, K_MERGE, hipLaunchKernelGGL((k_merge<false, SEG, 64>), grid_for(P.n, 64), dim3(64), 0,
                                              s->stream, P, merge_args(s), d.ctr, diag, c, 1, 0));
    } else if (bonds) {
        LAUNCH(s, K_MERGE, hipLaunchKernelGGL((k_merge<false, SEG, MERGE_TPB, true>), grid_for(P.n, MERGE_TPB),
                                              dim3(MERGE_TPB), 0, s->stream, P, merge_args(s), d.ctr, diag,
                                              c, 1, fr));
    } else {
        LAUNCH(s, K_MERGE, hipLaunchKernelGGL((k_merge<false, SEG, MERGE_TPB>), grid_for(P.n, MERGE_TPB),
                                              dim3(MERGE_TPB), 0, s->stream, P, merge_args(s), d.ctr, diag,
                                              c, 1, fr));
    }
    return force_impl<SEG>(s, last);
}

hipError_t run_lean_step(dem_scene* s, bool last) {
    if (last) s->diag_valid = false;
    return s->batched ? lean_impl<true>(s, last) : lean_impl<false>(s, last);
}

hipError_t run_rows_finish(dem_scene* s, bool rollback) {
    hipLaunchKernelGGL(k_rows_finish, grid_for(s->n, TPB), dim3(TPB), 0, s->stream, s->n, s->d.sel,
                       s->d.tcnt, s->d.ctr, rollback ? 1 : 0);
    return hipGetLastError();
}

hipError_t run_debug_cells(dem_scene* s, int32_t* kid, uint32_t* hid, double* kc, uint32_t* nbr) {
    hipLaunchKernelGGL(k_debug_cells, grid_for(s->P.n, TPB), dim3(TPB), 0, s->stream, s->P,
                       s->d.pos[s->cur], kid, hid, kc, nbr);
    return hipGetLastError();
}

hipError_t run_debug_sort(dem_scene* s) { return sort_particles(s, false, false); }

hipError_t run_debug_cp_list(dem_scene* s, int32_t** pairs_dev, int64_t* count) {
    DevState& d = s->d;
    const Params& P = s->P;
    hipError_t e = sort_particles(s, false, false);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_cp_list<false>), grid_for(P.n, TPB), dim3(TPB), 0, s->stream, P,
                       d.pos[s->cur], d.sidx, d.bstart, d.eid, d.scratch_u32, nullptr, nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = launch_scan(d.scan, d.scratch_u32, d.scratch2_u32, P.n, false, d.total, s->stream);
    if (e != hipSuccess) return e;
    uint64_t total = 0;
    if ((e = hipMemcpyAsync(&total, d.total, sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return e;
    int32_t* pairs = nullptr;
    if ((e = hipMalloc(&pairs, 2 * sizeof(int32_t) * (total ? total : 1))) != hipSuccess) return e;
    hipLaunchKernelGGL((k_cp_list<true>), grid_for(P.n, TPB), dim3(TPB), 0, s->stream, P,
                       d.pos[s->cur], d.sidx, d.bstart, d.eid, nullptr, d.scratch2_u32, pairs);
    if ((e = hipGetLastError()) != hipSuccess) { hipFree(pairs); return e; }
    *pairs_dev = pairs;
    *count = (int64_t)total;
    return hipSuccess;
}

hipError_t run_read_contacts(dem_scene* s, dem_contact_t** out_dev, uint8_t** bonded_dev,
                             int64_t* count) {
    DevState& d = s->d;
    const Params& P = s->P;
    hipLaunchKernelGGL((k_lo_rows<false>), grid_for(P.n, TPB), dim3(TPB), 0, s->stream, P, d.sel,
                       d.rcnt[0], d.rcnt[1], d.pq[0], d.pq[1], d.st[0], d.st[1], d.eid,
                       d.scratch_u32, nullptr, nullptr, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = launch_scan(d.scan, d.scratch_u32, d.scratch2_u32, P.n, false, d.total, s->stream)) != hipSuccess)
        return e;
    uint64_t total = 0;
    if ((e = hipMemcpyAsync(&total, d.total, sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return e;
    dem_contact_t* out = nullptr;
    uint8_t* bd = nullptr;
    if ((e = hipMalloc(&out, sizeof(dem_contact_t) * (total ? total : 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&bd, total ? total : 1)) != hipSuccess) { hipFree(out); return e; }
    hipLaunchKernelGGL((k_lo_rows<true>), grid_for(P.n, TPB), dim3(TPB), 0, s->stream, P, d.sel,
                       d.rcnt[0], d.rcnt[1], d.pq[0], d.pq[1], d.st[0], d.st[1], d.eid,
                       nullptr, d.scratch2_u32, out, bd);
    if ((e = hipGetLastError()) != hipSuccess) { hipFree(out); hipFree(bd); return e; }
    *out_dev = out;
    *bonded_dev = bd;
    *count = (int64_t)total;
    return hipSuccess;
}

hipError_t run_write_contacts(dem_scene* s, const uint64_t* dst_dev, const uint32_t* word_dev,
                              const double* state_dev, int64_t m) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_write_rows, grid_for(m, TPB), dim3(TPB), 0, s->stream, s->P, m, dst_dev,
                       word_dev, state_dev, s->d.pq[0], s->d.st[0]);
    return hipGetLastError();
}

hipError_t run_adopt(dem_scene* dst, int64_t dst_off, const dem_scene* src, int64_t src_off,
                     int64_t n, int32_t nwall, uint32_t* overflow_dev) {
    AdoptArgs a;
    a.n = n; a.src_off = src_off; a.dst_off = dst_off; a.Ns = src->n; a.Nd = dst->n;
    a.Ks = src->K; a.Kd = dst->K; a.nwall = nwall; a.src_cur = src->cur; a.dst_cur = dst->cur;
    const DevState& S = src->d;
    DevState& D = dst->d;
    a.spos = S.pos[src->cur]; a.svel = S.vel[src->cur]; a.somg = S.omg[src->cur]; a.squat = S.quat[src->cur];
    a.sacc = S.acc; a.sdw = S.dw; a.swshear = S.wshear; a.swact = S.wact;
    a.dpos = D.pos[dst->cur]; a.dvel = D.vel[dst->cur]; a.domg = D.omg[dst->cur]; a.dquat = D.quat[dst->cur];
    a.dacc = D.acc; a.ddw = D.dw; a.dwshear = D.wshear; a.dwact = D.wact;
    a.ssel = S.sel; a.dsel = D.sel;
    a.srcnt0 = S.rcnt[0]; a.srcnt1 = S.rcnt[1]; a.spq0 = S.pq[0]; a.spq1 = S.pq[1];
    a.drcnt0 = D.rcnt[0]; a.drcnt1 = D.rcnt[1]; a.dpq0 = D.pq[0];
    a.sst0 = S.st[0]; a.sst1 = S.st[1]; a.dst0 = D.st[0];
    a.overflow = overflow_dev;
    hipLaunchKernelGGL(k_adopt, grid_for(n, TPB), dim3(TPB), 0, dst->stream, a);
    return hipGetLastError();
}

hipError_t run_init_state(dem_scene* s, const dem_grain_in_t* rows, const double* r,
                          const double* I, const double* r0f, const int64_t* seg_off, int64_t gbase) {
    DevState& d = s->d;
    InitArgs a;
    a.n = s->n; a.gbase = gbase; a.nseg = s->nseg;
    a.rows = rows; a.r = r; a.I = I; a.r0f = r0f;
    a.seg_off = s->batched ? seg_off : nullptr;
    a.eid = d.eid;
    a.pos = d.pos[s->cur]; a.vel = d.vel[s->cur]; a.omg = d.omg[s->cur]; a.quat = d.quat[s->cur];
    a.r0 = d.r0; a.sidx0 = d.sidxb[0]; a.sidx1 = d.sidxb[1];
    a.sid = s->batched ? d.sid : nullptr;
    hipLaunchKernelGGL(k_init_state, grid_for(s->n, TPB), dim3(TPB), 0, s->stream, a);
    return hipGetLastError();
}

hipError_t run_copy(const void* src, void* dst, int64_t nbytes, hipStream_t stream) {
    const int64_t n = nbytes / (int64_t)sizeof(double4);
    // one element per thread in a single pass: on MI355X this moves 1 GiB faster than a
    // grid-stride loop over a few thousand blocks (about 6.2 against 4.8 TB/s read + write)
    hipLaunchKernelGGL(k_copy, grid_for(n, TPB), dim3(TPB), 0, stream, (const double4*)src, (double4*)dst, n);
    return hipGetLastError();
}

}  // namespace dem
