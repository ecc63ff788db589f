// scan.hip -- single-pass exclusive prefix sum of uint32 with decoupled look-back.
//
// Replaces CUDA.jl accumulate!(+, ...) of the reference (src/executor.jl:50,62).  One launch per
// scan: each workgroup claims a tile id from a counter (so a tile only ever waits on tiles that
// were claimed earlier and are therefore running), publishes its aggregate, then the whole
// workgroup looks back over up to 256 predecessors per round (every tile is resident at once on a
// large scan, so the look-back depth, not bandwidth, sets the time).  The last workgroup to finish
// clears the tile status words and counters, so no separate reset launch (and no memset node) is
// needed.
//
// The counting sort's scan (k_scan_tiles) is cheaper still: the histogram pass already knows each
// tile's total, so a tile only adds up the totals before it.
#include "dem_internal.hpp"

#include <algorithm>
#include <mutex>
#include <string>

namespace dem {

namespace {

constexpr int SCAN_THREADS = 256;
// Values per thread: short tiles keep a small scan's per-tile chain short; long tiles keep a large
// scan's look-back shallow.
constexpr int SCAN_ITEMS_SMALL = 8;   // 2048 values per tile
constexpr int SCAN_ITEMS_LARGE = 32;  // 8192 values per tile
constexpr int64_t SCAN_LARGE_FROM = (int64_t)1 << 21;
constexpr uint64_t FLAG_AGG = 1ull << 32;
constexpr uint64_t FLAG_INC = 2ull << 32;

__device__ __forceinline__ uint64_t load_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

template <int SCAN_ITEMS>
__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_exclusive(uint32_t* __restrict__ in, uint32_t* __restrict__ out, int64_t n, int zero_input,
                 uint64_t* __restrict__ status, uint32_t* __restrict__ counters,
                 uint64_t* __restrict__ total, uint32_t ntiles) {
    constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
    __shared__ uint32_t s_vals[SCAN_TILE + SCAN_TILE / 32];
    __shared__ uint32_t s_wave[SCAN_THREADS / 64];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_last;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(&counters[0], 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const int64_t base = (int64_t)tile * SCAN_TILE;

    // Coalesced load into LDS (padded by one word every 32 to avoid bank conflicts).  All loads
    // are issued before the input is cleared: a store to an address waits for the load before it.
    uint32_t v[SCAN_ITEMS];
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int64_t g = base + it * SCAN_THREADS + tid;
        v[it] = g < n ? in[g] : 0u;
    }
    if (zero_input) {
#pragma unroll
        for (int it = 0; it < SCAN_ITEMS; ++it) {
            const int64_t g = base + it * SCAN_THREADS + tid;
            if (g < n) in[g] = 0u;
        }
    }
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = it * SCAN_THREADS + tid;
        s_vals[idx + (idx >> 5)] = v[it];
    }
    __syncthreads();

    // Each thread owns SCAN_ITEMS consecutive values (re-read from LDS after the look-back).
    uint32_t sum = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = tid * SCAN_ITEMS + it;
        sum += s_vals[idx + (idx >> 5)];
    }
    const uint32_t incl = wave_incl_scan(sum);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0, agg = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        const uint32_t t = s_wave[w];
        if (w < wave) wave_off += t;
        agg += t;
    }
    const uint32_t thread_excl = wave_off + incl - sum;

    // Publish, then look back over 256 predecessors per round with the whole workgroup.
    __shared__ uint32_t s_first;
    __shared__ uint32_t s_part[SCAN_THREADS / 64];
    uint32_t prefix = 0;
    if (tile == 0) {
        if (tid == 0) store_status(&status[0], FLAG_INC | (uint64_t)agg);
    } else {
        if (tid == 0) store_status(&status[tile], FLAG_AGG | (uint64_t)agg);
        int64_t pred = (int64_t)tile - 1;
        while (true) {
            const int64_t t = pred - tid;
            uint64_t st = t >= 0 ? load_status(&status[t]) : (FLAG_INC | 0ull);
            // every predecessor was claimed before this tile and publishes without waiting
            while ((st >> 32) == 0) {
                __builtin_amdgcn_s_sleep(1);
                st = load_status(&status[t]);
            }
            if (tid == 0) s_first = SCAN_THREADS;
            __syncthreads();
            if ((st >> 32) == 2) atomicMin(&s_first, (uint32_t)tid);  // nearest inclusive prefix
            __syncthreads();
            const uint32_t first = s_first;
            const uint32_t v = (uint32_t)tid <= first ? (uint32_t)st : 0u;
            const uint32_t ws = (uint32_t)wave_sum_u64(v);
            if (lane == 0) s_part[wave] = ws;
            __syncthreads();
            uint32_t round = 0;
#pragma unroll
            for (int w = 0; w < SCAN_THREADS / 64; ++w) round += s_part[w];
            prefix += round;
            if (first < (uint32_t)SCAN_THREADS) break;
            pred -= SCAN_THREADS;
            __syncthreads();  // s_first and s_part are rewritten by the next round
        }
        if (tid == 0) store_status(&status[tile], FLAG_INC | (uint64_t)(uint32_t)(prefix + agg));
    }

    uint32_t run = prefix + thread_excl;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = tid * SCAN_ITEMS + it;
        const uint32_t x = s_vals[idx + (idx >> 5)];
        s_vals[idx + (idx >> 5)] = run;
        run += x;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = it * SCAN_THREADS + tid;
        const int64_t g = base + idx;
        if (g < n) out[g] = s_vals[idx + (idx >> 5)];
    }

    // The last workgroup to finish resets the status words and counters for the next scan.
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&counters[1], 1u);
        s_last = (done == ntiles - 1) ? 1u : 0u;
        if (tile == ntiles - 1 && total) *total = (uint64_t)(uint32_t)(prefix + agg);
    }
    __syncthreads();
    if (s_last) {
        for (uint32_t t = tid; t < ntiles; t += SCAN_THREADS) store_status(&status[t], 0ull);
        if (tid == 0) {
            __hip_atomic_store(&counters[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&counters[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Exclusive scan of bucket counts whose per-tile totals are already known (the counting sort's
// histogram pass adds each wave's counts into a copy of tagg[bucket >> tile shift]).  Every tile
// sums the totals of the tiles before it, so there is no look-back, no tile counter and no
// fence; the counts are cleared as they are read.  tagg is cleared by the sort's scatter pass,
// which runs after this kernel.
template <int SCAN_ITEMS>
__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_tiles(uint32_t* __restrict__ in, uint32_t* __restrict__ out, int64_t n,
             const uint32_t* __restrict__ tagg, const unsigned long long* __restrict__ gate) {
    if (gate && *gate == 0ull) return;  // the step reuses the near lists: no sort
    constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
    __shared__ uint32_t s_vals[SCAN_TILE + SCAN_TILE / 32];
    __shared__ uint32_t s_wave[SCAN_THREADS / 64];
    __shared__ uint32_t s_pre[SCAN_THREADS / 64];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const uint32_t tile = blockIdx.x;
    const int64_t base = (int64_t)tile * SCAN_TILE;

    uint32_t v[SCAN_ITEMS];
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int64_t g = base + it * SCAN_THREADS + tid;
        v[it] = g < n ? in[g] : 0u;
    }
    // the totals of the tiles before this one, eight independent loads per thread in flight (a
    // large grid's last tiles sum thousands of them)
    uint32_t pre = 0;
    const uint32_t lim = tile * TAGG_COPIES;
    uint32_t k = tid;
    for (; k + 7 * SCAN_THREADS < lim; k += 8 * SCAN_THREADS) {
        uint32_t a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = tagg[k + u * SCAN_THREADS];
#pragma unroll
        for (int u = 0; u < 8; ++u) pre += a[u];
    }
    for (; k < lim; k += SCAN_THREADS) pre += tagg[k];
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int64_t g = base + it * SCAN_THREADS + tid;
        if (g < n) in[g] = 0u;
    }
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = it * SCAN_THREADS + tid;
        s_vals[idx + (idx >> 5)] = v[it];
    }
    const uint32_t pw = (uint32_t)wave_sum_u64(pre);
    if (lane == 0) s_pre[wave] = pw;
    __syncthreads();

    uint32_t sum = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = tid * SCAN_ITEMS + it;
        sum += s_vals[idx + (idx >> 5)];
    }
    const uint32_t incl = wave_incl_scan(sum);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        run += s_pre[w];
        if (w < wave) run += s_wave[w];
    }
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = tid * SCAN_ITEMS + it;
        const uint32_t x = s_vals[idx + (idx >> 5)];
        s_vals[idx + (idx >> 5)] = run;
        run += x;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = it * SCAN_THREADS + tid;
        const int64_t g = base + idx;
        if (g < n) out[g] = s_vals[idx + (idx >> 5)];
    }
}

}  // namespace

static int64_t tile_of(int64_t n) {
    return (int64_t)SCAN_THREADS * (n >= SCAN_LARGE_FROM ? SCAN_ITEMS_LARGE : SCAN_ITEMS_SMALL);
}

int64_t scan_tiles(int64_t n) { return (n + tile_of(n) - 1) / tile_of(n); }

int scan_tile_shift(int64_t n) { return n >= SCAN_LARGE_FROM ? 13 : 11; }

int64_t scan_tiles_max(int64_t n) {
    return std::max(scan_tiles(std::min(n, SCAN_LARGE_FROM - 1)), scan_tiles(n));
}

hipError_t launch_scan(ScanWorkspace& ws, uint32_t* in, uint32_t* out, int64_t n, bool zero_input,
                       uint64_t* total, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t ntiles = scan_tiles(n);
    if (ntiles > ws.max_tiles) return hipErrorInvalidValue;
    if (n >= SCAN_LARGE_FROM)
        hipLaunchKernelGGL(k_scan_exclusive<SCAN_ITEMS_LARGE>, dim3((uint32_t)ntiles), dim3(SCAN_THREADS),
                           0, stream, in, out, n, zero_input ? 1 : 0, ws.status, ws.counters, total,
                           (uint32_t)ntiles);
    else
        hipLaunchKernelGGL(k_scan_exclusive<SCAN_ITEMS_SMALL>, dim3((uint32_t)ntiles), dim3(SCAN_THREADS),
                           0, stream, in, out, n, zero_input ? 1 : 0, ws.status, ws.counters, total,
                           (uint32_t)ntiles);
    return hipGetLastError();
}

// ---- the allocation-free scan (dem_scan_noalloc) -------------------------------------------
// Its look-back scratch lives in this file-scope array: words [0, 32) hold the two tile counters
// (the first 8 bytes; the rest pads the status words to a 256-byte boundary), words from 32 on
// hold one status word per tile.  Each GPU holds its own copy for the life of the process.
constexpr int64_t NOALLOC_MAX_VALUES = (int64_t)1 << 30;
constexpr int64_t NOALLOC_HEAD_WORDS = 32;
constexpr int64_t NOALLOC_MAX_TILES = ((int64_t)1 << 30) / (SCAN_THREADS * SCAN_ITEMS_LARGE);
static_assert(NOALLOC_MAX_VALUES >= SCAN_LARGE_FROM, "the largest scan uses long tiles");
static __device__ __align__(256) uint64_t scratch[NOALLOC_HEAD_WORDS + NOALLOC_MAX_TILES];

namespace {

// Clears the scratch words a scan of ntiles tiles reads (counters and status words), and writes
// a zero total for an empty scan.  A kernel rather than hipMemsetAsync, so that a caller may
// capture the call into a graph.
__global__ void __launch_bounds__(SCAN_THREADS)
k_noalloc_clear(uint64_t* __restrict__ words, int64_t nwords, uint64_t* __restrict__ zero_total) {
    const int64_t i = (int64_t)blockIdx.x * SCAN_THREADS + threadIdx.x;
    if (i < nwords) words[i] = 0ull;
    if (zero_total && i == 0) *zero_total = 0ull;
}

// The address of this device's copy of scratch, resolved once per device (the host-side name
// `scratch` is not a device address).
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
uint64_t* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

hipError_t noalloc_scratch(int device, uint64_t** out) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_noalloc_ptr[device] = (uint64_t*)p;
    }
    *out = g_noalloc_ptr[device];
    return hipSuccess;
}

}  // namespace

hipError_t launch_tile_scan(ScanWorkspace& ws, uint32_t* in, uint32_t* out, int64_t n,
                            hipStream_t stream, const unsigned long long* gate) {
    if (n <= 0) return hipSuccess;
    const int64_t ntiles = scan_tiles(n);
    if (ntiles > ws.max_tiles) return hipErrorInvalidValue;
    if (n >= SCAN_LARGE_FROM)
        hipLaunchKernelGGL(k_scan_tiles<SCAN_ITEMS_LARGE>, dim3((uint32_t)ntiles), dim3(SCAN_THREADS), 0,
                           stream, in, out, n, ws.tagg, gate);
    else
        hipLaunchKernelGGL(k_scan_tiles<SCAN_ITEMS_SMALL>, dim3((uint32_t)ntiles), dim3(SCAN_THREADS), 0,
                           stream, in, out, n, ws.tagg, gate);
    return hipGetLastError();
}

}  // namespace dem

extern "C" int32_t dem_scan_noalloc(const uint32_t* in, uint32_t* out, int64_t n, uint64_t* total,
                                    void* stream) {
    using namespace dem;
    if (n < 0 || n > NOALLOC_MAX_VALUES || (n > 0 && (!in || !out))) {
        set_last_error_global("dem_scan_noalloc: need 0 <= n <= 2^30 and device buffers in, out");
        return DEM_ERR_INVALID;
    }
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e == hipSuccess && (device < 0 || device >= NOALLOC_MAX_DEVICES)) e = hipErrorInvalidDevice;
    uint64_t* words = nullptr;
    if (e == hipSuccess) e = noalloc_scratch(device, &words);
    const hipStream_t st = (hipStream_t)stream;
    const int64_t ntiles = n > 0 ? scan_tiles(n) : 0;
    if (e == hipSuccess) {
        const int64_t nwords = NOALLOC_HEAD_WORDS + ntiles;
        hipLaunchKernelGGL(k_noalloc_clear, dim3((uint32_t)((nwords + SCAN_THREADS - 1) / SCAN_THREADS)),
                           dim3(SCAN_THREADS), 0, st, words, nwords, n == 0 ? total : nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess && n > 0) {
        ScanWorkspace ws;
        ws.counters = (uint32_t*)words;
        ws.status = words + NOALLOC_HEAD_WORDS;
        ws.max_tiles = NOALLOC_MAX_TILES;
        e = launch_scan(ws, const_cast<uint32_t*>(in), out, n, false, total, st);
    }
    if (e != hipSuccess) {
        set_last_error_global(std::string("dem_scan_noalloc: ") + hipGetErrorString(e));
        return DEM_ERR_HIP;
    }
    return DEM_OK;
}
