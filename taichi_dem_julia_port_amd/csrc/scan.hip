// scan.hip -- single-pass exclusive prefix sum of uint32 with decoupled look-back.
//
// Replaces CUDA.jl accumulate!(+, ...) of the reference (src/executor.jl:50,62).  One launch per
// scan: each workgroup claims a tile id from a counter (so a tile only ever waits on tiles that
// were claimed earlier and are therefore running), publishes its aggregate, then one wavefront
// looks back over up to 64 predecessors per round.  The last workgroup to finish clears the tile
// status words and counters, so no separate reset launch (and no memset node) is needed.
#include "dem_internal.hpp"

namespace dem {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 4096 values per tile
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

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_exclusive(uint32_t* __restrict__ in, uint32_t* __restrict__ out, int64_t n, int zero_input,
                 uint64_t* __restrict__ status, uint32_t* __restrict__ counters,
                 uint64_t* __restrict__ total, uint32_t ntiles) {
    __shared__ uint32_t s_vals[SCAN_TILE + SCAN_TILE / 32];
    __shared__ uint32_t s_wave[SCAN_THREADS / 64];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_prefix;
    __shared__ uint32_t s_last;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(&counters[0], 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const int64_t base = (int64_t)tile * SCAN_TILE;

    // Coalesced load into LDS (padded by one word every 32 to avoid bank conflicts).
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = it * SCAN_THREADS + tid;
        const int64_t g = base + idx;
        uint32_t v = 0;
        if (g < n) {
            v = in[g];
            if (zero_input) in[g] = 0u;
        }
        s_vals[idx + (idx >> 5)] = v;
    }
    __syncthreads();

    // Each thread owns SCAN_ITEMS consecutive values.
    uint32_t local[SCAN_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = tid * SCAN_ITEMS + it;
        local[it] = s_vals[idx + (idx >> 5)];
        sum += local[it];
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

    // Publish and look back.
    if (wave == 0) {
        uint32_t prefix = 0;
        if (tile == 0) {
            if (lane == 0) store_status(&status[0], FLAG_INC | (uint64_t)agg);
        } else {
            if (lane == 0) store_status(&status[tile], FLAG_AGG | (uint64_t)agg);
            int64_t pred = (int64_t)tile - 1;
            while (true) {
                const int64_t t = pred - lane;
                uint64_t st = t >= 0 ? load_status(&status[t]) : (FLAG_INC | 0ull);
                // Wait until every predecessor in the window has published something.
                while (__any((st >> 32) == 0)) {
                    __builtin_amdgcn_s_sleep(1);
                    if ((st >> 32) == 0) st = load_status(&status[t]);
                }
                const uint64_t inc_mask = __ballot((st >> 32) == 2);
                uint32_t v = (uint32_t)st;
                if (inc_mask) {
                    const int first = __ffsll((unsigned long long)inc_mask) - 1;  // nearest inclusive
                    if (lane > first) v = 0;
                    prefix += (uint32_t)wave_sum_u64(v);
                    break;
                }
                prefix += (uint32_t)wave_sum_u64(v);
                pred -= 64;
            }
            if (lane == 0) store_status(&status[tile], FLAG_INC | (uint64_t)(uint32_t)(prefix + agg));
        }
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    const uint32_t prefix = s_prefix;

    uint32_t run = prefix + thread_excl;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int idx = tid * SCAN_ITEMS + it;
        s_vals[idx + (idx >> 5)] = run;
        run += local[it];
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

}  // namespace

int64_t scan_tiles(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

hipError_t launch_scan(ScanWorkspace& ws, uint32_t* in, uint32_t* out, int64_t n, bool zero_input,
                       uint64_t* total, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int64_t ntiles = scan_tiles(n);
    if (ntiles > ws.max_tiles) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_exclusive, dim3((uint32_t)ntiles), dim3(SCAN_THREADS), 0, stream, in,
                       out, n, zero_input ? 1 : 0, ws.status, ws.counters, total, (uint32_t)ntiles);
    return hipGetLastError();
}

}  // namespace dem
