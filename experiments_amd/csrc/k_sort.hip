// k_sort.hip -- key/value LSD radix sort, exclusive scan and bound search for
// gfx950.  Replaces btRadixSort32CL (opencl/broadphase_benchmark/btRadixSort32CL.cpp:161-535,
// kernels primitives/AdlPrimitives/Sort/RadixSort32Kernels.cl), btPrefixScanCL
// (btPrefixScanCL.cpp:46-99) and btBoundSearchCL (btBoundSearchCL.cpp:78-136).
//
// Sort: one histogram kernel counts every digit of every pass in a single
// read of the keys; then one "onesweep" scatter kernel per pass ranks a
// tile of 4096 elements (1024 for sorts of up to 256k elements, so that
// they still span many workgroups) in LDS (wave ballots give the stable rank of each key
// among equal digits), obtains the tile's global digit offsets through a
// decoupled look-back over the preceding tiles, stages the tile in digit
// order in LDS and writes it out in contiguous runs.  Tiles take their index
// from an atomic counter in launch order, so every look-back waits only on
// tiles that are already running.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "dmath.h"
#include "rb_internal.h"

namespace {

constexpr int SORT_BLOCK = 256;            // 4 waves
constexpr int SORT_ITEMS = 16;             // keys per lane
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;  // 4096
constexpr int SORT_ITEMS_SMALL = 4;        // keys per lane of the small tile
constexpr int SORT_SMALL_MAX_N = 1 << 18;  // sorts up to this size use the small tile
constexpr int WAVES = SORT_BLOCK / 64;
static_assert(SORT_TILE == RB_SORT_TILE, "host-side tile count");
static_assert(SORT_BLOCK * SORT_ITEMS_SMALL == RB_SORT_TILE_SMALL, "host-side tile count");
constexpr int MAX_PASSES = 8;

constexpr int LB = 8;  // look-back words read per round trip (per digit)
constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_INC = 2u << 30;
constexpr uint32_t VAL_MASK = (1u << 30) - 1u;

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// The zero fills a sort starts from (digit counts, tile counters) in one
// launch; the histogram kernel clears the look-back status words.
__global__ void k_sort_clear(uint32_t *hist, size_t n_hist, uint32_t *tile_ctr, size_t n_ctr) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = i0; i < n_hist; i += stride) hist[i] = 0u;
    for (size_t i = i0; i < n_ctr; i += stride) tile_ctr[i] = 0u;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

template <int RBITS, int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_histogram(const uint2 *kv, int n, int passes,
                                                          int sort_bits, uint32_t *hist,
                                                          const int *n_dev, uint32_t *status,
                                                          int tiles) {
    constexpr int RADIX = 1 << RBITS;
    constexpr int TILE = SORT_BLOCK * ITEMS;
    if (n_dev && *n_dev < n) n = *n_dev;
    if ((int)blockIdx.x * TILE >= n) return;
    // this tile's look-back status words of every pass (the tiles beyond the
    // element count neither publish nor are read)
    for (int i = threadIdx.x; i < passes * RADIX; i += SORT_BLOCK)
        status[((size_t)(i / RADIX) * tiles + blockIdx.x) * RADIX + (i % RADIX)] = 0u;
    __shared__ uint32_t h[MAX_PASSES * RADIX];
    for (int i = threadIdx.x; i < passes * RADIX; i += SORT_BLOCK) h[i] = 0;
    __syncthreads();
    int base = blockIdx.x * TILE;
    uint32_t keys[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        int i = base + k * SORT_BLOCK + threadIdx.x;
        keys[k] = i < n ? kv[i].x : 0u;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        int i = base + k * SORT_BLOCK + threadIdx.x;
        if (i < n) {
            uint32_t key = keys[k];
            for (int p = 0; p < passes; p++) {
                int shift = p * RBITS;
                int bits = sort_bits - shift < RBITS ? sort_bits - shift : RBITS;
                uint32_t d = (key >> shift) & ((1u << bits) - 1u);
                atomicAdd(&h[p * RADIX + d], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * RADIX; i += SORT_BLOCK)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

template <int RBITS, int ITEMS>
__global__ __launch_bounds__(SORT_BLOCK) void k_onesweep(const uint2 *__restrict__ src,
                                                         uint2 *__restrict__ dst, int n, int shift,
                                                         int bits, const uint32_t *hist,
                                                         uint32_t *status, uint32_t *tile_ctr,
                                                         uint32_t *err, uint32_t err_bit,
                                                         const int *n_dev) {
    constexpr int RADIX = 1 << RBITS;
    constexpr int TILE = SORT_BLOCK * ITEMS;
    if (n_dev && *n_dev < n) n = *n_dev;
    __shared__ uint32_t s_wave_hist[WAVES][RADIX];
    __shared__ uint32_t s_digit_start[RADIX];   // tile-local start of each digit
    __shared__ uint32_t s_global[RADIX];        // global base + tile prefix of each digit
    __shared__ uint2 s_stage[TILE];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_scan[RADIX];

    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    for (int i = tid; i < WAVES * RADIX; i += SORT_BLOCK) (&s_wave_hist[0][0])[i] = 0;
    __syncthreads();
    const int tile = (int)s_tile;
    if (tile * TILE >= n) return;  // beyond the device-side count: nothing looks back here
    const int base = tile * TILE + wave * (ITEMS * 64);
    const uint32_t dmask = (1u << bits) - 1u;

    uint2 item[ITEMS];
    uint32_t rank[ITEMS];
    // every load of the tile is in flight before the first ranking step
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        int i = base + r * 64 + lane;
        item[r] = i < n ? src[i] : make_uint2(0xffffffffu, 0xffffffffu);
    }
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        int i = base + r * 64 + lane;
        bool valid = i < n;
        uint32_t d = (item[r].x >> shift) & dmask;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RBITS; b++) {
            uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t below = (uint32_t)__popcll(peers & rb::lanemask_lt());
        uint32_t before = 0;
        if (valid) before = s_wave_hist[wave][d];
        rb::wave_sync();
        if (valid && below == 0) s_wave_hist[wave][d] = before + (uint32_t)__popcll(peers);
        rb::wave_sync();
        rank[r] = before + below;
    }
    __syncthreads();

    // per digit: prefix over waves, tile total, publish and look back
    for (int d = tid; d < RADIX; d += SORT_BLOCK) {
        uint32_t run = 0;
        for (int w = 0; w < WAVES; w++) {
            uint32_t c = s_wave_hist[w][d];
            s_wave_hist[w][d] = run;
            run += c;
        }
        s_scan[d] = run;
        uint32_t *st = status + (size_t)tile * RADIX + d;
        if (tile == 0) {
            __hip_atomic_store(st, FLAG_INC | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_global[d] = 0;
        } else {
            __hip_atomic_store(st, FLAG_AGG | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // the predecessors' words LB at a time (independent loads, one
            // round trip), summed back to the nearest inclusive one; an
            // unpublished word is read again
            uint32_t prefix = 0, spins = 0;
            int t = tile - 1;
            while (true) {
                uint32_t v[LB];
#pragma unroll
                for (int k = 0; k < LB; k++)
                    v[k] = t - k >= 0 ? __hip_atomic_load(status + (size_t)(t - k) * RADIX + d,
                                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                      : FLAG_INC;  // (never reached: tile 0 is inclusive)
                int k = 0;
                bool done = false;
                for (; k < LB; k++) {
                    if ((v[k] & ~VAL_MASK) == 0) break;  // not published yet
                    prefix += v[k] & VAL_MASK;
                    if (v[k] & FLAG_INC) {
                        done = true;
                        break;
                    }
                }
                if (done) break;
                t -= k;
                if (k < LB && ++spins > (1u << 26)) { atomicOr(err, err_bit); break; }
            }
            __hip_atomic_store(st, FLAG_INC | (prefix + run), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            s_global[d] = prefix;
        }
    }
    __syncthreads();
    // tile-local digit starts and the global digit bases of this pass: block
    // exclusive scans over the digits (one digit per thread)
    {
        uint32_t c = tid < RADIX ? s_scan[tid] : 0u;
        uint32_t g = tid < RADIX ? hist[tid] : 0u;
        uint32_t ci = wave_incl_scan(c), gi = wave_incl_scan(g);
        __shared__ uint32_t s_wc[WAVES], s_wg[WAVES];
        if (lane == 63) { s_wc[wave] = ci; s_wg[wave] = gi; }
        __syncthreads();
        uint32_t oc = 0, og = 0;
        for (int w = 0; w < wave; w++) { oc += s_wc[w]; og += s_wg[w]; }
        if (tid < RADIX) {
            s_digit_start[tid] = oc + ci - c;
            s_global[tid] += og + gi - g;
        }
    }
    __syncthreads();
    // stage in digit order
    for (int r = 0; r < ITEMS; r++) {
        int i = base + r * 64 + lane;
        if (i < n) {
            uint32_t d = (item[r].x >> shift) & dmask;
            uint32_t lp = s_digit_start[d] + s_wave_hist[wave][d] + rank[r];
            s_stage[lp] = item[r];
        }
    }
    __syncthreads();
    int tile_n = n - tile * TILE;
    if (tile_n > TILE) tile_n = TILE;
    for (int k = tid; k < tile_n; k += SORT_BLOCK) {
        uint2 e = s_stage[k];
        uint32_t d = (e.x >> shift) & dmask;
        uint32_t g = s_global[d] + ((uint32_t)k - s_digit_start[d]);
        dst[g] = e;
    }
}

// ---- single-pass exclusive scan of uint32 with decoupled look-back ----
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
static_assert(SCAN_TILE == RB_SCAN_TILE, "host-side status size");
constexpr uint64_t SFLAG_AGG = 1ull << 62;
constexpr uint64_t SFLAG_INC = 2ull << 62;
constexpr uint64_t SVAL_MASK = (1ull << 62) - 1ull;

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan(const uint32_t *in, uint32_t *out, int n,
                                                     uint32_t *total, uint64_t *status,
                                                     uint32_t *tile_ctr, const int *n_dev,
                                                     uint32_t *err, uint32_t err_bit,
                                                     uint32_t *idx_out,
                                                     const uint32_t *idx_base) {
    if (n_dev && *n_dev < n) n = *n_dev;
    __shared__ uint32_t s_wave[SCAN_BLOCK / 64];
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_prefix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    const int tile = (int)s_tile;
    if (n <= 0) {
        if (total && tile == 0 && tid == 0) *total = 0;
        return;
    }
    if (tile * SCAN_TILE >= n) return;
    const int base = tile * SCAN_TILE + tid * SCAN_ITEMS;  // blocked arrangement
    uint32_t v[SCAN_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int i = base + k;
        v[k] = i < n ? in[i] : 0u;
        sum += v[k];
    }
    uint32_t incl = wave_incl_scan(sum);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0, tile_sum = 0;
    for (int w = 0; w < SCAN_BLOCK / 64; w++) {
        uint32_t c = s_wave[w];
        if (w < wave) wave_off += c;
        tile_sum += c;
    }
    if (wave == 0) {
        uint64_t *st = status + tile;
        if (tile == 0) {
            if (lane == 0) {
                __hip_atomic_store(st, SFLAG_INC | tile_sum, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                s_prefix = 0;
            }
        } else {
            if (lane == 0)
                __hip_atomic_store(st, SFLAG_AGG | tile_sum, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            // lane k reads the word of tile t - k: 64 predecessors per round
            // trip, summed back to the nearest inclusive one; from an
            // unpublished word on, the words are read again
            uint64_t prefix = 0;
            uint32_t spins = 0;
            int t = tile - 1;
            while (true) {
                const int tt = t - lane;
                const uint64_t w = tt >= 0 ? __hip_atomic_load(status + tt, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT)
                                           : SFLAG_INC;  // (never reached: tile 0 is inclusive)
                const uint64_t unpub = __ballot((w & ~SVAL_MASK) == 0);
                const uint64_t inc = __ballot((w & SFLAG_INC) != 0);
                const uint64_t stop = unpub | inc;
                const int j = stop ? __ffsll((long long)stop) - 1 : 64;
                const bool j_inc = j < 64 && ((inc >> j) & 1ull);
                uint64_t mine = (lane < j || (lane == j && j_inc)) ? (w & SVAL_MASK) : 0ull;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) mine += __shfl_xor(mine, o, 64);
                prefix += mine;
                if (j_inc) break;
                t -= j;
                if (j < 64 && ++spins > (1u << 26)) {
                    if (lane == 0) atomicOr(err, err_bit);
                    break;
                }
            }
            if (lane == 0) {
                __hip_atomic_store(st, SFLAG_INC | (prefix + tile_sum), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                s_prefix = prefix;
            }
        }
    }
    __syncthreads();
    uint32_t run = (uint32_t)s_prefix + wave_off + (incl - sum);
    const uint32_t ibase = idx_out && idx_base ? *idx_base : 0u;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int i = base + k;
        if (out && i < n) out[i] = run;
        if (idx_out && v[k]) idx_out[ibase + run] = (uint32_t)i;  // flags: the set indices, in order
        run += v[k];
    }
    int last_tile = (n - 1) / SCAN_TILE;
    if (total && tile == last_tile && tid == SCAN_BLOCK - 1) *total = run;
}

__global__ void k_scan_empty(uint32_t *total) { if (total) *total = 0; }

__global__ void k_copy_kv(const uint2 *src, uint2 *dst, int n, const int *n_dev) {
    if (n_dev && *n_dev < n) n = *n_dev;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_bound_search(const uint2 *kv, int n, uint32_t *dst, int n_dst, int mode) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    uint32_t prev = i > 0 ? kv[i - 1].x : 0xffffffffu;
    uint32_t cur = i < n ? kv[i].x : 0xffffffffu;
    bool start = i < n && (i == 0 || prev != cur);
    bool end = i > 0 && (i == n || prev != cur);
    if (mode == 0) {
        if (start && cur < (uint32_t)n_dst) dst[cur] = (uint32_t)i;
    } else if (mode == 1) {
        if (end && prev < (uint32_t)n_dst) dst[prev] = (uint32_t)i;
    } else {
        // count = (end index) - (start index), built from two wrapping adds
        if (start && cur < (uint32_t)n_dst) atomicAdd(&dst[cur], 0u - (uint32_t)i);
        if (end && prev < (uint32_t)n_dst) atomicAdd(&dst[prev], (uint32_t)i);
    }
}

template <int RBITS, int ITEMS>
hipError_t radix_sort_tiles(uint2 *kv, int n, int sort_bits, const SortWork &w, hipStream_t s,
                            const int *n_dev) {
    constexpr int RADIX = 1 << RBITS;
    constexpr int TILE = SORT_BLOCK * ITEMS;
    const int passes = (sort_bits + RBITS - 1) / RBITS;
    const int tiles = (n + TILE - 1) / TILE;
    if (passes > MAX_PASSES || tiles > w.max_tiles) return hipErrorInvalidValue;
    // (the error word after the tile counters is sticky: not cleared)
    if (!w.precleared)
        RB_LAUNCH(k_sort_clear, dim3(8), dim3(256), 0, s, w.hist, (size_t)passes * RADIX,
                  w.tile_ctr, (size_t)MAX_PASSES);
    RB_LAUNCH((k_histogram<RBITS, ITEMS>), dim3(tiles), dim3(SORT_BLOCK), 0, s, kv, n, passes,
                       sort_bits, w.hist, n_dev, w.status, tiles);
    // look-back timeouts: a sticky bit in the caller's overflow word, else
    // in the word after the tile counters
    uint32_t *err = w.err ? reinterpret_cast<uint32_t *>(w.err) : w.tile_ctr + MAX_PASSES;
    const uint32_t err_bit = w.err ? (uint32_t)MI_RB_OVERFLOW_LOOKBACK : 1u;
    uint2 *src = kv, *dst = w.tmp;
    for (int p = 0; p < passes; p++) {
        int shift = p * RBITS;
        int bits = sort_bits - shift < RBITS ? sort_bits - shift : RBITS;
        RB_LAUNCH((k_onesweep<RBITS, ITEMS>), dim3(tiles), dim3(SORT_BLOCK), 0, s, src, dst, n,
                           shift, bits, w.hist + (size_t)p * RADIX,
                           w.status + (size_t)p * tiles * RADIX, w.tile_ctr + p, err, err_bit,
                           n_dev);
        uint2 *t = src; src = dst; dst = t;
    }
    if (src != kv)  // odd pass count: copy back with a kernel (graph-capture friendly)
        RB_LAUNCH(k_copy_kv, dim3(256), dim3(256), 0, s, src, kv, n, n_dev);
    return rb_launch_status();
}

// The tile size: 1024 elements for sorts of up to SORT_SMALL_MAX_N (more
// workgroups, fewer keys ranked per lane) when the workspace holds that many
// tiles, else 4096; tile_keys (1024 or 4096) overrides the size rule.
template <int RBITS>
hipError_t radix_sort_impl(uint2 *kv, int n, int sort_bits, const SortWork &w, hipStream_t s,
                           const int *n_dev, int tile_keys) {
    const int small_tiles = (n + RB_SORT_TILE_SMALL - 1) / RB_SORT_TILE_SMALL;
    const bool small = tile_keys == RB_SORT_TILE_SMALL ||
                       (tile_keys != RB_SORT_TILE && n <= SORT_SMALL_MAX_N);
    if (small && small_tiles <= w.max_tiles)
        return radix_sort_tiles<RBITS, SORT_ITEMS_SMALL>(kv, n, sort_bits, w, s, n_dev);
    return radix_sort_tiles<RBITS, SORT_ITEMS>(kv, n, sort_bits, w, s, n_dev);
}

}  // namespace

namespace {
thread_local const char *g_bad_launch = nullptr;
}

bool rb_launch_ok(dim3 grid, dim3 block, const char *kernel) {
    const unsigned long long threads = (unsigned long long)block.x * block.y * block.z;
    const bool ok = grid.x >= 1 && grid.y >= 1 && grid.z >= 1 && grid.x <= 0x7fffffffu &&
                    grid.y <= 65535u && grid.z <= 65535u && threads >= 1 && threads <= 1024;
    if (!ok) {
        std::fprintf(stderr, "mi_rb: launch of %s refused: grid (%u, %u, %u), block (%u, %u, %u)\n",
                     kernel, grid.x, grid.y, grid.z, block.x, block.y, block.z);
        g_bad_launch = kernel;
    }
    return ok;
}

hipError_t rb_launch_status() {
    hipError_t e = hipGetLastError();
    if (g_bad_launch) {
        g_bad_launch = nullptr;
        return hipErrorInvalidConfiguration;
    }
    return e;
}

size_t sort_status_words(int max_n, int digit_bits) {
    int tiles = rb_sort_tiles(max_n);  // small tiles
    if (tiles < 1) tiles = 1;
    return (size_t)MAX_PASSES * tiles * (1u << digit_bits);
}

hipError_t launch_radix_sort(uint2 *kv, int n, int sort_bits, int digit_bits, const SortWork &w,
                             hipStream_t s, const int *n_dev, int tile_keys) {
    if (n <= 1 || sort_bits <= 0) return hipSuccess;
    switch (digit_bits) {
        case 4: return radix_sort_impl<4>(kv, n, sort_bits, w, s, n_dev, tile_keys);
        case 6: return radix_sort_impl<6>(kv, n, sort_bits, w, s, n_dev, tile_keys);
        case 7: return radix_sort_impl<7>(kv, n, sort_bits, w, s, n_dev, tile_keys);
        case 8: return radix_sort_impl<8>(kv, n, sort_bits, w, s, n_dev, tile_keys);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_exclusive_scan(const uint32_t *in, uint32_t *out, int n, uint32_t *total,
                                 uint32_t *status, hipStream_t s, const int *n_dev, int *err,
                                 uint32_t *idx_out, const uint32_t *idx_base, bool precleared) {
    if (n <= 0) {
        RB_LAUNCH(k_scan_empty, dim3(1), dim3(1), 0, s, total);
        return rb_launch_status();
    }
    int tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    // status: [0] tile counter, then 64-bit words (8-byte aligned)
    uint32_t *ctr = status;
    uint64_t *st = reinterpret_cast<uint64_t *>(status + 2);
    if (!precleared)
        RB_LAUNCH(k_fill_u32, dim3(64), dim3(256), 0, s, status, 0u, (size_t)(2 + 2 * tiles));
    // a look-back timeout sets MI_RB_OVERFLOW_LOOKBACK in *err, else the
    // word after the tile counter (cleared with the status words)
    uint32_t *e = err ? reinterpret_cast<uint32_t *>(err) : status + 1;
    const uint32_t bit = err ? (uint32_t)MI_RB_OVERFLOW_LOOKBACK : 1u;
    RB_LAUNCH(k_scan, dim3(tiles), dim3(SCAN_BLOCK), 0, s, in, out, n, total, st, ctr,
                       n_dev, e, bit, idx_out, idx_base);
    return rb_launch_status();
}

hipError_t launch_compact_flags(const uint32_t *flags, uint32_t *idx_out, int n, uint32_t *total,
                                uint32_t *status, hipStream_t s, const int *n_dev, int *err,
                                const uint32_t *idx_base, bool precleared) {
    return launch_exclusive_scan(flags, nullptr, n, total, status, s, n_dev, err, idx_out,
                                 idx_base, precleared);
}

hipError_t launch_bound_search(const uint2 *kv, int n, uint32_t *dst, int n_dst, int mode,
                               hipStream_t s) {
    if (mode == 2)
        RB_LAUNCH(k_fill_u32, dim3(256), dim3(256), 0, s, dst, 0u, (size_t)n_dst);
    RB_LAUNCH(k_bound_search, dim3((n + 1 + 255) / 256), dim3(256), 0, s, kv, n, dst,
                       n_dst, mode);
    return rb_launch_status();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    RB_LAUNCH(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, v, n);
    return rb_launch_status();
}

// ---------------------------------------------------------------- C ABI
struct RadixSortHandle {
    SortWork w;
    uint32_t *scan_status;
};

extern "C" mi_rb_status mi_rb_radix_sort_create(int max_elements, void **out_sort) {
    if (!out_sort || max_elements <= 0) return MI_RB_INVALID_ARG;
    RadixSortHandle *h = new RadixSortHandle();
    h->w.max_n = max_elements;
    h->w.max_tiles = rb_sort_tiles(max_elements);
    size_t status_words = sort_status_words(max_elements, 8);
    int scan_tiles = (max_elements + SCAN_TILE - 1) / SCAN_TILE;
    if (hipMalloc(&h->w.tmp, sizeof(uint2) * (size_t)max_elements) != hipSuccess ||
        hipMalloc(&h->w.hist, sizeof(uint32_t) * MAX_PASSES * 256) != hipSuccess ||
        hipMalloc(&h->w.status, sizeof(uint32_t) * status_words) != hipSuccess ||
        hipMalloc(&h->w.tile_ctr, sizeof(uint32_t) * (MAX_PASSES + 1)) != hipSuccess ||
        hipMalloc(&h->scan_status, sizeof(uint32_t) * (size_t)(4 + 2 * scan_tiles)) != hipSuccess) {
        delete h;
        return MI_RB_DEVICE;
    }
    *out_sort = h;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_radix_sort_destroy(void *sort) {
    RadixSortHandle *h = static_cast<RadixSortHandle *>(sort);
    if (!h) return MI_RB_INVALID_ARG;
    hipFree(h->w.tmp);
    hipFree(h->w.hist);
    hipFree(h->w.status);
    hipFree(h->w.tile_ctr);
    hipFree(h->scan_status);
    delete h;
    return MI_RB_OK;
}

extern "C" mi_rb_status mi_rb_radix_sort_kv(void *sort, void *d_keyvalue, int n, int sort_bits,
                                            void *stream) {
    RadixSortHandle *h = static_cast<RadixSortHandle *>(sort);
    if (!h || !d_keyvalue || n < 0 || n > h->w.max_n || sort_bits < 0 || sort_bits > 32)
        return MI_RB_INVALID_ARG;
    hipError_t e = launch_radix_sort(static_cast<uint2 *>(d_keyvalue), n, sort_bits, 8, h->w,
                                     static_cast<hipStream_t>(stream));
    return e == hipSuccess ? MI_RB_OK : MI_RB_DEVICE;
}

// ---- the same sort with its workspace in a file-scope device array ----
// Layout of `scratch` (32-bit words): the ping-pong copy of the pairs
// (2 per element), the digit counts of every pass, the look-back status words
// (passes x tiles x 256) and the tile counters with the look-back error word.
constexpr int NA_MAX_N = MI_RB_SORT_NOALLOC_MAX;
constexpr int NA_TILES = (NA_MAX_N + SORT_TILE - 1) / SORT_TILE;
constexpr int NA_PASSES = 4;  // 32 key bits in 8-bit digits
constexpr size_t NA_TMP = 0;
constexpr size_t NA_HIST = NA_TMP + 2 * (size_t)NA_MAX_N;
constexpr size_t NA_STATUS = NA_HIST + (size_t)MAX_PASSES * 256;
constexpr size_t NA_CTR = NA_STATUS + (size_t)NA_PASSES * NA_TILES * 256;
constexpr size_t NA_WORDS = NA_CTR + 64;
static_assert(NA_HIST % 64 == 0 && NA_STATUS % 64 == 0 && NA_CTR % 64 == 0, "256-byte sections");

static __device__ __align__(256) uint32_t scratch[NA_WORDS];

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;
uint32_t *g_na_ptr[NA_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
uint32_t *noalloc_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_na_ptr[dev] = static_cast<uint32_t *>(p);
    }
    return g_na_ptr[dev];
}
}  // namespace

extern "C" mi_rb_status mi_rb_radix_sort_kv_noalloc(void *d_keyvalue, int n, int sort_bits,
                                                    void *stream) {
    if (!d_keyvalue || n < 0 || sort_bits < 0 || sort_bits > 32) return MI_RB_INVALID_ARG;
    if (n > NA_MAX_N) return MI_RB_CAPACITY;
    uint32_t *base = noalloc_scratch();
    if (!base) return MI_RB_DEVICE;
    SortWork w;
    w.tmp = reinterpret_cast<uint2 *>(base + NA_TMP);
    w.hist = base + NA_HIST;
    w.status = base + NA_STATUS;
    w.tile_ctr = base + NA_CTR;
    w.scan_status = nullptr;
    w.err = nullptr;
    w.max_n = NA_MAX_N;
    w.max_tiles = NA_TILES;
    // the look-back error word keeps the previous call's value: clear it too
    if (n > 1 && sort_bits > 0) {
        RB_LAUNCH(k_fill_u32, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                           w.tile_ctr + MAX_PASSES, 0u, (size_t)1);
        if (rb_launch_status() != hipSuccess) return MI_RB_DEVICE;
    }
    hipError_t e = launch_radix_sort(static_cast<uint2 *>(d_keyvalue), n, sort_bits, 8, w,
                                     static_cast<hipStream_t>(stream));
    return e == hipSuccess ? MI_RB_OK : MI_RB_DEVICE;
}

extern "C" mi_rb_status mi_rb_exclusive_scan_u32(void *sort, const void *d_in, void *d_out, int n,
                                                 void *d_total, void *stream) {
    RadixSortHandle *h = static_cast<RadixSortHandle *>(sort);
    if (!h || (!d_in && n) || (!d_out && n) || n < 0 || n > h->w.max_n) return MI_RB_INVALID_ARG;
    hipError_t e = launch_exclusive_scan(static_cast<const uint32_t *>(d_in),
                                         static_cast<uint32_t *>(d_out), n,
                                         static_cast<uint32_t *>(d_total), h->scan_status,
                                         static_cast<hipStream_t>(stream));
    return e == hipSuccess ? MI_RB_OK : MI_RB_DEVICE;
}

extern "C" mi_rb_status mi_rb_bound_search(const void *d_keyvalue, int n, void *d_dst, int n_dst,
                                           int mode, void *stream) {
    if ((!d_keyvalue && n) || !d_dst || n < 0 || n_dst <= 0 || mode < 0 || mode > 2)
        return MI_RB_INVALID_ARG;
    hipError_t e = launch_bound_search(static_cast<const uint2 *>(d_keyvalue), n,
                                       static_cast<uint32_t *>(d_dst), n_dst, mode,
                                       static_cast<hipStream_t>(stream));
    return e == hipSuccess ? MI_RB_OK : MI_RB_DEVICE;
}
