// Device-wide primitives: exclusive scan and stable LSD radix sort.
//
// Sizes may live in device memory (n_dev) so that a whole timestep can be
// enqueued without reading counts back to the host; grids are sized from the
// host-side capacity and blocks past the live count exit early.
#include "mwd_internal.h"

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 4096

constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / MWD_WAVE;
constexpr int SORT_ROUNDS = 4;
constexpr int SORT_TILE = SORT_THREADS * SORT_ROUNDS;  // 1024

__device__ __forceinline__ int64_t live_n(int64_t n_cap, const int64_t* n_dev) {
    if (!n_dev) return n_cap;
    int64_t n = *n_dev;
    return n < n_cap ? n : n_cap;
}

// Inclusive scan across a 256-thread block; returns the inclusive value.
__device__ __forceinline__ int64_t block_inclusive_scan(int64_t v, int64_t* lds_wave /*[4]*/, int64_t* total) {
    const int lane = threadIdx.x & (MWD_WAVE - 1);
    const int wave = threadIdx.x / MWD_WAVE;
    int64_t x = v;
#pragma unroll
    for (int d = 1; d < MWD_WAVE; d <<= 1) {
        int64_t y = __shfl_up(x, d, MWD_WAVE);
        if (lane >= d) x += y;
    }
    if (lane == MWD_WAVE - 1) lds_wave[wave] = x;
    __syncthreads();
    int64_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w) {
        int64_t t = lds_wave[w];
        if (w < wave) off += t;
        tot += t;
    }
    __syncthreads();
    if (total) *total = tot;
    return x + off;
}

// ------------------------------------------------- single-pass scan
// Decoupled look-back (Merrill & Garland, "Single-pass Parallel Prefix Scan
// with Decoupled Look-back", 2016). Blocks past the live tile count exit at
// once; each live block claims a tile through an atomic ticket, publishes the
// tile's aggregate, then its inclusive prefix once known; a successor's first
// wave reads 64 predecessors' status words at a time and adds them up to the
// nearest inclusive one. A status word packs a 2-bit flag above a 62-bit
// value, so a reader needs only that word: loads and stores are relaxed
// device-scope atomics and no fence is needed (a device-scope fence would
// write back or invalidate the XCD's L2).
//
// Scratch: word 0 holds [call:32 | ticket:32]; two status buffers of S words
// follow. A call uses buffer (call & 1) and clears the other for the next
// call; the block that draws the last ticket resets the ticket and advances
// the call number in the same word. So there is one atomic per live tile and
// nothing to clear between calls.
constexpr uint64_t ST_AGG = 1ull << 62;
constexpr uint64_t ST_INC = 2ull << 62;
constexpr uint64_t ST_VAL = (1ull << 62) - 1;

__device__ __forceinline__ uint64_t status_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_dlb(const int32_t* __restrict__ in, int64_t* __restrict__ out,
                                                           int64_t* total, int64_t n_cap, const int64_t* n_dev,
                                                           uint64_t* ctr, uint64_t* status2, int64_t S) {
    __shared__ int64_t lw[SORT_WAVES];
    __shared__ int64_t s_prefix;
    __shared__ uint64_t s_claim;
    const int64_t n = live_n(n_cap, n_dev);
    int64_t ntl = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (ntl < 1) ntl = 1;  // one block runs even when nothing is live: it clears and advances
    if ((int64_t)blockIdx.x >= ntl) return;
    if (threadIdx.x == 0) {
        const uint64_t c = atomicAdd((unsigned long long*)ctr, 1ull);
        s_claim = c;
        if ((int64_t)(c & 0xffffffffull) == ntl - 1) status_store(ctr, ((c >> 32) + 1) << 32);  // last ticket
    }
    __syncthreads();
    const uint64_t claim = s_claim;
    const int64_t tile = (int64_t)(claim & 0xffffffffull);
    const int par = (int)((claim >> 32) & 1);
    uint64_t* status = status2 + par * S;
    uint64_t* other = status2 + (1 - par) * S;
    const int64_t base = tile * SCAN_TILE;
    if (base < n) {
        int32_t v[SCAN_ITEMS];
        int64_t acc = 0;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            const int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + k;
            v[k] = (i < n) ? in[i] : 0;
            acc += v[k];
        }
        int64_t agg;
        const int64_t incl = block_inclusive_scan(acc, lw, &agg);
        if (threadIdx.x < MWD_WAVE) {
            const int lane = threadIdx.x;
            int64_t prefix = 0;
            if (tile == 0) {
                if (lane == 0) status_store(&status[0], ST_INC | (uint64_t)agg);
            } else {
                if (lane == 0) status_store(&status[tile], ST_AGG | (uint64_t)agg);
                int64_t j = tile - 1;  // window: predecessors j, j-1, ..., j-63
                for (;;) {
                    const int64_t t = j - lane;
                    const uint64_t sw = t >= 0 ? status_load(&status[t]) : ST_INC;
                    const uint64_t f = sw & ~ST_VAL;
                    const uint64_t inc = __ballot(f == ST_INC);
                    const uint64_t none = __ballot(f == 0);
                    // lanes up to and including the nearest inclusive predecessor
                    const uint64_t upto = inc ? (inc & (~inc + 1)) * 2 - 1 : ~0ull;
                    if (none & upto) continue;  // a predecessor in range has not published yet
                    int64_t x = ((upto >> lane) & 1) ? (int64_t)(sw & ST_VAL) : 0;
#pragma unroll
                    for (int d = MWD_WAVE / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, MWD_WAVE);
                    prefix += x;
                    if (inc) break;
                    j -= MWD_WAVE;
                }
                if (lane == 0) status_store(&status[tile], ST_INC | (uint64_t)(prefix + agg));
            }
            if (lane == 0) {
                s_prefix = prefix;
                if (total && base + SCAN_TILE >= n) *total = prefix + agg;  // the last live tile
            }
        }
        __syncthreads();
        int64_t run = incl - acc + s_prefix;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            const int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + k;
            if (i < n) out[i] = run;
            run += v[k];
        }
    } else if (total && threadIdx.x == 0) {
        *total = 0;  // nothing live (only tile 0 exists then)
    }
    // clear the other buffer for the next call
    for (int64_t j = tile * SCAN_THREADS + threadIdx.x; j < S; j += ntl * SCAN_THREADS) other[j] = 0;
}

// --------------------------------------------------------------- radix sort
// LSD radix sort with digits of up to BITS bits (2^BITS bins); the digit of a
// pass is (key >> shift) & (2^width - 1).
// The histogram covers the live tiles only: digit d of tile t sits at
// hist[d * nlive + t] for digits d <= mask, and block 0 stores the entry count
// (mask + 1) * nlive in *hcount for the scan, so the scan's cost follows the
// live count rather than the capacity.
__device__ __forceinline__ int64_t live_tiles(int64_t n) { return (n + SORT_TILE - 1) / SORT_TILE; }

template <int BITS>
__global__ void __launch_bounds__(SORT_THREADS) k_radix_hist(const uint32_t* __restrict__ keys, int64_t n_cap,
                                                             const int64_t* n_dev, int shift, uint32_t mask,
                                                             int32_t* __restrict__ hist, int64_t* hcount) {
    constexpr int NB = 1 << BITS;
    __shared__ int32_t h[NB];
    const int64_t n = live_n(n_cap, n_dev);
    const int64_t nlive = live_tiles(n);
    if (blockIdx.x == 0 && threadIdx.x == 0) *hcount = (int64_t)(mask + 1) * nlive;
    if ((int64_t)blockIdx.x >= nlive) return;  // uniform per block
    for (int i = threadIdx.x; i < NB; i += SORT_THREADS) h[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
    for (int k = 0; k < SORT_ROUNDS; ++k) {
        const int64_t i = base + (int64_t)k * SORT_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & mask], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d <= (int)mask; d += SORT_THREADS) hist[(int64_t)d * nlive + blockIdx.x] = h[d];
}

// Stable scatter. Item order inside a tile is wave-major, then round, then
// lane; ranks among equal digits follow that order, so the sort is stable.
template <int BITS>
__global__ void __launch_bounds__(SORT_THREADS) k_radix_scatter(const uint32_t* __restrict__ keys,
                                                                const int32_t* __restrict__ vals,
                                                                uint32_t* __restrict__ keys_out,
                                                                int32_t* __restrict__ vals_out, int64_t n_cap,
                                                                const int64_t* n_dev, int shift, int width,
                                                                const int64_t* __restrict__ hist_off) {
    constexpr int NB = 1 << BITS;
    __shared__ int32_t cnt[SORT_WAVES][NB];
    const int lane = threadIdx.x & (MWD_WAVE - 1);
    const int wave = threadIdx.x / MWD_WAVE;
    for (int i = threadIdx.x; i < SORT_WAVES * NB; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int64_t n = live_n(n_cap, n_dev);
    const int64_t nlive = live_tiles(n);
    if ((int64_t)blockIdx.x >= nlive) return;  // uniform per block
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE + (int64_t)wave * (MWD_WAVE * SORT_ROUNDS);
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (MWD_WAVE - lane));
    const uint32_t mask = (1u << width) - 1u;
    uint32_t k[SORT_ROUNDS];
    int32_t v[SORT_ROUNDS];
    int32_t r[SORT_ROUNDS];
#pragma unroll
    for (int rd = 0; rd < SORT_ROUNDS; ++rd) {
        const int64_t i = base + (int64_t)rd * MWD_WAVE + lane;
        const bool act = i < n;
        k[rd] = act ? keys[i] : 0u;
        v[rd] = act ? vals[i] : 0;
        const uint32_t d = (k[rd] >> shift) & mask;
        uint64_t match = __ballot(act);
        for (int b = 0; b < width; ++b) {
            const uint64_t m = __ballot((d >> b) & 1u);
            match &= ((d >> b) & 1u) ? m : ~m;
        }
        const int below = __popcll(match & lt);
        const int tot = __popcll(match);
        const int pre = cnt[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (act && below == 0) cnt[wave][d] = pre + tot;
        __builtin_amdgcn_wave_barrier();
        r[rd] = act ? pre + below : -1;
    }
    __syncthreads();
    for (int d = threadIdx.x; d < NB; d += SORT_THREADS) {  // wave offsets per digit
        int32_t run = 0;
#pragma unroll
        for (int w = 0; w < SORT_WAVES; ++w) {
            const int32_t t = cnt[w][d];
            cnt[w][d] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int rd = 0; rd < SORT_ROUNDS; ++rd) {
        if (r[rd] < 0) continue;
        const uint32_t d = (k[rd] >> shift) & mask;
        const int64_t pos = hist_off[(int64_t)d * nlive + blockIdx.x] + cnt[wave][d] + r[rd];
        keys_out[pos] = k[rd];
        vals_out[pos] = v[rd];
    }
}

__global__ void k_set_i64(int64_t* p, int64_t v) { *p = v; }

__global__ void k_fill2_i32(int32_t* p, int32_t* q, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) {
        p[i] = v;
        q[i] = v;
    }
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

}  // namespace

int64_t mwd_scan_tmp_len(int64_t n_cap) { return 2 * ((n_cap + SCAN_TILE - 1) / SCAN_TILE + 1) + 1; }
int64_t mwd_sort_tiles(int64_t n_cap) { return (n_cap + SORT_TILE - 1) / SORT_TILE; }

int mwd_scan_i32(const int32_t* in, int64_t* out, int64_t* total, int64_t n_cap, const int64_t* n_dev,
                 int64_t* tmp, int64_t tmp_len, hipStream_t s) {
    if (n_cap <= 0) {
        if (total) {
            hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, s, total, (int64_t)0);
            MWD_LAUNCH_CHECK();
        }
        return MWD_OK;
    }
    const int64_t nb = (n_cap + SCAN_TILE - 1) / SCAN_TILE;
    // scratch layout: [0] call/ticket word, then two status buffers of S words
    const int64_t S = (tmp_len - 1) / 2;
    if (nb > S) {
        mwd_set_error("scan scratch too small");
        return MWD_E_CAPACITY;
    }
    hipLaunchKernelGGL(k_scan_dlb, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, in, out, total, n_cap, n_dev,
                       reinterpret_cast<uint64_t*>(tmp), reinterpret_cast<uint64_t*>(tmp + 1), S);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_radix_sort(uint32_t* keys, int32_t* vals, uint32_t* keys_alt, int32_t* vals_alt, int64_t n_cap,
                   const int64_t* n_dev, int key_bits, int32_t* hist, int64_t* hist_off, int64_t* hcount,
                   int64_t* tmp, int64_t tmp_len, hipStream_t s, uint32_t** keys_out, int32_t** vals_out) {
    uint32_t* ka = keys;
    int32_t* va = vals;
    uint32_t* kb = keys_alt;
    int32_t* vb = vals_alt;
    if (key_bits < 1) key_bits = 1;
    const int passes = (key_bits + MWD_SORT_MAX_BITS - 1) / MWD_SORT_MAX_BITS;
    const int width = (key_bits + passes - 1) / passes;
    if (n_cap > 0) {
        const int64_t nt = mwd_sort_tiles(n_cap);
        const bool wide = width > 8;
        const int64_t bins = wide ? (1 << MWD_SORT_MAX_BITS) : 256;
        for (int p = 0; p < passes; ++p) {
            const int shift = width * p;
            const int w = (shift + width <= key_bits) ? width : key_bits - shift;
            const uint32_t mask = (1u << w) - 1u;
            if (wide)
                hipLaunchKernelGGL(k_radix_hist<MWD_SORT_MAX_BITS>, dim3((unsigned)nt), dim3(SORT_THREADS), 0, s, ka,
                                   n_cap, n_dev, shift, mask, hist, hcount);
            else
                hipLaunchKernelGGL(k_radix_hist<8>, dim3((unsigned)nt), dim3(SORT_THREADS), 0, s, ka, n_cap, n_dev,
                                   shift, mask, hist, hcount);
            MWD_LAUNCH_CHECK();
            int rc = mwd_scan_i32(hist, hist_off, nullptr, bins * nt, hcount, tmp, tmp_len, s);
            if (rc) return rc;
            if (wide)
                hipLaunchKernelGGL(k_radix_scatter<MWD_SORT_MAX_BITS>, dim3((unsigned)nt), dim3(SORT_THREADS), 0, s,
                                   ka, va, kb, vb, n_cap, n_dev, shift, w, hist_off);
            else
                hipLaunchKernelGGL(k_radix_scatter<8>, dim3((unsigned)nt), dim3(SORT_THREADS), 0, s, ka, va, kb, vb,
                                   n_cap, n_dev, shift, w, hist_off);
            MWD_LAUNCH_CHECK();
            uint32_t* tk = ka; ka = kb; kb = tk;
            int32_t* tv = va; va = vb; vb = tv;
        }
    }
    // the result sits in (keys, vals) after an even number of passes, in the
    // alternate buffers after an odd number
    *keys_out = ka;
    *vals_out = va;
    return MWD_OK;
}

int mwd_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return MWD_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_fill2_i32(int32_t* p, int32_t* q, int32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return MWD_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill2_i32, dim3((unsigned)blocks), dim3(256), 0, s, p, q, v, n);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_set_i64(int64_t* p, int64_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, s, p, v);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}
