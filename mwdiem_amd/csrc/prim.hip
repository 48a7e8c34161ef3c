// Device-wide primitives: exclusive scan and stable LSD radix sort.
//
// Sizes may live in device memory (n_dev) so that a whole timestep can be
// enqueued without reading counts back to the host; grids are sized from the
// host-side capacity and blocks past the live count exit early.
#include "mwd_internal.h"
#include "mwd_scan.h"

namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / MWD_WAVE;
constexpr int SORT_ROUNDS = 4;
constexpr int SORT_TILE = SORT_THREADS * SORT_ROUNDS;  // 1024

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

struct CopyJob {
    const uint64_t* src[MWD_MAX_COPY];
    uint64_t* dst[MWD_MAX_COPY];
    int64_t end[MWD_MAX_COPY];  // running word counts: array i holds words [end[i-1], end[i])
    int n;
};

// All arrays of the job as one grid-strided range of words.
__global__ void k_copy_arrays(CopyJob job) {
    const int64_t total = job.end[job.n - 1];
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += st) {
        int i = 0;
        while (w >= job.end[i]) ++i;
        const int64_t o = w - (i ? job.end[i - 1] : 0);
        job.dst[i][o] = job.src[i][o];
    }
}

}  // namespace

int mwd_launch_copy_arrays(int n, void* const* dst, const void* const* src, const int64_t* words, hipStream_t s) {
    if (n <= 0) return MWD_OK;
    CopyJob job;
    int64_t acc = 0;
    for (int i = 0; i < n; ++i) {
        job.src[i] = static_cast<const uint64_t*>(src[i]);
        job.dst[i] = static_cast<uint64_t*>(dst[i]);
        acc += words[i];
        job.end[i] = acc;
    }
    job.n = n;
    if (acc == 0) return MWD_OK;
    int64_t blocks = (acc + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_copy_arrays, dim3((unsigned)blocks), dim3(256), 0, s, job);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int64_t mwd_scan_tmp_len(int64_t n_cap) {
    const int64_t smallest_tile = SCAN_THREADS;  // one item per thread
    return 2 * ((n_cap + smallest_tile - 1) / smallest_tile + 1) + 1;
}
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
    return mwd_scan_io(ScanIO{in, out}, n_cap, n_dev, total, tmp, tmp_len, s);
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
