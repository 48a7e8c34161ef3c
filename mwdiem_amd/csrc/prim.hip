// Device-wide primitives: exclusive scan and stable LSD radix sort.
//
// Sizes may live in device memory (n_dev) so that a whole timestep can be
// enqueued without reading counts back to the host; grids are sized from the
// host-side capacity and blocks past the live count exit early.
#include "mwd_internal.h"

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 2048

constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / MWD_WAVE;
constexpr int SORT_ROUNDS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ROUNDS;  // 4096
constexpr int RADIX = 256;

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

__global__ void k_scan_reduce(const int32_t* __restrict__ in, int64_t n_cap, const int64_t* n_dev,
                              int64_t* __restrict__ bsum) {
    __shared__ int64_t lw[SORT_WAVES];
    const int64_t n = live_n(n_cap, n_dev);
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t acc = 0;
    if (base < n) {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            int64_t i = base + (int64_t)k * SCAN_THREADS + threadIdx.x;
            if (i < n) acc += in[i];
        }
    }
    int64_t tot;
    block_inclusive_scan(acc, lw, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of nb block sums in place by one 1024-thread block.
__global__ void k_scan_blocks(int64_t* bsum, int64_t nb, int64_t* total) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = t * per;
    int64_t s = 0;
    for (int64_t i = 0; i < per; ++i) {
        int64_t j = b0 + i;
        if (j < nb) s += bsum[j];
    }
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        int64_t y = (t >= d) ? part[t - d] : 0;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    int64_t run = (t == 0) ? 0 : part[t - 1];
    for (int64_t i = 0; i < per; ++i) {
        int64_t j = b0 + i;
        if (j < nb) {
            int64_t v = bsum[j];
            bsum[j] = run;
            run += v;
        }
    }
    if (t == 1023 && total) *total = part[1023];
}

__global__ void k_scan_apply(const int32_t* __restrict__ in, int64_t* __restrict__ out, int64_t n_cap,
                             const int64_t* n_dev, const int64_t* __restrict__ bsum) {
    __shared__ int64_t lw[SORT_WAVES];
    const int64_t n = live_n(n_cap, n_dev);
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    // blocked arrangement: thread t owns items [base + t*ITEMS, +ITEMS)
    int32_t v[SCAN_ITEMS];
    int64_t acc = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + k;
        v[k] = (i < n) ? in[i] : 0;
        acc += v[k];
    }
    int64_t incl = block_inclusive_scan(acc, lw, nullptr);
    int64_t run = incl - acc + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        int64_t i = base + (int64_t)threadIdx.x * SCAN_ITEMS + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

// --------------------------------------------------------------- radix sort
__global__ void k_radix_hist(const uint32_t* __restrict__ keys, int64_t n_cap, const int64_t* n_dev, int shift,
                             int32_t* __restrict__ hist, int64_t ntiles) {
    __shared__ int32_t h[RADIX];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t n = live_n(n_cap, n_dev);
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
    if (base < n) {
        for (int k = 0; k < SORT_ROUNDS; ++k) {
            int64_t i = base + (int64_t)k * SORT_THREADS + threadIdx.x;
            if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xff], 1);
        }
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter. Item order inside a tile is wave-major, then round, then
// lane; ranks among equal digits follow that order, so the sort is stable.
__global__ void k_radix_scatter(const uint32_t* __restrict__ keys, const int32_t* __restrict__ vals,
                                uint32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out, int64_t n_cap,
                                const int64_t* n_dev, int shift, const int64_t* __restrict__ hist_off,
                                int64_t ntiles) {
    __shared__ int32_t cnt[SORT_WAVES][RADIX];
    const int lane = threadIdx.x & (MWD_WAVE - 1);
    const int wave = threadIdx.x / MWD_WAVE;
    for (int i = threadIdx.x; i < SORT_WAVES * RADIX; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int64_t n = live_n(n_cap, n_dev);
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE + (int64_t)wave * (MWD_WAVE * SORT_ROUNDS);
    if ((int64_t)blockIdx.x * SORT_TILE >= n) return;  // uniform per block
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (MWD_WAVE - lane));
    uint32_t k[SORT_ROUNDS];
    int32_t v[SORT_ROUNDS];
    int32_t r[SORT_ROUNDS];
#pragma unroll
    for (int rd = 0; rd < SORT_ROUNDS; ++rd) {
        const int64_t i = base + (int64_t)rd * MWD_WAVE + lane;
        const bool act = i < n;
        k[rd] = act ? keys[i] : 0u;
        v[rd] = act ? vals[i] : 0;
        const uint32_t d = (k[rd] >> shift) & 0xffu;
        uint64_t match = __ballot(act);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
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
    // wave offsets per digit
    {
        const int d = threadIdx.x;  // SORT_THREADS == RADIX
        int32_t run = 0;
#pragma unroll
        for (int w = 0; w < SORT_WAVES; ++w) {
            int32_t t = cnt[w][d];
            cnt[w][d] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int rd = 0; rd < SORT_ROUNDS; ++rd) {
        if (r[rd] < 0) continue;
        const uint32_t d = (k[rd] >> shift) & 0xffu;
        const int64_t pos = hist_off[(int64_t)d * ntiles + blockIdx.x] + cnt[wave][d] + r[rd];
        keys_out[pos] = k[rd];
        vals_out[pos] = v[rd];
    }
}

__global__ void k_set_i64(int64_t* p, int64_t v) { *p = v; }

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

}  // namespace

int64_t mwd_scan_tmp_len(int64_t n_cap) { return (n_cap + SCAN_TILE - 1) / SCAN_TILE + 1; }
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
    if (nb > tmp_len) {
        mwd_set_error("scan scratch too small");
        return MWD_E_CAPACITY;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, in, n_cap, n_dev, tmp);
    MWD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, tmp, nb, total);
    MWD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, in, out, n_cap, n_dev, tmp);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

int mwd_radix_sort(uint32_t* keys, int32_t* vals, uint32_t* keys_alt, int32_t* vals_alt, int64_t n_cap,
                   const int64_t* n_dev, int passes, int32_t* hist, int64_t* hist_off, int64_t* tmp,
                   int64_t tmp_len, hipStream_t s, uint32_t** keys_out, int32_t** vals_out) {
    uint32_t* ka = keys;
    int32_t* va = vals;
    uint32_t* kb = keys_alt;
    int32_t* vb = vals_alt;
    if (n_cap > 0) {
        const int64_t nt = mwd_sort_tiles(n_cap);
        for (int p = 0; p < passes; ++p) {
            const int shift = 8 * p;
            hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nt), dim3(RADIX), 0, s, ka, n_cap, n_dev, shift, hist,
                               nt);
            MWD_LAUNCH_CHECK();
            int rc = mwd_scan_i32(hist, hist_off, nullptr, RADIX * nt, nullptr, tmp, tmp_len, s);
            if (rc) return rc;
            hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nt), dim3(SORT_THREADS), 0, s, ka, va, kb, vb, n_cap,
                               n_dev, shift, hist_off, nt);
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

int mwd_set_i64(int64_t* p, int64_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, s, p, v);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}
