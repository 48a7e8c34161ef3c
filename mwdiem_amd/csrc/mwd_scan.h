// Single-pass exclusive scan over int32 counts, fused with what produces the
// counts and what consumes the offsets.
//
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
//
// IO is a small struct with
//   __device__ V load(int64_t i) const;                   count of item i
//   __device__ void store(int64_t i, int64_t off, V v) const;
//                                                         its exclusive offset
//   __device__ void finish(int64_t total) const;          once, with the total
// load/store are called for the live items i < n only. V is int32_t or
// int64_t (an int64 count may pack two counts, as hi << 32 | lo, whose sums
// stay below 2^62).
#pragma once
#include <type_traits>
#include <utility>
#include "mwd_internal.h"

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_WAVES = SCAN_THREADS / MWD_WAVE;
// Items per thread: 16 (4,096-item tiles) for long scans, 4 (1,024-item
// tiles) below a million items, where more tiles give more blocks in flight.
constexpr int SCAN_ITEMS_LONG = 16;
constexpr int SCAN_ITEMS_SHORT = 4;
constexpr int64_t SCAN_SHORT_MAX = 1 << 20;

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
    for (int w = 0; w < SCAN_WAVES; ++w) {
        int64_t t = lds_wave[w];
        if (w < wave) off += t;
        tot += t;
    }
    __syncthreads();
    if (total) *total = tot;
    return x + off;
}

constexpr uint64_t ST_AGG = 1ull << 62;
constexpr uint64_t ST_INC = 2ull << 62;
constexpr uint64_t ST_VAL = (1ull << 62) - 1;

__device__ __forceinline__ uint64_t status_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tile items are loaded and stored in striped order (item k*256 + t by
// thread t: consecutive lanes touch consecutive items) and scanned in blocked
// order (items 16t .. 16t+15 by thread t), transposed through LDS. The LDS
// index is padded by one word per 32 to keep the blocked accesses off a
// single bank.
__device__ __forceinline__ int scan_pad(int i) { return i + (i >> 5); }

// LDS of one block of a scan: the tile's counts and offsets (padded, see
// scan_pad), the waves' partial sums and the tile's prefix and ticket.
template <class V, int SCAN_ITEMS>
struct ScanLds {
    static constexpr int TILE = SCAN_THREADS * SCAN_ITEMS;
    static constexpr int LEN = TILE + TILE / 32;
    int64_t lw[SCAN_WAVES];
    int64_t prefix;
    uint64_t claim;
    V v[LEN];
    int64_t off[LEN];
};

// One block's part of a scan: claim a tile, scan it, look back, store. `bid`
// is the block's index among the scan's blocks (those at or past the live
// tile count return at once).
template <class IO, int SCAN_ITEMS, class V>
__device__ __forceinline__ void scan_block(const IO& io, int64_t n_cap, const int64_t* n_dev, int64_t* total,
                                           uint64_t* ctr, uint64_t* status2, int64_t S, int64_t bid,
                                           ScanLds<V, SCAN_ITEMS>& L) {
    constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
    const int64_t n = live_n(n_cap, n_dev);
    int64_t ntl = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (ntl < 1) ntl = 1;  // one block runs even when nothing is live: it clears and advances
    if (bid >= ntl) return;
    if (threadIdx.x == 0) {
        const uint64_t c = atomicAdd((unsigned long long*)ctr, 1ull);
        L.claim = c;
        if ((int64_t)(c & 0xffffffffull) == ntl - 1) status_store(ctr, ((c >> 32) + 1) << 32);  // last ticket
    }
    __syncthreads();
    const uint64_t claim = L.claim;
    const int64_t tile = (int64_t)(claim & 0xffffffffull);
    const int par = (int)((claim >> 32) & 1);
    uint64_t* status = status2 + par * S;
    uint64_t* other = status2 + (1 - par) * S;
    const int64_t base = tile * SCAN_TILE;
    if (base < n) {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {  // striped loads
            const int il = k * SCAN_THREADS + threadIdx.x;
            const int64_t i = base + il;
            L.v[scan_pad(il)] = (i < n) ? io.load(i) : 0;
        }
        __syncthreads();
        V v[SCAN_ITEMS];
        int64_t acc = 0;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {  // blocked values
            v[k] = L.v[scan_pad(threadIdx.x * SCAN_ITEMS + k)];
            acc += v[k];
        }
        int64_t agg;
        const int64_t incl = block_inclusive_scan(acc, L.lw, &agg);
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
                L.prefix = prefix;
                if (base + SCAN_TILE >= n) {  // the last live tile
                    if (total) *total = prefix + agg;
                    io.finish(prefix + agg);
                }
            }
        }
        __syncthreads();
        int64_t run = incl - acc + L.prefix;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {  // blocked offsets
            L.off[scan_pad(threadIdx.x * SCAN_ITEMS + k)] = run;
            run += v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {  // striped stores
            const int il = k * SCAN_THREADS + threadIdx.x;
            const int64_t i = base + il;
            if (i < n) io.store(i, L.off[scan_pad(il)], L.v[scan_pad(il)]);
        }
    } else if (threadIdx.x == 0) {  // nothing live (only tile 0 exists then)
        if (total) *total = 0;
        io.finish(0);
    }
    // clear the other buffer for the next call
    for (int64_t j = tile * SCAN_THREADS + threadIdx.x; j < S; j += ntl * SCAN_THREADS) other[j] = 0;
}

template <class IO>
using ScanV = std::decay_t<decltype(static_cast<const IO*>(nullptr)->load(int64_t(0)))>;

template <class IO, int SCAN_ITEMS>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_io(IO io, int64_t n_cap, const int64_t* n_dev, int64_t* total,
                                                          uint64_t* ctr, uint64_t* status2, int64_t S) {
    __shared__ ScanLds<ScanV<IO>, SCAN_ITEMS> L;
    scan_block<IO, SCAN_ITEMS>(io, n_cap, n_dev, total, ctr, status2, S, (int64_t)blockIdx.x, L);
}

// Two independent scans in one launch: the first nbA blocks run scan A, the
// others scan B, each with its own scratch. They share the launch (and the
// LDS, as a union), so two short scans cost about one.
template <class IOA, int IA, class IOB, int IB>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_io2(IOA ioa, int64_t na_cap, const int64_t* na_dev,
                                                           uint64_t* ctra, uint64_t* sta, int64_t Sa, int64_t nba,
                                                           IOB iob, int64_t nb_cap, const int64_t* nb_dev,
                                                           uint64_t* ctrb, uint64_t* stb, int64_t Sb) {
    __shared__ union U {
        ScanLds<ScanV<IOA>, IA> a;
        ScanLds<ScanV<IOB>, IB> b;
    } L;
    if ((int64_t)blockIdx.x < nba)
        scan_block<IOA, IA>(ioa, na_cap, na_dev, nullptr, ctra, sta, Sa, (int64_t)blockIdx.x, L.a);
    else
        scan_block<IOB, IB>(iob, nb_cap, nb_dev, nullptr, ctrb, stb, Sb, (int64_t)blockIdx.x - nba, L.b);
}

// out[i] = exclusive prefix of in[i]
struct ScanIO {
    const int32_t* in;
    int64_t* out;
    __device__ __forceinline__ int32_t load(int64_t i) const { return in[i]; }
    __device__ __forceinline__ void store(int64_t i, int64_t off, int32_t) const { out[i] = off; }
    __device__ __forceinline__ void finish(int64_t) const {}
};

// ITEMS: items per thread; 0 picks by length (16 for long scans, 4 below a
// million items). An IO whose load does much latency-bound work per item
// asks for 1, so that each thread has one item and more blocks run.
template <class IO, int ITEMS = 0>
int mwd_scan_io(const IO& io, int64_t n_cap, const int64_t* n_dev, int64_t* total, int64_t* tmp, int64_t tmp_len,
                hipStream_t s) {
    const int items = ITEMS > 0 ? ITEMS : (n_cap <= SCAN_SHORT_MAX ? SCAN_ITEMS_SHORT : SCAN_ITEMS_LONG);
    const int64_t tile = (int64_t)SCAN_THREADS * items;
    const int64_t nb = n_cap > 0 ? (n_cap + tile - 1) / tile : 1;
    // scratch layout: [0] call/ticket word, then two status buffers of S words
    const int64_t S = (tmp_len - 1) / 2;
    if (nb > S) {
        mwd_set_error("scan scratch too small");
        return MWD_E_CAPACITY;
    }
    uint64_t* ctr = reinterpret_cast<uint64_t*>(tmp);
    uint64_t* st2 = reinterpret_cast<uint64_t*>(tmp + 1);
    const int64_t n = n_cap > 0 ? n_cap : 0;
    if constexpr (ITEMS > 0) {
        hipLaunchKernelGGL((k_scan_io<IO, ITEMS>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, io, n, n_dev, total,
                           ctr, st2, S);
    } else {
        if (items == SCAN_ITEMS_SHORT)
            hipLaunchKernelGGL((k_scan_io<IO, SCAN_ITEMS_SHORT>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, io, n,
                               n_dev, total, ctr, st2, S);
        else
            hipLaunchKernelGGL((k_scan_io<IO, SCAN_ITEMS_LONG>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, io, n,
                               n_dev, total, ctr, st2, S);
    }
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

// Scans A and B in one launch (k_scan_io2), A with IA items per thread and B
// with IB; each takes its scratch as mwd_scan_io does.
template <class IOA, int IA, class IOB, int IB>
int mwd_scan_io2(const IOA& ioa, int64_t na_cap, const int64_t* na_dev, int64_t* tmpa, int64_t tmpa_len,
                 const IOB& iob, int64_t nb_cap, const int64_t* nb_dev, int64_t* tmpb, int64_t tmpb_len,
                 hipStream_t s) {
    const int64_t ta = (int64_t)SCAN_THREADS * IA, tb = (int64_t)SCAN_THREADS * IB;
    const int64_t nba = na_cap > 0 ? (na_cap + ta - 1) / ta : 1;
    const int64_t nbb = nb_cap > 0 ? (nb_cap + tb - 1) / tb : 1;
    const int64_t Sa = (tmpa_len - 1) / 2, Sb = (tmpb_len - 1) / 2;
    if (nba > Sa || nbb > Sb) {
        mwd_set_error("scan scratch too small");
        return MWD_E_CAPACITY;
    }
    hipLaunchKernelGGL((k_scan_io2<IOA, IA, IOB, IB>), dim3((unsigned)(nba + nbb)), dim3(SCAN_THREADS), 0, s, ioa,
                       na_cap > 0 ? na_cap : 0, na_dev, reinterpret_cast<uint64_t*>(tmpa),
                       reinterpret_cast<uint64_t*>(tmpa + 1), Sa, nba, iob, nb_cap > 0 ? nb_cap : 0, nb_dev,
                       reinterpret_cast<uint64_t*>(tmpb), reinterpret_cast<uint64_t*>(tmpb + 1), Sb);
    MWD_LAUNCH_CHECK();
    return MWD_OK;
}

}  // namespace
