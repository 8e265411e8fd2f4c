// Shared device helpers for the gfx950 interval-set kernels.
//
// Wavefront size is 64 on CDNA4; every helper here is written for 64-lane
// waves and for power-of-two "lane groups" of G <= 64 lanes inside a wave
// (several short rows per wave).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define STX_WAVE 64

namespace stx {

// ---- status codes (mirrored in include/subsetix_hip.h) --------------------
enum : int32_t {
    STX_OK = 0,
    STX_EINVAL = -1,      // bad argument
    STX_ECAPACITY = -2,   // output capacity exceeded (device flag)
    STX_EALIGN = -3,      // multilevel alignment violated
    STX_EHIP = -4,        // HIP runtime error
    STX_ESPIN = -5,       // look-back wait exceeded its bound
};

// Device-side error word bits (written with atomicOr into a status word).
enum : uint32_t {
    STX_FLAG_CAPACITY = 1u,
    STX_FLAG_SPIN = 2u,
    STX_FLAG_ALIGN = 4u,
    STX_FLAG_ROWS = 8u,
    STX_FLAG_DOMAIN = 16u,  // cells outside the [0, w) domain a segment table was built for
};

// Byte offset of element `idx` for a raw buffer load, computed in uint32 so
// that fields of up to 4 GiB address every element without signed overflow.
// A negative index (-1 for a missing neighbour, -vector width for a missing
// row) becomes an offset within 64 bytes of 2^32, above the record size of
// every buffer resource built here (< 0xffffff00 bytes), so the load returns
// 0: the zero exterior of the stencils and the absent children of transfers.
__device__ __forceinline__ int32_t buffer_byte_offset(int32_t idx, uint32_t elem_bytes) {
    return (int32_t)((uint32_t)idx * elem_bytes);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (STX_WAVE - 1)); }
__device__ __forceinline__ int wave_id() { return (int)(threadIdx.x / STX_WAVE); }

// Ballot restricted to the G-lane group of the calling lane, shifted down so
// that bit k is group lane k.
template <int G>
__device__ __forceinline__ uint64_t group_ballot(bool pred) {
    uint64_t m = __ballot(pred);
    if (G == 64) return m;
    const int grp = lane_id() / G;
    return (m >> (grp * G)) & ((1ull << G) - 1ull);
}

// Number of set bits of `mask` strictly below group lane `gl`.
__device__ __forceinline__ int rank_below(uint64_t mask, int gl) {
    return __popcll(gl ? (mask & ((~0ull) >> (64 - gl))) : 0ull);
}

// Inclusive max-scan over a G-lane group.
template <int G>
__device__ __forceinline__ int group_scan_max(int v) {
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        int o = __shfl_up(v, d, G);
        if ((lane_id() & (G - 1)) >= d) v = v > o ? v : o;
    }
    return v;
}

// Inclusive sum-scan over a G-lane group.
template <int G, typename T>
__device__ __forceinline__ T group_scan_sum(T v) {
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        T o = __shfl_up(v, d, G);
        if ((lane_id() & (G - 1)) >= d) v += o;
    }
    return v;
}

template <int G>
__device__ __forceinline__ int group_bcast(int v, int src_gl) {
    return __shfl(v, src_gl, G);
}

__device__ __forceinline__ int32_t floordiv(int32_t a, int32_t b) {
    int32_t q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

// ---- element transforms ----------------------------------------------------
// Monotone maps applied to an operand's intervals on load; they keep each
// row's intervals sorted by begin.  A transformed interval with b >= e is
// "dead" (only the clip transform produces those).
enum : int32_t { XF_NONE = 0, XF_WIDEN = 1, XF_COARSEN = 2, XF_CLIP = 3, XF_SCALE = 4 };

struct XForm {
    int32_t mode;
    int32_t p;
};

__device__ __forceinline__ void xform_apply(const XForm& xf, int32_t& b, int32_t& e) {
    if (xf.mode == XF_WIDEN) {
        b -= xf.p;
        e += xf.p;
    } else if (xf.mode == XF_COARSEN) {
        b = floordiv(b, xf.p);
        e = floordiv(e + xf.p - 1, xf.p);
    } else if (xf.mode == XF_CLIP) {
        b = b < 0 ? 0 : b;
        e = e > xf.p ? xf.p : e;
    } else if (xf.mode == XF_SCALE) {
        b *= xf.p;
        e *= xf.p;
    }
}

// ---- decoupled look-back ---------------------------------------------------
// Scratch layout shared by every single-pass kernel (caller-provided, zeroed
// once at allocation): a 64-byte header (unused) then one 64-bit state word
// per tile.  A word packs (epoch << 34) | (flag << 32) | value, flag 1 =
// aggregate, 2 = inclusive prefix.  Every launch passes a fresh epoch (the
// host counts launches per scratch buffer), so words left by earlier launches
// read as "not published" and nothing has to be cleared between launches.
// Tiles are numbered by blockIdx: workgroups are dispatched in increasing id
// order, so every tile a block waits on belongs to a block that has already
// been dispatched; the wait is also bounded (spin limit), so a block always
// exits.
struct LookbackScratch {
    unsigned long long* state;
    int64_t max_tiles;
};

__device__ __forceinline__ LookbackScratch lookback_scratch(void* base, int64_t max_tiles) {
    LookbackScratch s;
    s.state = (unsigned long long*)((char*)base + 64);
    s.max_tiles = max_tiles;
    return s;
}

// A state word carries its own payload, so no other memory has to be ordered
// with it: relaxed device-scope atomics are enough.  (Acquire/release at
// device scope would write back / invalidate the whole L2 of the XCD on every
// poll, because the L2s of the XCDs are not coherent with each other.)
__device__ __forceinline__ void state_store(unsigned long long* p, uint32_t epoch, uint32_t flag, uint32_t value) {
    unsigned long long w = ((unsigned long long)((epoch << 2) | flag) << 32) | (unsigned long long)value;
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long state_load(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publishes tile `tile`'s aggregate ahead of its look-back (one thread):
// tile 0's aggregate is already its inclusive prefix.
__device__ __forceinline__ void lookback_publish(LookbackScratch s, uint32_t tile, uint32_t agg, uint32_t epoch) {
    state_store(&s.state[tile], epoch, tile == 0 ? 2u : 1u, agg);
}

// Called by all 64 lanes of ONE wave of the block.  Publishes `agg` for tile
// `tile` (unless PUBLISHED: lookback_publish stored it already) and returns
// the exclusive prefix of all earlier tiles.  One poll reads K blocks of 64
// consecutive predecessors (block k, lane l: tile pos - 64 k - l; each block
// one coalesced load), all issued before any is used, and folds the blocks
// nearest first: aggregates are summed until a block holds an inclusive
// prefix.  A block with a predecessor not yet published (before the first
// prefix) stops the fold; the blocks folded so far stay folded and the next
// poll starts behind them.  After `spin_limit` polls that found no progress
// the block records STX_FLAG_SPIN in `status` and proceeds (its result is
// then invalid, but every wave still exits).  A poll that finds nothing
// ready sleeps SLEEP x 64 cycles first, and with PEEK a wave first reads only
// the nearest unfolded predecessor's word (one 8-byte load) until that one
// has published: long-lived tiles then poll less and leave the memory system
// to the tiles still loading (the 1024-thread set tiles on 128 stacked
// pairs: union 159 -> 153 us with a 64 x 64-cycle sleep, 150 us with the
// peek and a 16 x 64-cycle sleep; sleeps of 96 or 127 were slower, as was a
// doubling backoff).
template <int K = 1, bool PUBLISHED = false, int SLEEP = 1, bool PEEK = false>
__device__ __forceinline__ uint32_t lookback_exclusive(LookbackScratch s, uint32_t tile, uint32_t agg, uint32_t epoch,
                                                       uint32_t* status, int64_t spin_limit) {
    const int lane = lane_id();
    if (tile == 0) {
        if (!PUBLISHED && lane == 0) state_store(&s.state[0], epoch, 2u, agg);
        return 0u;
    }
    if (!PUBLISHED && lane == 0) state_store(&s.state[tile], epoch, 1u, agg);
    uint32_t excl = 0;
    int64_t pos = (int64_t)tile - 1;  // nearest predecessor not yet folded in
    int64_t spins = 0;
    while (pos >= 0) {
        if (PEEK) {
            const unsigned long long w0 = state_load(&s.state[pos]);
            const uint32_t hi0 = (uint32_t)(w0 >> 32);
            if ((hi0 >> 2) != epoch || (hi0 & 3u) == 0u) {  // not published yet
                if (++spins > spin_limit) {
                    if (lane == 0) atomicOr(status, STX_FLAG_SPIN);
                    break;
                }
                __builtin_amdgcn_s_sleep(SLEEP);
                continue;
            }
        }
        unsigned long long w[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int64_t p = pos - 64 * k - lane;
            // before tile 0: an empty inclusive prefix
            w[k] = p >= 0 ? state_load(&s.state[p]) : ((unsigned long long)((epoch << 2) | 2u) << 32);
        }
        bool done = false;
        int folded = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t hi = (uint32_t)(w[k] >> 32);
            const uint32_t flag = (hi >> 2) == epoch ? (hi & 3u) : 0u;
            const uint64_t has_prefix = __ballot(flag == 2u);
            const uint64_t not_ready = __ballot(flag == 0u);
            const int first_prefix = has_prefix ? __ffsll((long long)has_prefix) - 1 : 64;
            const uint64_t upto = first_prefix >= 63 ? ~0ull : ((1ull << (first_prefix + 1)) - 1ull);
            if (not_ready & upto) break;
            uint32_t contrib = lane <= first_prefix ? (uint32_t)w[k] : 0u;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) contrib += __shfl_xor(contrib, d, 64);
            excl += contrib;
            folded = k + 1;
            if (first_prefix < 64) {
                done = true;
                break;
            }
        }
        if (done) break;
        pos -= 64 * folded;
        if (folded == 0) {
            if (++spins > spin_limit) {
                if (lane == 0) atomicOr(status, STX_FLAG_SPIN);
                break;
            }
            __builtin_amdgcn_s_sleep(SLEEP);
        }
    }
    if (lane == 0) state_store(&s.state[tile], epoch, 2u, excl + agg);
    return excl;
}

// Bytes of look-back scratch needed for `ntiles` tiles.
__host__ __device__ inline int64_t lookback_bytes(int64_t ntiles) { return 64 + 8 * ntiles; }

// Count-balanced tiles over U rows.  h(r) = floor(c(r) / T) + floor(r / M)
// with c non-decreasing, c(0) = 0, c(r) the weight of the rows below r: tile
// k starts at the first row r with h(r) >= k.  Two starts in a row differ in
// h by one, so a tile's rows weigh less than T plus its last row, and there
// are at most M of them; a row weighing more than T gives empty tiles
// before it.  Entry k of the table (k = 0 .. ntiles) is the 8-byte word
// {start row, 0}, starts past h(U) being U: with upper half 0 a table left
// in the look-back scratch never reads as a published state (epochs start
// at 1), whatever later launch uses those words.  Thread i writes the starts
// that fall on row i (i <= U) and entry i past h(U); *status gets
// STX_FLAG_CAPACITY when h(U) >= ntiles (the table was sized for less).
template <class HFn>
__device__ __forceinline__ void balanced_table_entry(int64_t i, int64_t U, int64_t ntiles, HFn h, int2* tab,
                                                     uint32_t* status) {
    const int64_t hU = h(U);
    if (i <= U) {
        const int64_t hi = h(i), lo = i > 0 ? h(i - 1) : -1;
        for (int64_t k = lo + 1; k <= hi && k <= ntiles; ++k) tab[k] = make_int2((int32_t)i, 0);
    }
    if (i > hU && i <= ntiles) tab[i] = make_int2((int32_t)U, 0);
    if (i == 0 && hU >= ntiles) atomicOr(status, STX_FLAG_CAPACITY);
}

// Byte offset of a tile table placed after the look-back states of ntiles
// tiles, and the scratch bytes of both.
__host__ __device__ inline int64_t balanced_table_offset(int64_t ntiles) { return (lookback_bytes(ntiles) + 15) / 16 * 16; }
__host__ __device__ inline int64_t balanced_table_scratch_bytes(int64_t ntiles) {
    return balanced_table_offset(ntiles) + 8 * (ntiles + 1);
}

// ---- wave scans through DPP ------------------------------------------------
// Hillis-Steele inside each 16-lane row (row_shr 1, 2, 4, 8), then
// row_bcast:15 / row_bcast:31 carry the row totals across the wave (GFX9
// DPP forms, present on CDNA).  Lanes without a source take `id`, the
// identity of the operation.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int32_t dpp_mov(int32_t id, int32_t v) {
    return __builtin_amdgcn_update_dpp(id, v, CTRL, ROW_MASK, 0xf, false);
}

__device__ __forceinline__ int32_t wave_incl_sum(int32_t x) {
    x += dpp_mov<0x111>(0, x);
    x += dpp_mov<0x112>(0, x);
    x += dpp_mov<0x114>(0, x);
    x += dpp_mov<0x118>(0, x);
    x += dpp_mov<0x142, 0xa>(0, x);
    x += dpp_mov<0x143, 0xc>(0, x);
    return x;
}

__device__ __forceinline__ int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
__device__ __forceinline__ int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

__device__ __forceinline__ int32_t wave_incl_max(int32_t x) {
    x = imax(x, dpp_mov<0x111>(INT32_MIN, x));
    x = imax(x, dpp_mov<0x112>(INT32_MIN, x));
    x = imax(x, dpp_mov<0x114>(INT32_MIN, x));
    x = imax(x, dpp_mov<0x118>(INT32_MIN, x));
    x = imax(x, dpp_mov<0x142, 0xa>(INT32_MIN, x));
    x = imax(x, dpp_mov<0x143, 0xc>(INT32_MIN, x));
    return x;
}

__device__ __forceinline__ int32_t wave_incl_min(int32_t x) {
    x = imin(x, dpp_mov<0x111>(INT32_MAX, x));
    x = imin(x, dpp_mov<0x112>(INT32_MAX, x));
    x = imin(x, dpp_mov<0x114>(INT32_MAX, x));
    x = imin(x, dpp_mov<0x118>(INT32_MAX, x));
    x = imin(x, dpp_mov<0x142, 0xa>(INT32_MAX, x));
    x = imin(x, dpp_mov<0x143, 0xc>(INT32_MAX, x));
    return x;
}

// Value of lane 63 (the wave total after an inclusive scan).
__device__ __forceinline__ int32_t wave_last(int32_t x) { return __builtin_amdgcn_readlane(x, 63); }

// Inclusive scans inside each 16-lane row (row_shr 1, 2, 4, 8).
__device__ __forceinline__ int32_t row16_incl_sum(int32_t x) {
    x += dpp_mov<0x111>(0, x);
    x += dpp_mov<0x112>(0, x);
    x += dpp_mov<0x114>(0, x);
    x += dpp_mov<0x118>(0, x);
    return x;
}

__device__ __forceinline__ int32_t row16_incl_max(int32_t x) {
    x = imax(x, dpp_mov<0x111>(INT32_MIN, x));
    x = imax(x, dpp_mov<0x112>(INT32_MIN, x));
    x = imax(x, dpp_mov<0x114>(INT32_MIN, x));
    x = imax(x, dpp_mov<0x118>(INT32_MIN, x));
    return x;
}

__device__ __forceinline__ int32_t row16_incl_min(int32_t x) {
    x = imin(x, dpp_mov<0x111>(INT32_MAX, x));
    x = imin(x, dpp_mov<0x112>(INT32_MAX, x));
    x = imin(x, dpp_mov<0x114>(INT32_MAX, x));
    x = imin(x, dpp_mov<0x118>(INT32_MAX, x));
    return x;
}

// Block-wide exclusive sum over one value per thread with one barrier:
// `lds` holds one slot per wave and is used once per kernel.  After the
// barrier lane w of every wave holds wave w's total, and one row or wave
// scan of those gives the prefix and the block total.
template <int NW>
__device__ __forceinline__ int32_t block_excl_sum(int32_t v, int32_t* lds, int32_t* total) {
    static_assert(NW >= 1 && NW <= 64, "at most 64 waves");
    const int wave = wave_id(), lane = lane_id();
    const int32_t incl = wave_incl_sum(v);
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    int32_t t = 0;
    if (lane < NW) t = lds[lane];
    t = NW <= 16 ? row16_incl_sum(t) : wave_incl_sum(t);
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int32_t before = wv > 0 ? __builtin_amdgcn_readlane(t, wv > 0 ? wv - 1 : 0) : 0;
    *total = __builtin_amdgcn_readlane(t, NW - 1);
    return before + incl - v;
}

// Block-wide exclusive max (INT32_MIN before thread 0), one barrier; `lds`
// holds one slot per wave and is used once per kernel.
template <int NW>
__device__ __forceinline__ int32_t block_excl_max(int32_t v, int32_t* lds) {
    static_assert(NW >= 1 && NW <= 64, "at most 64 waves");
    const int wave = wave_id(), lane = lane_id();
    const int32_t incl = wave_incl_max(v);
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    int32_t t = INT32_MIN;
    if (lane < wv) t = lds[lane];  // the waves before this one
    t = NW <= 16 ? row16_incl_max(t) : wave_incl_max(t);
    const int32_t before = __builtin_amdgcn_readlane(t, NW <= 16 ? 15 : 63);
    // wave_shr:1 - the inclusive value of the lane below (lane 0: identity)
    return imax(before, dpp_mov<0x138>(INT32_MIN, incl));
}

// Block-wide exclusive scan of one value per thread (blockDim.x <= 1024).
// Returns the exclusive prefix; *total receives the block sum.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds /* >= 17 */, T* total) {
    const int lane = lane_id(), wave = wave_id();
    const int nwaves = (int)(blockDim.x + STX_WAVE - 1) / STX_WAVE;
    T incl = group_scan_sum<64, T>(v);
    __syncthreads();
    if (lane == 63) lds[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = 0;
        for (int w = 0; w < nwaves; ++w) {
            T t = lds[w];
            lds[w] = acc;
            acc += t;
        }
        lds[16] = acc;
    }
    __syncthreads();
    T res = lds[wave] + incl - v;
    *total = lds[16];
    return res;
}

}  // namespace stx
