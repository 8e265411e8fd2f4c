// Single-pass count -> scan -> write over rows.
//
// Every CSR-producing kernel in this library (set algebra, merges, clipping,
// threshold tagging) has the same shape: each output row produces a number of
// intervals that is only known after walking the row.  One launch does all
// three steps: a block takes the next tile of rows (dynamic ticket), counts
// them with one G-lane group per row, scans the tile, obtains the tile's
// global offset by decoupled look-back over the tiles before it, writes the
// row offsets and finally walks its rows again to write the intervals (the
// second walk reads what the first one pulled into L1/L2).
#pragma once

#include "stx_common.h"
#include "stx_host.h"

namespace stx {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / STX_WAVE;
constexpr int kRowsPerGroupIter = 4;  // rows each group handles per tile

template <int G>
struct TileShape {
    static constexpr int groups_per_wave = STX_WAVE / G;
    static constexpr int rows = kWaves * groups_per_wave * kRowsPerGroupIter;
};

// RowOp must provide
//   template <int G, bool WRITE>
//   __device__ int32_t run(int64_t row, int32_t* ob, int32_t* oe, int64_t base) const;
// executed by all G lanes of a group, returning the row's interval count.
template <int G, class RowOp>
__global__ __launch_bounds__(kBlock) void fused_rows_kernel(RowOp op, const int32_t* __restrict__ d_nrows,
                                                            int64_t nrows_bound, uint32_t ntiles,
                                                            int32_t* __restrict__ out_off, int32_t* __restrict__ ob,
                                                            int32_t* __restrict__ oe, int64_t capacity,
                                                            int32_t* __restrict__ d_total, uint32_t* status,
                                                            void* scratch, uint32_t epoch, int64_t spin_limit) {
    constexpr int RPT = TileShape<G>::rows;
    __shared__ int32_t s_cnt[RPT];
    __shared__ int32_t s_pref[RPT];
    __shared__ int32_t s_scan[17];
    __shared__ uint32_t s_excl;

    LookbackScratch lb = lookback_scratch(scratch, ntiles);
    const uint32_t tile = blockIdx.x;
    const int64_t U = d_nrows ? (int64_t)*d_nrows : nrows_bound;
    const int64_t row0 = (int64_t)tile * RPT;
    const int wave = wave_id();
    const int grp = lane_id() / G;
    const int gl = lane_id() & (G - 1);

    for (int k = 0; k < kRowsPerGroupIter; ++k) {
        const int local = (wave * kRowsPerGroupIter + k) * TileShape<G>::groups_per_wave + grp;
        const int64_t u = row0 + local;
        int32_t cnt = 0;
        if (u < U) cnt = op.template run<G, false>(u, nullptr, nullptr, 0);
        if (gl == 0) s_cnt[local] = cnt;
    }
    __syncthreads();
    int32_t tile_total = 0;
    int32_t v = (int)threadIdx.x < RPT ? s_cnt[threadIdx.x] : 0;
    int32_t ex = block_exclusive_scan<int32_t>(v, s_scan, &tile_total);
    if ((int)threadIdx.x < RPT) s_pref[threadIdx.x] = ex;
    if (wave == 0) {
        uint32_t e = lookback_exclusive(lb, tile, (uint32_t)tile_total, epoch, status, spin_limit);
        if (lane_id() == 0) s_excl = e;
    }
    __syncthreads();
    const int64_t excl = s_excl;
    if ((int)threadIdx.x < RPT) {
        const int64_t u = row0 + threadIdx.x;
        if (u <= U) out_off[u] = (int32_t)(excl + s_pref[threadIdx.x]);
        if (u == U && d_total) *d_total = (int32_t)(excl + s_pref[threadIdx.x]);
    }
    if (threadIdx.x == 0 && row0 + RPT == U) {
        out_off[U] = (int32_t)(excl + tile_total);
        if (d_total) *d_total = (int32_t)(excl + tile_total);
    }
    if (excl + tile_total > capacity) {
        if (threadIdx.x == 0 && tile_total > 0) atomicOr(status, STX_FLAG_CAPACITY);
    } else {
        for (int k = 0; k < kRowsPerGroupIter; ++k) {
            const int local = (wave * kRowsPerGroupIter + k) * TileShape<G>::groups_per_wave + grp;
            const int64_t u = row0 + local;
            if (u < U && s_cnt[local] > 0) op.template run<G, true>(u, ob, oe, excl + s_pref[local]);
        }
    }
}

struct FusedLaunch {
    const int32_t* d_nrows;
    int64_t nrows_bound;
    int32_t* out_off;
    int32_t* ob;
    int32_t* oe;
    int64_t capacity;
    int32_t* d_total;
    uint32_t* status;
    void* scratch;
    int64_t scratch_bytes;
    uint32_t epoch;
    hipStream_t stream;
};

template <int G>
inline uint32_t fused_rows_tiles(int64_t nrows_bound) {
    constexpr int RPT = TileShape<G>::rows;
    return (uint32_t)((nrows_bound + RPT) / RPT);  // one spare slot for out_off[U]
}

inline int64_t fused_rows_scratch_bytes(int64_t nrows_bound, int32_t group) {
    uint32_t t = group >= 64 ? fused_rows_tiles<64>(nrows_bound)
                             : group >= 16 ? fused_rows_tiles<16>(nrows_bound) : fused_rows_tiles<4>(nrows_bound);
    return lookback_bytes(t);
}

template <int G, class RowOp>
inline int32_t launch_fused_rows(const RowOp& op, const FusedLaunch& L, const char* what) {
    if (L.nrows_bound < 0) return stx_fail(STX_EINVAL, "fused rows: negative row count");
    const uint32_t ntiles = fused_rows_tiles<G>(L.nrows_bound);
    if (lookback_bytes(ntiles) > L.scratch_bytes) return stx_fail(STX_EINVAL, "fused rows: scratch too small");
    hipLaunchKernelGGL((fused_rows_kernel<G, RowOp>), dim3(ntiles), dim3(kBlock), 0, L.stream, op, L.d_nrows,
                       L.nrows_bound, ntiles, L.out_off, L.ob, L.oe, L.capacity, L.d_total, L.status, L.scratch,
                       L.epoch, stx_spin_limit());
    return stx_launch_status(what);
}

template <class RowOp>
inline int32_t launch_fused_rows_g(int32_t group, const RowOp& op, const FusedLaunch& L, const char* what) {
    if (group >= 64) return launch_fused_rows<64>(op, L, what);
    if (group >= 16) return launch_fused_rows<16>(op, L, what);
    return launch_fused_rows<4>(op, L, what);
}

}  // namespace stx
