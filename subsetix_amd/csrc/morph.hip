// Box dilation of an interval set in a fixed number of launches
// (dilate_interval_set, morphology.py:211-265).
//
// The reference builds the dilation as the union of (2 hx + 1)(2 hy + 1)
// translated copies.  Here the output is formed directly, with every count
// kept on the device until the end:
//   1. rows that hold intervals are compacted (stx_compact_rows);
//   2. output rows: every compacted input row r covers [r - hy, r + hy]; row
//      k contributes the part of its window not covered by row k - 1
//      (count, exclusive scan, write), and each output row records the
//      range [klo, khi] of input rows whose window holds it;
//   3. per output row, one thread merges its (at most 2 hy + 1) input rows,
//      each interval widened by hx, with a k-way union (begin order, running
//      largest end; touching runs join) - a count pass, an exclusive scan of
//      the counts into the output offsets, and a write pass.
// Three scans use the look-back scratch with epochs epoch, epoch + 1 and
// epoch + 2.
#include <limits.h>

#include "stx_common.h"
#include "stx_host.h"

using namespace stx;

extern "C" int32_t stx_scan_i32(const int32_t* in, const int32_t* d_n, int64_t n_bound, int32_t* out, int32_t* d_total,
                                uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch,
                                hipStream_t stream);
extern "C" int32_t stx_compact_rows(const int32_t* rows, const int32_t* off, int64_t R, int32_t shift,
                                    int32_t* rows_out, int32_t* off_out, int32_t* d_count, int32_t* tmp,
                                    uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch,
                                    hipStream_t stream);

namespace {

constexpr int kT = 256;
constexpr int kMaxHalo = 4;  // window of at most 9 rows per output row

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

// Output rows contributed by compacted row k: [lo, r_k + hy] with lo past
// the window of row k - 1.
// (64-bit: row ids near the int32 limits cannot wrap the loops below)
__device__ __forceinline__ int64_t window_lo(const int32_t* rows_c, int64_t k, int32_t hy) {
    int64_t lo = (int64_t)rows_c[k] - hy;
    if (k > 0) {
        const int64_t prev = (int64_t)rows_c[k - 1] + hy + 1;
        lo = prev > lo ? prev : lo;
    }
    return lo;
}

__global__ void win_rows_count_kernel(const int32_t* __restrict__ rows_c, const int32_t* __restrict__ d_rc,
                                      int64_t Rb, int32_t hy, int32_t* __restrict__ cnt) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= Rb || k >= *d_rc) return;
    cnt[k] = (int32_t)((int64_t)rows_c[k] + hy - window_lo(rows_c, k, hy) + 1);
}

__global__ void win_rows_write_kernel(const int32_t* __restrict__ rows_c, const int32_t* __restrict__ d_rc,
                                      int64_t Rb, int32_t hy, const int32_t* __restrict__ pos,
                                      int32_t* __restrict__ rows_out, int32_t* __restrict__ klo,
                                      int32_t* __restrict__ khi) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t rc = *d_rc;
    if (k >= Rb || k >= rc) return;
    const int64_t lo = window_lo(rows_c, k, hy), hi = (int64_t)rows_c[k] + hy;
    int64_t last = k;  // last input row with r <= y + hy (rows strictly increase)
    for (int64_t y = lo; y <= hi; ++y) {
        while (last + 1 < rc && rows_c[last + 1] <= y + hy) ++last;
        const int64_t u = (int64_t)pos[k] + (y - lo);
        rows_out[u] = (int32_t)y;
        klo[u] = (int32_t)k;  // row k - 1's window ends below lo
        khi[u] = (int32_t)last;
    }
}

// k-way union of the input rows [klo, khi] (each interval widened by hx) of
// output row u: count (WRITE = false) or write at out_off[u].
template <int KMAX, bool WRITE>
__global__ void win_union_kernel(const int32_t* __restrict__ off_c, const int32_t* __restrict__ beg,
                                 const int32_t* __restrict__ end, const int32_t* __restrict__ d_ny, int64_t NYb,
                                 const int32_t* __restrict__ klo, const int32_t* __restrict__ khi, int32_t hx,
                                 int32_t* __restrict__ cnt, const int32_t* __restrict__ out_off,
                                 int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t capacity,
                                 uint32_t* __restrict__ status) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= NYb || u >= *d_ny) return;
    const int32_t k0 = klo[u], K = khi[u] - k0 + 1;
    int32_t j[KMAX], je[KMAX], hb[KMAX];  // cursor, end, head begin (INT_MAX when done)
#pragma unroll
    for (int t = 0; t < KMAX; ++t) {
        j[t] = t < K ? off_c[k0 + t] : 0;
        je[t] = t < K ? off_c[k0 + t + 1] : 0;
        hb[t] = j[t] < je[t] ? beg[j[t]] : INT_MAX;
    }
    int64_t w = WRITE ? (int64_t)out_off[u] : 0;
    int32_t n = 0, s = 0, m = INT_MIN;
    bool open = false;
    while (true) {
        int32_t best = INT_MAX, tb = -1;
#pragma unroll
        for (int t = 0; t < KMAX; ++t)
            if (hb[t] < best) { best = hb[t]; tb = t; }
        if (tb < 0) break;
        int32_t e = 0;
#pragma unroll
        for (int t = 0; t < KMAX; ++t) {
            if (t == tb) {
                e = end[j[t]];
                ++j[t];
                hb[t] = j[t] < je[t] ? beg[j[t]] : INT_MAX;
            }
        }
        const int32_t b = best - hx;
        e += hx;
        if (!open || b > m) {
            if (open) {
                if (WRITE) {
                    if (w < capacity) { ob[w] = s; oe[w] = m; } else atomicOr(status, STX_FLAG_CAPACITY);
                    ++w;
                }
                ++n;
            }
            s = b;
            m = e;
            open = true;
        } else if (e > m) {
            m = e;
        }
    }
    if (open) {
        if (WRITE) {
            if (w < capacity) { ob[w] = s; oe[w] = m; } else atomicOr(status, STX_FLAG_CAPACITY);
        }
        ++n;
    }
    if (!WRITE) cnt[u] = n;
}

template <int KMAX>
void launch_union(bool write, const int32_t* off_c, const int32_t* beg, const int32_t* end, const int32_t* d_ny,
                  int64_t NYb, const int32_t* klo, const int32_t* khi, int32_t hx, int32_t* cnt, const int32_t* out_off,
                  int32_t* ob, int32_t* oe, int64_t capacity, uint32_t* status, hipStream_t stream) {
    if (write)
        hipLaunchKernelGGL((win_union_kernel<KMAX, true>), dim3(nblk(NYb)), dim3(kT), 0, stream, off_c, beg, end, d_ny,
                           NYb, klo, khi, hx, cnt, out_off, ob, oe, capacity, status);
    else
        hipLaunchKernelGGL((win_union_kernel<KMAX, false>), dim3(nblk(NYb)), dim3(kT), 0, stream, off_c, beg, end,
                           d_ny, NYb, klo, khi, hx, cnt, out_off, ob, oe, capacity, status);
}

// ---------------------------------------------------------------------------
// Tiled dilation over dense rows.
//
// The rows lo - hy .. hi + hy of the output are numbered i = 0 .. nt - 1
// (row lo - hy + i) and the input is read through dense offsets D over the
// rows lo - 2 hy + j (j = 0 .. nt + 2 hy), so output row i is the union of
// the input rows j = i .. i + 2 hy, each interval widened by hx.  A tile
// takes nr consecutive output rows [i0, i0 + nr): it stages the input rows
// [i0, i0 + nr + 2 hy) in LDS as one keyed sequence (key = (local row <<
// shift) + (x - xmin), begins sorted since rows are), and forms the window
// union by doubling on that sequence: W1 = the staged rows, W2q(l) = Wq(l) U
// Wq(l + q) for q = 1, 2, .. < p (p the largest power of two <= 2 hy + 1),
// then out(l) = Wp(l) U Wp(l + 2 hy + 1 - p) for l < nr.  Every step is one
// merge-path union of the sequence with its own suffix from row q, the
// suffix's keys lowered by q rows (begins stay sorted; a run never crosses a
// row since 2^shift exceeds every key offset), compacted into the other of
// two LDS buffers.  The tile's count then goes through the look-back, and
// one coalesced copy decodes the keys into the output.  A tile whose input
// or one of whose steps exceeds an LDS buffer, or whose keys do not fit 30
// bits, merges each of its output rows in one thread instead (the 2 hy + 1
// input rows as cursors, a count pass and a write pass).
constexpr int kDTB = 1024;    // threads per tile
constexpr int kDVT = 8;       // merge positions per thread and step
constexpr int kDCap = 4096;   // intervals of one LDS buffer (one step merges <= 2 kDCap = kDTB * kDVT)
constexpr int kDRows = 1040;  // staged row starts (nr + 2 hy + 1 <= 1024 + 8 + 1)
static_assert(2 * kDCap == kDTB * kDVT, "a step merges a buffer with its suffix");

__device__ __forceinline__ int32_t dsk(int32_t i) { return i + (i >> 4); }  // one spare slot per 16: fewer bank conflicts
constexpr int dsk_size(int n) { return n + n / 16 + 1; }

// Union of the dense input rows j = i .. i + K - 1, widened by hx: the
// interval count (WRITE = false) or the intervals written from ob/oe + w.
template <int KMAX, bool WRITE>
__device__ int32_t dense_window_union(const int32_t* __restrict__ D, const int32_t* __restrict__ beg,
                                      const int32_t* __restrict__ end, int64_t i, int K, int32_t hx,
                                      int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t w, int64_t capacity,
                                      uint32_t* __restrict__ status) {
    int32_t j[KMAX], je[KMAX], hb[KMAX];  // cursor, end, head begin (INT_MAX when done)
#pragma unroll
    for (int t = 0; t < KMAX; ++t) {
        j[t] = t < K ? D[i + t] : 0;
        je[t] = t < K ? D[i + t + 1] : 0;
        hb[t] = j[t] < je[t] ? beg[j[t]] : INT_MAX;
    }
    int32_t n = 0, s = 0, m = INT_MIN;
    bool open = false;
    while (true) {
        int32_t best = INT_MAX, tb = -1;
#pragma unroll
        for (int t = 0; t < KMAX; ++t)
            if (hb[t] < best) { best = hb[t]; tb = t; }
        if (tb < 0) break;
        int32_t e = 0;
#pragma unroll
        for (int t = 0; t < KMAX; ++t) {
            if (t == tb) {
                e = end[j[t]];
                ++j[t];
                hb[t] = j[t] < je[t] ? beg[j[t]] : INT_MAX;
            }
        }
        const int32_t b = best - hx;
        e += hx;
        if (!open || b > m) {
            if (open) {
                if (WRITE) {
                    if (w < capacity) { ob[w] = s; oe[w] = m; } else atomicOr(status, STX_FLAG_CAPACITY);
                    ++w;
                }
                ++n;
            }
            s = b;
            m = e;
            open = true;
        } else if (e > m) {
            m = e;
        }
    }
    if (open) {
        if (WRITE) {
            if (w < capacity) { ob[w] = s; oe[w] = m; } else atomicOr(status, STX_FLAG_CAPACITY);
        }
        ++n;
    }
    return n;
}

// One doubling step, called by every thread of the block: Y = X[0, nA) U
// (X[b0, b1) lowered by dq rows), compacted, with the row starts in rs
// (those of X on entry, read by the caller before the call) rewritten for Y
// over the local rows 0 .. Rt.  Returns Y's count, or -1 when it would not
// fit a buffer (uniform; Y and rs are then left as they are).
__device__ __forceinline__ int32_t union_step(const int2* __restrict__ X, int2* __restrict__ Y, int32_t* rs,
                                              int32_t nA, int32_t b0, int32_t b1, int dq, uint32_t shift, int Rt,
                                              int32_t* wmax, int32_t* wsum) {
    constexpr int kTW = kDTB / STX_WAVE;
    const int tid = threadIdx.x;
    const int32_t delta = (int32_t)((uint32_t)dq << shift);
    const int32_t nB = b1 - b0;
    const int32_t P = nA + nB;
    const int32_t p0 = tid * kDVT;
    const uint32_t valid = p0 >= P ? 0u : (P - p0 >= kDVT ? (1u << kDVT) - 1u : (1u << (P - p0)) - 1u);
    int32_t rb[kDVT], re[kDVT];
    uint32_t emit = 0;
    int32_t lm = INT32_MIN;  // largest end key before the current position (this thread's part)
    {
        int32_t ia = nA, jb = nB;  // positions past the end read sentinels only
        if (p0 < P) {
            int32_t l2 = p0 - nB > 0 ? p0 - nB : 0, h2 = p0 < nA ? p0 : nA;
            while (l2 < h2) {
                const int32_t mid = (l2 + h2) >> 1;
                if (X[dsk(mid)].x <= X[dsk(b0 + p0 - 1 - mid)].x - delta) l2 = mid + 1;
                else h2 = mid;
            }
            ia = l2;
            jb = p0 - l2;
        }
        int2 ca = ia < nA ? X[dsk(ia)] : make_int2(INT32_MAX, INT32_MIN);
        int2 cb = make_int2(INT32_MAX, INT32_MIN);
        if (jb < nB) {
            cb = X[dsk(b0 + jb)];
            cb.x -= delta;
            cb.y -= delta;
        }
#pragma unroll
        for (int v = 0; v < kDVT; ++v) {
            const bool isA = ca.x <= cb.x;
            const int32_t x = isA ? ca.x : cb.x;
            const int32_t y = isA ? ca.y : cb.y;
            rb[v] = x;
            re[v] = lm;
            emit |= (x > lm ? 1u : 0u) << v;
            lm = imax(lm, y);
            if (isA) {
                ++ia;
                ca = ia < nA ? X[dsk(ia)] : make_int2(INT32_MAX, INT32_MIN);
            } else {
                ++jb;
                if (jb < nB) {
                    cb = X[dsk(b0 + jb)];
                    cb.x -= delta;
                    cb.y -= delta;
                } else {
                    cb = make_int2(INT32_MAX, INT32_MIN);
                }
            }
        }
    }
    // a position starts a run when its begin is above every end before it;
    // begins grow along the walk, so those at or below the carry form a
    // prefix of the thread's positions
    const int32_t carry = block_excl_max<kTW>(lm, wmax);
    int32_t below = 0;
#pragma unroll
    for (int v = 0; v < kDVT; ++v) below += rb[v] <= carry ? 1 : 0;
    emit &= ~((1u << below) - 1u);
    emit &= valid;
    int32_t step_total = 0;
    const int32_t my_base = block_excl_sum<kTW>(__popc(emit), wsum, &step_total);
    if (step_total > kDCap) return -1;
    // outputs, and Y's row starts: a row's first position always starts a
    // run (rows never touch), so the row of the last output before this
    // thread is that of the largest end before it, the carry; rows
    // (previous output's row, r] start at the output of row r
    int32_t w = my_base;
    int32_t prow = carry >= 0 ? carry >> shift : -1;
#pragma unroll
    for (int v = 0; v < kDVT; ++v) {
        if (emit & (1u << v)) {
            Y[dsk(w)].x = rb[v];
            if (w > 0) Y[dsk(w - 1)].y = imax(carry, re[v]);
            const int32_t r = rb[v] >> shift;
            for (int32_t rr = prow + 1; rr <= r; ++rr) rs[rr] = w;
            prow = r;
            ++w;
        }
    }
    if (step_total == 0) {
        for (int r = tid; r <= Rt; r += kDTB) rs[r] = 0;
    } else if (p0 < P && P - p0 <= kDVT) {
        // the owner of the last position closes the last run; the rows after
        // the last output's start at the end (lm lies in the last row)
        Y[dsk(step_total - 1)].y = imax(carry, lm);
        for (int32_t rr = (lm >> shift) + 1; rr <= Rt; ++rr) rs[rr] = step_total;
    }
    __syncthreads();  // Y and its row starts complete; every read of X done
    return step_total;
}

// The window union of a staged tile: doubling steps W2q(l) = Wq(l) U
// Wq(l + q) while 2q <= K, then out(l) = Wp(l) U Wp(l + K - p), l < nr.
// Returns the count (the result in buf[*cur]) or -1 on overflow.
__device__ __forceinline__ int32_t window_union_steps(int2 (*buf)[dsk_size(kDCap)], int& cur, int32_t* rs, int32_t nX,
                                                      int K, int nr, int Rt, uint32_t shift, int32_t* wmax,
                                                      int32_t* wsum) {
    int p = 1;
    while (2 * p <= K) p *= 2;
    for (int q = 1;; q *= 2) {
        const bool last = q >= p;
        const int dq = last ? K - p : q;  // rows the suffix is lowered by
        const int32_t nA = last ? rs[nr] : nX;
        const int32_t b0 = rs[dq];
        const int32_t b1 = last ? rs[nr + dq] : nX;
        nX = union_step(buf[cur], buf[cur ^ 1], rs, nA, b0, b1, dq, shift, Rt, wmax, wsum);
        if (nX < 0) return -1;
        cur ^= 1;
        if (last) return nX;
    }
}

template <int KMAX>
__global__ __launch_bounds__(kDTB) __attribute__((amdgpu_waves_per_eu(8))) void dilate_tile_kernel(
    const int32_t* __restrict__ D, const int32_t* __restrict__ beg, const int32_t* __restrict__ end, int64_t nt,
    int32_t hx, int32_t hy, const int2* __restrict__ tab, uint32_t ntiles, int32_t* __restrict__ out_off,
    int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t capacity, int32_t* __restrict__ d_total,
    uint32_t* status, void* scratch, uint32_t epoch, int64_t spin_limit) {
    constexpr int kTW = kDTB / STX_WAVE;
    constexpr int kPerThread = kDCap / kDTB;
    __shared__ __align__(16) int2 s_buf[2][dsk_size(kDCap)];
    __shared__ int32_t s_rs[kDRows];
    __shared__ int32_t s_red[2 * kTW];
    __shared__ int32_t s_wmax[kTW];
    __shared__ int32_t s_wsum[kTW];
    __shared__ uint32_t s_excl;
    LookbackScratch lb = lookback_scratch(scratch, ntiles);
    const uint32_t tile = blockIdx.x;
    const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    const int K = 2 * hy + 1;
    // output rows [i0, i0 + nr) from the count-balanced table (tiles past
    // the last row start at nt, and no tile waits on them; an empty tile
    // before a row too large for one tile publishes 0)
    const int64_t i0 = tab[tile].x;
    if (i0 >= nt && tile > 0) return;
    const int64_t n_rows = (int64_t)tab[tile + 1].x - i0;
    if (n_rows > kDTB && tid == 0) atomicOr(status, STX_FLAG_CAPACITY);  // a table not made for these rows
    const int nr = (int)(n_rows < kDTB ? n_rows : kDTB);
    const int Rt = nr + 2 * hy;  // staged input rows [i0, i0 + Rt)
    const int32_t base = D[i0];
    const int32_t nIn = D[i0 + Rt] - base;
    bool lds_ok = nr > 0 && nIn <= kDCap;  // uniform
    for (int r = tid; r <= Rt; r += kDTB) s_rs[r] = D[i0 + r] - base;
    // this thread's intervals: kPerThread consecutive ones from q0
    const int32_t q0 = tid * kPerThread;
    const int32_t n_mine = lds_ok ? (nIn - q0 < 0 ? 0 : (nIn - q0 > kPerThread ? kPerThread : nIn - q0)) : 0;
    int32_t xb[kPerThread], xe[kPerThread];
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    if (n_mine == kPerThread) {
        int4 vb, ve;
        __builtin_memcpy(&vb, __builtin_assume_aligned(beg + base + q0, 4), 16);
        __builtin_memcpy(&ve, __builtin_assume_aligned(end + base + q0, 4), 16);
        xb[0] = vb.x, xb[1] = vb.y, xb[2] = vb.z, xb[3] = vb.w;
        xe[0] = ve.x, xe[1] = ve.y, xe[2] = ve.z, xe[3] = ve.w;
    } else {
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) {
            xb[u] = u < n_mine ? beg[base + q0 + u] : INT32_MAX;
            xe[u] = u < n_mine ? end[base + q0 + u] : INT32_MIN;
        }
    }
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
        lo = imin(lo, xb[u]);
        hi = imax(hi, xe[u]);
    }
    lo = wave_last(wave_incl_min(lo));
    hi = wave_last(wave_incl_max(hi));
    if (lane == 0) {
        s_red[wave] = lo;
        s_red[kTW + wave] = hi;
    }
    __syncthreads();  // row starts, per-wave ranges
    {
        int32_t l = INT32_MAX, h = INT32_MIN;
        if (lane < kTW) {
            l = s_red[lane];
            h = s_red[kTW + lane];
        }
        lo = __builtin_amdgcn_readlane(row16_incl_min(l), 15);
        hi = __builtin_amdgcn_readlane(row16_incl_max(h), 15);
    }
    // widened values span [lo - hx, hi + hx]; keys below 2^30 leave room for
    // the row shifts
    const int64_t range = nIn > 0 ? (int64_t)hi - (int64_t)lo + 2 * (int64_t)hx : 0;
    const uint32_t shift = range > 0 ? 64u - (uint32_t)__clzll((unsigned long long)range) : 0u;
    lds_ok = lds_ok && range >= 0 && shift < 31 && ((int64_t)Rt << shift) < (1ll << 30);
    const int32_t xmin = (int32_t)((uint32_t)lo - (uint32_t)hx);
    int32_t total = 0;
    int cur = 0;
    if (lds_ok) {
        // stage: interval q of local row k -> key (k << shift) + (x - xmin)
        if (n_mine > 0) {
            int k = 0, kh = Rt;  // s_rs[k] <= q0 < s_rs[k + 1]
            while (kh - k > 1) {
                const int mid = (k + kh) >> 1;
                if (s_rs[mid] <= q0) k = mid;
                else kh = mid;
            }
            int32_t nxt = s_rs[k + 1];
#pragma unroll
            for (int u = 0; u < kPerThread; ++u) {
                const int32_t q = q0 + u;
                if (u < n_mine) {
                    while (q >= nxt) nxt = s_rs[++k + 1];  // s_rs[Rt] = nIn > q
                    const uint32_t off = ((uint32_t)k << shift) - (uint32_t)xmin;
                    s_buf[0][dsk(q)] = make_int2((int32_t)((uint32_t)xb[u] - (uint32_t)hx + off),
                                                 (int32_t)((uint32_t)xe[u] + (uint32_t)hx + off));
                }
            }
        }
        __syncthreads();  // staged keys
        total = window_union_steps(s_buf, cur, s_rs, nIn, K, nr, Rt, shift, s_wmax, s_wsum);
        lds_ok = total >= 0;
    }
    int32_t ex = 0;  // row-merge path: this thread's row's exclusive output count
    if (!lds_ok) {
        __syncthreads();  // s_red is read by the range combine above
        const int32_t c = tid < nr ? dense_window_union<KMAX, false>(D, beg, end, i0 + tid, K, hx, nullptr, nullptr, 0,
                                                                     0, status)
                                   : 0;
        ex = block_excl_sum<kTW>(c, s_red, &total);
    }
    if (wave == 0) {
        const uint32_t e = lookback_exclusive<1, false>(lb, tile, (uint32_t)total, epoch, status, spin_limit);
        if (lane == 0) s_excl = e;
    }
    __syncthreads();
    const int64_t excl = (int64_t)s_excl;
    const bool over = excl + total > capacity;
    if (over && tid == 0 && total > 0) atomicOr(status, STX_FLAG_CAPACITY);
    if (lds_ok) {
        const int2* X = s_buf[cur];
        if (!over) {
            for (int32_t k = tid; k < total; k += kDTB) {
                const int2 o = X[dsk(k)];
                const uint32_t boff = ((uint32_t)(o.x >> shift) << shift) - (uint32_t)xmin;
                ob[excl + k] = (int32_t)((uint32_t)o.x - boff);
                oe[excl + k] = (int32_t)((uint32_t)o.y - boff);
            }
        }
        for (int r = tid; r < nr; r += kDTB) out_off[i0 + r] = (int32_t)(excl + s_rs[r]);
    } else if (tid < nr) {
        out_off[i0 + tid] = (int32_t)(excl + ex);
        if (!over) dense_window_union<KMAX, true>(D, beg, end, i0 + tid, K, hx, ob, oe, excl + ex, capacity, status);
    }
    if (i0 + nr == nt && tid == 0) {
        out_off[nt] = (int32_t)(excl + total);
        if (d_total) *d_total = (int32_t)(excl + total);
    }
}

// One complement step, called by every thread of the block: for each local
// row r < Rc of X (rows starting at rs[r]), the gaps between the runs of its
// non-empty intervals inside the row's band [0, W) (keys (r << shift) + hx
// .. (r << shift) + W + hx), each widened by hw, in order, into Y; only rows
// in [ok_lo, ok_hi) produce gaps.  Slots: row r's intervals k at k + r, then
// its row-end slot at rs[r + 1] + r (the gap after the row's last run).  A
// gap ends at an interval's begin (or the band's end) and starts at the
// largest end before it (or the band's start): the thread's own part of
// that maximum is kept per slot and the part before the thread (the carry)
// folded in after one block max.  Rewrites rs for Y (rows 0 .. Rc).
// Returns Y's count or -1 when it would not fit a buffer (uniform).
__device__ __forceinline__ int32_t complement_step(const int2* __restrict__ X, int2* __restrict__ Y, int32_t* rs,
                                                   int Rc, int ok_lo, int ok_hi, int32_t W, int32_t hx, int32_t hw,
                                                   uint32_t shift, int32_t* wmax, int32_t* wsum) {
    constexpr int kTW = kDTB / STX_WAVE;
    const int tid = threadIdx.x;
    const int32_t P = rs[Rc] + Rc;
    const int32_t p0 = tid * kDVT;
    int32_t gs[kDVT], g1[kDVT];  // gap start (own part of the maximum) and end, keys
    uint32_t cand = 0, first = 0, rowend = 0;
    int32_t lm = INT32_MIN;  // largest end key of this thread's non-empty intervals so far
    int r0 = 0;
    if (p0 < P) {
        int k = 0, kh = Rc;  // row of slot p0: rs[k] + k <= p0 < rs[k + 1] + k + 1
        while (kh - k > 1) {
            const int mid = (k + kh) >> 1;
            if (rs[mid] + mid <= p0) k = mid;
            else kh = mid;
        }
        r0 = k;
    }
    {
        int r = r0;
        int32_t rstart = rs[r] + r, rnext = rs[r + 1];
#pragma unroll
        for (int v = 0; v < kDVT; ++v) {
            const int32_t p = p0 + v;
            const int32_t lo_key = (int32_t)(((uint32_t)r << shift) + (uint32_t)hx);
            const int32_t hi_key = lo_key + W;
            const bool ok = r >= ok_lo && r < ok_hi;
            gs[v] = imax(lm, lo_key);
            g1[v] = hi_key;
            if (p < P) {
                first |= (p == rstart ? 1u : 0u) << v;
                if (p - r < rnext) {  // interval slot
                    const int2 it = X[dsk(p - r)];
                    if (it.y > it.x) {
                        g1[v] = imin(it.x, hi_key);
                        cand |= (ok ? 1u : 0u) << v;
                        lm = imax(lm, it.y);
                    }
                } else {  // row-end slot
                    cand |= (ok ? 1u : 0u) << v;
                    rowend |= 1u << v;
                    ++r;
                    rstart = p + 1;
                    rnext = r < Rc ? rs[r + 1] : 0;
                }
            }
        }
    }
    const int32_t carry = block_excl_max<kTW>(lm, wmax);
    uint32_t emit = 0;
#pragma unroll
    for (int v = 0; v < kDVT; ++v) {
        gs[v] = imax(gs[v], carry);
        emit |= ((cand & (1u << v)) && g1[v] > gs[v] ? 1u : 0u) << v;
    }
    int32_t step_total = 0;
    const int32_t my_base = block_excl_sum<kTW>(__popc(emit), wsum, &step_total);
    if (step_total > kDCap) return -1;
    int32_t w = my_base;
    int r = r0;
#pragma unroll
    for (int v = 0; v < kDVT; ++v) {
        if (first & (1u << v)) rs[r] = w;
        if (emit & (1u << v)) {
            Y[dsk(w)] = make_int2(gs[v] - hw, g1[v] + hw);
            ++w;
        }
        if (rowend & (1u << v)) ++r;
    }
    if (tid == 0) rs[Rc] = step_total;
    __syncthreads();  // Y and its row starts complete; every read of X done
    return step_total;
}

// Next non-empty gap inside [0, W) of one input row (interval cursor jj
// before jje, pe the largest clipped end so far; -1 once the gap after the
// row's last interval has been produced): [gb, g1), or gb = INT_MAX.
__device__ __forceinline__ void next_gap(const int32_t* __restrict__ beg, const int32_t* __restrict__ end, int32_t& jj,
                                         int32_t jje, int32_t& pe, int32_t W, int32_t& gb, int32_t& g1) {
    while (true) {
        if (jj < jje) {
            const int32_t b = imin(imax(beg[jj], 0), W), e = imin(imax(end[jj], 0), W);
            ++jj;
            if (e > b) {
                const bool gap = b > pe;
                const int32_t s = pe;
                pe = imax(pe, e);
                if (gap) {
                    gb = s;
                    g1 = b;
                    return;
                }
            }
        } else if (pe >= 0) {
            const int32_t s = pe;
            pe = -1;
            if (W > s) {
                gb = s;
                g1 = W;
                return;
            }
        } else {
            gb = INT_MAX;
            return;
        }
    }
}

// Erosion of output row i (0 <= i < H), one thread: [0, W) minus the union
// of the gaps inside [0, W) of the rows i - hy .. i + hy that lie in [0, H)
// (dense rows j = i .. i + K - 1 of D, row j holding row j - hy), each gap
// widened by hx.  The count (WRITE = false) or the intervals from ob/oe + w.
template <int KMAX, bool WRITE>
__device__ int32_t erode_row(const int32_t* __restrict__ D, const int32_t* __restrict__ beg,
                             const int32_t* __restrict__ end, int64_t i, int K, int64_t H, int32_t hy, int32_t W,
                             int32_t hx, int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t w,
                             int64_t capacity, uint32_t* __restrict__ status) {
    int32_t j[KMAX], je[KMAX], pe[KMAX], gb[KMAX], g1[KMAX];
#pragma unroll
    for (int t = 0; t < KMAX; ++t) {
        const int64_t y = i + t - hy;
        const bool in = t < K && y >= 0 && y < H;
        j[t] = in ? D[i + t] : 0;
        je[t] = in ? D[i + t + 1] : 0;
        pe[t] = in ? 0 : -1;
        gb[t] = INT_MAX;
        g1[t] = 0;
        if (in) next_gap(beg, end, j[t], je[t], pe[t], W, gb[t], g1[t]);
    }
    int32_t n = 0, op = 0;  // op: where the next output piece may start
    int32_t s = 0, m = 0;
    bool open = false;
    while (true) {
        int32_t best = INT_MAX, tb = -1;
#pragma unroll
        for (int t = 0; t < KMAX; ++t)
            if (gb[t] < best) { best = gb[t]; tb = t; }
        int32_t b = 0, e = 0;
        if (tb >= 0) {
#pragma unroll
            for (int t = 0; t < KMAX; ++t) {
                if (t == tb) {
                    b = gb[t] - hx;
                    e = g1[t] + hx;
                    next_gap(beg, end, j[t], je[t], pe[t], W, gb[t], g1[t]);
                }
            }
        }
        if (open && (tb < 0 || b > m)) {
            // run [s, m) of the union closes: the piece of [0, W) before it
            const int32_t cs = imin(imax(s, 0), W);
            if (cs > op) {
                if (WRITE) {
                    if (w < capacity) { ob[w] = op; oe[w] = cs; } else atomicOr(status, STX_FLAG_CAPACITY);
                    ++w;
                }
                ++n;
            }
            op = imax(op, imin(imax(m, 0), W));
            open = false;
        }
        if (tb < 0) break;
        if (!open) {
            s = b;
            m = e;
            open = true;
        } else if (e > m) {
            m = e;
        }
    }
    if (W > op) {
        if (WRITE) {
            if (w < capacity) { ob[w] = op; oe[w] = W; } else atomicOr(status, STX_FLAG_CAPACITY);
        }
        ++n;
    }
    return n;
}

// Erosion over the rows 0 .. H - 1 (erode_interval_set): U - clip(dilate(U -
// clip(S))), U = [0, W) x [0, H), in one pass per tile of output rows.  A tile
// stages its rows' input window (rows i0 - hy .. i0 + nr + hy; values
// clipped to [0, W], keys (row << shift) + x + hx with 2^shift > W + 2 hx),
// takes the complement inside [0, W) of the rows that lie in [0, H) with its
// gaps widened by hx (complement_step), the window union by doubling
// (window_union_steps), and the complement of that inside [0, W) (the
// erosion).  A tile that leaves the LDS layout erodes each row in one thread
// (erode_row).  Output: offsets of every row 0 .. H - 1.
template <int KMAX>
__global__ __launch_bounds__(kDTB) __attribute__((amdgpu_waves_per_eu(8))) void erode_tile_kernel(
    const int32_t* __restrict__ D, const int32_t* __restrict__ beg, const int32_t* __restrict__ end, int64_t H,
    int32_t W, int32_t hx, int32_t hy, uint32_t shift, const int2* __restrict__ tab, uint32_t ntiles, int32_t* __restrict__ out_off,
    int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t capacity, int32_t* __restrict__ d_total,
    uint32_t* status, void* scratch, uint32_t epoch, int64_t spin_limit) {
    constexpr int kTW = kDTB / STX_WAVE;
    constexpr int kPerThread = kDCap / kDTB;
    __shared__ __align__(16) int2 s_buf[2][dsk_size(kDCap)];
    __shared__ int32_t s_rs[kDRows];
    __shared__ int32_t s_wmax[kTW];
    __shared__ int32_t s_wsum[kTW];
    __shared__ int32_t s_fb[kTW];
    __shared__ uint32_t s_excl;
    LookbackScratch lb = lookback_scratch(scratch, ntiles);
    const uint32_t tile = blockIdx.x;
    const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    const int K = 2 * hy + 1;
    // rows [i0, i0 + nr) from the count-balanced table, as for the dilation
    const int64_t i0 = tab[tile].x;
    if (i0 >= H && tile > 0) return;
    const int64_t n_rows = (int64_t)tab[tile + 1].x - i0;
    if (n_rows > kDTB && tid == 0) atomicOr(status, STX_FLAG_CAPACITY);
    const int nr = (int)(n_rows < kDTB ? n_rows : kDTB);
    const int Rt = nr + 2 * hy;  // staged rows i0 - hy .. i0 + nr + hy - 1
    const int32_t base = D[i0];
    const int32_t nIn = D[i0 + Rt] - base;
    bool lds_ok = nr > 0 && nIn <= kDCap && ((int64_t)Rt << shift) < (1ll << 30);  // uniform
    for (int r = tid; r <= Rt; r += kDTB) s_rs[r] = D[i0 + r] - base;
    const int32_t q0 = tid * kPerThread;
    const int32_t n_mine = lds_ok ? (nIn - q0 < 0 ? 0 : (nIn - q0 > kPerThread ? kPerThread : nIn - q0)) : 0;
    int32_t xb[kPerThread], xe[kPerThread];
    if (n_mine == kPerThread) {
        int4 vb, ve;
        __builtin_memcpy(&vb, __builtin_assume_aligned(beg + base + q0, 4), 16);
        __builtin_memcpy(&ve, __builtin_assume_aligned(end + base + q0, 4), 16);
        xb[0] = vb.x, xb[1] = vb.y, xb[2] = vb.z, xb[3] = vb.w;
        xe[0] = ve.x, xe[1] = ve.y, xe[2] = ve.z, xe[3] = ve.w;
    } else {
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) {
            xb[u] = u < n_mine ? beg[base + q0 + u] : 0;
            xe[u] = u < n_mine ? end[base + q0 + u] : 0;
        }
    }
    __syncthreads();  // row starts
    int32_t total = 0;
    int cur = 0;
    if (lds_ok) {
        if (n_mine > 0) {
            int k = 0, kh = Rt;  // s_rs[k] <= q0 < s_rs[k + 1]
            while (kh - k > 1) {
                const int mid = (k + kh) >> 1;
                if (s_rs[mid] <= q0) k = mid;
                else kh = mid;
            }
            int32_t nxt = s_rs[k + 1];
#pragma unroll
            for (int u = 0; u < kPerThread; ++u) {
                const int32_t q = q0 + u;
                if (u < n_mine) {
                    while (q >= nxt) nxt = s_rs[++k + 1];
                    const int32_t off = (int32_t)(((uint32_t)k << shift) + (uint32_t)hx);
                    s_buf[0][dsk(q)] = make_int2(imin(imax(xb[u], 0), W) + off, imin(imax(xe[u], 0), W) + off);
                }
            }
        }
        __syncthreads();  // staged keys
        // local row r holds row i0 + r - hy: rows in [0, H) take part
        const int ok_lo = (int)(hy - i0 > 0 ? hy - i0 : 0);
        const int ok_hi = (int)(H + hy - i0 < Rt ? H + hy - i0 : Rt);
        int32_t n = complement_step(s_buf[0], s_buf[1], s_rs, Rt, ok_lo, ok_hi, W, hx, hx, shift, s_wmax, s_wsum);
        cur = 1;
        if (n >= 0 && K > 1) n = window_union_steps(s_buf, cur, s_rs, n, K, nr, Rt, shift, s_wmax, s_wsum);
        if (n >= 0) {
            n = complement_step(s_buf[cur], s_buf[cur ^ 1], s_rs, nr, 0, nr, W, hx, 0, shift, s_wmax, s_wsum);
            cur ^= 1;
        }
        lds_ok = n >= 0;
        total = n;
    }
    int32_t ex = 0;  // row-merge path: this thread's row's exclusive output count
    if (!lds_ok) {
        const int32_t c = tid < nr ? erode_row<KMAX, false>(D, beg, end, i0 + tid, K, H, hy, W, hx, nullptr, nullptr,
                                                            0, 0, status)
                                   : 0;
        ex = block_excl_sum<kTW>(c, s_fb, &total);
    }
    if (wave == 0) {
        const uint32_t e = lookback_exclusive<1, false>(lb, tile, (uint32_t)total, epoch, status, spin_limit);
        if (lane == 0) s_excl = e;
    }
    __syncthreads();
    const int64_t excl = (int64_t)s_excl;
    const bool over = excl + total > capacity;
    if (over && tid == 0 && total > 0) atomicOr(status, STX_FLAG_CAPACITY);
    if (lds_ok) {
        const int2* X = s_buf[cur];
        if (!over) {
            for (int32_t k = tid; k < total; k += kDTB) {
                const int2 o = X[dsk(k)];
                const int32_t boff = (int32_t)(((uint32_t)(o.x >> shift) << shift) + (uint32_t)hx);
                ob[excl + k] = o.x - boff;
                oe[excl + k] = o.y - boff;
            }
        }
        for (int r = tid; r < nr; r += kDTB) out_off[i0 + r] = (int32_t)(excl + s_rs[r]);
    } else if (tid < nr) {
        out_off[i0 + tid] = (int32_t)(excl + ex);
        if (!over) erode_row<KMAX, true>(D, beg, end, i0 + tid, K, H, hy, W, hx, ob, oe, excl + ex, capacity, status);
    }
    if (i0 + nr == H && tid == 0) {
        out_off[H] = (int32_t)(excl + total);
        if (d_total) *d_total = (int32_t)(excl + total);
    }
}

// The caller's first and last row id against the set's (STX_FLAG_ROWS if
// they differ: rows outside the dense range would be dropped silently).
__global__ void row_bounds_check_kernel(const int32_t* __restrict__ rows, int64_t R, int32_t lo, int32_t hi,
                                        uint32_t* status) {
    if (threadIdx.x == 0 && R > 0 && (rows[0] != lo || rows[R - 1] != hi)) atomicOr(status, STX_FLAG_ROWS);
}

__global__ void dense_rows_kernel(int32_t* __restrict__ rows, int64_t n, int32_t y0) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rows[i] = (int32_t)((int64_t)y0 + i);
}

}  // namespace

// Count-balanced row tiles over dense offsets D (D row j holds row j - hy
// of the tile rows' numbering, so tile row i is D row i + hy): with c(i) =
// the intervals of D rows hy .. hy + i - 1 plus per_row for each of those
// rows (the erosion's complement adds up to one gap per row) and M = kDTB,
// h(i) = floor(c(i) / T + i / M), tile k starts at the first row with h >=
// k.  Between a tile's first and last row h grows by less than one, so its
// rows hold fewer than T intervals plus its last row, and there are at most
// M of them; unlike floor(c / T) + floor(i / M) no tile is cut every M rows
// (that cut 8 % more tiles on uniform stacks).  A tile's staged window adds
// hy rows on each side.  Table entries as in balanced_table_entry.
__global__ void dense_tile_table_kernel(const int32_t* __restrict__ D, int64_t U, int32_t hy, int32_t per_row,
                                        int64_t T, int64_t ntiles, int2* __restrict__ tab, uint32_t* status) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c0 = D[hy];
    auto h = [&](int64_t r) { return (((int64_t)D[r + hy] - c0 + (int64_t)per_row * r) * kDTB + r * T) / (T * kDTB); };
    balanced_table_entry(i, U, ntiles, h, tab, status);
}

// Table target for tiles staged to `fill` of a buffer: the tile's own rows
// get t = fill x kDCap less `halo` x the expected 2 hy halo rows of `mean`
// intervals, at least a quarter buffer (which bounds the tile count for the
// scratch size), raised by 1 / (1 - t / (M mean)) for the rows' own share of
// h, so that tiles of rows at the mean length hold t intervals.
inline int64_t dense_tile_target(double fill, double halo, int32_t hy, double mean) {
    double t = fill * kDCap - halo * 2.0 * hy * mean;
    if (t < kDCap / 4) t = kDCap / 4;
    const double rows_share = t / (kDTB * (mean > 1e-3 ? mean : 1e-3));
    return (int64_t)(rows_share < 0.5 ? t / (1.0 - rows_share) : 2.0 * t);
}

// Tiles of the table: h(U) + 1 for rows of total weight `weight`.
inline int64_t dense_tiles(int64_t weight, int64_t T, int64_t U) { return (weight * kDTB + U * T) / (T * kDTB) + 1; }

// Writes the table of `ntiles` tiles over U rows after the look-back states
// in the scratch; returns it.
static int2* dense_tile_table(const int32_t* D, int64_t U, int32_t hy, int32_t per_row, int64_t T, int64_t ntiles,
                              void* scratch, uint32_t* status, hipStream_t stream) {
    int2* tab = (int2*)((char*)scratch + balanced_table_offset(ntiles));
    const int64_t n = (U > ntiles ? U : ntiles) + 1;
    hipLaunchKernelGGL(dense_tile_table_kernel, dim3(nblk(n)), dim3(kT), 0, stream, D, U, hy, per_row, T, ntiles, tab,
                       status);
    return tab;
}

extern "C" int64_t stx_scan_scratch_bytes(int64_t n_bound);
extern "C" int32_t stx_dense_offsets(const int32_t* rows, const int32_t* off, int64_t R, int64_t y0, int64_t n,
                                     int32_t* out, hipStream_t stream);

extern "C" {

int32_t stx_dilate_max_halo(void) { return kMaxHalo; }

int64_t stx_dilate_tmp_ints(int64_t R, int32_t hy) {
    const int64_t NYb = R * (2 * (int64_t)hy + 1);
    // rows_c, off_c, compact tmp, d_rc, row counts, row positions, klo, khi, union counts
    return R + (R + 1) + 2 * (R + 1) + 1 + (R + 1) + (R + 1) + NYb + NYb + (NYb + 1);
}

int32_t stx_dilate(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg, const int32_t* end,
                   int32_t hx, int32_t hy, int32_t* rows_out, int32_t* off_out, int32_t* beg_out, int32_t* end_out,
                   int64_t row_cap, int64_t capacity, int32_t* d_ny, int32_t* d_total, int32_t* tmp, int64_t tmp_ints,
                   uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    if (R < 0 || hx < 0 || hy < 1 || hy > kMaxHalo) return stx_fail(STX_EINVAL, "dilate: need R >= 0, hx >= 0, 1 <= hy <= 4");
    const int64_t NYb = R * (2 * (int64_t)hy + 1);
    if (row_cap < NYb) return stx_fail(STX_EINVAL, "dilate: row capacity below rows * (2 hy + 1)");
    if (tmp_ints < stx_dilate_tmp_ints(R, hy)) return stx_fail(STX_EINVAL, "dilate: tmp too small");
    int32_t* rows_c = tmp;
    int32_t* off_c = rows_c + R;
    int32_t* ctmp = off_c + (R + 1);
    int32_t* d_rc = ctmp + 2 * (R + 1);
    int32_t* rcnt = d_rc + 1;
    int32_t* rpos = rcnt + (R + 1);
    int32_t* klo = rpos + (R + 1);
    int32_t* khi = klo + NYb;
    int32_t* ucnt = khi + NYb;
    int32_t rc = stx_compact_rows(rows, off, R, 0, rows_c, off_c, d_rc, ctmp, status, scratch, scratch_bytes, epoch,
                                  stream);
    if (rc) return rc;
    if (R > 0)
        hipLaunchKernelGGL(win_rows_count_kernel, dim3(nblk(R)), dim3(kT), 0, stream, rows_c, d_rc, R, hy, rcnt);
    rc = stx_scan_i32(rcnt, d_rc, R, rpos, d_ny, status, scratch, scratch_bytes, epoch + 1, stream);
    if (rc) return rc;
    if (R > 0)
        hipLaunchKernelGGL(win_rows_write_kernel, dim3(nblk(R)), dim3(kT), 0, stream, rows_c, d_rc, R, hy, rpos,
                           rows_out, klo, khi);
    const int K = 2 * hy + 1;
    if (NYb > 0) {
        if (K <= 3) launch_union<3>(false, off_c, beg, end, d_ny, NYb, klo, khi, hx, ucnt, nullptr, nullptr, nullptr, 0, status, stream);
        else if (K <= 5) launch_union<5>(false, off_c, beg, end, d_ny, NYb, klo, khi, hx, ucnt, nullptr, nullptr, nullptr, 0, status, stream);
        else launch_union<9>(false, off_c, beg, end, d_ny, NYb, klo, khi, hx, ucnt, nullptr, nullptr, nullptr, 0, status, stream);
    }
    rc = stx_scan_i32(ucnt, d_ny, NYb, off_out, d_total, status, scratch, scratch_bytes, epoch + 2, stream);
    if (rc) return rc;
    if (NYb > 0) {
        if (K <= 3) launch_union<3>(true, off_c, beg, end, d_ny, NYb, klo, khi, hx, nullptr, off_out, beg_out, end_out, capacity, status, stream);
        else if (K <= 5) launch_union<5>(true, off_c, beg, end, d_ny, NYb, klo, khi, hx, nullptr, off_out, beg_out, end_out, capacity, status, stream);
        else launch_union<9>(true, off_c, beg, end, d_ny, NYb, klo, khi, hx, nullptr, off_out, beg_out, end_out, capacity, status, stream);
    }
    return stx_launch_status("dilate");
}

int64_t stx_dilate_dense_tmp_ints(int64_t nt, int32_t hy) {
    // dense input offsets, dense output offsets, dense row ids, compaction tmp
    return (nt + 2 * (int64_t)hy + 1) + (nt + 1) + nt + 2 * (nt + 1);
}

int64_t stx_dilate_dense_scratch_bytes(int64_t nt, int64_t N) {
    const int64_t lb = balanced_table_scratch_bytes(dense_tiles(N > 0 ? N : 0, kDCap / 4, nt));
    const int64_t sc = stx_scan_scratch_bytes(nt);
    return lb > sc ? lb : sc;
}

int32_t stx_dilate_dense(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg, const int32_t* end,
                         int64_t N, int32_t hx, int32_t hy, int32_t row_lo, int32_t row_hi, int32_t* rows_out,
                         int32_t* off_out, int32_t* beg_out, int32_t* end_out, int64_t row_cap, int64_t capacity,
                         int32_t* d_ny, int32_t* d_total, int32_t* tmp, int64_t tmp_ints, uint32_t* status,
                         void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    if (R < 1 || N < 0 || hx < 0 || hy < 1 || hy > kMaxHalo || row_hi < row_lo)
        return stx_fail(STX_EINVAL, "dilate_dense: need R >= 1, N >= 0, hx >= 0, 1 <= hy <= 4, row_lo <= row_hi");
    if ((int64_t)row_lo - hy < INT32_MIN || (int64_t)row_hi + hy > INT32_MAX)
        return stx_fail(STX_EINVAL, "dilate_dense: dilated row ids leave the int32 range");
    const int64_t span = (int64_t)row_hi - row_lo + 1;
    const int64_t nt = span + 2 * (int64_t)hy;
    if (row_cap < nt) return stx_fail(STX_EINVAL, "dilate_dense: row capacity below the dense row count");
    if (tmp_ints < stx_dilate_dense_tmp_ints(nt, hy)) return stx_fail(STX_EINVAL, "dilate_dense: tmp too small");
    if (scratch_bytes < stx_dilate_dense_scratch_bytes(nt, N))
        return stx_fail(STX_EINVAL, "dilate_dense: scratch too small");
    int32_t* D = tmp;
    int32_t* off_d = D + (nt + 2 * (int64_t)hy + 1);
    int32_t* rows_d = off_d + (nt + 1);
    int32_t* ctmp = rows_d + nt;
    hipLaunchKernelGGL(row_bounds_check_kernel, dim3(1), dim3(64), 0, stream, rows, R, row_lo, row_hi, status);
    int32_t rc = stx_dense_offsets(rows, off, R, (int64_t)row_lo - 2 * hy, nt + 2 * hy, D, stream);
    if (rc) return rc;
    // count-balanced tiles staging about kDFill of a buffer, so that the
    // steps (the first may grow the count) rarely overflow: with tiles of
    // equal rows planned at the mean row length, the 64-instance stack took
    // 0.23 / 0.27 / 0.29 ms at r = 1 / 2 / 4 with 0.7, 0.55 / 0.40 / 0.28 ms
    // with 0.8 (tiles falling back to row merges), 0.24 / 0.27 / 0.31 ms
    // with 0.6
    constexpr double kDFill = 0.7;
    const int64_t T = dense_tile_target(kDFill, 1.0, hy, (double)N / (double)span);
    const int64_t ntiles = dense_tiles(N, T, nt);
    if (balanced_table_scratch_bytes(ntiles) > scratch_bytes)
        return stx_fail(STX_EINVAL, "dilate_dense: scratch too small");
    const int2* tab = dense_tile_table(D, nt, hy, 0, T, ntiles, scratch, status, stream);
    const int K = 2 * hy + 1;
#define STX_DILATE_LAUNCH(KM)                                                                                          \
    hipLaunchKernelGGL((dilate_tile_kernel<KM>), dim3((unsigned)ntiles), dim3(kDTB), 0, stream, D, beg, end, nt, hx, hy, \
                       tab, (uint32_t)ntiles, off_d, beg_out, end_out, capacity, d_total, status, scratch, epoch,       \
                       stx_spin_limit())
    if (K <= 3) STX_DILATE_LAUNCH(3);
    else if (K <= 5) STX_DILATE_LAUNCH(5);
    else STX_DILATE_LAUNCH(9);
#undef STX_DILATE_LAUNCH
    rc = stx_launch_status("dilate_tile_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(dense_rows_kernel, dim3(nblk(nt)), dim3(kT), 0, stream, rows_d, nt, (int32_t)(row_lo - hy));
    return stx_compact_rows(rows_d, off_d, nt, 0, rows_out, off_out, d_ny, ctmp, status, scratch, scratch_bytes,
                            epoch + 1, stream);
}

int64_t stx_erode_dense_tmp_ints(int64_t H, int32_t hy) { return H + 2 * (int64_t)hy + 1; }

int64_t stx_erode_dense_scratch_bytes(int64_t H, int64_t N) {
    return balanced_table_scratch_bytes(dense_tiles((N > 0 ? N : 0) + (H > 0 ? H : 0), kDCap / 4, H > 1 ? H : 1));
}

int32_t stx_erode_dense(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg, const int32_t* end,
                        int64_t N, int32_t hx, int32_t hy, int32_t W, int64_t H, int32_t* off_out, int32_t* beg_out,
                        int32_t* end_out, int64_t capacity, int32_t* d_total, int32_t* tmp, int64_t tmp_ints,
                        uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    if (R < 0 || N < 0 || hx < 0 || hy < 0 || hy > kMaxHalo || W < 1 || H < 1 || H > INT32_MAX)
        return stx_fail(STX_EINVAL, "erode_dense: need R, N, hx >= 0, 0 <= hy <= 4, W >= 1, 1 <= H < 2^31");
    if (tmp_ints < stx_erode_dense_tmp_ints(H, hy)) return stx_fail(STX_EINVAL, "erode_dense: tmp too small");
    if (scratch_bytes < stx_erode_dense_scratch_bytes(H, N))
        return stx_fail(STX_EINVAL, "erode_dense: scratch too small");
    int32_t* D = tmp;  // row j of D holds row j - hy, j = 0 .. H + 2 hy - 1
    int32_t rc = stx_dense_offsets(rows, off, R, -(int64_t)hy, H + 2 * hy, D, stream);
    if (rc) return rc;
    const int64_t band = (int64_t)W + 2 * (int64_t)hx;  // key offsets 0 .. W + 2 hx
    uint32_t shift = 0;
    while (shift < 40 && ((int64_t)1 << shift) <= band) ++shift;
    // the complement holds about one interval more per row than the set;
    // its widened gaps only merge from there, so the staged count is the
    // tile's largest (0.9 of a buffer: 0.22 / 0.24 / 0.25 ms against 0.25 /
    // 0.27 / 0.31 ms at 0.7 on the 64-instance stack, r = 1 / 2 / 4, tiles
    // of equal rows); count-balanced tiles, a row weighing its intervals + 1,
    // the halo allowance 1.5 x the mean halo: at this fill the longer halo
    // rows of denser regions matter (with 1 x, the tiles of the dense
    // instances of a stack alternating 16 / 64 intervals per row overflowed
    // at r = 4: 3.3 ms against 0.21 ms; uniform stacks lose about 5 % at r =
    // 4 to the smaller tiles)
    constexpr double kDFill = 0.9;
    const int64_t T = dense_tile_target(kDFill, 1.5, hy, (double)N / (double)H + 1.0);
    const int64_t ntiles = dense_tiles(N + H, T, H);
    if (balanced_table_scratch_bytes(ntiles) > scratch_bytes)
        return stx_fail(STX_EINVAL, "erode_dense: scratch too small");
    const int2* tab = dense_tile_table(D, H, hy, 1, T, ntiles, scratch, status, stream);
    const int K = 2 * hy + 1;
#define STX_ERODE_LAUNCH(KM)                                                                                         \
    hipLaunchKernelGGL((erode_tile_kernel<KM>), dim3((unsigned)ntiles), dim3(kDTB), 0, stream, D, beg, end, H, W, hx, \
                       hy, shift, tab, (uint32_t)ntiles, off_out, beg_out, end_out, capacity, d_total, status, \
                       scratch, epoch, stx_spin_limit())
    if (K <= 1) STX_ERODE_LAUNCH(1);
    else if (K <= 3) STX_ERODE_LAUNCH(3);
    else if (K <= 5) STX_ERODE_LAUNCH(5);
    else STX_ERODE_LAUNCH(9);
#undef STX_ERODE_LAUNCH
    return stx_launch_status("erode_tile_kernel");
}

}  // extern "C"
