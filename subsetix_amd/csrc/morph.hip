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

}  // namespace

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

}  // extern "C"
