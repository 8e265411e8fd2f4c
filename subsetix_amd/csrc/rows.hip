// Row-structure kernels: CSR validation, interval -> row expansion, dense row
// lookup (clip), empty-row compaction, coarse-row maps (restriction), set
// prolongation, translation and fills.
#include "stx_common.h"
#include "stx_host.h"

using namespace stx;

namespace {

constexpr int kT = 256;

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

// flags: bit0 off[0] != 0, bit1 off[R] != N, bit2 rows not strictly
// increasing, bit3 offsets decreasing, bit4 rows != arange(R)
__global__ void validate_kernel(const int32_t* __restrict__ off, int64_t R, const int32_t* __restrict__ rows, int64_t N,
                                uint32_t* flags) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t f = 0;
    if (r == 0) {
        if (off[0] != 0) f |= 1u;
        if ((int64_t)off[R] != N) f |= 2u;
    }
    if (r < R) {
        if (r + 1 < R && rows[r + 1] <= rows[r]) f |= 4u;
        if (off[r + 1] < off[r]) f |= 8u;
        if (rows[r] != (int32_t)r) f |= 16u;
    }
    if (f) atomicOr(flags, f);
}

__global__ void interval_rows_kernel(const int32_t* __restrict__ off, int64_t R, const int32_t* __restrict__ rows, int64_t N,
                                     int32_t* __restrict__ out_row, int32_t* __restrict__ out_ord) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int64_t lo = 0, hi = R;  // last r with off[r] <= i
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    // skip empty rows that share the same offset
    while (lo + 1 < R && off[lo + 1] <= i) ++lo;
    if (out_row) out_row[i] = rows[lo];
    if (out_ord) out_ord[i] = (int32_t)lo;
}

__global__ void rows_lookup_kernel(const int32_t* __restrict__ rows, int64_t R, int32_t y0, int64_t H,
                                   int32_t* __restrict__ map) {
    int64_t y = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= H) return;
    int32_t v = y0 + (int32_t)y;
    int64_t lo = 0, hi = R;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (rows[mid] < v) lo = mid + 1; else hi = mid;
    }
    map[y] = (lo < R && rows[lo] == v) ? (int32_t)lo : -1;
}

// Dense offsets over the rows y0 .. y0 + n - 1: out[j] = off[k], k the first
// row index with rows[k] >= y0 + j, i.e. the number of intervals (counted
// from off[0]) in rows below y0 + j; rows the set does not hold come out
// empty.  n + 1 entries.
__global__ void dense_offsets_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ off, int64_t R,
                                     int64_t y0, int64_t n, int32_t* __restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    const int64_t y = y0 + j;
    int64_t lo = 0, hi = R;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)rows[mid] < y) lo = mid + 1; else hi = mid;
    }
    out[j] = off[lo];
}

__global__ void nonempty_flag_kernel(const int32_t* __restrict__ off, int64_t R, int32_t* __restrict__ flag) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) flag[r] = off[r + 1] > off[r] ? 1 : 0;
}

__global__ void compact_scatter_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ off, int64_t R,
                                       const int32_t* __restrict__ flag, const int32_t* __restrict__ pos, int32_t shift,
                                       int32_t* __restrict__ rows_out, int32_t* __restrict__ off_out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R && flag[r]) {
        rows_out[pos[r]] = rows[r] + shift;
        off_out[pos[r]] = off[r];
    }
    if (r == R) off_out[pos[R]] = off[R];
}

// Restriction rows.  A coarse row c = floor(row / ratio) is kept when one of
// its fine rows holds an interval (or, with all_rows, whenever any fine row
// maps to it).  flag marks the first qualifying fine row of each coarse row.
__global__ void coarse_flag_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ off, int64_t R,
                                   int32_t ratio, int32_t all_rows, int32_t* __restrict__ flag) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int32_t c = floordiv(rows[r], ratio);
    const bool ok = all_rows || off[r + 1] > off[r];
    bool first = true;
    for (int64_t q = r - 1; q >= 0; --q) {  // at most ratio-1 steps (rows strictly increasing)
        if (floordiv(rows[q], ratio) != c) break;
        if (all_rows || off[q + 1] > off[q]) { first = false; break; }
    }
    flag[r] = (ok && first) ? 1 : 0;
}

// maps: ratio arrays of length cap: maps[ph*cap + cidx] = fine row index of
// row c*ratio + ph that holds intervals (or -1).
__global__ void coarse_scatter_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ off, int64_t R,
                                      int32_t ratio, const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                                      int32_t* __restrict__ crow, int32_t* __restrict__ maps, int64_t cap) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R || !flag[r]) return;
    const int32_t c = floordiv(rows[r], ratio);
    const int64_t cidx = pos[r];
    crow[cidx] = c;
    if (maps == nullptr) return;
    for (int32_t ph = 0; ph < ratio; ++ph) maps[(int64_t)ph * cap + cidx] = -1;
    for (int64_t q = r; q < R; ++q) {
        if (floordiv(rows[q], ratio) != c) break;
        if (off[q + 1] > off[q]) maps[(int64_t)(rows[q] - c * ratio) * cap + cidx] = (int32_t)q;
    }
}

// Prolongation of a set with no empty rows (multilevel.py:213-264): coarse
// row k (id r) becomes fine rows r*R .. r*R+R-1, each holding the row's
// intervals scaled by R, in order.  Thread per coarse interval.
__global__ void prolong_set_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ off, int64_t K,
                                   const int32_t* __restrict__ beg, const int32_t* __restrict__ end, int64_t N,
                                   int32_t ratio, int32_t* __restrict__ frows, int32_t* __restrict__ foff,
                                   int32_t* __restrict__ fbeg, int32_t* __restrict__ fend, int32_t* __restrict__ fbase) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        int64_t lo = 0, hi = K;
        while (hi - lo > 1) {
            int64_t mid = (lo + hi) >> 1;
            if (off[mid] <= i) lo = mid; else hi = mid;
        }
        const int64_t k = lo;
        const int64_t n_k = off[k + 1] - off[k];
        const int64_t m = i - off[k];
        const int32_t b = beg[i] * ratio, e = end[i] * ratio;
        for (int32_t d = 0; d < ratio; ++d) {
            const int64_t o = (int64_t)off[k] * ratio + d * n_k + m;
            fbeg[o] = b;
            fend[o] = e;
            if (fbase) fbase[o] = (int32_t)i;
        }
    }
    if (i < K) {
        for (int32_t d = 0; d < ratio; ++d) {
            frows[i * ratio + d] = rows[i] * ratio + d;
            foff[i * ratio + d] = (int32_t)((int64_t)off[i] * ratio + d * (int64_t)(off[i + 1] - off[i]));
        }
        if (i == K - 1) foff[K * ratio] = (int32_t)((int64_t)off[K] * ratio);
    }
}

// Rows of an all-empty prolongation: sorted {r*R + d}.
__global__ void prolong_rows_kernel(const int32_t* __restrict__ rows, int64_t K, int32_t ratio, int32_t* __restrict__ frows) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < K * ratio) frows[t] = rows[t / ratio] * ratio + (int32_t)(t % ratio);
}

__global__ void fill_kernel(int32_t* __restrict__ p, int64_t n, int32_t v, int32_t step) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v + (int32_t)k * step;
}

__global__ void add_kernel(const int32_t* __restrict__ src, int32_t* __restrict__ dst, int64_t n, int32_t v) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k] + v;
}


// Point location (replaces locate_cells, interval_field.py:159-297): one
// lane per query; binary search of the query row among the row ids, then of
// the column among that row's intervals.  out_interval[q] = index of the
// interval holding (row, col), or -1 (the reference's result);
// out_cell[q] (optional, needs co) = index of the cell's value, -1 when the
// row exists but the cell is inactive, -2 when the row is absent.
__global__ void locate_cells_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ off,
                                    const int32_t* __restrict__ beg, const int32_t* __restrict__ end,
                                    const int32_t* __restrict__ co, int64_t R, const int32_t* __restrict__ qrow,
                                    const int32_t* __restrict__ qcol, int64_t nq, int32_t* __restrict__ out_interval,
                                    int32_t* __restrict__ out_cell) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += (int64_t)gridDim.x * blockDim.x) {
        const int32_t y = qrow[q], x = qcol[q];
        int64_t lo = 0, hi = R;  // first row id >= y
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (rows[mid] < y) lo = mid + 1; else hi = mid;
        }
        int32_t found = -1, cell = -2;
        if (lo < R && rows[lo] == y) {
            cell = -1;
            int32_t a = off[lo], b = off[lo + 1];  // first interval with end > x
            while (a < b) {
                const int32_t mid = a + ((b - a) >> 1);
                if (end[mid] <= x) a = mid + 1; else b = mid;
            }
            if (a < off[lo + 1] && beg[a] <= x) {
                found = a;
                if (co) cell = co[a] + (x - beg[a]);
            }
        }
        if (out_interval) out_interval[q] = found;
        if (out_cell) out_cell[q] = cell;
    }
}

}  // namespace

extern "C" int32_t stx_scan_i32(const int32_t* in, const int32_t* d_n, int64_t n_bound, int32_t* out, int32_t* d_total,
                                uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch,
                                hipStream_t stream);

extern "C" {

int32_t stx_validate_csr(const int32_t* off, int64_t R, const int32_t* rows, int64_t N, uint32_t* flags, hipStream_t stream) {
    if (R < 0 || N < 0) return stx_fail(STX_EINVAL, "validate: bad sizes");
    hipLaunchKernelGGL(validate_kernel, dim3(nblk(R + 1)), dim3(kT), 0, stream, off, R, rows, N, flags);
    return stx_launch_status("validate_csr");
}

int32_t stx_interval_rows(const int32_t* off, int64_t R, const int32_t* rows, int64_t N, int32_t* out_row,
                          int32_t* out_ord, hipStream_t stream) {
    if (N <= 0) return STX_OK;
    if (R <= 0) return stx_fail(STX_EINVAL, "interval_rows: intervals without rows");
    hipLaunchKernelGGL(interval_rows_kernel, dim3(nblk(N)), dim3(kT), 0, stream, off, R, rows, N, out_row, out_ord);
    return stx_launch_status("interval_rows");
}

int32_t stx_rows_lookup(const int32_t* rows, int64_t R, int32_t y0, int64_t H, int32_t* map, hipStream_t stream) {
    if (H <= 0) return STX_OK;
    hipLaunchKernelGGL(rows_lookup_kernel, dim3(nblk(H)), dim3(kT), 0, stream, rows, R, y0, H, map);
    return stx_launch_status("rows_lookup");
}

int32_t stx_dense_offsets(const int32_t* rows, const int32_t* off, int64_t R, int64_t y0, int64_t n, int32_t* out,
                          hipStream_t stream) {
    if (R < 0 || n < 0) return stx_fail(STX_EINVAL, "dense_offsets: negative size");
    hipLaunchKernelGGL(dense_offsets_kernel, dim3(nblk(n + 1)), dim3(kT), 0, stream, rows, off, R, y0, n, out);
    return stx_launch_status("dense_offsets");
}

// Drops the rows without intervals.  tmp: 2*(R+1) int32.  rows_out/off_out
// sized R / R+1; *d_count receives the kept row count.
int32_t stx_compact_rows(const int32_t* rows, const int32_t* off, int64_t R, int32_t shift, int32_t* rows_out,
                         int32_t* off_out, int32_t* d_count, int32_t* tmp, uint32_t* status, void* scratch,
                         int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    if (R < 0) return stx_fail(STX_EINVAL, "compact_rows: bad size");
    int32_t* flag = tmp;
    int32_t* pos = tmp + (R + 1);
    if (R > 0) hipLaunchKernelGGL(nonempty_flag_kernel, dim3(nblk(R)), dim3(kT), 0, stream, off, R, flag);
    int32_t rc = stx_scan_i32(flag, nullptr, R, pos, d_count, status, scratch, scratch_bytes, epoch, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(compact_scatter_kernel, dim3(nblk(R + 1)), dim3(kT), 0, stream, rows, off, R, flag, pos, shift,
                       rows_out, off_out);
    return stx_launch_status("compact_rows");
}

// Coarse rows of a restriction by `ratio` and, when maps != NULL, the fine
// source row of every phase.  tmp: 2*(R+1) int32; crow / maps rows sized
// cap >= R.  *d_count receives the coarse row count.
int32_t stx_coarse_rows(const int32_t* rows, const int32_t* off, int64_t R, int32_t ratio, int32_t all_rows,
                        int32_t* crow, int32_t* maps, int64_t cap, int32_t* d_count, int32_t* tmp, uint32_t* status,
                        void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    if (R < 0 || ratio < 1 || cap < R) return stx_fail(STX_EINVAL, "coarse_rows: bad arguments");
    int32_t* flag = tmp;
    int32_t* pos = tmp + (R + 1);
    if (R > 0) hipLaunchKernelGGL(coarse_flag_kernel, dim3(nblk(R)), dim3(kT), 0, stream, rows, off, R, ratio, all_rows, flag);
    int32_t rc = stx_scan_i32(flag, nullptr, R, pos, d_count, status, scratch, scratch_bytes, epoch, stream);
    if (rc) return rc;
    if (R > 0)
        hipLaunchKernelGGL(coarse_scatter_kernel, dim3(nblk(R)), dim3(kT), 0, stream, rows, off, R, ratio, flag, pos, crow,
                           maps, cap);
    return stx_launch_status("coarse_rows");
}

// Set prolongation of a set whose rows all hold intervals (K rows, N
// intervals) -> K*ratio rows, N*ratio intervals.  fbase optional.
int32_t stx_prolong_set(const int32_t* rows, const int32_t* off, int64_t K, const int32_t* beg, const int32_t* end,
                        int64_t N, int32_t ratio, int32_t* frows, int32_t* foff, int32_t* fbeg, int32_t* fend,
                        int32_t* fbase, hipStream_t stream) {
    if (ratio < 1 || K < 0 || N < 0) return stx_fail(STX_EINVAL, "prolong_set: bad arguments");
    int64_t n = N > K ? N : K;
    if (n > 0)
        hipLaunchKernelGGL(prolong_set_kernel, dim3(nblk(n)), dim3(kT), 0, stream, rows, off, K, beg, end, N, ratio, frows,
                           foff, fbeg, fend, fbase);
    return stx_launch_status("prolong_set");
}

int32_t stx_prolong_rows(const int32_t* rows, int64_t K, int32_t ratio, int32_t* frows, hipStream_t stream) {
    if (K * ratio > 0)
        hipLaunchKernelGGL(prolong_rows_kernel, dim3(nblk(K * ratio)), dim3(kT), 0, stream, rows, K, ratio, frows);
    return stx_launch_status("prolong_rows");
}

int32_t stx_fill_i32(int32_t* p, int64_t n, int32_t v, int32_t step, hipStream_t stream) {
    if (n > 0) hipLaunchKernelGGL(fill_kernel, dim3(nblk(n)), dim3(kT), 0, stream, p, n, v, step);
    return stx_launch_status("fill_i32");
}

int32_t stx_add_i32(const int32_t* src, int32_t* dst, int64_t n, int32_t v, hipStream_t stream) {
    if (n > 0) hipLaunchKernelGGL(add_kernel, dim3(nblk(n)), dim3(kT), 0, stream, src, dst, n, v);
    return stx_launch_status("add_i32");
}


int32_t stx_locate_cells(const int32_t* rows, const int32_t* off, const int32_t* beg, const int32_t* end,
                         const int32_t* co, int64_t R, const int32_t* qrow, const int32_t* qcol, int64_t nq,
                         int32_t* out_interval, int32_t* out_cell, hipStream_t stream) {
    if (R < 0 || nq < 0 || (out_cell && !co)) return stx_fail(STX_EINVAL, "locate_cells: bad arguments");
    if (nq > 0) {
        const int64_t blocks = nblk(nq) < 4096 ? nblk(nq) : 4096;
        hipLaunchKernelGGL(locate_cells_kernel, dim3((unsigned)blocks), dim3(kT), 0, stream, rows, off, beg, end, co, R,
                           qrow, qcol, nq, out_interval, out_cell);
    }
    return stx_launch_status("locate_cells");
}

}  // extern "C"
