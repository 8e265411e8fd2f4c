// Inter-level transfers onto a target subset, fused with the gather and the
// scatter of the reference's synchronisation (fields.py:570-612 restricts the
// whole fine field, gathers the refine cells, scatters them; here one pass
// reads the children and writes the coarse cells in place).
//
//   prolong_into:  dst(y, x) = src(floor(y/R), floor(x/R))           (T cells)
//   restrict_into: dst(y, x) = reduce_{dy,dx < R} src(yR+dy, xR+dx)  (T cells)
//
// T is a set of cells that lie in the destination geometry D; source cells
// missing from the source geometry S read as 0 (restrict) / leave dst as is
// (prolong).  One lane group per T row, lanes striding over the row's cells;
// each lane keeps monotone cursors into T, D and the S row(s).  Row ids may
// carry per-instance bases (level L+1 base = R * level L base), so a stack of
// independent instances is transferred by one launch.
#include "stx_common.h"
#include "stx_host.h"

using namespace stx;

namespace {

constexpr int kT = 256;

__device__ __forceinline__ int64_t find_row(const int32_t* __restrict__ rows, int64_t R, int32_t v) {
    int64_t lo = 0, hi = R;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (rows[mid] < v) lo = mid + 1; else hi = mid;
    }
    return (lo < R && rows[lo] == v) ? lo : -1;
}

struct Geom {
    const int32_t* rows;
    const int32_t* off;
    const int32_t* beg;
    const int32_t* end;
    const int32_t* co;
    int64_t R;
};

struct Cursor {
    int32_t j, stop;
};

__device__ __forceinline__ Cursor row_cursor(const Geom& g, int32_t row_id) {
    const int64_t k = find_row(g.rows, g.R, row_id);
    if (k < 0) return Cursor{0, 0};
    return Cursor{g.off[k], g.off[k + 1]};
}

// cell index of column x in the cursor's row, or -1
__device__ __forceinline__ int64_t cell_at(Cursor& c, const Geom& g, int32_t x) {
    while (c.j < c.stop && g.end[c.j] <= x) ++c.j;
    if (c.j < c.stop && g.beg[c.j] <= x) return (int64_t)g.co[c.j] + (x - g.beg[c.j]);
    return -1;
}

template <int G, typename T>
__global__ __launch_bounds__(kT) void prolong_into_kernel(Geom tg, Geom dg, Geom sg, int32_t R, const T* __restrict__ src,
                                                          T* __restrict__ dst) {
    const int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (k >= tg.R) return;
    const int gl = threadIdx.x & (G - 1);
    const int32_t t0 = tg.off[k], t1 = tg.off[k + 1];
    if (t0 >= t1) return;
    const int32_t y = tg.rows[k];
    const int64_t cs = tg.co[t0], ce = tg.co[t1];
    Cursor dc = row_cursor(dg, y);
    Cursor sc = row_cursor(sg, floordiv(y, R));
    int32_t i = t0;
    for (int64_t c = cs + gl; c < ce; c += G) {
        while (tg.co[i + 1] <= c) ++i;
        const int32_t x = tg.beg[i] + (int32_t)(c - tg.co[i]);
        const int64_t di = cell_at(dc, dg, x);
        const int64_t si = cell_at(sc, sg, floordiv(x, R));
        if (di >= 0 && si >= 0) dst[di] = src[si];
    }
}

template <int G, typename T, typename Acc>
__global__ __launch_bounds__(kT) void restrict_into_kernel(Geom tg, Geom dg, Geom sg, int32_t R, Acc norm,
                                                           const T* __restrict__ src, T* __restrict__ dst) {
    const int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (k >= tg.R) return;
    const int gl = threadIdx.x & (G - 1);
    const int32_t t0 = tg.off[k], t1 = tg.off[k + 1];
    if (t0 >= t1) return;
    const int32_t y = tg.rows[k];
    const int64_t cs = tg.co[t0], ce = tg.co[t1];
    Cursor dc = row_cursor(dg, y);
    constexpr int kMaxR = 8;
    Cursor sc[kMaxR];
#pragma unroll
    for (int d = 0; d < kMaxR; ++d) sc[d] = d < R ? row_cursor(sg, y * R + d) : Cursor{0, 0};
    int32_t i = t0;
    for (int64_t c = cs + gl; c < ce; c += G) {
        while (tg.co[i + 1] <= c) ++i;
        const int32_t x = tg.beg[i] + (int32_t)(c - tg.co[i]);
        Acc acc = (Acc)0;
#pragma unroll
        for (int d = 0; d < kMaxR; ++d) {
            if (d >= R) break;
            for (int32_t e = 0; e < R; ++e) {
                const int64_t si = cell_at(sc[d], sg, x * R + e);
                if (si >= 0) acc = acc + (Acc)src[si];
            }
        }
        const int64_t di = cell_at(dc, dg, x);
        if (di >= 0) dst[di] = (T)(acc * norm);
    }
}

template <int G, typename T>
void launch_prolong(const Geom& t, const Geom& d, const Geom& s, int32_t R, const void* src, void* dst, hipStream_t st) {
    const unsigned grid = (unsigned)((t.R * G + kT - 1) / kT);
    hipLaunchKernelGGL((prolong_into_kernel<G, T>), dim3(grid), dim3(kT), 0, st, t, d, s, R, (const T*)src, (T*)dst);
}

template <int G, typename T, typename Acc>
void launch_restrict(const Geom& t, const Geom& d, const Geom& s, int32_t R, Acc norm, const void* src, void* dst,
                     hipStream_t st) {
    const unsigned grid = (unsigned)((t.R * G + kT - 1) / kT);
    hipLaunchKernelGGL((restrict_into_kernel<G, T, Acc>), dim3(grid), dim3(kT), 0, st, t, d, s, R, norm, (const T*)src,
                       (T*)dst);
}

}  // namespace

extern "C" {

// Geometry arguments: rows, off, beg, end, co (cell offsets), R (row count).
// dtype 4: float (restrict accumulates in double when acc64 != 0, else
// float); 8: double.  mean = sum * (1/R^2).
int32_t stx_transfer(int32_t mode, int32_t dtype, int32_t acc64, int32_t group, int32_t ratio, const int32_t* t_rows,
                     const int32_t* t_off, const int32_t* t_beg, const int32_t* t_end, const int32_t* t_co, int64_t t_R,
                     const int32_t* d_rows, const int32_t* d_off, const int32_t* d_beg, const int32_t* d_end,
                     const int32_t* d_co, int64_t d_R, const int32_t* s_rows, const int32_t* s_off, const int32_t* s_beg,
                     const int32_t* s_end, const int32_t* s_co, int64_t s_R, const void* src, void* dst,
                     hipStream_t stream) {
    if (ratio < 1 || ratio > 8) return stx_fail(STX_EINVAL, "transfer: ratio must be in [1, 8]");
    if (t_R <= 0) return STX_OK;
    Geom t{t_rows, t_off, t_beg, t_end, t_co, t_R};
    Geom d{d_rows, d_off, d_beg, d_end, d_co, d_R};
    Geom s{s_rows, s_off, s_beg, s_end, s_co, s_R};
    const int G = group >= 64 ? 64 : group >= 16 ? 16 : 4;
#define STX_G(CALL) \
    if (G == 64) { CALL(64); } else if (G == 16) { CALL(16); } else { CALL(4); }
    if (mode == 0) {
        if (dtype == 4) {
#define P4(g) launch_prolong<g, float>(t, d, s, ratio, src, dst, stream)
            STX_G(P4)
#undef P4
        } else if (dtype == 8) {
#define P8(g) launch_prolong<g, double>(t, d, s, ratio, src, dst, stream)
            STX_G(P8)
#undef P8
        } else {
            return stx_fail(STX_EINVAL, "transfer: dtype must be 4 or 8");
        }
    } else if (mode == 1) {
        const double normd = 1.0 / (double)(ratio * ratio);
        if (dtype == 4 && acc64) {
#define R4D(g) launch_restrict<g, float, double>(t, d, s, ratio, normd, src, dst, stream)
            STX_G(R4D)
#undef R4D
        } else if (dtype == 4) {
            const float normf = (float)normd;
#define R4(g) launch_restrict<g, float, float>(t, d, s, ratio, normf, src, dst, stream)
            STX_G(R4)
#undef R4
        } else if (dtype == 8) {
#define R8(g) launch_restrict<g, double, double>(t, d, s, ratio, normd, src, dst, stream)
            STX_G(R8)
#undef R8
        } else {
            return stx_fail(STX_EINVAL, "transfer: dtype must be 4 or 8");
        }
    } else {
        return stx_fail(STX_EINVAL, "transfer: mode must be 0 (prolong) or 1 (restrict)");
    }
#undef STX_G
    return stx_launch_status("transfer");
}

}  // extern "C"
