// Field kernels on interval geometries: prolongation / restriction between
// levels, the first-order upwind stencil, interval subset gather / scatter,
// value-run extraction, FP64 mass and threshold tagging.
//
// Layout: an IntervalField stores one value per cell, row by row and interval
// by interval (interval_field.py:19-89); `co` is the exclusive scan of the
// interval lengths (cell offsets), so the cells of one row are contiguous.
// The kernels therefore work per ROW with a G-lane group striding over the
// row's cells; a lane keeps monotone cursors (own interval, interval of the
// row above / below) so each cell costs O(1) index loads.
#include <math.h>

#include "stx_common.h"
#include "stx_fused.h"
#include "stx_host.h"

using namespace stx;

namespace {

constexpr int kT = 256;

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

template <int G>
__device__ __forceinline__ int64_t group_index() {
    return ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
}

// ---------------------------------------------------------------------------
// Prolongation (multilevel.py:365-407, kernels.py:384-455) of a coarse field
// whose rows all hold intervals (empty rows dropped): fine row (k, d) holds
// coarse row k's intervals scaled by R, so fine cell (k, d, R*q + dx) copies
// coarse cell q of row k.
template <int G, typename T>
__global__ __launch_bounds__(kT) void prolong_field_kernel(const int32_t* __restrict__ c_off, int64_t K,
                                                           const int32_t* __restrict__ c_co, const T* __restrict__ cv,
                                                           int32_t R, T* __restrict__ fv) {
    const int64_t k = group_index<G>();
    if (k >= K) return;
    const int gl = threadIdx.x & (G - 1);
    const int64_t cs = c_co[c_off[k]];
    const int64_t cw = (int64_t)c_co[c_off[k + 1]] - cs;
    const int64_t fbase = cs * R * R;
    for (int64_t q = gl; q < cw; q += G) {
        const T v = cv[cs + q];
        for (int32_t d = 0; d < R; ++d) {
            T* dst = fv + fbase + d * R * cw + (int64_t)R * q;
            for (int32_t dx = 0; dx < R; ++dx) dst[dx] = v;
        }
    }
}

// Restriction (multilevel.py:410-479, kernels.py:458-598) for an ALIGNED fine
// field: coarse row k (cells [ccs, ccs+ccw)) has fine rows k*R + d whose cells
// start at fcs(d) = f_co[f_off[k*R + d]] and hold R*ccw cells in the same
// interval order.  Children are visited dy-major then dx, as the reference.
// reducer: 0 mean, 1 sum, 2 min, 3 max.  Acc is the accumulation type.
template <int G, typename Tin, typename Acc, typename Tout>
__global__ __launch_bounds__(kT) void restrict_field_kernel(const int32_t* __restrict__ c_off, int64_t K,
                                                            const int32_t* __restrict__ c_co,
                                                            const int32_t* __restrict__ f_off,
                                                            const int32_t* __restrict__ f_co, const Tin* __restrict__ fv,
                                                            int32_t R, int32_t reducer, Acc norm, Tout* __restrict__ cv) {
    const int64_t k = group_index<G>();
    if (k >= K) return;
    const int gl = threadIdx.x & (G - 1);
    const int64_t ccs = c_co[c_off[k]];
    const int64_t ccw = (int64_t)c_co[c_off[k + 1]] - ccs;
    for (int64_t q = gl; q < ccw; q += G) {
        Acc acc = (Acc)0, mn = (Acc)0, mx = (Acc)0;
        bool have = false;
        for (int32_t dy = 0; dy < R; ++dy) {
            const int64_t fcs = f_co[f_off[k * R + dy]];
            const Tin* src = fv + fcs + (int64_t)R * q;
            for (int32_t dx = 0; dx < R; ++dx) {
                const Acc v = (Acc)src[dx];
                if (reducer == 2) {
                    if (!have || v < mn) mn = v;
                } else if (reducer == 3) {
                    if (!have || v > mx) mx = v;
                } else {
                    acc = acc + v;
                }
                have = true;
            }
        }
        Acc r = reducer == 0 ? acc * norm : reducer == 1 ? acc : reducer == 2 ? mn : mx;
        cv[ccs + q] = (Tout)r;
    }
}

// Alignment check of restrict_field (multilevel.py:169-210, :433-451): the
// fine rows must be exactly {c*R + d} for every coarse row c, and the fine
// interval of ordinal m in row c*R+d must be [b*R, e*R) of coarse interval
// (c, m).  Sets STX_FLAG_ALIGN.
__global__ void restrict_align_kernel(const int32_t* __restrict__ c_rows, const int32_t* __restrict__ c_off, int64_t K,
                                      const int32_t* __restrict__ c_beg, const int32_t* __restrict__ c_end,
                                      const int32_t* __restrict__ f_rows, const int32_t* __restrict__ f_off,
                                      const int32_t* __restrict__ f_beg, const int32_t* __restrict__ f_end, int32_t R,
                                      uint32_t* status) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    bool bad = false;
    const int32_t n = c_off[k + 1] - c_off[k];
    for (int32_t d = 0; d < R && !bad; ++d) {
        const int64_t fr = k * R + d;
        if (f_rows[fr] != c_rows[k] * R + d) { bad = true; break; }
        if (f_off[fr + 1] - f_off[fr] < n) { bad = true; break; }
        for (int32_t m = 0; m < n; ++m) {
            const int32_t ci = c_off[k] + m, fi = f_off[fr] + m;
            if (f_beg[fi] != c_beg[ci] * R || f_end[fi] != c_end[ci] * R) { bad = true; break; }
        }
    }
    if (bad) atomicOr(status, STX_FLAG_ALIGN);
}

// ---------------------------------------------------------------------------
// Upwind stencil (interval_stencil.py:16-113): zero exterior, one group per
// row.  Row ordinals of the rows above / below come from the sorted row ids.
template <typename T>
struct RowCursor {
    int32_t j, stop;  // current interval index, row end
};

template <typename T>
__device__ __forceinline__ T sample_advance(RowCursor<T>& cur, int32_t x, const int32_t* __restrict__ beg,
                                            const int32_t* __restrict__ end, const int32_t* __restrict__ co,
                                            const T* __restrict__ u) {
    while (cur.j < cur.stop && end[cur.j] <= x) ++cur.j;
    if (cur.j < cur.stop && beg[cur.j] <= x) return u[co[cur.j] + (x - beg[cur.j])];
    return (T)0;
}

// BATCH: every row belongs to an instance (row_inst[k]) with its own width
// inst_w[i] and coefficients inst_coef[5*i .. 5*i+4] = a, b, dt, dx, dy; row
// ids of different instances are separated by gaps, so the row above / below
// of a row is found by id equality alone.  An instance with dx == 0 is
// retired: its rows are skipped (their output cells are not written).
template <int G, typename T, bool BATCH>
__global__ __launch_bounds__(kT) void upwind_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ off,
                                                    int64_t R, const int32_t* __restrict__ beg,
                                                    const int32_t* __restrict__ end, const int32_t* __restrict__ co,
                                                    const T* __restrict__ u, T* __restrict__ out, int32_t width,
                                                    int32_t height, T a, T b, T dt, T dx, T dy,
                                                    const int32_t* __restrict__ row_inst,
                                                    const int32_t* __restrict__ inst_w,
                                                    const T* __restrict__ inst_coef) {
    const int64_t k = group_index<G>();
    if (k >= R) return;
    const int gl = threadIdx.x & (G - 1);
    if (BATCH) {
        const int32_t inst = row_inst[k];
        width = inst_w[inst];
        height = INT32_MAX;
        a = inst_coef[5 * inst + 0];
        b = inst_coef[5 * inst + 1];
        dt = inst_coef[5 * inst + 2];
        dx = inst_coef[5 * inst + 3];
        dy = inst_coef[5 * inst + 4];
        if (dx == (T)0) return;
    }
    const int32_t o0 = off[k], o1 = off[k + 1];
    if (o0 >= o1) return;
    const int32_t rid = rows[k];
    const int64_t cs = co[o0];
    const int64_t cw = (int64_t)co[o1] - cs;
    // neighbour rows (rows are strictly increasing)
    const bool has_dn = k > 0 && rows[k - 1] == rid - 1 && rid - 1 >= 0;
    const bool has_up = k + 1 < R && rows[k + 1] == rid + 1 && rid + 1 < height;
    RowCursor<T> dn{has_dn ? off[k - 1] : 0, has_dn ? off[k] : 0};
    RowCursor<T> up{has_up ? off[k + 1] : 0, has_up ? off[k + 2] : 0};
    int32_t i = o0;
    for (int64_t q = gl; q < cw; q += G) {
        const int64_t c = cs + q;
        while (co[i + 1] <= c) ++i;
        const int32_t x = beg[i] + (int32_t)(c - co[i]);
        const T center = u[c];
        T left = (T)0, right = (T)0;
        if (x - 1 >= 0 && x - 1 < width) {
            if (c > co[i]) left = u[c - 1];
            else if (i > o0 && end[i - 1] == x) left = u[c - 1];
        }
        if (x + 1 >= 0 && x + 1 < width) {
            if (c + 1 < co[i + 1]) right = u[c + 1];
            else if (i + 1 < o1 && beg[i + 1] == x + 1) right = u[c + 1];
        }
        T down = (T)0, upv = (T)0;
        if (x >= 0 && x < width) {
            down = sample_advance(dn, x, beg, end, co, u);
            upv = sample_advance(up, x, beg, end, co, u);
        }
        const T dudx = (a >= (T)0) ? (center - left) / dx : (right - center) / dx;
        const T dudy = (b >= (T)0) ? (center - down) / dy : (upv - center) / dy;
        const T t1 = a * dudx;
        const T t2 = b * dudy;
        out[c] = center - dt * (t1 + t2);
    }
}

// Row-id range check for the stencil (interval_stencil.py:183-191).
__global__ void rows_in_range_kernel(const int32_t* __restrict__ rows, int64_t R, int32_t height, uint32_t* status) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < R && (rows[k] < 0 || rows[k] >= height)) atomicOr(status, STX_FLAG_ROWS);
}

// ---------------------------------------------------------------------------
// FP64 mass: deterministic two-level sum (fixed chunking, fixed tree).
constexpr int kMassChunk = 4096;

template <typename T>
__global__ __launch_bounds__(kT) void mass_partial_kernel(const T* __restrict__ u, int64_t n, double* __restrict__ part) {
    __shared__ double s[kT];
    const int64_t base = (int64_t)blockIdx.x * kMassChunk;
    double acc = 0.0;
    for (int k = threadIdx.x; k < kMassChunk; k += kT) {
        const int64_t idx = base + k;
        if (idx < n) acc += (double)u[idx];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kT) void mass_final_kernel(const double* __restrict__ part, int64_t m, double scale,
                                                        double* __restrict__ out) {
    __shared__ double s[kT];
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < m; k += kT) acc += part[k];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0] * scale;
}

// ---------------------------------------------------------------------------
// Interval subset gather / scatter (fields.py:127-294) for subset <= super:
// a group per subset interval walks the super intervals of the same row
// that cover it.  sub_srcrow: index of the subset row in the super rows.
template <int G, typename T, bool GATHER>
__global__ __launch_bounds__(kT) void subset_copy_kernel(const int32_t* __restrict__ sub_beg,
                                                         const int32_t* __restrict__ sub_end,
                                                         const int32_t* __restrict__ sub_co,
                                                         const int32_t* __restrict__ sub_srcrow, int64_t nsub,
                                                         const int32_t* __restrict__ sup_off,
                                                         const int32_t* __restrict__ sup_beg,
                                                         const int32_t* __restrict__ sup_end,
                                                         const int32_t* __restrict__ sup_co, T* __restrict__ sup_v,
                                                         T* __restrict__ sub_v) {
    const int64_t i = group_index<G>();
    if (i >= nsub) return;
    const int gl = threadIdx.x & (G - 1);
    const int32_t sr = sub_srcrow[i];
    if (sr < 0) return;
    const int32_t start = sub_beg[i], stop = sub_end[i];
    if (stop <= start) return;
    int32_t lo = sup_off[sr], hi = sup_off[sr + 1];
    // first super interval with end > start
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (sup_end[mid] <= start) lo = mid + 1; else hi = mid;
    }
    const int32_t rstop = sup_off[sr + 1];
    const int64_t dbase = sub_co[i];
    for (int32_t j = lo; j < rstop; ++j) {
        const int32_t b = sup_beg[j], e = sup_end[j];
        if (b >= stop) break;
        const int32_t s0 = start > b ? start : b;
        const int32_t s1 = stop < e ? stop : e;
        for (int32_t x = s0 + gl; x < s1; x += G) {
            const int64_t si = (int64_t)sup_co[j] + (x - b);
            const int64_t di = dbase + (x - start);
            if (GATHER) sub_v[di] = sup_v[si];
            else sup_v[si] = sub_v[di];
        }
    }
}

// Subset rows located in the super rows (-1 when absent -> STX_FLAG_ROWS).
__global__ void locate_rows_kernel(const int32_t* __restrict__ sub_rows, const int32_t* __restrict__ sub_off, int64_t Rs,
                                   const int32_t* __restrict__ sup_rows, int64_t Rp, int32_t* __restrict__ srcrow_per_iv,
                                   uint32_t* status) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= Rs) return;
    const int32_t v = sub_rows[r];
    int64_t lo = 0, hi = Rp;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (sup_rows[mid] < v) lo = mid + 1; else hi = mid;
    }
    const bool found = lo < Rp && sup_rows[lo] == v;
    const int32_t n0 = sub_off[r], n1 = sub_off[r + 1];
    if (!found && n1 > n0) atomicOr(status, STX_FLAG_ROWS);
    for (int32_t k = n0; k < n1; ++k) srcrow_per_iv[k] = found ? (int32_t)lo : -1;
}

template <typename T>
__global__ void fill_value_kernel(T* __restrict__ p, int64_t n, T v) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

__global__ void cell_offsets_kernel(const int32_t* __restrict__ beg, const int32_t* __restrict__ end, int64_t n,
                                    int32_t* __restrict__ len) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) len[k] = end[k] - beg[k];
}

// ---------------------------------------------------------------------------
// Threshold tagging (regrid.py:20-124, 175-274): maximal runs of v >= t on a
// dense W x H field, one wave per row, 64 columns per ballot; optionally the
// value is the gradient magnitude (central differences x0.5 inside,
// one-sided at the edges, sqrtf(gx*gx + gy*gy)) computed on the fly.
template <bool GRAD>
struct ThresholdRow {
    const float* data;
    int32_t width, height;
    float t;

    __device__ __forceinline__ float value(int64_t y, int32_t x) const {
        const int64_t idx = y * width + x;
        if (!GRAD) return data[idx];
        const float c = data[idx];
        float gx = 0.0f, gy = 0.0f;
        if (width > 1) {
            if (x == 0) gx = data[idx + 1] - c;
            else if (x == width - 1) gx = c - data[idx - 1];
            else gx = (data[idx + 1] - data[idx - 1]) * 0.5f;
        }
        if (height > 1) {
            if (y == 0) gy = data[idx + width] - c;
            else if (y == height - 1) gy = c - data[idx - width];
            else gy = (data[idx + width] - data[idx - width]) * 0.5f;
        }
        const float gx2 = gx * gx;
        const float gy2 = gy * gy;
        return sqrtf(gx2 + gy2);
    }

    template <int G, bool WRITE>
    __device__ int32_t run(int64_t y, int32_t* ob, int32_t* oe, int64_t base) const {
        const int gl = lane_id() & (G - 1);
        int32_t starts = 0, stops = 0;
        bool prev = false;
        for (int32_t x0 = 0; x0 < width; x0 += G) {
            const int32_t x = x0 + gl;
            const bool on = x < width && value(y, x) >= t;
            const uint64_t m = group_ballot<G>(on);
            const uint64_t pm = (m << 1) | (prev ? 1ull : 0ull);
            const int32_t nvalid = width - x0 < G ? width - x0 : G;
            const uint64_t lim = nvalid >= 64 ? ~0ull : ((1ull << nvalid) - 1ull);
            // only columns inside the row: a run reaching the row end is
            // closed once, after the loop (not also at column `width`)
            const uint64_t st = m & ~pm & lim;
            const uint64_t sp = pm & ~m & lim;  // a run ended just before this column
            if (WRITE) {
                if ((st >> gl) & 1ull) ob[base + starts + rank_below(st, gl)] = x;
                if ((sp >> gl) & 1ull) oe[base + stops + rank_below(sp, gl)] = x;
            }
            starts += __popcll(st);
            stops += __popcll(sp);
            prev = (m >> (nvalid - 1)) & 1ull;
        }
        if (WRITE && prev && gl == 0) oe[base + stops] = width;
        return starts;
    }
};

// Dense gradient magnitude (regrid.py:175-223), same arithmetic order as the
// threshold variant.
__global__ void gradient_kernel(const float* __restrict__ data, int32_t width, int32_t height, float* __restrict__ out) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)width * height) return;
    ThresholdRow<true> g{data, width, height, 0.0f};
    out[idx] = g.value(idx / width, (int32_t)(idx % width));
}

// Runs of cells equal to `target` inside each interval (fields.py:41-100):
// a run never crosses an interval boundary; rows are the field's rows.
template <typename T>
struct ValueRunsRow {
    const int32_t* off;
    const int32_t* beg;
    const int32_t* co;
    const T* v;
    T target;

    template <int G, bool WRITE>
    __device__ int32_t run(int64_t k, int32_t* ob, int32_t* oe, int64_t base) const {
        const int gl = lane_id() & (G - 1);
        const int32_t o0 = off[k], o1 = off[k + 1];
        if (o0 >= o1) return 0;
        const int64_t cs = co[o0], ce = co[o1];
        int32_t starts = 0, ends = 0;
        int32_t i = o0;
        for (int64_t c0 = cs; c0 < ce; c0 += G) {
            const int64_t c = c0 + gl;
            bool st = false, en = false;
            int32_t x = 0;
            if (c < ce) {
                while (co[i + 1] <= c) ++i;
                x = beg[i] + (int32_t)(c - co[i]);
                const bool m = v[c] == target;
                const bool pm = c > co[i] && v[c - 1] == target;
                const bool nm = c + 1 < co[i + 1] && v[c + 1] == target;
                st = m && !pm;
                en = m && !nm;
            }
            const uint64_t ms = group_ballot<G>(st);
            const uint64_t me = group_ballot<G>(en);
            if (WRITE) {
                if (st) ob[base + starts + rank_below(ms, gl)] = x;
                if (en) oe[base + ends + rank_below(me, gl)] = x + 1;
            }
            starts += __popcll(ms);
            ends += __popcll(me);
        }
        return starts;
    }
};

// Gradient tagging on an interval field (the sparse generalisation of
// regrid.py:175-223 + 20-124): per cell, central differences x0.5 when both
// neighbours along an axis are in the set, one-sided when only one is, 0
// when none (on a dense W x H field this is exactly the reference formula);
// output: maximal runs of tagged cells that are adjacent in x, per row.
template <typename T>
struct GradTagRow {
    const int32_t* rows;
    const int32_t* off;
    const int32_t* beg;
    const int32_t* end;
    const int32_t* co;
    int64_t R;
    const T* u;
    T t;

    __device__ __forceinline__ bool sample(RowCursor<T>& cur, int32_t x, T& v) const {
        while (cur.j < cur.stop && end[cur.j] <= x) ++cur.j;
        if (cur.j < cur.stop && beg[cur.j] <= x) {
            v = u[co[cur.j] + (x - beg[cur.j])];
            return true;
        }
        return false;
    }

    template <int G, bool WRITE>
    __device__ int32_t run(int64_t k, int32_t* ob, int32_t* oe, int64_t base) const {
        const int gl = lane_id() & (G - 1);
        const int32_t o0 = off[k], o1 = off[k + 1];
        if (o0 >= o1) return 0;
        const int32_t rid = rows[k];
        const bool has_dn = k > 0 && rows[k - 1] == rid - 1;
        const bool has_up = k + 1 < R && rows[k + 1] == rid + 1;
        RowCursor<T> dn{has_dn ? off[k - 1] : 0, has_dn ? off[k] : 0};
        RowCursor<T> up{has_up ? off[k + 1] : 0, has_up ? off[k + 2] : 0};
        const int64_t cs = co[o0], ce = co[o1];
        int32_t starts = 0, ends = 0;
        int32_t i = o0;
        bool carry_tag = false;  // tag of the last cell of the previous chunk
        int32_t carry_x = 0;
        for (int64_t c0 = cs; c0 < ce; c0 += G) {
            const int64_t c = c0 + gl;
            bool tag = false;
            int32_t x = 0;
            if (c < ce) {
                while (co[i + 1] <= c) ++i;
                x = beg[i] + (int32_t)(c - co[i]);
                const T cv = u[c];
                const bool hl = c > co[i] || (i > o0 && end[i - 1] == x);
                const bool hr = c + 1 < co[i + 1] || (i + 1 < o1 && beg[i + 1] == x + 1);
                T gx = (T)0, gy = (T)0;
                if (hl && hr) gx = (u[c + 1] - u[c - 1]) * (T)0.5;
                else if (hr) gx = u[c + 1] - cv;
                else if (hl) gx = cv - u[c - 1];
                T dv = (T)0, uv = (T)0;
                const bool hd = sample(dn, x, dv);
                const bool hu = sample(up, x, uv);
                if (hd && hu) gy = (uv - dv) * (T)0.5;
                else if (hu) gy = uv - cv;
                else if (hd) gy = cv - dv;
                const T gx2 = gx * gx;
                const T gy2 = gy * gy;
                tag = sqrt(gx2 + gy2) >= t;
            }
            // previous cell in row order: the previous lane, or the last
            // cell of the previous chunk
            const bool prev_tag = __shfl_up(tag ? 1 : 0, 1, G) != 0;
            const int32_t prev_x = __shfl_up(x, 1, G);
            bool p_tag = gl == 0 ? carry_tag : prev_tag;
            const int32_t p_x = gl == 0 ? carry_x : prev_x;
            if (c == cs) p_tag = false;
            const bool valid = c < ce;
            const bool adj = p_x == x - 1;
            const bool st = valid && tag && !(p_tag && adj);
            const bool en = valid && p_tag && !(tag && adj);  // a run ended at p_x + 1
            const uint64_t ms = group_ballot<G>(st);
            const uint64_t me = group_ballot<G>(en);
            if (WRITE) {
                if (st) ob[base + starts + rank_below(ms, gl)] = x;
                if (en) oe[base + ends + rank_below(me, gl)] = p_x + 1;
            }
            starts += __popcll(ms);
            ends += __popcll(me);
            const int last = (int)((ce - c0 < G ? ce - c0 : G) - 1);
            carry_tag = __shfl(tag ? 1 : 0, last, G) != 0;
            carry_x = __shfl(x, last, G);
        }
        if (WRITE && carry_tag && gl == 0) oe[base + ends] = carry_x + 1;
        return starts;
    }
};

}  // namespace

// ===========================================================================

template <typename T>
static int32_t launch_prolong(int32_t group, const int32_t* c_off, int64_t K, const int32_t* c_co, const T* cv, int32_t R,
                              T* fv, hipStream_t stream) {
    if (K <= 0) return STX_OK;
    const int G = group >= 64 ? 64 : group >= 16 ? 16 : 4;
    const unsigned grid = (unsigned)((K * G + kT - 1) / kT);
    if (G == 64) hipLaunchKernelGGL((prolong_field_kernel<64, T>), dim3(grid), dim3(kT), 0, stream, c_off, K, c_co, cv, R, fv);
    else if (G == 16) hipLaunchKernelGGL((prolong_field_kernel<16, T>), dim3(grid), dim3(kT), 0, stream, c_off, K, c_co, cv, R, fv);
    else hipLaunchKernelGGL((prolong_field_kernel<4, T>), dim3(grid), dim3(kT), 0, stream, c_off, K, c_co, cv, R, fv);
    return stx_launch_status("prolong_field");
}

template <int G, typename Tin, typename Acc, typename Tout>
static void launch_restrict_g(unsigned grid, hipStream_t stream, const int32_t* c_off, int64_t K, const int32_t* c_co,
                              const int32_t* f_off, const int32_t* f_co, const void* fv, int32_t R, int32_t reducer,
                              Acc norm, void* cv) {
    hipLaunchKernelGGL((restrict_field_kernel<G, Tin, Acc, Tout>), dim3(grid), dim3(kT), 0, stream, c_off, K, c_co, f_off,
                       f_co, (const Tin*)fv, R, reducer, norm, (Tout*)cv);
}

template <typename Tin, typename Acc, typename Tout>
static void launch_restrict(int32_t group, hipStream_t stream, const int32_t* c_off, int64_t K, const int32_t* c_co,
                            const int32_t* f_off, const int32_t* f_co, const void* fv, int32_t R, int32_t reducer,
                            Acc norm, void* cv) {
    const int G = group >= 64 ? 64 : group >= 16 ? 16 : 4;
    const unsigned grid = (unsigned)((K * G + kT - 1) / kT);
    if (G == 64) launch_restrict_g<64, Tin, Acc, Tout>(grid, stream, c_off, K, c_co, f_off, f_co, fv, R, reducer, norm, cv);
    else if (G == 16) launch_restrict_g<16, Tin, Acc, Tout>(grid, stream, c_off, K, c_co, f_off, f_co, fv, R, reducer, norm, cv);
    else launch_restrict_g<4, Tin, Acc, Tout>(grid, stream, c_off, K, c_co, f_off, f_co, fv, R, reducer, norm, cv);
}

template <int G, typename T>
static void launch_upwind_g(hipStream_t stream, const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg,
                            const int32_t* end, const int32_t* co, const void* u, void* out, int32_t W, int32_t H,
                            double a, double b, double dt, double dx, double dy) {
    const unsigned grid = (unsigned)((R * G + kT - 1) / kT);
    hipLaunchKernelGGL((upwind_kernel<G, T, false>), dim3(grid), dim3(kT), 0, stream, rows, off, R, beg, end, co,
                       (const T*)u, (T*)out, W, H, (T)a, (T)b, (T)dt, (T)dx, (T)dy, nullptr, nullptr, nullptr);
}

template <int G, typename T>
static void launch_upwind_batched_g(hipStream_t stream, const int32_t* rows, const int32_t* off, int64_t R,
                                    const int32_t* beg, const int32_t* end, const int32_t* co, const void* u, void* out,
                                    const int32_t* row_inst, const int32_t* inst_w, const void* inst_coef) {
    const unsigned grid = (unsigned)((R * G + kT - 1) / kT);
    hipLaunchKernelGGL((upwind_kernel<G, T, true>), dim3(grid), dim3(kT), 0, stream, rows, off, R, beg, end, co,
                       (const T*)u, (T*)out, 0, 0, (T)0, (T)0, (T)0, (T)1, (T)1, row_inst, inst_w,
                       (const T*)inst_coef);
}

template <typename T, bool GATHER>
static void launch_subset(int32_t group, hipStream_t stream, const int32_t* sb, const int32_t* se, const int32_t* sco,
                          const int32_t* srow, int64_t n, const int32_t* poff, const int32_t* pb, const int32_t* pe,
                          const int32_t* pco, void* pv, void* sv) {
    const int G = group >= 64 ? 64 : group >= 16 ? 16 : 4;
    const unsigned grid = (unsigned)((n * G + kT - 1) / kT);
    if (G == 64)
        hipLaunchKernelGGL((subset_copy_kernel<64, T, GATHER>), dim3(grid), dim3(kT), 0, stream, sb, se, sco, srow, n, poff,
                           pb, pe, pco, (T*)pv, (T*)sv);
    else if (G == 16)
        hipLaunchKernelGGL((subset_copy_kernel<16, T, GATHER>), dim3(grid), dim3(kT), 0, stream, sb, se, sco, srow, n, poff,
                           pb, pe, pco, (T*)pv, (T*)sv);
    else
        hipLaunchKernelGGL((subset_copy_kernel<4, T, GATHER>), dim3(grid), dim3(kT), 0, stream, sb, se, sco, srow, n, poff,
                           pb, pe, pco, (T*)pv, (T*)sv);
}

extern "C" {
// dtype: 4 = float32, 8 = float64
int32_t stx_prolong_field(int32_t dtype, int32_t group, const int32_t* c_off, int64_t K, const int32_t* c_co,
                          const void* cv, int32_t ratio, void* fv, hipStream_t stream) {
    if (ratio < 1 || K < 0) return stx_fail(STX_EINVAL, "prolong_field: bad arguments");
    if (dtype == 4) return launch_prolong<float>(group, c_off, K, c_co, (const float*)cv, ratio, (float*)fv, stream);
    if (dtype == 8) return launch_prolong<double>(group, c_off, K, c_co, (const double*)cv, ratio, (double*)fv, stream);
    return stx_fail(STX_EINVAL, "prolong_field: dtype must be 4 or 8");
}

int32_t stx_restrict_align(const int32_t* c_rows, const int32_t* c_off, int64_t K, const int32_t* c_beg,
                           const int32_t* c_end, const int32_t* f_rows, const int32_t* f_off, const int32_t* f_beg,
                           const int32_t* f_end, int32_t ratio, uint32_t* status, hipStream_t stream) {
    if (K <= 0) return STX_OK;
    hipLaunchKernelGGL(restrict_align_kernel, dim3(nblk(K)), dim3(kT), 0, stream, c_rows, c_off, K, c_beg, c_end, f_rows,
                       f_off, f_beg, f_end, ratio, status);
    return stx_launch_status("restrict_align");
}

// kind: 0 = f32 in / f32 accumulate / f32 out (the reference kernel),
//       1 = f32 in / f64 accumulate / f32 out, 2 = f64 throughout.
int32_t stx_restrict_field(int32_t kind, int32_t group, const int32_t* c_off, int64_t K, const int32_t* c_co,
                           const int32_t* f_off, const int32_t* f_co, const void* fv, int32_t ratio, int32_t reducer,
                           void* cv, hipStream_t stream) {
    if (ratio < 1 || K < 0 || reducer < 0 || reducer > 3) return stx_fail(STX_EINVAL, "restrict_field: bad arguments");
    if (K == 0) return STX_OK;
    const int32_t samples = ratio * ratio;
    if (kind == 0) {
        const float norm = reducer == 0 ? (float)(1.0 / samples) : 1.0f;
        launch_restrict<float, float, float>(group, stream, c_off, K, c_co, f_off, f_co, fv, ratio, reducer, norm, cv);
    } else if (kind == 1) {
        const double norm = reducer == 0 ? 1.0 / samples : 1.0;
        launch_restrict<float, double, float>(group, stream, c_off, K, c_co, f_off, f_co, fv, ratio, reducer, norm, cv);
    } else if (kind == 2) {
        const double norm = reducer == 0 ? 1.0 / samples : 1.0;
        launch_restrict<double, double, double>(group, stream, c_off, K, c_co, f_off, f_co, fv, ratio, reducer, norm, cv);
    } else {
        return stx_fail(STX_EINVAL, "restrict_field: kind must be 0, 1 or 2");
    }
    return stx_launch_status("restrict_field");
}

int32_t stx_rows_in_range(const int32_t* rows, int64_t R, int32_t height, uint32_t* status, hipStream_t stream) {
    if (R <= 0) return STX_OK;
    hipLaunchKernelGGL(rows_in_range_kernel, dim3(nblk(R)), dim3(kT), 0, stream, rows, R, height, status);
    return stx_launch_status("rows_in_range");
}

// dtype 4: all arithmetic in float (a, b, dt, dx, dy rounded to float, as
// the reference passes np.float32 scalars); dtype 8: double.
int32_t stx_upwind(int32_t dtype, int32_t group, const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg,
                   const int32_t* end, const int32_t* co, const void* u, void* out, int32_t width, int32_t height,
                   double a, double b, double dt, double dx, double dy, hipStream_t stream) {
    if (width <= 0 || height <= 0) return stx_fail(STX_EINVAL, "upwind: width and height must be positive");
    if (R <= 0) return STX_OK;
    const int G = group >= 64 ? 64 : group >= 16 ? 16 : 4;
    if (dtype == 4) {
        if (G == 64) launch_upwind_g<64, float>(stream, rows, off, R, beg, end, co, u, out, width, height, a, b, dt, dx, dy);
        else if (G == 16) launch_upwind_g<16, float>(stream, rows, off, R, beg, end, co, u, out, width, height, a, b, dt, dx, dy);
        else launch_upwind_g<4, float>(stream, rows, off, R, beg, end, co, u, out, width, height, a, b, dt, dx, dy);
    } else if (dtype == 8) {
        if (G == 64) launch_upwind_g<64, double>(stream, rows, off, R, beg, end, co, u, out, width, height, a, b, dt, dx, dy);
        else if (G == 16) launch_upwind_g<16, double>(stream, rows, off, R, beg, end, co, u, out, width, height, a, b, dt, dx, dy);
        else launch_upwind_g<4, double>(stream, rows, off, R, beg, end, co, u, out, width, height, a, b, dt, dx, dy);
    } else {
        return stx_fail(STX_EINVAL, "upwind: dtype must be 4 or 8");
    }
    return stx_launch_status("upwind");
}

// Batched upwind over a stack of instances (see upwind_kernel, BATCH);
// inst_coef holds 5 values of the field type per instance.
int32_t stx_upwind_batched(int32_t dtype, int32_t group, const int32_t* rows, const int32_t* off, int64_t R,
                           const int32_t* beg, const int32_t* end, const int32_t* co, const void* u, void* out,
                           const int32_t* row_inst, const int32_t* inst_w, const void* inst_coef, hipStream_t stream) {
    if (R <= 0) return STX_OK;
    const int G = group >= 64 ? 64 : group >= 16 ? 16 : 4;
    if (dtype == 4) {
        if (G == 64) launch_upwind_batched_g<64, float>(stream, rows, off, R, beg, end, co, u, out, row_inst, inst_w, inst_coef);
        else if (G == 16) launch_upwind_batched_g<16, float>(stream, rows, off, R, beg, end, co, u, out, row_inst, inst_w, inst_coef);
        else launch_upwind_batched_g<4, float>(stream, rows, off, R, beg, end, co, u, out, row_inst, inst_w, inst_coef);
    } else if (dtype == 8) {
        if (G == 64) launch_upwind_batched_g<64, double>(stream, rows, off, R, beg, end, co, u, out, row_inst, inst_w, inst_coef);
        else if (G == 16) launch_upwind_batched_g<16, double>(stream, rows, off, R, beg, end, co, u, out, row_inst, inst_w, inst_coef);
        else launch_upwind_batched_g<4, double>(stream, rows, off, R, beg, end, co, u, out, row_inst, inst_w, inst_coef);
    } else {
        return stx_fail(STX_EINVAL, "upwind_batched: dtype must be 4 or 8");
    }
    return stx_launch_status("upwind_batched");
}

// part: >= ceil(n / 4096) doubles.  out: one double (sum * scale).
int64_t stx_mass_parts(int64_t n) { return n > 0 ? (n + kMassChunk - 1) / kMassChunk : 1; }

int32_t stx_mass(int32_t dtype, const void* u, int64_t n, double scale, double* part, double* out, hipStream_t stream) {
    const int64_t m = stx_mass_parts(n);
    if (n > 0) {
        if (dtype == 4)
            hipLaunchKernelGGL(mass_partial_kernel<float>, dim3((unsigned)m), dim3(kT), 0, stream, (const float*)u, n, part);
        else if (dtype == 8)
            hipLaunchKernelGGL(mass_partial_kernel<double>, dim3((unsigned)m), dim3(kT), 0, stream, (const double*)u, n, part);
        else
            return stx_fail(STX_EINVAL, "mass: dtype must be 4 or 8");
    } else {
        hipLaunchKernelGGL(fill_value_kernel<double>, dim3(1), dim3(1), 0, stream, part, (int64_t)1, 0.0);
    }
    hipLaunchKernelGGL(mass_final_kernel, dim3(1), dim3(kT), 0, stream, part, m, scale, out);
    return stx_launch_status("mass");
}

int32_t stx_locate_rows(const int32_t* sub_rows, const int32_t* sub_off, int64_t Rs, const int32_t* sup_rows, int64_t Rp,
                        int32_t* srcrow_per_iv, uint32_t* status, hipStream_t stream) {
    if (Rs <= 0) return STX_OK;
    hipLaunchKernelGGL(locate_rows_kernel, dim3(nblk(Rs)), dim3(kT), 0, stream, sub_rows, sub_off, Rs, sup_rows, Rp,
                       srcrow_per_iv, status);
    return stx_launch_status("locate_rows");
}

// mode 0 gather (super -> subset), 1 scatter (subset -> super); esize 1, 4, 8.
int32_t stx_subset_copy(int32_t mode, int32_t esize, int32_t group, const int32_t* sub_beg, const int32_t* sub_end,
                        const int32_t* sub_co, const int32_t* sub_srcrow, int64_t nsub, const int32_t* sup_off,
                        const int32_t* sup_beg, const int32_t* sup_end, const int32_t* sup_co, void* sup_v, void* sub_v,
                        hipStream_t stream) {
    if (nsub <= 0) return STX_OK;
    const bool g = mode == 0;
#define STX_SUBSET(T)                                                                                                  \
    if (g) launch_subset<T, true>(group, stream, sub_beg, sub_end, sub_co, sub_srcrow, nsub, sup_off, sup_beg, sup_end, \
                                  sup_co, sup_v, sub_v);                                                               \
    else launch_subset<T, false>(group, stream, sub_beg, sub_end, sub_co, sub_srcrow, nsub, sup_off, sup_beg, sup_end,  \
                                 sup_co, sup_v, sub_v);
    if (esize == 1) { STX_SUBSET(int8_t) }
    else if (esize == 4) { STX_SUBSET(int32_t) }
    else if (esize == 8) { STX_SUBSET(int64_t) }
    else return stx_fail(STX_EINVAL, "subset_copy: element size must be 1, 4 or 8");
#undef STX_SUBSET
    return stx_launch_status("subset_copy");
}

// Fill n elements of size esize (1, 4, 8) with the bit pattern of `bits`.
int32_t stx_fill_value(int32_t esize, void* p, int64_t n, int64_t bits, hipStream_t stream) {
    if (n <= 0) return STX_OK;
    if (esize == 1) hipLaunchKernelGGL(fill_value_kernel<int8_t>, dim3(nblk(n)), dim3(kT), 0, stream, (int8_t*)p, n, (int8_t)bits);
    else if (esize == 4) hipLaunchKernelGGL(fill_value_kernel<int32_t>, dim3(nblk(n)), dim3(kT), 0, stream, (int32_t*)p, n, (int32_t)bits);
    else if (esize == 8) hipLaunchKernelGGL(fill_value_kernel<int64_t>, dim3(nblk(n)), dim3(kT), 0, stream, (int64_t*)p, n, (int64_t)bits);
    else return stx_fail(STX_EINVAL, "fill_value: element size must be 1, 4 or 8");
    return stx_launch_status("fill_value");
}

// interval lengths (for the cell-offset scan)
int32_t stx_interval_lengths(const int32_t* beg, const int32_t* end, int64_t n, int32_t* len, hipStream_t stream) {
    if (n <= 0) return STX_OK;
    hipLaunchKernelGGL(cell_offsets_kernel, dim3(nblk(n)), dim3(kT), 0, stream, beg, end, n, len);
    return stx_launch_status("interval_lengths");
}

// Threshold tagging: runs of (gradient ? |grad v| : v) >= t on a dense
// width x height float field; CSR with dense rows.  Fused count/scan/write.
int32_t stx_threshold(int32_t gradient, const float* data, int32_t width, int32_t height, float t, int32_t* out_off,
                      int32_t* out_beg, int32_t* out_end, int64_t capacity, int32_t* d_total, uint32_t* status,
                      void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    if (width <= 0 || height <= 0) return stx_fail(STX_EINVAL, "threshold: width and height must be positive");
    FusedLaunch L{nullptr, height, out_off, out_beg, out_end, capacity, d_total, status, scratch, scratch_bytes, epoch, stream};
    if (gradient) return launch_fused_rows<64>(ThresholdRow<true>{data, width, height, t}, L, "threshold_gradient");
    return launch_fused_rows<64>(ThresholdRow<false>{data, width, height, t}, L, "threshold");
}

int64_t stx_threshold_scratch_bytes(int32_t height) { return fused_rows_scratch_bytes(height, 64); }

int32_t stx_value_runs(int32_t esize, int32_t group, const int32_t* off, int64_t R, const int32_t* beg, const int32_t* co,
                       const void* values, int64_t target_bits, int32_t* out_off, int32_t* out_beg, int32_t* out_end,
                       int64_t capacity, int32_t* d_total, uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch,
                       hipStream_t stream) {
    FusedLaunch L{nullptr, R, out_off, out_beg, out_end, capacity, d_total, status, scratch, scratch_bytes, epoch, stream};
    if (esize == 1)
        return launch_fused_rows_g(group, ValueRunsRow<int8_t>{off, beg, co, (const int8_t*)values, (int8_t)target_bits},
                                   L, "value_runs");
    if (esize == 4)
        return launch_fused_rows_g(group, ValueRunsRow<int32_t>{off, beg, co, (const int32_t*)values, (int32_t)target_bits},
                                   L, "value_runs");
    return stx_fail(STX_EINVAL, "value_runs: element size must be 1 or 4");
}

// Gradient tagging on an interval field (GradTagRow); dtype 4 / 8.
int32_t stx_grad_tag(int32_t dtype, int32_t group, const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg,
                     const int32_t* end, const int32_t* co, const void* u, double t, int32_t* out_off, int32_t* out_beg,
                     int32_t* out_end, int64_t capacity, int32_t* d_total, uint32_t* status, void* scratch,
                     int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    FusedLaunch L{nullptr, R, out_off, out_beg, out_end, capacity, d_total, status, scratch, scratch_bytes, epoch, stream};
    if (dtype == 4)
        return launch_fused_rows_g(group, GradTagRow<float>{rows, off, beg, end, co, R, (const float*)u, (float)t}, L,
                                   "grad_tag");
    if (dtype == 8)
        return launch_fused_rows_g(group, GradTagRow<double>{rows, off, beg, end, co, R, (const double*)u, t}, L,
                                   "grad_tag");
    return stx_fail(STX_EINVAL, "grad_tag: dtype must be 4 or 8");
}

int32_t stx_gradient(const float* data, int32_t width, int32_t height, float* out, hipStream_t stream) {
    if (width <= 0 || height <= 0) return stx_fail(STX_EINVAL, "gradient: width and height must be positive");
    hipLaunchKernelGGL(gradient_kernel, dim3(nblk((int64_t)width * height)), dim3(kT), 0, stream, data, width, height, out);
    return stx_launch_status("gradient");
}

}  // extern "C"
