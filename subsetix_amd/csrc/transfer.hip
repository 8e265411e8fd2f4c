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

// ---------------------------------------------------------------------------
// Segment form (ratio 2, and the ratio-1 copy).  A per-(target, destination,
// source) table cuts the target's intervals into segments over which the
// destination interval and the source interval(s) do not change, so inside a
// segment, with c the target's own cell index:
//   restrict: dst cell = dD + c, child (d, e) = sD_d + 2 c + e (d, e < 2),
//             mask bit 2 d + e: the child exists (missing children read 0,
//             which leaves an FP sum unchanged, exactly as skipping them);
//   prolong:  dst cell = dD + c, src cell = sB + floor((xD + c) / 2);
//   copy:     dst cell = dD + c, src cell = sD + c.
// Touching intervals of a row have contiguous cells, so one base per source
// row covers a segment even where it spans them.  Record (8 int32):
//   restrict: c0, c1, dD, mask, sD0, sD1, 0, 0
//   prolong:  c0, c1, dD, xD, sB, 0, 0, 0
//   copy:     c0, c1, dD, 0, sD, 0, 0, 0
// (dD = kNone: destination cell absent; prolong sB / copy sD = kNone: source
// absent, the destination cell keeps its value).
constexpr int32_t kNone = INT32_MIN;
constexpr int kRec = 8;
constexpr int kThreads = 256;
constexpr int kTileCells = 2048;  // 256 threads x 8 cells
constexpr int kSegCap = 512;
constexpr int kXcds = 8;
// Block -> tile order: the XCDs take chunks of kOrderChunk consecutive tiles
// in turn (block b runs on XCD b % 8), so the rows above and below a tile are
// mostly read through the same L2 while the whole chip sweeps the field front
// to back (measured: same FP32 step time, 10 % faster FP64 steps than one
// contiguous eighth per XCD).  Grids are padded to a multiple of
// kOrderSpan, which makes the order a permutation of the padded grid.
constexpr int64_t kOrderChunk = 8;
constexpr int64_t kOrderSpan = kOrderChunk * kXcds;
__device__ __forceinline__ int64_t xcd_chunk_order(int64_t block) {
    const int64_t xcd = block % kXcds, nth = block / kXcds;
    return ((nth / kOrderChunk) * kXcds + xcd) * kOrderChunk + nth % kOrderChunk;
}


__device__ __forceinline__ int32_t first_ending_after(const int32_t* __restrict__ end, int32_t j, int32_t jend,
                                                      int32_t x) {
    int32_t lo = j, hi = jend;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (end[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int32_t ceil2(int32_t v) { return floordiv(v + 1, 2); }

// Smallest projected boundary of interval [b, e) above x.  Restrict (fine
// interval seen in coarse columns): floor(b/2), ceil(b/2), floor(e/2),
// ceil(e/2).  Prolong (coarse interval in fine columns): 2b, 2e.  Copy (same
// columns): b, e.
template <int MODE>
__device__ __forceinline__ int32_t next_bound(int32_t b, int32_t e, int32_t x) {
    int32_t best = INT32_MAX;
    if (MODE == 2) {
        if (b > x) best = b;
        else if (e > x) best = e;
    } else if (MODE == 1) {
        const int32_t c[4] = {floordiv(b, 2), ceil2(b), floordiv(e, 2), ceil2(e)};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (c[k] > x && c[k] < best) best = c[k];
    } else {
        if (2 * b > x) best = 2 * b;
        else if (2 * e > x) best = 2 * e;
    }
    return best;
}

template <int MODE, bool WRITE>
__global__ void transfer_segments_kernel(Geom tg, const int32_t* __restrict__ t_ord, int64_t tN, Geom dg, Geom sg,
                                         int32_t* __restrict__ counts, const int32_t* __restrict__ seg_off,
                                         int32_t* __restrict__ segs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tN) return;
    const int32_t y = tg.rows[t_ord[i]];
    const int32_t bT = tg.beg[i], eT = tg.end[i], cT = tg.co[i];
    // destination row y
    int32_t jd = 0, jde = 0;
    {
        const int64_t k = find_row(dg.rows, dg.R, y);
        if (k >= 0) {
            jde = dg.off[k + 1];
            jd = first_ending_after(dg.end, dg.off[k], jde, bT);
        }
    }
    // source rows: restrict 2y, 2y + 1; prolong floor(y / 2); copy y
    constexpr int NS = MODE == 1 ? 2 : 1;
    int32_t js[NS], jse[NS];
#pragma unroll
    for (int d = 0; d < NS; ++d) {
        js[d] = jse[d] = 0;
        const int32_t sy = MODE == 1 ? 2 * y + d : (MODE == 2 ? y : floordiv(y, 2));
        const int64_t k = find_row(sg.rows, sg.R, sy);
        if (k >= 0) {
            jse[d] = sg.off[k + 1];
            // first source interval whose projected end is above bT
            const int32_t px = MODE == 1 ? 2 * bT : (MODE == 2 ? bT : floordiv(bT, 2));
            js[d] = first_ending_after(sg.end, sg.off[k], jse[d], px);
        }
    }
    int32_t x = bT, n = 0;
    const int64_t base = WRITE ? (int64_t)seg_off[i] : 0;
    while (x < eT) {
        while (jd < jde && dg.end[jd] <= x) ++jd;
        int32_t nx = eT;
        if (jd < jde) {
            const int32_t bd = x < dg.beg[jd] ? dg.beg[jd] : dg.end[jd];
            nx = bd < nx ? bd : nx;
        }
#pragma unroll
        for (int d = 0; d < NS; ++d) {
            // skip source intervals whose projection ends at or before x
            while (js[d] < jse[d] &&
                   (MODE == 1 ? ceil2(sg.end[js[d]]) <= x : (MODE == 2 ? sg.end[js[d]] : 2 * sg.end[js[d]]) <= x))
                ++js[d];
            for (int q = 0; q < 2 && js[d] + q < jse[d]; ++q) {  // the interval and the next (touching) one
                const int32_t nb = next_bound<MODE>(sg.beg[js[d] + q], sg.end[js[d] + q], x);
                nx = nb < nx ? nb : nx;
            }
        }
        if (WRITE) {
            int32_t* rec = segs + kRec * (base + n);
            const int32_t c0 = cT + (x - bT);
            rec[0] = c0;
            rec[1] = cT + (nx - bT);
            rec[2] = (jd < jde && dg.beg[jd] <= x) ? (int32_t)((int64_t)dg.co[jd] - dg.beg[jd] + x - c0) : kNone;
            if (MODE == 1) {
                int32_t mask = 0;
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    int64_t sbase = kNone;
                    for (int e = 0; e < 2; ++e) {
                        const int32_t pos = 2 * x + e;
                        for (int q = 0; q < 2 && js[d] + q < jse[d]; ++q) {
                            const int32_t j = js[d] + q;
                            if (sg.beg[j] <= pos && pos < sg.end[j]) {
                                mask |= 1 << (2 * d + e);
                                sbase = (int64_t)sg.co[j] - sg.beg[j];
                                break;
                            }
                        }
                    }
                    rec[4 + d] = sbase == kNone ? 0 : (int32_t)(sbase + 2 * (int64_t)x - 2 * (int64_t)c0);
                }
                rec[3] = mask;
                rec[6] = 0;
                rec[7] = 0;
            } else if (MODE == 2) {
                // source cell of target cell c: rec[4] + c (kNone: column absent)
                int64_t sb = kNone;
                for (int q = 0; q < 2 && js[0] + q < jse[0]; ++q) {
                    const int32_t j = js[0] + q;
                    if (sg.beg[j] <= x && x < sg.end[j]) {
                        sb = (int64_t)sg.co[j] - sg.beg[j];
                        break;
                    }
                }
                rec[3] = 0;
                rec[4] = sb == kNone ? kNone : (int32_t)(sb + x - c0);
                rec[5] = 0;
                rec[6] = 0;
                rec[7] = 0;
            } else {
                const int32_t sx = floordiv(x, 2);
                int64_t sb = kNone;
                for (int q = 0; q < 2 && js[0] + q < jse[0]; ++q) {
                    const int32_t j = js[0] + q;
                    if (sg.beg[j] <= sx && sx < sg.end[j]) {
                        sb = (int64_t)sg.co[j] - sg.beg[j];
                        break;
                    }
                }
                rec[3] = x - c0;
                rec[4] = sb == kNone ? kNone : (int32_t)sb;
                rec[5] = 0;
                rec[6] = 0;
                rec[7] = 0;
            }
        }
        ++n;
        x = nx;
    }
    if (!WRITE) counts[i] = n;
}

__global__ void transfer_plan_kernel(const int32_t* __restrict__ segs, int64_t nseg, int64_t C, int64_t ntiles,
                                     int32_t* __restrict__ plan) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int64_t cf = t * kTileCells;
    const int64_t cl = (cf + kTileCells < C ? cf + kTileCells : C) - 1;
    auto locate = [&](int64_t c) {
        int64_t lo = 0, hi = nseg;
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (segs[kRec * mid] <= c) lo = mid; else hi = mid;
        }
        return lo;
    };
    plan[2 * t] = (int32_t)locate(cf);
    plan[2 * t + 1] = (int32_t)(locate(cl) + 1);
}

struct TSeg {
    int32_t c0, c1, dD, a3, a4, a5;
};

template <typename T>
struct VecT;
template <>
struct VecT<float> {
    typedef float __attribute__((ext_vector_type(4))) type;
    static constexpr int n = 4;
};
template <>
struct VecT<double> {
    typedef double __attribute__((ext_vector_type(2))) type;
    static constexpr int n = 2;
};

template <typename T>
__device__ __forceinline__ T gatherT(__amdgpu_buffer_rsrc_t r, int32_t idx);
template <>
__device__ __forceinline__ float gatherT<float>(__amdgpu_buffer_rsrc_t r, int32_t idx) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, buffer_byte_offset(idx, 4u), 0, 0));
}
template <>
__device__ __forceinline__ double gatherT<double>(__amdgpu_buffer_rsrc_t r, int32_t idx) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, buffer_byte_offset(idx, 8u), 0, 0));
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_of(const void* p, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int32_t)bytes, 0x00020000);
}

template <typename V, typename T>
__device__ __forceinline__ V load_dw(const T* p) {
    V v;
    __builtin_memcpy(&v, __builtin_assume_aligned(p, 4), sizeof(V));
    return v;
}

template <typename V, typename T>
__device__ __forceinline__ void store_dw(T* p, const V& v) {
    __builtin_memcpy(__builtin_assume_aligned(p, 4), &v, sizeof(V));
}

// One group of G = 16 / sizeof(T) target cells [c, c + G).
template <int MODE, typename T, typename Acc, class SG>
__device__ __forceinline__ void transfer_group(const SG& sg, int32_t s0, int32_t s1, int32_t c, int32_t CT,
                                               __amdgpu_buffer_rsrc_t rs, const T* __restrict__ src,
                                               T* __restrict__ dst, Acc norm) {
    constexpr int G = VecT<T>::n;
    int32_t lo = s0, hi = s1;
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (sg.c0(mid) <= c) lo = mid; else hi = mid;
    }
    int32_t s = lo;
    TSeg g = sg.get(s);
    if (MODE == 1 && c + G <= CT && c + G <= g.c1 && g.dD != kNone && g.a3 == 0xF) {
        // all children present: two rows of 2G consecutive children
        using V2 = T __attribute__((ext_vector_type(2 * G)));
        const V2 r0 = load_dw<V2>(src + (g.a4 + 2 * c));
        const V2 r1 = load_dw<V2>(src + (g.a5 + 2 * c));
        typename VecT<T>::type ov;
#pragma unroll
        for (int v = 0; v < G; ++v) {
            Acc acc = (Acc)0;
            acc = acc + (Acc)r0[2 * v];
            acc = acc + (Acc)r0[2 * v + 1];
            acc = acc + (Acc)r1[2 * v];
            acc = acc + (Acc)r1[2 * v + 1];
            ov[v] = (T)(acc * norm);
        }
        store_dw(dst + (g.dD + c), ov);
        return;
    }
    if (MODE == 2 && c + G <= CT && c + G <= g.c1 && g.dD != kNone && g.a4 != kNone) {
        store_dw(dst + (g.dD + c), load_dw<typename VecT<T>::type>(src + (g.a4 + c)));
        return;
    }
#pragma unroll
    for (int v = 0; v < G; ++v) {
        const int32_t cc = c + v;
        if (cc >= CT) break;
        if (cc >= g.c1) {
            ++s;
            g = sg.get(s);
        }
        if (g.dD == kNone) continue;
        if (MODE == 1) {
            Acc acc = (Acc)0;
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const int32_t base = (d == 0 ? g.a4 : g.a5) + 2 * cc;
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const bool ok = (g.a3 >> (2 * d + e)) & 1;
                    acc = acc + (Acc)gatherT<T>(rs, ok ? base + e : -1);  // absent children read 0
                }
            }
            dst[g.dD + cc] = (T)(acc * norm);
        } else if (MODE == 2) {
            if (g.a4 == kNone) continue;
            dst[g.dD + cc] = src[g.a4 + cc];
        } else {
            if (g.a4 == kNone) continue;
            dst[g.dD + cc] = src[g.a4 + floordiv(g.a3 + cc, 2)];
        }
    }
}

struct GSegs {
    const int4* __restrict__ p;
    __device__ __forceinline__ int32_t c0(int32_t s) const { return p[2 * s].x; }
    __device__ __forceinline__ TSeg get(int32_t s) const {
        const int4 a = p[2 * s], b = p[2 * s + 1];
        return TSeg{a.x, a.y, a.z, a.w, b.x, b.y};
    }
};

struct LSegs {
    const int4* p;
    int32_t base;
    __device__ __forceinline__ int32_t c0(int32_t s) const { return p[2 * (s - base)].x; }
    __device__ __forceinline__ TSeg get(int32_t s) const {
        const int4 a = p[2 * (s - base)], b = p[2 * (s - base) + 1];
        return TSeg{a.x, a.y, a.z, a.w, b.x, b.y};
    }
};

// `split` blocks share one tile of the plan (a small target would otherwise
// give too few blocks to fill the chip): block b works on part b % split of
// tile b / split, searching the tile's segment range as a whole block would.
template <int MODE, typename T, typename Acc>
__global__ __launch_bounds__(kThreads) void transfer_seg_kernel(const int32_t* __restrict__ plan, int64_t ntiles,
                                                                int32_t CT, const int32_t* __restrict__ segs,
                                                                const T* __restrict__ src, int64_t src_cells,
                                                                T* __restrict__ dst, Acc norm, int32_t split) {
    __shared__ int4 s_seg[2 * kSegCap];
    const int64_t b = xcd_chunk_order(blockIdx.x);
    const int64_t t = b / split;
    if (t >= ntiles) return;
    const int32_t s0 = plan[2 * t], s1 = plan[2 * t + 1];
    const int32_t M = s1 - s0;
    const int32_t part = kTileCells / split;
    const int32_t cb = (int32_t)(t * kTileCells) + (int32_t)(b % split) * part;  // first cell of this block
    const int32_t ce = cb + part < CT ? cb + part : CT;
    const int4* gs = (const int4*)segs;
    const __amdgpu_buffer_rsrc_t rs = rsrc_of(src, src_cells * (int64_t)sizeof(T));
    constexpr int G = VecT<T>::n;
    if (M <= kSegCap) {
        for (int q = threadIdx.x; q < 2 * M; q += kThreads) s_seg[q] = gs[2 * (int64_t)s0 + q];
        __syncthreads();
        const LSegs sg{s_seg, s0};
        for (int32_t c = cb + (int32_t)threadIdx.x * G; c < ce; c += kThreads * G)
            transfer_group<MODE, T, Acc>(sg, s0, s1, c, CT, rs, src, dst, norm);
    } else {
        const GSegs sg{gs};
        for (int32_t c = cb + (int32_t)threadIdx.x * G; c < ce; c += kThreads * G)
            transfer_group<MODE, T, Acc>(sg, s0, s1, c, CT, rs, src, dst, norm);
    }
}

template <int MODE, typename T, typename Acc>
void launch_transfer_seg(const int32_t* plan, int64_t ntiles, int64_t CT, const int32_t* segs, const void* src,
                         int64_t src_cells, void* dst, Acc norm, hipStream_t stream) {
    const int32_t split = ntiles < 1024 ? 4 : 1;  // fewer tiles than 4 per CU: quarter tiles
    const int64_t blocks = ntiles * split;
    const int64_t grid = (blocks + kOrderSpan - 1) / kOrderSpan * kOrderSpan;
    hipLaunchKernelGGL((transfer_seg_kernel<MODE, T, Acc>), dim3((unsigned)grid), dim3(kThreads), 0, stream, plan,
                       ntiles, (int32_t)CT, segs, (const T*)src, src_cells, (T*)dst, norm, split);
}

inline unsigned nblocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

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

extern "C" int32_t stx_scan_i32(const int32_t* in, const int32_t* d_n, int64_t n_bound, int32_t* out, int32_t* d_total,
                                uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch,
                                hipStream_t stream);

extern "C" {

int64_t stx_transfer_tile_cells(void) { return kTileCells; }

// Segment table of a transfer onto target T (t_ord: row index of every target
// interval, tN intervals): mode 0 ratio-2 prolong, 1 ratio-2 restrict, 2 copy
// (ratio 1: dst(y, x) = src(y, x)).
int32_t stx_transfer_seg_count(int32_t mode, const int32_t* t_rows, const int32_t* t_off, const int32_t* t_beg,
                               const int32_t* t_end, const int32_t* t_co, int64_t t_R, const int32_t* t_ord,
                               int64_t tN, const int32_t* d_rows, const int32_t* d_off, const int32_t* d_beg,
                               const int32_t* d_end, const int32_t* d_co, int64_t d_R, const int32_t* s_rows,
                               const int32_t* s_off, const int32_t* s_beg, const int32_t* s_end, const int32_t* s_co,
                               int64_t s_R, int32_t* counts, int32_t* seg_off, int32_t* d_total, uint32_t* status,
                               void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    if (mode < 0 || mode > 2) return stx_fail(STX_EINVAL, "transfer_seg: mode must be 0, 1 or 2");
    if (tN < 0) return stx_fail(STX_EINVAL, "transfer_seg: bad sizes");
    Geom t{t_rows, t_off, t_beg, t_end, t_co, t_R};
    Geom d{d_rows, d_off, d_beg, d_end, d_co, d_R};
    Geom s{s_rows, s_off, s_beg, s_end, s_co, s_R};
    if (tN > 0) {
        if (mode == 2)
            hipLaunchKernelGGL((transfer_segments_kernel<2, false>), dim3(nblocks256(tN)), dim3(256), 0, stream, t,
                               t_ord, tN, d, s, counts, nullptr, nullptr);
        else if (mode == 1)
            hipLaunchKernelGGL((transfer_segments_kernel<1, false>), dim3(nblocks256(tN)), dim3(256), 0, stream, t,
                               t_ord, tN, d, s, counts, nullptr, nullptr);
        else
            hipLaunchKernelGGL((transfer_segments_kernel<0, false>), dim3(nblocks256(tN)), dim3(256), 0, stream, t,
                               t_ord, tN, d, s, counts, nullptr, nullptr);
    }
    int32_t rc = stx_launch_status("transfer_seg_count");
    if (rc) return rc;
    return stx_scan_i32(counts, nullptr, tN, seg_off, d_total, status, scratch, scratch_bytes, epoch, stream);
}

int32_t stx_transfer_seg_build(int32_t mode, const int32_t* t_rows, const int32_t* t_off, const int32_t* t_beg,
                               const int32_t* t_end, const int32_t* t_co, int64_t t_R, const int32_t* t_ord,
                               int64_t tN, const int32_t* d_rows, const int32_t* d_off, const int32_t* d_beg,
                               const int32_t* d_end, const int32_t* d_co, int64_t d_R, const int32_t* s_rows,
                               const int32_t* s_off, const int32_t* s_beg, const int32_t* s_end, const int32_t* s_co,
                               int64_t s_R, const int32_t* seg_off, int64_t nseg, int32_t* segs, int64_t CT,
                               int32_t* plan, int64_t ntiles, hipStream_t stream) {
    if (mode < 0 || mode > 2) return stx_fail(STX_EINVAL, "transfer_seg: mode must be 0, 1 or 2");
    if (ntiles != (CT + kTileCells - 1) / kTileCells) return stx_fail(STX_EINVAL, "transfer_seg: tile count");
    if (2 * CT + 16 > 0x7fffffffLL) return stx_fail(STX_EINVAL, "transfer_seg: too many target cells");
    Geom t{t_rows, t_off, t_beg, t_end, t_co, t_R};
    Geom d{d_rows, d_off, d_beg, d_end, d_co, d_R};
    Geom s{s_rows, s_off, s_beg, s_end, s_co, s_R};
    if (tN > 0) {
        if (mode == 2)
            hipLaunchKernelGGL((transfer_segments_kernel<2, true>), dim3(nblocks256(tN)), dim3(256), 0, stream, t,
                               t_ord, tN, d, s, nullptr, seg_off, segs);
        else if (mode == 1)
            hipLaunchKernelGGL((transfer_segments_kernel<1, true>), dim3(nblocks256(tN)), dim3(256), 0, stream, t,
                               t_ord, tN, d, s, nullptr, seg_off, segs);
        else
            hipLaunchKernelGGL((transfer_segments_kernel<0, true>), dim3(nblocks256(tN)), dim3(256), 0, stream, t,
                               t_ord, tN, d, s, nullptr, seg_off, segs);
    }
    if (ntiles > 0) {
        if (nseg <= 0) return stx_fail(STX_EINVAL, "transfer_seg: cells without segments");
        hipLaunchKernelGGL(transfer_plan_kernel, dim3(nblocks256(ntiles)), dim3(256), 0, stream, segs, nseg, CT, ntiles,
                           plan);
    }
    return stx_launch_status("transfer_seg_build");
}

// Transfer with a segment table (mode as the table's): dtype 4 / 8, acc64 as
// stx_transfer (restrict only).  src_cells: number of cells of the source
// field (reads beyond are zero).
int32_t stx_transfer_seg(int32_t mode, int32_t dtype, int32_t acc64, const int32_t* plan, int64_t ntiles, int64_t CT,
                         const int32_t* segs, const void* src, int64_t src_cells, void* dst, hipStream_t stream) {
    if (ntiles <= 0 || CT <= 0) return STX_OK;
    if (ntiles != (CT + kTileCells - 1) / kTileCells) return stx_fail(STX_EINVAL, "transfer_seg: plan size mismatch");
    if (src_cells * (dtype == 8 ? 8 : 4) > 0xffffff00LL)
        return stx_fail(STX_EINVAL, "transfer_seg: source field larger than 4 GiB");
    if (mode == 1) {
        if (dtype == 4 && acc64)
            launch_transfer_seg<1, float, double>(plan, ntiles, CT, segs, src, src_cells, dst, 0.25, stream);
        else if (dtype == 4)
            launch_transfer_seg<1, float, float>(plan, ntiles, CT, segs, src, src_cells, dst, 0.25f, stream);
        else if (dtype == 8)
            launch_transfer_seg<1, double, double>(plan, ntiles, CT, segs, src, src_cells, dst, 0.25, stream);
        else
            return stx_fail(STX_EINVAL, "transfer_seg: dtype must be 4 or 8");
    } else if (mode == 0) {
        if (dtype == 4)
            launch_transfer_seg<0, float, float>(plan, ntiles, CT, segs, src, src_cells, dst, 1.0f, stream);
        else if (dtype == 8)
            launch_transfer_seg<0, double, double>(plan, ntiles, CT, segs, src, src_cells, dst, 1.0, stream);
        else
            return stx_fail(STX_EINVAL, "transfer_seg: dtype must be 4 or 8");
    } else if (mode == 2) {
        if (dtype == 4)
            launch_transfer_seg<2, float, float>(plan, ntiles, CT, segs, src, src_cells, dst, 1.0f, stream);
        else if (dtype == 8)
            launch_transfer_seg<2, double, double>(plan, ntiles, CT, segs, src, src_cells, dst, 1.0, stream);
        else
            return stx_fail(STX_EINVAL, "transfer_seg: dtype must be 4 or 8");
    } else {
        return stx_fail(STX_EINVAL, "transfer_seg: mode must be 0, 1 or 2");
    }
    return stx_launch_status("transfer_seg");
}

}  // extern "C"
