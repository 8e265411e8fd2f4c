// Segment-based first-order upwind stencil on an IntervalField
// (interval_stencil.py:67-160, kernels.py:1108-1199).
//
// Cells are stored row by row, interval by interval.  A per-geometry table,
// built once and reused for every step on that geometry, cuts every interval
// into SEGMENTS: maximal runs of cells over which the interval of the row
// below and the interval of the row above do not change.  Inside a segment
// the vertical neighbours of cell c are cells c + dd and c + ud (constant
// deltas, or none), and the horizontal neighbours are c - 1 / c + 1 except at
// the segment ends, where a flag says whether the adjacent cell is the
// neighbour (same interval, or a touching interval of the same row).  The
// reference instead binary-searches the neighbour rows for every cell.
//
// Segments are also cut at the domain edges x = 0 and x = w (the row's width:
// the call's width, or the instance width for a stack of instances), so the
// zero-exterior range checks fold into per-segment flags.  Segment record (8
// int32): c0, c_end, dd, ud, flags, instance, x0 (column of c0), row
// ordinal.  flags bit 0: c0 has a
// left neighbour (cell c0 - 1); bit 1: the segment lies in [0, w), so inner
// cells have both horizontal neighbours; bit 2: c_end - 1 has a right
// neighbour (cell c_end).  dd / ud = kNone when there is no vertical
// neighbour (row absent, column not held, or segment outside [0, w)).
//
// The update kernel cuts the cells into tiles of kTileCells; a thread owns
// 32 bytes of consecutive cells per array (8 floats or 4 doubles: two vector
// groups, each one segment lookup, vector loads and stores), so a block has
// 256 threads for float and 512 for double; the block stages the tile's
// segment records in LDS.  Tiles are mapped to
// blocks in XCD chunks (xcd_chunk_order) so the rows above and below a tile
// are mostly read through the same L2.
//
// Precondition (checked by the Python layer unless disabled): row ids lie in
// [0, height), so adjacency of row ids decides the vertical neighbours.
#include <limits.h>

#include <cmath>
#include <limits>

#include "stx_common.h"
#include "stx_host.h"

using namespace stx;

namespace {

constexpr int kThreads = 256;
constexpr int kCellsPerThread = 8;
constexpr int kTileCells = kThreads * kCellsPerThread;
constexpr int kSegCap = 512;
// Threads of an update block: a tile of kTileCells at 32 bytes of cells per
// thread (measured on the composite step: FP64 with 8 cells per thread
// 201 us, with 4 cells 193 us; FP32 with 4 cells 124 us, with 8 cells 118 us,
// with 16 cells 121 us).
template <typename T>
constexpr int seg_threads() {
    return kTileCells * (int)sizeof(T) / 32;
}
constexpr int kXcds = 8;
// Block -> tile order: the XCDs take chunks of kOrderChunk consecutive tiles
// in turn (block b runs on XCD b % 8), so the rows above and below a tile are
// mostly read through the same L2 while the whole chip sweeps the field front
// to back.  Measured on the composite step (4096^2 coarse, 8192^2 fine
// discs): chunks of 32 tiles take 109.7 us (FP32) and 182 us (FP64) per
// step, chunks of 8 112.8 and 195.5 us, chunks of 64 or 128 111.4 and
// 184 us, chunks of 2 121.6 and 210.8 us; one contiguous eighth per XCD was
// slower still for FP64.  Grids are padded to a multiple of kOrderSpan,
// which makes the order a permutation of the padded grid.
constexpr int64_t kOrderChunk = 32;
constexpr int64_t kOrderSpan = kOrderChunk * kXcds;
__device__ __forceinline__ int64_t xcd_chunk_order(int64_t block) {
    const int64_t xcd = block % kXcds, nth = block / kXcds;
    return ((nth / kOrderChunk) * kXcds + xcd) * kOrderChunk + nth % kOrderChunk;
}

constexpr int32_t kNone = INT32_MIN;
constexpr int kRec = 8;

inline unsigned nblocks(int64_t n) { return (unsigned)((n + 255) / 256); }

// First interval of [j, jend) whose end is > x (jend when none).
__device__ __forceinline__ int32_t first_ending_after(const int32_t* __restrict__ end, int32_t j, int32_t jend,
                                                      int32_t x) {
    int32_t lo = j, hi = jend;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (end[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Segments of interval i: count (WRITE = false, counts[i]) or records
// (WRITE = true, at seg_off[i]).  Breakpoints: the boundaries of the
// intervals of the rows directly below and above, and the domain edges 0 and
// w (the row's width), so that the x-range checks of a segment's cells are
// constant and can be folded into its flags.
template <bool WRITE>
__global__ void segments_kernel(const int32_t* __restrict__ rows, const int32_t* __restrict__ off, int64_t R,
                                const int32_t* __restrict__ beg, const int32_t* __restrict__ end,
                                const int32_t* __restrict__ co, const int32_t* __restrict__ ord, int64_t N,
                                int32_t width, const int32_t* __restrict__ row_inst,
                                const int32_t* __restrict__ inst_w, int32_t* __restrict__ counts,
                                const int32_t* __restrict__ seg_off, int32_t* __restrict__ segs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t k = ord[i];
    const int32_t rid = rows[k];
    const int32_t o0 = off[k], o1 = off[k + 1];
    const int32_t bi = beg[i], ei = end[i], ci = co[i];
    int32_t jd = 0, jde = 0, ju = 0, jue = 0;
    if (k > 0 && rows[k - 1] == rid - 1) {
        jde = o0;
        jd = first_ending_after(end, off[k - 1], jde, bi);
    }
    if (k + 1 < R && rows[k + 1] == rid + 1) {
        jue = off[k + 2];
        ju = first_ending_after(end, o1, jue, bi);
    }
    const bool fl_first = i > o0 && end[i - 1] == bi;
    const bool fr_last = i + 1 < o1 && beg[i + 1] == ei;
    int32_t inst = 0, w = width;
    if (row_inst) {
        inst = row_inst[k];
        w = inst_w[inst];
    }
    int32_t x = bi;
    int32_t n = 0;
    const int64_t base = WRITE ? (int64_t)seg_off[i] : 0;
    while (x < ei) {
        while (jd < jde && end[jd] <= x) ++jd;
        while (ju < jue && end[ju] <= x) ++ju;
        int32_t nx = ei;
        if (jd < jde) {
            const int32_t bd = x < beg[jd] ? beg[jd] : end[jd];
            nx = bd < nx ? bd : nx;
        }
        if (ju < jue) {
            const int32_t bu = x < beg[ju] ? beg[ju] : end[ju];
            nx = bu < nx ? bu : nx;
        }
        if (x < 0 && 0 < nx) nx = 0;
        if (x < w && w < nx) nx = w;
        if (WRITE) {
            int32_t* rec = segs + kRec * (base + n);
            const int64_t cself = (int64_t)ci - bi;  // cell of column x is cself + x
            const bool inside = x >= 0 && nx <= w;   // all of [x, nx) or none of it in [0, w)
            const bool lf = (x == bi ? fl_first : true) && x - 1 >= 0 && x - 1 < w;
            const bool rl = (nx == ei ? fr_last : true) && nx >= 0 && nx < w;
            rec[0] = (int32_t)(cself + x);
            rec[1] = (int32_t)(cself + nx);
            rec[2] = (inside && jd < jde && beg[jd] <= x) ? (int32_t)(((int64_t)co[jd] - beg[jd]) - cself) : kNone;
            rec[3] = (inside && ju < jue && beg[ju] <= x) ? (int32_t)(((int64_t)co[ju] - beg[ju]) - cself) : kNone;
            rec[4] = (lf ? 1 : 0) | (inside ? 2 : 0) | (rl ? 4 : 0);
            rec[5] = inst;
            rec[6] = x;
            rec[7] = (int32_t)k;
        }
        ++n;
        x = nx;
    }
    if (!WRITE) counts[i] = n;
}

// plan[2 t], plan[2 t + 1] = segments [s0, s1) overlapping tile t.
__global__ void seg_plan_kernel(const int32_t* __restrict__ segs, int64_t nseg, int64_t C, int64_t ntiles,
                                int32_t* __restrict__ plan) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int64_t cf = t * kTileCells;
    const int64_t cl = (cf + kTileCells < C ? cf + kTileCells : C) - 1;
    auto locate = [&](int64_t c) {  // last segment with c0 <= c
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

struct Seg {
    int32_t c0, cend, dd, ud, flags, inst;
};

struct GlobalSegs {
    const int4* __restrict__ p;
    __device__ __forceinline__ int32_t c0(int32_t s) const { return p[2 * s].x; }
    __device__ __forceinline__ Seg get(int32_t s) const {
        const int4 a = p[2 * s], b = p[2 * s + 1];
        return Seg{a.x, a.y, a.z, a.w, b.x, b.y};
    }
};

struct LdsSegs {
    const int4* p;
    int32_t base;
    __device__ __forceinline__ int32_t c0(int32_t s) const { return p[2 * (s - base)].x; }
    __device__ __forceinline__ Seg get(int32_t s) const {
        const int4 a = p[2 * (s - base)], b = p[2 * (s - base) + 1];
        return Seg{a.x, a.y, a.z, a.w, b.x, b.y};
    }
};

template <typename T>
struct Vec;
template <>
struct Vec<float> {
    typedef float __attribute__((ext_vector_type(4))) type;
    static constexpr int n = 4;
};
template <>
struct Vec<double> {
    typedef double __attribute__((ext_vector_type(2))) type;
    static constexpr int n = 2;
};

template <typename T>
__device__ __forceinline__ T upwind_cell(T center, T left, T right, T down, T up, T a, T b, T dt, T dx, T dy) {
    const T nx = (a >= (T)0) ? center - left : right - center;
    const T ny = (b >= (T)0) ? center - down : up - center;
    const T dudx = nx / dx;
    const T dudy = ny / dy;
    const T t1 = a * dudx;
    const T t2 = b * dudy;
    return center - dt * (t1 + t2);
}

// 1 / d when d is a positive power of two whose reciprocal is a normal
// number (then x / d == x * (1 / d) exactly: both are the correctly rounded
// value of the same exact scaling), else 0.  From the bits: 2^e -> 2^-e.
__device__ __forceinline__ float pow2_recip(float d) {
    const uint32_t u = __float_as_uint(d);
    const uint32_t e = u >> 23;  // sign bit 0 and mantissa 0 checked below
    return ((u & 0x007fffffu) == 0u && e >= 1u && e <= 253u) ? __uint_as_float((254u - e) << 23) : 0.0f;
}
__device__ __forceinline__ double pow2_recip(double d) {
    const uint64_t u = __double_as_longlong(d);
    const uint64_t e = u >> 52;
    return ((u & 0x000fffffffffffffull) == 0ull && e >= 1ull && e <= 2045ull) ? __longlong_as_double((int64_t)((2046ull - e) << 52))
                                                                              : 0.0;
}

// upwind_cell with the divisions by dx / dy as products with their exact
// reciprocals rx / ry (powers of two only, see pow2_recip).
template <typename T>
__device__ __forceinline__ T upwind_cell_p2(T center, T left, T right, T down, T up, T a, T b, T dt, T rx, T ry) {
    const T nx = (a >= (T)0) ? center - left : right - center;
    const T ny = (b >= (T)0) ? center - down : up - center;
    const T dudx = nx * rx;
    const T dudy = ny * ry;
    const T t1 = a * dudx;
    const T t2 = b * dudy;
    return center - dt * (t1 + t2);
}

// Gathers go through a buffer resource over the whole field: a 32-bit byte
// offset per load, and an index of -1 (or >= C) reads 0 (out-of-range buffer
// loads return zero), which is exactly the zero exterior of the stencil.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const void* u, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(u), (short)0, (int32_t)bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T gather(__amdgpu_buffer_rsrc_t r, int32_t idx);
template <>
__device__ __forceinline__ float gather<float>(__amdgpu_buffer_rsrc_t r, int32_t idx) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, buffer_byte_offset(idx, 4u), 0, 0));
}
template <>
__device__ __forceinline__ double gather<double>(__amdgpu_buffer_rsrc_t r, int32_t idx) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, buffer_byte_offset(idx, 8u), 0, 0));
}

template <typename T>
__device__ __forceinline__ typename Vec<T>::type gather_vec(__amdgpu_buffer_rsrc_t r, int32_t idx) {
    return __builtin_bit_cast(typename Vec<T>::type,
                              __builtin_amdgcn_raw_buffer_load_b128(r, buffer_byte_offset(idx, (uint32_t)sizeof(T)), 0, 0));
}

// Coefficients a, b, dt, dx, dy of instance `inst` of a stack.  The first
// active lane's instance is read through the scalar cache, off the vector
// memory path that the field streams through (a wave rarely spans two
// instances); lanes of another instance read theirs.
template <typename T>
__device__ __forceinline__ void instance_coef(const T* __restrict__ inst_coef, int32_t inst, T* k) {
    typedef const __attribute__((address_space(4))) T ConstT;
    const int32_t i0 = __builtin_amdgcn_readfirstlane(inst);
    ConstT* q0 = (ConstT*)(inst_coef) + 5 * (int64_t)i0;
#pragma unroll
    for (int j = 0; j < 5; ++j) k[j] = q0[j];
    if (inst != i0) {
        const T* q = inst_coef + 5 * (int64_t)inst;
#pragma unroll
        for (int j = 0; j < 5; ++j) k[j] = q[j];
    }
}

// Cells [c, c + G) crossing a segment end or the field end: a per-cell walk.
// The own values and the cells beside the group come from cv / lv / rv (a
// cell's left and right neighbours, when in the set, are the cells before
// and after it in storage); the rows below and above are gathered per cell.
// k0: coefficients of segment s's instance, read again only for a cell of
// another instance.
template <typename T, bool BATCH, class SG>
__device__ __forceinline__ void upwind_cells(const SG& sg, int32_t s, int32_t c, int32_t C,
                                             __amdgpu_buffer_rsrc_t rs, T* __restrict__ out,
                                             const typename Vec<T>::type& cv, T lv, T rv, const T* k0,
                                             const T* __restrict__ inst_coef) {
    constexpr int G = Vec<T>::n;
    Seg g = sg.get(s);
    const int32_t inst0 = g.inst;
#pragma unroll
    for (int v = 0; v < G; ++v) {
        const int32_t cc = c + v;
        if (cc < C) {
            if (cc >= g.cend) {
                ++s;
                g = sg.get(s);
            }
            T k[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) k[j] = k0[j];
            if (BATCH && g.inst != inst0) {
                const T* q = inst_coef + 5 * (int64_t)g.inst;
#pragma unroll
                for (int j = 0; j < 5; ++j) k[j] = q[j];
            }
            if (BATCH && k[3] == (T)0) continue;  // retired instance
            const bool lok = (cc == g.c0) ? (g.flags & 1) : (g.flags & 2);
            const bool rok = (cc + 1 == g.cend) ? (g.flags & 4) : (g.flags & 2);
            const T left = lok ? (v == 0 ? lv : cv[v - 1]) : (T)0;
            const T right = rok ? (v == G - 1 ? rv : cv[v + 1]) : (T)0;
            const T down = gather<T>(rs, g.dd != kNone ? cc + g.dd : -1);
            const T up = gather<T>(rs, g.ud != kNone ? cc + g.ud : -1);
            out[cc] = upwind_cell<T>(cv[v], left, right, down, up, k[0], k[1], k[2], k[3], k[4]);
        }
    }
}

// One group of G = 16 / sizeof(T) cells [c, c + G), c a multiple of G, in
// two steps.  issue(): the group's segment (binary search over the tile's
// segments), its rows below and above (one dword-aligned 16-byte load each;
// reads past the field return 0) and its instance's coefficients.
// finish(): a group inside one segment (the fast path) computes from its own
// values cv (with lv / rv, the cells beside it, loaded at the kernel's
// start) and stores one vector; other groups walk their cells.
template <typename T>
struct UpwindGroup {
    using VT = typename Vec<T>::type;
    static constexpr int G = Vec<T>::n;
    int32_t s;
    uint32_t bits;  // 1: fast path, 2: left neighbour of the first cell, 4: right of the last, 8: inner
    VT dv, uv;
    T k[5];

    template <bool BATCH, class SG>
    __device__ __forceinline__ void issue(const SG& sg, int32_t s0, int32_t s1, int32_t c, int32_t C,
                                          __amdgpu_buffer_rsrc_t rs, T a, T b, T dt, T dx, T dy,
                                          const T* __restrict__ inst_coef) {
        int32_t lo = s0, hi = s1;
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (sg.c0(mid) <= c) lo = mid; else hi = mid;
        }
        s = lo;
        const Seg g = sg.get(lo);
        const bool inner = g.flags & 2;
        const bool lfirst = c == g.c0 ? (g.flags & 1) : inner;
        const bool rlast = c + G == g.cend ? (g.flags & 4) : inner;
        const bool fast = c + G <= C && c + G <= g.cend;
        bits = (fast ? 1u : 0u) | (lfirst ? 2u : 0u) | (rlast ? 4u : 0u) | (inner ? 8u : 0u);
        dv = gather_vec<T>(rs, g.dd != kNone ? c + g.dd : -G);  // -G: no row below, reads 0
        uv = gather_vec<T>(rs, g.ud != kNone ? c + g.ud : -G);
        if (BATCH) {
            instance_coef<T>(inst_coef, g.inst, k);
        } else {
            k[0] = a;
            k[1] = b;
            k[2] = dt;
            k[3] = dx;
            k[4] = dy;
        }
    }

    template <bool BATCH, class SG>
    __device__ __forceinline__ void finish(const SG& sg, int32_t c, int32_t C, __amdgpu_buffer_rsrc_t rs,
                                           T* __restrict__ out, const VT& cv, T lv, T rv,
                                           const T* __restrict__ inst_coef) const {
        if (!(bits & 1u)) {
            upwind_cells<T, BATCH>(sg, s, c, C, rs, out, cv, lv, rv, k, inst_coef);
            return;
        }
        const bool inner = bits & 8u;
        VT ov;
        const T rx = pow2_recip(k[3]), ry = pow2_recip(k[4]);
        if (rx != (T)0 && ry != (T)0) {
#pragma unroll
            for (int v = 0; v < G; ++v) {
                const bool lok = v == 0 ? (bool)(bits & 2u) : inner;
                const bool rok = v == G - 1 ? (bool)(bits & 4u) : inner;
                const T left = lok ? (v == 0 ? lv : cv[v - 1]) : (T)0;
                const T right = rok ? (v == G - 1 ? rv : cv[v + 1]) : (T)0;
                ov[v] = upwind_cell_p2<T>(cv[v], left, right, dv[v], uv[v], k[0], k[1], k[2], rx, ry);
            }
        } else {
#pragma unroll
            for (int v = 0; v < G; ++v) {
                const bool lok = v == 0 ? (bool)(bits & 2u) : inner;
                const bool rok = v == G - 1 ? (bool)(bits & 4u) : inner;
                const T left = lok ? (v == 0 ? lv : cv[v - 1]) : (T)0;
                const T right = rok ? (v == G - 1 ? rv : cv[v + 1]) : (T)0;
                ov[v] = upwind_cell<T>(cv[v], left, right, dv[v], uv[v], k[0], k[1], k[2], k[3], k[4]);
            }
        }
        if (!BATCH || k[3] != (T)0) *(VT*)(out + c) = ov;  // dx = 0: retired instance
    }
};

// A thread's NG groups, c = ct + (q TH + tid) G: lanes of a wave own
// consecutive groups, so every vector load of the wave covers one contiguous
// 1 KiB span.  FP32 issues every group's loads before finishing any (their
// latencies overlap); FP64 finishes each group in turn, since holding both
// groups' state needs more than 64 VGPRs and halves the resident waves
// (measured on 512 stacked instances, 327M cells: FP32 level step 751 ->
// 707 us with the overlap, FP64 1279 -> 1402 us).
template <typename T, bool BATCH, class SG>
__device__ __forceinline__ void upwind_thread(const SG& sg, int32_t s0, int32_t s1, int32_t ct, int32_t C,
                                              __amdgpu_buffer_rsrc_t rs, const T* __restrict__ u,
                                              T* __restrict__ out, T a, T b, T dt, T dx, T dy,
                                              const T* __restrict__ inst_coef, const typename Vec<T>::type* cv,
                                              const T* lv, const T* rv) {
    constexpr int G = Vec<T>::n;
    constexpr int TH = seg_threads<T>();
    constexpr int NG = kTileCells / TH / G;
    if constexpr (sizeof(T) == 4) {
        UpwindGroup<T> grp[NG];
#pragma unroll
        for (int q = 0; q < NG; ++q)
            grp[q].template issue<BATCH>(sg, s0, s1, ct + (q * TH + (int32_t)threadIdx.x) * G, C, rs, a, b, dt, dx,
                                         dy, inst_coef);
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int32_t c = ct + (q * TH + (int32_t)threadIdx.x) * G;
            if (c < C)
                grp[q].template finish<BATCH>(sg, c, C, rs, out, cv[q], lv[q], rv[q], inst_coef);
        }
    } else {
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int32_t c = ct + (q * TH + (int32_t)threadIdx.x) * G;
            if (c < C) {
                UpwindGroup<T> grp;
                grp.template issue<BATCH>(sg, s0, s1, c, C, rs, a, b, dt, dx, dy, inst_coef);
                grp.template finish<BATCH>(sg, c, C, rs, out, cv[q], lv[q], rv[q], inst_coef);
            }
        }
    }
}

// One block per tile of kTileCells cells.  Every thread first loads its
// groups' own values and the cells beside them (they do not depend on the
// segments, so their HBM latency overlaps the segment staging); the tile's
// segments then go to LDS (or, past kSegCap, are searched in global memory).
template <typename T, bool BATCH>
__global__ __launch_bounds__(seg_threads<T>()) void upwind_seg_kernel(const int32_t* __restrict__ plan, int64_t ntiles,
                                                              int32_t C, const int32_t* __restrict__ segs,
                                                              const T* __restrict__ u, T* __restrict__ out, T a,
                                                              T b, T dt, T dx, T dy,
                                                              const T* __restrict__ inst_coef) {
    __shared__ int4 s_seg[2 * kSegCap];
    const int64_t t = xcd_chunk_order(blockIdx.x);
    if (t >= ntiles) return;
    constexpr int G = Vec<T>::n;
    constexpr int TH = seg_threads<T>();
    constexpr int NG = kTileCells / TH / G;
    const int32_t ct = (int32_t)(t * kTileCells);
    const __amdgpu_buffer_rsrc_t rs = field_rsrc(u, (int64_t)C * (int64_t)sizeof(T));
    typename Vec<T>::type cv[NG];
    T lv[NG], rv[NG];
#pragma unroll
    for (int q = 0; q < NG; ++q) {  // reads past C return 0
        const int32_t c = ct + (q * TH + (int32_t)threadIdx.x) * G;
        cv[q] = gather_vec<T>(rs, c);
        lv[q] = gather<T>(rs, c - 1);
        rv[q] = gather<T>(rs, c + G);
    }
    const int32_t s0 = plan[2 * t], s1 = plan[2 * t + 1];
    const int32_t M = s1 - s0;
    const int4* gs = (const int4*)segs;
    if (M <= kSegCap) {
        for (int q = threadIdx.x; q < 2 * M; q += TH) s_seg[q] = gs[2 * (int64_t)s0 + q];
        __syncthreads();
        upwind_thread<T, BATCH>(LdsSegs{s_seg, s0}, s0, s1, ct, C, rs, u, out, a, b, dt, dx, dy, inst_coef, cv, lv,
                                rv);
    } else {
        upwind_thread<T, BATCH>(GlobalSegs{gs}, s0, s1, ct, C, rs, u, out, a, b, dt, dx, dy, inst_coef, cv, lv, rv);
    }
}

// ---------------------------------------------------------------------------
// Gradient tags over a segment table (the regrid tag of a level whose cells
// all lie in [0, w) of their row: every segment is then "inside" and its
// flags and vertical deltas are exactly the in-set neighbours that the
// row-group kernel GradTagRow finds by walking the rows).  Three streaming
// passes over bit masks, 32 cells per word:
//   1. tag_bits: per cell, tag = |grad u| >= t (central differences where
//      both neighbours exist, one-sided otherwise) and lex = the cell has a
//      left neighbour; two words per 32 cells;
//   2. tag_runs: run starts S = tag & ~(left tag & lex), run ends
//      E = tag & ~(right tag & right exists), popcount of S per word (then an
//      exclusive scan of the counts);
//   3. tag_write: the r-th start / end of the field in storage order is the
//      r-th run's begin / end (x from the cell's segment: x0 + c - c0), and
//      the row offsets are the start counts before each row's first cell.
// Tag test |grad u| >= t as gx^2 + gy^2 >= smin, smin = sqrt_threshold(t):
// a correctly rounded square root is monotone, so sqrt(s) >= t exactly when
// s >= smin (NaN s and NaN t fail both forms).
template <typename T>
__device__ __forceinline__ bool grad_tag_cell(T center, T left, T right, T down, T up, bool hl, bool hr, bool hd,
                                              bool hu, T smin) {
    T gx = (T)0, gy = (T)0;
    if (hl && hr) gx = (right - left) * (T)0.5;
    else if (hr) gx = right - center;
    else if (hl) gx = center - left;
    if (hd && hu) gy = (up - down) * (T)0.5;
    else if (hu) gy = up - center;
    else if (hd) gy = center - down;
    const T gx2 = gx * gx;
    const T gy2 = gy * gy;
    return gx2 + gy2 >= smin;
}

// Tag and left-neighbour bits of one group [c, c + G) (bit v = cell c + v),
// in two steps as UpwindGroup: issue() finds the segment and loads the rows
// below and above; finish() computes from the own values cv and the cells
// beside the group (lv, rv), or walks the cells of a group that crosses a
// segment end or the field end.
template <typename T>
struct TagGroup {
    using VT = typename Vec<T>::type;
    static constexpr int G = Vec<T>::n;
    int32_t s;
    uint32_t bits;  // 1: fast path, 2: left neighbour of the first cell, 4: right of the last, 8: inner,
                    // 16: row below, 32: row above
    VT dv, uv;

    template <class SG>
    __device__ __forceinline__ void issue(const SG& sg, int32_t s0, int32_t s1, int32_t c, int32_t C,
                                          __amdgpu_buffer_rsrc_t rs) {
        int32_t lo = s0, hi = s1;
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (sg.c0(mid) <= c) lo = mid; else hi = mid;
        }
        s = lo;
        const Seg g = sg.get(lo);
        const bool inner = g.flags & 2;
        const bool lfirst = c == g.c0 ? (g.flags & 1) : inner;
        const bool rlast = c + G == g.cend ? (g.flags & 4) : inner;
        const bool fast = c + G <= C && c + G <= g.cend;
        bits = (fast ? 1u : 0u) | (lfirst ? 2u : 0u) | (rlast ? 4u : 0u) | (inner ? 8u : 0u) |
               (g.dd != kNone ? 16u : 0u) | (g.ud != kNone ? 32u : 0u);
        dv = gather_vec<T>(rs, g.dd != kNone ? c + g.dd : -G);  // -G: no row below, reads 0
        uv = gather_vec<T>(rs, g.ud != kNone ? c + g.ud : -G);
    }

    template <class SG>
    __device__ __forceinline__ void finish(const SG& sg, int32_t c, int32_t C, __amdgpu_buffer_rsrc_t rs,
                                           const T* __restrict__ u, const VT& cv, T lv, T rv, T smin, uint32_t& tags,
                                           uint32_t& lex, bool& outside) const {
        tags = 0u;
        lex = 0u;
        if (bits & 1u) {
            const bool inner = bits & 8u;
            outside |= !inner;
            const bool hd = bits & 16u, hu = bits & 32u;
#pragma unroll
            for (int v = 0; v < G; ++v) {
                const bool lok = v == 0 ? (bool)(bits & 2u) : inner;
                const bool rok = v == G - 1 ? (bool)(bits & 4u) : inner;
                const T left = v == 0 ? lv : cv[v - 1];
                const T right = v == G - 1 ? rv : cv[v + 1];
                if (grad_tag_cell<T>(cv[v], left, right, dv[v], uv[v], lok, rok, hd, hu, smin)) tags |= 1u << v;
                if (lok) lex |= 1u << v;
            }
            return;
        }
        int32_t si = s;
        Seg g = sg.get(si);
#pragma unroll
        for (int v = 0; v < G; ++v) {
            const int32_t cc = c + v;
            if (cc < C) {
                if (cc >= g.cend) {
                    ++si;
                    g = sg.get(si);
                }
                outside |= !(g.flags & 2);
                const bool lok = (cc == g.c0) ? (g.flags & 1) : (g.flags & 2);
                const bool rok = (cc + 1 == g.cend) ? (g.flags & 4) : (g.flags & 2);
                const T left = v == 0 ? lv : cv[v - 1];
                const T right = v == G - 1 ? rv : cv[v + 1];
                const T down = gather<T>(rs, g.dd != kNone ? cc + g.dd : -1);
                const T up = gather<T>(rs, g.ud != kNone ? cc + g.ud : -1);
                if (grad_tag_cell<T>(cv[v], left, right, down, up, lok, rok, g.dd != kNone, g.ud != kNone, smin))
                    tags |= 1u << v;
                if (lok) lex |= 1u << v;
            }
        }
    }
};

// One thread's NG groups, then the 32-cell words: 32 / G consecutive lanes
// hold one word, their bits are OR-reduced across the lanes and the first
// lane of each word writes it.  Every lane takes part in the reductions
// (lanes past the last cell contribute zero bits).  FP32 issues both groups'
// loads before finishing either; FP64 (four groups) one group at a time.
template <typename T, class SG>
__device__ __forceinline__ void tag_thread(const SG& sg, int32_t s0, int32_t s1, int32_t ct, int32_t C,
                                           __amdgpu_buffer_rsrc_t rs, const T* __restrict__ u, T smin,
                                           const typename Vec<T>::type* cv, const T* lv, const T* rv,
                                           uint32_t* __restrict__ tag_w, uint32_t* __restrict__ lex_w, int64_t W,
                                           bool& outside) {
    constexpr int G = Vec<T>::n;
    constexpr int LPW = 32 / G;  // lanes per word
    constexpr int NG = kCellsPerThread / G;
    const int lane = lane_id();
    auto word = [&](int32_t c, uint32_t tg, uint32_t lx) {
        const int sh = (lane % LPW) * G;
        uint32_t tw = tg << sh, lw = lx << sh;
#pragma unroll
        for (int d = 1; d < LPW; d <<= 1) {
            tw |= (uint32_t)__shfl_xor((int)tw, d, 64);
            lw |= (uint32_t)__shfl_xor((int)lw, d, 64);
        }
        const int64_t w = (int64_t)c >> 5;
        if (lane % LPW == 0 && w < W) {
            tag_w[w] = tw;
            lex_w[w] = lw;
        }
    };
    if constexpr (sizeof(T) == 4) {
        TagGroup<T> grp[NG];
#pragma unroll
        for (int q = 0; q < NG; ++q) grp[q].issue(sg, s0, s1, ct + (q * kThreads + (int32_t)threadIdx.x) * G, C, rs);
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int32_t c = ct + (q * kThreads + (int32_t)threadIdx.x) * G;
            uint32_t tg = 0u, lx = 0u;
            if (c < C) grp[q].finish(sg, c, C, rs, u, cv[q], lv[q], rv[q], smin, tg, lx, outside);
            word(c, tg, lx);
        }
    } else {
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int32_t c = ct + (q * kThreads + (int32_t)threadIdx.x) * G;
            uint32_t tg = 0u, lx = 0u;
            if (c < C) {
                TagGroup<T> grp;
                grp.issue(sg, s0, s1, c, C, rs);
                grp.finish(sg, c, C, rs, u, cv[q], lv[q], rv[q], smin, tg, lx, outside);
            }
            word(c, tg, lx);
        }
    }
}

// Pass 1.  Own values and the cells beside each group load first (they do
// not depend on the segments); the tile's segments then go to LDS.  smin:
// sqrt_threshold of the tag threshold.
template <typename T>
__global__ __launch_bounds__(kThreads) void tag_bits_kernel(const int32_t* __restrict__ plan, int64_t ntiles,
                                                            int32_t C, const int32_t* __restrict__ segs,
                                                            const T* __restrict__ u, T smin,
                                                            uint32_t* __restrict__ tag_w, uint32_t* __restrict__ lex_w,
                                                            int64_t W, uint32_t* __restrict__ status) {
    constexpr int G = Vec<T>::n;
    constexpr int NG = kCellsPerThread / G;
    __shared__ int4 s_seg[2 * kSegCap];
    const int64_t t = xcd_chunk_order(blockIdx.x);
    if (t >= ntiles) return;
    const int32_t ct = (int32_t)(t * kTileCells);
    const __amdgpu_buffer_rsrc_t rs = field_rsrc(u, (int64_t)C * (int64_t)sizeof(T));
    typename Vec<T>::type cv[NG];
    T lv[NG], rv[NG];
#pragma unroll
    for (int q = 0; q < NG; ++q) {  // reads past C return 0
        const int32_t c = ct + (q * kThreads + (int32_t)threadIdx.x) * G;
        cv[q] = gather_vec<T>(rs, c);
        lv[q] = gather<T>(rs, c - 1);
        rv[q] = gather<T>(rs, c + G);
    }
    const int32_t s0 = plan[2 * t], s1 = plan[2 * t + 1];
    const int32_t M = s1 - s0;
    const int4* gs = (const int4*)segs;
    bool outside = false;
    if (M <= kSegCap) {
        for (int q = threadIdx.x; q < 2 * M; q += kThreads) s_seg[q] = gs[2 * (int64_t)s0 + q];
        __syncthreads();
        tag_thread<T>(LdsSegs{s_seg, s0}, s0, s1, ct, C, rs, u, smin, cv, lv, rv, tag_w, lex_w, W, outside);
    } else {
        tag_thread<T>(GlobalSegs{gs}, s0, s1, ct, C, rs, u, smin, cv, lv, rv, tag_w, lex_w, W, outside);
    }
    if (outside) atomicOr(status, STX_FLAG_DOMAIN);
}

// Pass 2: run start / end words and the start count of every word.
__global__ void tag_runs_kernel(const uint32_t* __restrict__ tag_w, const uint32_t* __restrict__ lex_w, int64_t W,
                                uint32_t* __restrict__ st_w, uint32_t* __restrict__ en_w, int32_t* __restrict__ cnt) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const uint32_t T = tag_w[w], L = lex_w[w];
    const uint32_t tp = w > 0 ? tag_w[w - 1] >> 31 : 0u;
    const uint32_t tn = w + 1 < W ? tag_w[w + 1] & 1u : 0u;
    const uint32_t ln = w + 1 < W ? lex_w[w + 1] & 1u : 0u;
    const uint32_t S = T & ~(((T << 1) | tp) & L);
    const uint32_t rx = (L >> 1) | (ln << 31);  // cell has a right neighbour
    const uint32_t E = T & ~(((T >> 1) | (tn << 31)) & rx);
    st_w[w] = S;
    en_w[w] = E;
    cnt[w] = __popc(S);
}

// Column of cell c from its segment (plan tile, then a binary search).
__device__ __forceinline__ int32_t cell_x(const int32_t* __restrict__ plan, const int32_t* __restrict__ segs,
                                          int32_t c) {
    const int64_t t = c / kTileCells;
    int32_t lo = plan[2 * t], hi = plan[2 * t + 1];
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (segs[kRec * (int64_t)mid] <= c) lo = mid; else hi = mid;
    }
    return segs[kRec * (int64_t)lo + 6] + (c - segs[kRec * (int64_t)lo]);
}

// Pass 3a: begins / ends of the runs.  pref: exclusive scan of the start
// counts; a run still open at the first cell of word w (its last cell in an
// earlier word) shifts the end ranks of the word by one.
__global__ void tag_write_kernel(const uint32_t* __restrict__ tag_w, const uint32_t* __restrict__ st_w,
                                 const uint32_t* __restrict__ en_w, const int32_t* __restrict__ pref, int64_t W,
                                 const int32_t* __restrict__ plan, const int32_t* __restrict__ segs,
                                 int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t capacity,
                                 uint32_t* __restrict__ status) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    uint32_t S = st_w[w], E = en_w[w];
    if ((S | E) == 0u) return;
    const int64_t base = pref[w];
    const int64_t open = (w > 0 && (tag_w[w - 1] >> 31) && !(en_w[w - 1] >> 31)) ? 1 : 0;
    const int32_t c0 = (int32_t)(w << 5);
    for (int64_t r = base; S; ++r) {
        const int b = __ffs((int)S) - 1;
        S &= S - 1u;
        if (r < capacity) ob[r] = cell_x(plan, segs, c0 + b);
        else atomicOr(status, STX_FLAG_CAPACITY);
    }
    for (int64_t r = base - open; E; ++r) {
        const int b = __ffs((int)E) - 1;
        E &= E - 1u;
        if (r < capacity) oe[r] = cell_x(plan, segs, c0 + b) + 1;
        else atomicOr(status, STX_FLAG_CAPACITY);
    }
}

__global__ void zero_total_kernel(int32_t* __restrict__ p) { *p = 0; }

// Pass 3b: out_off[k] = starts before the first cell of row k (k = R: all).
__global__ void tag_rows_kernel(const int32_t* __restrict__ off, const int32_t* __restrict__ co, int64_t R,
                                int32_t C, const uint32_t* __restrict__ st_w, const int32_t* __restrict__ pref,
                                const int32_t* __restrict__ cnt, int64_t W, int32_t* __restrict__ out_off) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > R) return;
    const int32_t total = W > 0 ? pref[W - 1] + cnt[W - 1] : 0;
    const int32_t c = k == R ? C : co[off[k]];
    const int64_t w = (int64_t)c >> 5;
    const int b = c & 31;
    out_off[k] = w >= W ? total : pref[w] + __popc(st_w[w] & ((1u << b) - 1u));
}

template <typename T, bool BATCH>
int32_t launch_upwind_seg(const int32_t* plan, int64_t ntiles, int64_t C, const int32_t* segs, const void* u,
                          void* out, double a, double b, double dt, double dx, double dy, const void* inst_coef,
                          hipStream_t stream) {
    const int64_t grid = (ntiles + kOrderSpan - 1) / kOrderSpan * kOrderSpan;
    if (grid > 0x7fffffffLL) return stx_fail(STX_EINVAL, "upwind_seg: too many tiles");
    if ((uintptr_t)u % 16 != 0 || (uintptr_t)out % 16 != 0)
        return stx_fail(STX_EINVAL, "upwind_seg: values and output must be 16-byte aligned");
    hipLaunchKernelGGL((upwind_seg_kernel<T, BATCH>), dim3((unsigned)grid), dim3(seg_threads<T>()), 0, stream, plan, ntiles,
                       (int32_t)C, segs, (const T*)u, (T*)out, (T)a, (T)b, (T)dt, (T)dx, (T)dy, (const T*)inst_coef);
    return stx_launch_status("upwind_seg");
}

}  // namespace

extern "C" int32_t stx_scan_i32(const int32_t* in, const int32_t* d_n, int64_t n_bound, int32_t* out, int32_t* d_total,
                                uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch,
                                hipStream_t stream);

extern "C" {

int64_t stx_stencil_tile_cells(void) { return kTileCells; }

int64_t stx_stencil_plan_tiles(int64_t cells) { return (cells + kTileCells - 1) / kTileCells; }

int32_t stx_stencil_seg_count(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg,
                              const int32_t* end, const int32_t* co, const int32_t* ord, int64_t N, int32_t width,
                              const int32_t* row_inst, const int32_t* inst_w, int32_t* counts, int32_t* seg_off,
                              int32_t* d_total, uint32_t* status, void* scratch, int64_t scratch_bytes,
                              uint32_t epoch, hipStream_t stream) {
    if (N < 0 || R < 0) return stx_fail(STX_EINVAL, "stencil_seg_count: bad sizes");
    if (N > 0 && R == 0) return stx_fail(STX_EINVAL, "stencil_seg_count: intervals without rows");
    if ((row_inst == nullptr) != (inst_w == nullptr))
        return stx_fail(STX_EINVAL, "stencil_seg_count: row_inst and inst_w go together");
    if (row_inst == nullptr && width <= 0) return stx_fail(STX_EINVAL, "stencil_seg_count: width must be positive");
    if (N > 0)
        hipLaunchKernelGGL(segments_kernel<false>, dim3(nblocks(N)), dim3(256), 0, stream, rows, off, R, beg, end, co,
                           ord, N, width, row_inst, inst_w, counts, nullptr, nullptr);
    int32_t rc = stx_launch_status("stencil_seg_count");
    if (rc) return rc;
    return stx_scan_i32(counts, nullptr, N, seg_off, d_total, status, scratch, scratch_bytes, epoch, stream);
}

int32_t stx_stencil_seg_build(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg,
                              const int32_t* end, const int32_t* co, const int32_t* ord, int64_t N, int32_t width,
                              const int32_t* row_inst, const int32_t* inst_w, const int32_t* seg_off, int64_t nseg,
                              int32_t* segs, int64_t C, int32_t* plan, int64_t ntiles, hipStream_t stream) {
    if (N < 0 || R < 0 || nseg < 0 || ntiles < 0) return stx_fail(STX_EINVAL, "stencil_seg_build: bad sizes");
    if ((row_inst == nullptr) != (inst_w == nullptr))
        return stx_fail(STX_EINVAL, "stencil_seg_build: row_inst and inst_w go together");
    if (ntiles != (C + kTileCells - 1) / kTileCells) return stx_fail(STX_EINVAL, "stencil_seg_build: tile count");
    if (N > 0)
        hipLaunchKernelGGL(segments_kernel<true>, dim3(nblocks(N)), dim3(256), 0, stream, rows, off, R, beg, end, co,
                           ord, N, width, row_inst, inst_w, nullptr, seg_off, segs);
    if (ntiles > 0) {
        if (nseg <= 0) return stx_fail(STX_EINVAL, "stencil_seg_build: cells without segments");
        hipLaunchKernelGGL(seg_plan_kernel, dim3(nblocks(ntiles)), dim3(256), 0, stream, segs, nseg, C, ntiles, plan);
    }
    return stx_launch_status("stencil_seg_build");
}

int32_t stx_upwind_seg(int32_t dtype, const int32_t* plan, int64_t ntiles, int64_t C, const int32_t* segs,
                       const void* u, void* out, double a, double b, double dt, double dx, double dy,
                       const void* inst_coef, hipStream_t stream) {
    if (ntiles <= 0 || C <= 0) return STX_OK;
    if (C > 0x7fffffffLL - 2 * kTileCells) return stx_fail(STX_EINVAL, "upwind_seg: too many cells");
    if (C * (dtype == 8 ? 8 : 4) > 0xffffff00LL)
        return stx_fail(STX_EINVAL, "upwind_seg: field larger than 4 GiB (use the row-group kernel)");
    if (ntiles != (C + kTileCells - 1) / kTileCells) return stx_fail(STX_EINVAL, "upwind_seg: plan size mismatch");
    const bool batch = inst_coef != nullptr;
    if (dtype == 4)
        return batch ? launch_upwind_seg<float, true>(plan, ntiles, C, segs, u, out, a, b, dt, dx, dy, inst_coef, stream)
                     : launch_upwind_seg<float, false>(plan, ntiles, C, segs, u, out, a, b, dt, dx, dy,
                                                       inst_coef, stream);
    if (dtype == 8)
        return batch ? launch_upwind_seg<double, true>(plan, ntiles, C, segs, u, out, a, b, dt, dx, dy,
                                                       inst_coef, stream)
                     : launch_upwind_seg<double, false>(plan, ntiles, C, segs, u, out, a, b, dt, dx, dy,
                                                        inst_coef, stream);
    return stx_fail(STX_EINVAL, "upwind_seg: dtype must be 4 or 8");
}

int64_t stx_grad_tag_seg_words(int64_t cells) { return (cells + 31) / 32; }

int32_t stx_grad_tag_seg(int32_t dtype, const int32_t* plan, int64_t ntiles, int64_t C, const int32_t* segs,
                         const int32_t* off, const int32_t* co, int64_t R, const void* u, double t, uint32_t* words,
                         int64_t W, int32_t* out_off, int32_t* out_beg, int32_t* out_end, int64_t capacity,
                         int32_t* d_total, uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch,
                         hipStream_t stream) {
    if (C < 0 || R < 0 || ntiles < 0) return stx_fail(STX_EINVAL, "grad_tag_seg: bad sizes");
    if (W != (C + 31) / 32) return stx_fail(STX_EINVAL, "grad_tag_seg: word count must be ceil(cells / 32)");
    if (ntiles != (C + kTileCells - 1) / kTileCells) return stx_fail(STX_EINVAL, "grad_tag_seg: plan size mismatch");
    if (C > 0x7fffffffLL - 2 * kTileCells) return stx_fail(STX_EINVAL, "grad_tag_seg: too many cells");
    if (C * (dtype == 8 ? 8 : 4) > 0xffffff00LL)
        return stx_fail(STX_EINVAL, "grad_tag_seg: field larger than 4 GiB (use the row-group kernel)");
    if ((uintptr_t)u % 16 != 0) return stx_fail(STX_EINVAL, "grad_tag_seg: values must be 16-byte aligned");
    // words (6 W + 1): tag, lex, start, end, count (W each), prefix (W + 1)
    uint32_t* tag_w = words;
    uint32_t* lex_w = words + W;
    uint32_t* st_w = words + 2 * W;
    uint32_t* en_w = words + 3 * W;
    int32_t* cnt = (int32_t*)(words + 4 * W);
    int32_t* pref = (int32_t*)(words + 5 * W);
    if (W > 0) {
        const int64_t grid = (ntiles + kOrderSpan - 1) / kOrderSpan * kOrderSpan;
        if (dtype == 4)
            hipLaunchKernelGGL(tag_bits_kernel<float>, dim3((unsigned)grid), dim3(kThreads), 0, stream, plan, ntiles,
                               (int32_t)C, segs, (const float*)u, sqrt_threshold<float, uint32_t>((float)t), tag_w,
                               lex_w, W, status);
        else if (dtype == 8)
            hipLaunchKernelGGL(tag_bits_kernel<double>, dim3((unsigned)grid), dim3(kThreads), 0, stream, plan, ntiles,
                               (int32_t)C, segs, (const double*)u, sqrt_threshold<double, uint64_t>(t), tag_w, lex_w,
                               W, status);
        else
            return stx_fail(STX_EINVAL, "grad_tag_seg: dtype must be 4 or 8");
        hipLaunchKernelGGL(tag_runs_kernel, dim3(nblocks(W)), dim3(256), 0, stream, tag_w, lex_w, W, st_w, en_w, cnt);
        int32_t rc = stx_launch_status("grad_tag_seg");
        if (rc) return rc;
        rc = stx_scan_i32(cnt, nullptr, W, pref, d_total, status, scratch, scratch_bytes, epoch, stream);
        if (rc) return rc;
        hipLaunchKernelGGL(tag_write_kernel, dim3(nblocks(W)), dim3(256), 0, stream, tag_w, st_w, en_w, pref, W, plan,
                           segs, out_beg, out_end, capacity, status);
    } else {
        hipLaunchKernelGGL(zero_total_kernel, dim3(1), dim3(1), 0, stream, d_total);
    }
    hipLaunchKernelGGL(tag_rows_kernel, dim3(nblocks(R + 1)), dim3(256), 0, stream, off, co, R, (int32_t)C, st_w, pref,
                       cnt, W, out_off);
    return stx_launch_status("grad_tag_seg");
}

}  // extern "C"
