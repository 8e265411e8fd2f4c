// Interval-set algebra on int32 CSR sets for gfx950.
//
// Replaces the reference's one-thread-per-row serial walks
// (subsetix_cupy/kernels.py:11-236 count_segments / write_segments, and the
// sort + merge rebuild kernels.py:292-381) with a lane-parallel merge:
//
//   * Each output row is handled by a G-lane group (G = 64: one wavefront per
//     row; G < 64: several short rows per wavefront).  The group walks the
//     row's merged order (A and B merged by key) G positions at a time; lane l
//     finds the element at merged position l by a co-rank (merge-path)
//     binary search.
//   * union: a group max-scan of interval ends gives, for every element, the
//     largest end before it; an element starts an output run when its begin
//     is above that value (the reference's `seg_start <= active_end` merge
//     rule).  This handles operands whose intervals overlap (widened or
//     coarsened inputs), so the same code is the "rebuild" merge.
//   * intersection / difference: every output piece starts either at an A
//     element or at a B element; whether it does, and where the piece ends,
//     follows from the element's neighbours in merged order.  Exact for
//     operands whose rows are sorted and non-overlapping (the reference's
//     documented input contract, expressions.py:420-425).
//   * count, scan and write run in ONE kernel: a block counts its tile of
//     rows, takes its global output offset from a decoupled look-back over
//     earlier tiles, then writes (the second walk reads from L1/L2).
#include "stx_common.h"
#include "stx_fused.h"

#include <limits.h>

#include <mutex>

using namespace stx;

namespace {

struct Operand {
    const int32_t* beg;
    const int32_t* end;
    const int32_t* off;  // nullptr: operand is empty
    const int32_t* map;  // output row -> source row (-1 = empty); nullptr = identity
    XForm xf;
};

__device__ __forceinline__ void row_extent(const Operand& X, int64_t u, int32_t& x0, int32_t& nx) {
    x0 = 0;
    nx = 0;
    if (X.off == nullptr) return;
    int64_t src = X.map ? (int64_t)X.map[u] : u;
    if (src < 0) return;
    x0 = X.off[src];
    nx = X.off[src + 1] - x0;
}

__device__ __forceinline__ int32_t keyb(const Operand& X, int32_t k) {
    int32_t b = X.beg[k], e = 0;
    XForm xf = X.xf;
    if (xf.mode == XF_WIDEN) return b - xf.p;
    if (xf.mode == XF_COARSEN) return floordiv(b, xf.p);
    if (xf.mode == XF_CLIP) return b < 0 ? 0 : b;
    if (xf.mode == XF_SCALE) return b * xf.p;
    (void)e;
    return b;
}

__device__ __forceinline__ int32_t keye(const Operand& X, int32_t k) {
    int32_t e = X.end[k];
    XForm xf = X.xf;
    if (xf.mode == XF_WIDEN) return e + xf.p;
    if (xf.mode == XF_COARSEN) return floordiv(e + xf.p - 1, xf.p);
    if (xf.mode == XF_CLIP) return e > xf.p ? xf.p : e;
    if (xf.mode == XF_SCALE) return e * xf.p;
    return e;
}

// true if A element ia precedes B element jb in the merged order of `OP`
template <int OP>
__device__ __forceinline__ bool a_first(const Operand& A, int32_t ia, const Operand& B, int32_t jb) {
    if (OP == 2) return keyb(A, ia) < keye(B, jb);  // difference: A by begin, B by end, B first on ties
    return keyb(A, ia) <= keyb(B, jb);             // union / intersection: both by begin, A first
}

// Processes one output row with a G-lane group.  Returns the number of
// output intervals (uniform across the group).  With WRITE, writes them at
// ob/oe[base ...].
template <int G, int OP, bool WRITE>
__device__ __forceinline__ int32_t setop_row(const Operand& A, int32_t a0, int32_t na, const Operand& B, int32_t b0, int32_t nb,
                             int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t base) {
    const int gl = lane_id() & (G - 1);
    const int32_t n = na + nb;
    int32_t count = 0;
    int32_t carry_pm = INT_MIN;
    for (int32_t c = 0; c < n; c += G) {
        const int32_t l = c + gl;
        const bool valid = l < n;
        bool emit = false;
        int32_t pb = 0, pe = 0;
        int32_t eb = 0, ee = INT_MIN;
        bool alive = false;
        if (valid) {
            int32_t lo = l - nb > 0 ? l - nb : 0;
            int32_t hi = l < na ? l : na;
            while (lo < hi) {
                int32_t mid = (lo + hi) >> 1;
                if (a_first<OP>(A, a0 + mid, B, b0 + (l - 1 - mid))) lo = mid + 1;
                else hi = mid;
            }
            const int32_t i = lo, j = l - lo;
            const bool isA = (j >= nb) || (i < na && a_first<OP>(A, a0 + i, B, b0 + j));
            if (OP == 0) {
                if (isA) { eb = keyb(A, a0 + i); ee = keye(A, a0 + i); }
                else { eb = keyb(B, b0 + j); ee = keye(B, b0 + j); }
                const bool clip_a = A.xf.mode == XF_CLIP, clip_b = B.xf.mode == XF_CLIP;
                alive = !((isA ? clip_a : clip_b) && eb >= ee);
                if (!alive) ee = INT_MIN;
            } else if (OP == 1) {
                if (isA) {
                    const int32_t ab = keyb(A, a0 + i), ae = keye(A, a0 + i);
                    int32_t ce = INT_MIN;
                    if (j >= 1 && keye(B, b0 + j - 1) > ab) ce = keye(B, b0 + j - 1);
                    else if (j < nb && keyb(B, b0 + j) == ab) ce = keye(B, b0 + j);
                    if (ce != INT_MIN) { emit = true; pb = ab; pe = ae < ce ? ae : ce; }
                } else {
                    const int32_t bb = keyb(B, b0 + j), be = keye(B, b0 + j);
                    if (i >= 1) {
                        const int32_t pab = keyb(A, a0 + i - 1), pae = keye(A, a0 + i - 1);
                        if (pab < bb && bb < pae) { emit = true; pb = bb; pe = be < pae ? be : pae; }
                    }
                }
            } else {
                if (isA) {
                    const int32_t ab = keyb(A, a0 + i), ae = keye(A, a0 + i);
                    const bool covered = j < nb && keyb(B, b0 + j) <= ab;
                    if (!covered) {
                        emit = true;
                        pb = ab;
                        pe = ae;
                        if (j < nb) { const int32_t nbg = keyb(B, b0 + j); pe = pe < nbg ? pe : nbg; }
                    }
                } else {
                    const int32_t be = keye(B, b0 + j);
                    if (i >= 1) {
                        const int32_t pae = keye(A, a0 + i - 1);
                        const bool open = (j + 1 >= nb) || keyb(B, b0 + j + 1) > be;
                        if (be < pae && open) {
                            emit = true;
                            pb = be;
                            pe = pae;
                            if (j + 1 < nb) { const int32_t nbg = keyb(B, b0 + j + 1); pe = pe < nbg ? pe : nbg; }
                        }
                    }
                }
            }
        }
        if (OP == 0) {
            int32_t pm_incl = group_scan_max<G>(ee);
            pm_incl = pm_incl > carry_pm ? pm_incl : carry_pm;
            int32_t pm_excl = __shfl_up(pm_incl, 1, G);
            if (gl == 0) pm_excl = carry_pm;
            const bool start = valid && alive && eb > pm_excl;
            const uint64_t m = group_ballot<G>(start);
            if (WRITE && start) {
                const int32_t k = count + rank_below(m, gl);
                ob[base + k] = eb;
                if (k >= 1) oe[base + k - 1] = pm_excl;
            }
            count += __popcll(m);
            carry_pm = __shfl(pm_incl, G - 1, G);
        } else {
            const uint64_t m = group_ballot<G>(emit);
            if (WRITE && emit) {
                const int32_t k = count + rank_below(m, gl);
                ob[base + k] = pb;
                oe[base + k] = pe;
            }
            count += __popcll(m);
        }
    }
    if (OP == 0 && WRITE && count >= 1 && gl == 0) oe[base + count - 1] = carry_pm;
    return count;
}

template <int OP>
struct SetopRow {
    Operand A, B;
    template <int G, bool WRITE>
    __device__ int32_t run(int64_t u, int32_t* ob, int32_t* oe, int64_t base) const {
        int32_t a0, na, b0, nb;
        row_extent(A, u, a0, na);
        row_extent(B, u, b0, nb);
        return setop_row<G, OP, WRITE>(A, a0, na, B, b0, nb, ob, oe, base);
    }
};


// ---------------------------------------------------------------------------
// Tile kernel (identity row maps): LDS-staged merge path.
//
// A block owns RPT consecutive rows.  Their A and B intervals are contiguous
// in memory; the block copies them (transformed keys) into LDS with
// coalesced loads.  The tile's merged positions (row by row, A and B merged
// by key) are split evenly: thread t owns positions [t*VT, t*VT + VT).  A
// thread finds its first element with one binary search (row) and one
// co-rank search (inside the row), then walks its positions with a
// two-pointer merge, keeping each position's result in registers.  Union
// needs the largest end before each position (within the row): a segmented
// max-scan over the threads' aggregates.  Emission counts are scanned over
// the block, the tile offset comes from the decoupled look-back, and the
// intervals are written in tile order (contiguous).  A tile whose rows hold
// more than PCAP intervals falls back to one wavefront per row.
// Tile shapes: TB threads, VT merged positions per thread (capacity TB * VT
// positions per tile).  512 x 8 keeps 4 tiles (32 waves) on a CU; the
// narrower shapes amortise each thread's searches and the block scans over
// more positions with fewer barriers, at fewer waves per CU.
template <int VT>
struct MaskOf {
    typedef uint32_t type;
};
template <>
struct MaskOf<8> {
    typedef uint8_t type;
};
template <>
struct MaskOf<16> {
    typedef uint16_t type;
};
template <int TB>
constexpr int max_rows_per_tile() {
    return TB < 256 ? TB : 256;  // the fallback scans one row count per thread
}
template <int VT>
constexpr int tile_waves_per_eu() {
    return VT <= 8 ? 8 : (VT <= 16 ? 5 : 3);
}

struct SegMax {
    int32_t v;
    int32_t f;
};

__device__ __forceinline__ SegMax seg_combine(SegMax a, SegMax b) {  // a before b
    SegMax r;
    r.f = a.f | b.f;
    r.v = b.f ? b.v : (a.v > b.v ? a.v : b.v);
    return r;
}

// Block-wide exclusive segmented max-scan (one element per thread).
__device__ __forceinline__ SegMax block_seg_max_exclusive(SegMax x, SegMax* lds /* >= waves of the block */) {
    const int lane = lane_id(), wave = wave_id();
    SegMax inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        SegMax o;
        o.v = __shfl_up(inc.v, d, 64);
        o.f = __shfl_up(inc.f, d, 64);
        if (lane >= d) inc = seg_combine(o, inc);
    }
    __syncthreads();
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    SegMax carry{INT_MIN, 0};
    for (int w = 0; w < wave; ++w) carry = seg_combine(carry, lds[w]);
    SegMax ex;
    ex.v = __shfl_up(inc.v, 1, 64);
    ex.f = __shfl_up(inc.f, 1, 64);
    if (lane == 0) ex = SegMax{INT_MIN, 0};
    return seg_combine(carry, ex);
}

// s_k[idx] when ok, else (0, INT_MIN); the LDS read is unconditional (index
// 0 when not ok) so the select is on values, never on addresses.
__device__ __forceinline__ int2 lds_or_none(const int2* s, int32_t idx, bool ok) {
    const int2 v = s[ok ? idx : 0];
    return ok ? v : make_int2(0, INT_MIN);
}

template <int OP>
__device__ __forceinline__ bool key_a_first(int32_t akb, int32_t bkb, int32_t bke) {
    if (OP == 2) return akb < bke;
    return akb <= bkb;
}

// Copies intervals [0, n) of (b, e) into dst as (begin, end) pairs: each
// thread moves 4 consecutive intervals with one 16-byte load per array (the
// arrays are only 4-byte aligned) and two 16-byte LDS stores (dst 16-byte
// aligned); the last partial group element by element.
template <int TB>
__device__ __forceinline__ void stage_pairs(const int32_t* __restrict__ b, const int32_t* __restrict__ e, int32_t n,
                                            int2* dst, int tid) {
    for (int32_t k = tid * 4; k < n; k += TB * 4) {
        if (k + 4 <= n) {
            int4 vb, ve;
            __builtin_memcpy(&vb, __builtin_assume_aligned(b + k, 4), 16);
            __builtin_memcpy(&ve, __builtin_assume_aligned(e + k, 4), 16);
            int4* d = (int4*)(dst + k);
            d[0] = make_int4(vb.x, ve.x, vb.y, ve.y);
            d[1] = make_int4(vb.z, ve.z, vb.w, ve.w);
        } else {
            for (int32_t q = k; q < n; ++q) dst[q] = make_int2(b[q], e[q]);
        }
    }
}

// MODE: kFused = count, look-back, write in one launch.  Two-pass form:
// kCount publishes the tile's output count as its look-back state word
// (aggregate) and stops; prefix_states_kernel turns all aggregates into
// inclusive prefixes; kWrite reads tile - 1's prefix (one load, never a
// wait) and writes.  It trades a second read of the inputs for tiles that
// never wait on each other; every scratch word stays epoch-tagged.
constexpr int kFused = 0, kCount = 1, kWrite = 2;


template <int OP, int MODE, bool XF, int TB, int VT>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(tile_waves_per_eu<VT>()))) void setop_tile_kernel(Operand A, Operand B, int64_t U, int32_t RPT, uint32_t ntiles,
                                                            int32_t* __restrict__ out_off, int32_t* __restrict__ ob,
                                                            int32_t* __restrict__ oe, int64_t capacity,
                                                            int32_t* __restrict__ d_total, uint32_t* status,
                                                            void* scratch, uint32_t epoch, int64_t spin_limit) {
    constexpr int kTW = TB / STX_WAVE;
    constexpr int kPCAP = TB * VT;
    constexpr int kMaxRPT = max_rows_per_tile<TB>();
    typedef typename MaskOf<VT>::type MaskT;
    __shared__ __align__(16) int2 s_k[kPCAP];  // (begin, end) of the tile's A then B intervals
    __shared__ int32_t s_aS[kMaxRPT + 1];
    __shared__ int32_t s_bS[kMaxRPT + 1];
    __shared__ int32_t s_rs[kMaxRPT + 1];  // merged position where each row starts
    // per thread: emissions before its first position, and its emission mask
    // (fallback path: per-row counts, then their exclusive scan)
    __shared__ int32_t s_tbase[TB + 1];
    __shared__ MaskT s_tmask[TB];
    __shared__ int32_t s_scan[17];
    __shared__ SegMax s_seg[kTW];
    __shared__ uint32_t s_excl;

    LookbackScratch lb = lookback_scratch(scratch, ntiles);
    const uint32_t tile = blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t row0 = (int64_t)tile * RPT;
    const int nr = (int)(U - row0 < RPT ? (U - row0 > 0 ? U - row0 : 0) : RPT);

    // row0 <= U for every tile (ntiles = floor(U / RPT) + 1), so off[row0]
    // exists; the trailing tile (nr == 0 when RPT divides U) then sees P = 0.
    // The tile's extent comes straight from the offsets, so that the row
    // offsets and the intervals are loaded together before one barrier.
    const int32_t a_base = A.off[row0];
    const int32_t b_base = B.off ? B.off[row0] : 0;
    const int32_t nA = A.off[row0 + nr] - a_base;
    const int32_t nB = B.off ? B.off[row0 + nr] - b_base : 0;
    const int32_t P = nA + nB;
    // B is staged from an even slot so that its pairs are 16-byte aligned
    const int32_t bOff = (nA + 1) & ~1;
    const bool fits = bOff + nB <= kPCAP;
    for (int r = tid; r <= nr; r += TB) {
        int32_t as = A.off[row0 + r] - a_base;
        int32_t bs = B.off ? B.off[row0 + r] - b_base : 0;
        s_aS[r] = as;
        s_bS[r] = bs;
        s_rs[r] = as + bs;
    }

    int32_t tile_total = 0;
    // per-position results (registers)
    int32_t rb[VT], re[VT];
    uint32_t emit_mask = 0, start_mask = 0, alive_mask = 0, rowstart_mask = 0, rowend_mask = 0;
    SegMax carry{INT_MIN, 0};
    int32_t my_base = 0;

    if (fits) {
        if (XF) {
            for (int k = tid; k < nA; k += TB) s_k[k] = make_int2(keyb(A, a_base + k), keye(A, a_base + k));
            for (int k = tid; k < nB; k += TB) s_k[bOff + k] = make_int2(keyb(B, b_base + k), keye(B, b_base + k));
        } else {
            stage_pairs<TB>(A.beg + a_base, A.end + a_base, nA, s_k, tid);
            if (nB > 0) stage_pairs<TB>(B.beg + b_base, B.end + b_base, nB, s_k + bOff, tid);
        }
    }
    __syncthreads();  // row offsets and staged intervals
    if (fits) {
        const int32_t p0 = tid * VT;
        if (p0 < P) {
            // row of p0
            int lo = 0, hi = nr;
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (s_rs[mid] <= p0) lo = mid; else hi = mid;
            }
            int r = lo;
            int32_t aS = s_aS[r], na = s_aS[r + 1] - aS;
            int32_t bS = bOff + s_bS[r], nb = s_bS[r + 1] - s_bS[r];
            int32_t rs = s_rs[r];
            int32_t rowend = s_rs[r + 1];
            // co-rank inside the row
            int32_t l = p0 - rs;
            int32_t clo = l - nb > 0 ? l - nb : 0, chi = l < na ? l : na;
            while (clo < chi) {
                int32_t mid = (clo + chi) >> 1;
                const int2 kb = s_k[bS + (l - 1 - mid)];
                if (key_a_first<OP>(s_k[aS + mid].x, kb.x, kb.y)) clo = mid + 1;
                else chi = mid;
            }
            int32_t i = clo, j = l - clo;
            // The walk keeps the current and previous element of each operand
            // in registers (as scalars: selecting between int2 values makes
            // the compiler select between their addresses) and is branch-free:
            // each position picks its element by value, computes the emission
            // of both cases, and reads only the next element of the side it
            // consumes (one LDS read per position).
            int32_t cAx, cAy, cBx, cBy, pAx, pAy, pBx, pBy;
            {
                const int2 t0 = s_k[i < na ? aS + i : 0];
                const int2 t1 = s_k[j < nb ? bS + j : 0];
                const int2 t2 = s_k[i >= 1 ? aS + i - 1 : 0];
                const int2 t3 = s_k[j >= 1 ? bS + j - 1 : 0];
                cAx = i < na ? t0.x : 0;
                cAy = i < na ? t0.y : INT_MIN;
                cBx = j < nb ? t1.x : 0;
                cBy = j < nb ? t1.y : INT_MIN;
                pAx = i >= 1 ? t2.x : 0;
                pAy = i >= 1 ? t2.y : INT_MIN;
                pBx = j >= 1 ? t3.x : 0;
                pBy = j >= 1 ? t3.y : INT_MIN;
            }
#pragma unroll
            for (int v = 0; v < VT; ++v) {
                const int32_t p = p0 + v;
                rb[v] = 0;
                re[v] = INT_MIN;
                if (p < P) {
                    if (p >= rowend) {
                        do {  // next non-empty row
                            ++r;
                            aS = s_aS[r];
                            na = s_aS[r + 1] - aS;
                            bS = bOff + s_bS[r];
                            nb = s_bS[r + 1] - s_bS[r];
                            rs = rowend;
                            rowend = s_rs[r + 1];
                        } while (p >= rowend);
                        i = 0;
                        j = 0;
                        const int2 t0 = s_k[na > 0 ? aS : 0];
                        const int2 t1 = s_k[nb > 0 ? bS : 0];
                        cAx = na > 0 ? t0.x : 0;
                        cAy = na > 0 ? t0.y : INT_MIN;
                        cBx = nb > 0 ? t1.x : 0;
                        cBy = nb > 0 ? t1.y : INT_MIN;
                    }
                    rowstart_mask |= (p == rs ? 1u : 0u) << v;
                    rowend_mask |= (p == rowend - 1 ? 1u : 0u) << v;
                    const bool isA = (j >= nb) || (i < na && key_a_first<OP>(cAx, cBx, cBy));
                    // next element of the consumed side
                    const bool nvalid = isA ? (i + 1 < na) : (j + 1 < nb);
                    const int2 nraw = s_k[nvalid ? (isA ? aS + i + 1 : bS + j + 1) : 0];
                    const int32_t nxx = nvalid ? nraw.x : 0;
                    const int32_t nxy = nvalid ? nraw.y : INT_MIN;
                    if (OP == 0) {
                        const int32_t ex = isA ? cAx : cBx;
                        const int32_t ey = isA ? cAy : cBy;
                        const bool clip = XF && (isA ? A.xf.mode == XF_CLIP : B.xf.mode == XF_CLIP);
                        alive_mask |= ((clip && ex >= ey) ? 0u : 1u) << v;
                        rb[v] = ex;
                        re[v] = ey;
                    } else if (OP == 1) {
                        // A element: clipped by the B element covering its begin
                        const int32_t ceA = (j >= 1 && pBy > cAx) ? pBy : ((j < nb && cBx == cAx) ? cBy : INT_MIN);
                        const bool emA = ceA != INT_MIN;
                        // B element: inside the previous A element
                        const bool emB = i >= 1 && pAx < cBx && cBx < pAy;
                        emit_mask |= ((isA ? emA : emB) ? 1u : 0u) << v;
                        rb[v] = isA ? cAx : cBx;
                        re[v] = isA ? (cAy < ceA ? cAy : ceA) : (cBy < pAy ? cBy : pAy);
                    } else {
                        // A element not covered at its begin: runs to the next B begin
                        const bool emA = !(j < nb && cBx <= cAx);
                        const int32_t peA = (j < nb && cBx < cAy) ? cBx : cAy;
                        // B element ending inside the previous A element
                        const bool openB = (j + 1 >= nb) || nxx > cBy;
                        const bool emB = i >= 1 && cBy < pAy && openB;
                        const int32_t peB = (j + 1 < nb && nxx < pAy) ? nxx : pAy;
                        emit_mask |= ((isA ? emA : emB) ? 1u : 0u) << v;
                        rb[v] = isA ? cAx : cBy;
                        re[v] = isA ? peA : peB;
                    }
                    pAx = isA ? cAx : pAx;
                    pAy = isA ? cAy : pAy;
                    pBx = isA ? pBx : cBx;
                    pBy = isA ? pBy : cBy;
                    cAx = isA ? nxx : cAx;
                    cAy = isA ? nxy : cAy;
                    cBx = isA ? cBx : nxx;
                    cBy = isA ? cBy : nxy;
                    i += isA ? 1 : 0;
                    j += isA ? 0 : 1;
                }
            }
        }
        if (OP == 0) {
            // thread aggregate: max end since the last row start in the thread
            SegMax agg{INT_MIN, 0};
#pragma unroll
            for (int v = 0; v < VT; ++v) {
                if (rowstart_mask & (1u << v)) agg = SegMax{INT_MIN, 1};
                if (alive_mask & (1u << v)) agg.v = agg.v > re[v] ? agg.v : re[v];
            }
            carry = block_seg_max_exclusive(agg, s_seg);
            int32_t pm = carry.v;
#pragma unroll
            for (int v = 0; v < VT; ++v) {
                if (rowstart_mask & (1u << v)) pm = INT_MIN;
                if ((alive_mask & (1u << v)) && rb[v] > pm) start_mask |= 1u << v;
                if (alive_mask & (1u << v)) pm = pm > re[v] ? pm : re[v];
            }
            emit_mask = start_mask;
        }
        const int32_t my_count = __popc(emit_mask);
        my_base = block_exclusive_scan<int32_t>(my_count, s_scan, &tile_total);
        s_tbase[tid] = my_base;
        s_tmask[tid] = (MaskT)emit_mask;
    } else {
        // fallback: one wavefront per row, counts into s_tbase[r]
        const int wave = wave_id();
        for (int r = wave; r < nr; r += kTW) {
            int32_t a0, na, b0, nb;
            row_extent(A, row0 + r, a0, na);
            row_extent(B, row0 + r, b0, nb);
            int32_t c = setop_row<64, OP, false>(A, a0, na, B, b0, nb, nullptr, nullptr, 0);
            if (lane_id() == 0) s_tbase[r] = c;
        }
        __syncthreads();
        int32_t v = tid < nr ? s_tbase[tid] : 0;
        int32_t ex = 0;
        // nr <= kMaxRPT <= TB
        ex = block_exclusive_scan<int32_t>(v, s_scan, &tile_total);
        __syncthreads();
        if (tid < nr) s_tbase[tid] = ex;
        if (tid == 0) s_tbase[nr] = tile_total;
    }
    if (MODE == kCount) {
        if (tid == 0) state_store(&lb.state[tile], epoch, 1u, (uint32_t)tile_total);
        return;
    }
    // the aggregate goes out before the output staging, so that later tiles
    // find it sooner
    if (MODE == kFused && tid == 0) lookback_publish(lb, tile, (uint32_t)tile_total, epoch);
    if (fits) {
        // stage the tile's output intervals in LDS (the input stage is free
        // once every walk is done; emissions <= positions <= kPCAP) before
        // the look-back, so that the global writes are one coalesced copy
        __syncthreads();
        int32_t w = my_base;
        if (OP == 0) {
            // starts write begins; a start that is not the first of its row
            // closes the previous run; the last position of a row closes the
            // row's last run.
            int32_t pm = carry.v;
            const int32_t p0 = tid * VT;
#pragma unroll
            for (int v = 0; v < VT; ++v) {
                if (p0 + v >= P) break;
                if (rowstart_mask & (1u << v)) pm = INT_MIN;
                const bool st = start_mask & (1u << v);
                if (st) {
                    s_k[w].x = rb[v];
                    if (pm != INT_MIN) s_k[w - 1].y = pm;
                    ++w;
                }
                if (alive_mask & (1u << v)) pm = pm > re[v] ? pm : re[v];
                if ((rowend_mask & (1u << v)) && pm != INT_MIN) s_k[w - 1].y = pm;
            }
        } else {
#pragma unroll
            for (int v = 0; v < VT; ++v) {
                if (emit_mask & (1u << v)) {
                    s_k[w] = make_int2(rb[v], re[v]);
                    ++w;
                }
            }
        }
    }
    if (MODE == kFused && wave_id() == 0) {
        uint32_t e = lookback_exclusive<1, true>(lb, tile, (uint32_t)tile_total, epoch, status, spin_limit);
        if (lane_id() == 0) s_excl = e;
    }
    __syncthreads();
    const int64_t excl =
        MODE == kFused ? (int64_t)s_excl : (tile == 0 ? 0 : (int64_t)(uint32_t)state_load(&lb.state[tile - 1]));
    // row offsets
    for (int r = tid; r <= nr; r += TB) {
        if (r == nr && row0 + nr != U) break;
        int32_t local;
        if (fits) {
            // emissions before the row's first position q
            const int32_t q = s_rs[r];
            if (q >= P) {
                local = tile_total;
            } else {
                const int t = q / VT;
                local = s_tbase[t] + __popc((uint32_t)s_tmask[t] & ((1u << (q % VT)) - 1u));
            }
        } else {
            local = s_tbase[r];
        }
        out_off[row0 + r] = (int32_t)(excl + local);
        if (r == nr && d_total) *d_total = (int32_t)(excl + local);
    }
    if (nr == 0 && row0 == U && tid == 0) {
        out_off[U] = (int32_t)excl;
        if (d_total) *d_total = (int32_t)excl;
    }
    if (excl + tile_total > capacity) {
        if (tid == 0 && tile_total > 0) atomicOr(status, STX_FLAG_CAPACITY);
    } else if (fits) {
        for (int k = tid; k < tile_total; k += TB) {
            const int2 o = s_k[k];
            ob[excl + k] = o.x;
            oe[excl + k] = o.y;
        }
    } else {
        const int wave = wave_id();
        for (int r = wave; r < nr; r += kTW) {
            int32_t a0, na, b0, nb;
            row_extent(A, row0 + r, a0, na);
            row_extent(B, row0 + r, b0, nb);
            setop_row<64, OP, true>(A, a0, na, B, b0, nb, ob, oe, excl + s_tbase[r]);
        }
    }
}

// ---------------------------------------------------------------------------
// Flat tile kernel (identity row maps): the tile's rows merged as ONE sorted
// sequence.
//
// Tiles of whole rows as in setop_tile_kernel (up to one row per thread).
// On staging, every value becomes
//     key = (tile-local row << s) + (x - xmin)
// where xmin and s (2^s > xmax - xmin) come from a block min / max over the
// tile's begins and ends.  Each row then owns a disjoint band of keys, and
// the first begin of a row lies at least one unit above every end of the
// rows before it.  Consequences: the row-by-row merge of A and B is the merge
// of the two staged sequences; the union's per-row max-scan is a plain
// max-scan over the tile (a row's first interval always starts a run); and
// every intersection / difference rule sees an element of another row
// exactly as it would see "no element".  The walk therefore carries no row
// state.  Around each operand sit sentinels: (INT32_MIN, INT32_MIN) below the
// first element, and VT + 1 copies of an upper sentinel (x = INT32_MAX)
// above the last, so the walk needs no bounds tests and the positions past
// the tile's end consume sentinel copies only.  Output keys are decoded
// (row = key >> s) while they are copied out, and the row offsets follow from
// the row of each output and of its predecessor.
// A tile that does not fit (intervals beyond the LDS capacity, or
// nr << s >= 2^31) is processed one wavefront per row instead.

// Rows per flat tile: at most one per thread (the per-row fallback scans one
// row count per thread).
template <int TB>
constexpr int flat_rows_per_tile() {
    return TB;
}

template <bool XF, int VT>
__device__ __forceinline__ void flat_load(const Operand& A, const Operand& B, int32_t a_base, int32_t b_base,
                                          int32_t nA, int32_t nB, int32_t q0, int32_t* xb, int32_t* xe) {
    const int32_t P = nA + nB;
    if (q0 + VT <= nA || (q0 >= nA && q0 + VT <= P)) {
        const bool inA = q0 < nA;
        const int32_t* pb = inA ? A.beg + a_base + q0 : B.beg + b_base + (q0 - nA);
        const int32_t* pe = inA ? A.end + a_base + q0 : B.end + b_base + (q0 - nA);
#pragma unroll
        for (int c = 0; c < VT; c += 4) {
            int4 vb, ve;
            __builtin_memcpy(&vb, __builtin_assume_aligned(pb + c, 4), 16);
            __builtin_memcpy(&ve, __builtin_assume_aligned(pe + c, 4), 16);
            xb[c] = vb.x, xb[c + 1] = vb.y, xb[c + 2] = vb.z, xb[c + 3] = vb.w;
            xe[c] = ve.x, xe[c + 1] = ve.y, xe[c + 2] = ve.z, xe[c + 3] = ve.w;
        }
        if (XF) {
            const XForm xf = inA ? A.xf : B.xf;
#pragma unroll
            for (int u = 0; u < VT; ++u) xform_apply(xf, xb[u], xe[u]);
        }
    } else {
#pragma unroll
        for (int u = 0; u < VT; ++u) {
            const int32_t q = q0 + u;
            xb[u] = 0;
            xe[u] = 0;
            if (q < nA) {
                xb[u] = A.beg[a_base + q];
                xe[u] = A.end[a_base + q];
                if (XF) xform_apply(A.xf, xb[u], xe[u]);
            } else if (q < P) {
                xb[u] = B.beg[b_base + (q - nA)];
                xe[u] = B.end[b_base + (q - nA)];
                if (XF) xform_apply(B.xf, xb[u], xe[u]);
            }
        }
    }
}

// Look-back of the flat tiles: peek at the nearest predecessor, sleeping
// kFlatSleep x 64 cycles while it has not published (see lookback_exclusive).
constexpr int kFlatSleep = 16;

// Staged-key slot i lives at i + i / 16 in LDS: one spare slot per 16
// spreads the strided accesses of the staging and of the merge walks (lanes
// 4 to 8 slots apart) over all banks.
__device__ __forceinline__ int32_t sk(int32_t i) { return i + (i >> 4); }
constexpr int sk_size(int n) { return n + n / 16 + 1; }

// LDS of a flat tile's produce phase; the staged keys (s_k) live apart so
// that a block can keep one tile's outputs while it stages the next.
template <int TB>
struct FlatShared {
    // element starts of the tile's A rows [0, nr], then nA + those of its B
    // rows at [nr + 1, 2 nr + 1], then two INT32_MAX sentinels
    int32_t cS[2 * flat_rows_per_tile<TB>() + 4];
    int32_t red[2 * (TB / STX_WAVE)];
    int32_t wmax[TB / STX_WAVE];
    int32_t wsum[TB / STX_WAVE];
};

// A produced tile: its rows, output count and key transform (uniform).
struct FlatTile {
    int64_t row0;
    int32_t nr, total, xmin;
    uint32_t shift;
    bool flat;
};

// Stages tile `tile` (rows [row0, row0 + nr), nr <= flat_rows_per_tile) and
// walks it, leaving its outputs compacted in s_k (a tile processed one
// wavefront per row leaves its rows' exclusive output offsets there
// instead), and publishes the tile's output count as its look-back
// aggregate.  Called by every thread of the block.
template <int OP, bool XF, int TB, int VT>
__device__ __forceinline__ FlatTile flat_produce(const Operand& A, const Operand& B, int64_t row0, int nr, uint32_t tile,
                                                 int2* __restrict__ s_k, FlatShared<TB>& sh, LookbackScratch lb,
                                                 uint32_t epoch) {
    constexpr int kTW = TB / STX_WAVE;
    constexpr int kPCAP = TB * VT;
    constexpr int kMaxRPT = flat_rows_per_tile<TB>();
    static_assert(VT >= 4 && VT % 4 == 0 && VT < 32, "VT: multiple of 4, below 32");
    constexpr int32_t a0 = 2;  // first A slot; slot 1 holds A's lower sentinel
    // upper sentinels: x above every key; y below every key (the union's
    // max-scan ignores them), or above (difference: B merges by end)
    constexpr int32_t kHiY = OP == 2 ? INT32_MAX : INT32_MIN;
    const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    const int32_t a_base = A.off[row0];
    const int32_t b_base = B.off ? B.off[row0] : 0;
    const int32_t nA = A.off[row0 + nr] - a_base;
    const int32_t nB = B.off ? B.off[row0 + nr] - b_base : 0;
    const int32_t P = nA + nB;
    // B's first slot: past A's VT + 1 upper sentinels and B's lower one, even
    const int32_t b0 = (a0 + nA + VT + 3) & ~1;
    bool flat = b0 + nB + VT + 1 <= kPCAP;  // uniform
    uint32_t shift = 0;
    int32_t xmin = 0;
    if (flat) {
        for (int r = tid; r <= nr; r += TB) {
            sh.cS[r] = A.off[row0 + r] - a_base;
            sh.cS[nr + 1 + r] = nA + (B.off ? B.off[row0 + r] - b_base : 0);
        }
        if (tid < 2) sh.cS[2 * nr + 2 + tid] = INT32_MAX;
        // VT consecutive elements of the sequence A then B per thread
        const int32_t q0 = tid * VT;
        int32_t xb[VT], xe[VT];
        flat_load<XF, VT>(A, B, a_base, b_base, nA, nB, q0, xb, xe);
        int32_t lo = INT32_MAX, hi = INT32_MIN;
#pragma unroll
        for (int u = 0; u < VT; ++u) {
            if (q0 + u < P) {
                lo = imin(lo, imin(xb[u], xe[u]));
                hi = imax(hi, imax(xb[u], xe[u]));
            }
        }
        lo = wave_last(wave_incl_min(lo));
        hi = wave_last(wave_incl_max(hi));
        if (lane == 0) {
            sh.red[wave] = lo;
            sh.red[kTW + wave] = hi;
        }
        __syncthreads();  // row starts, block min / max
#pragma unroll
        for (int w = 0; w < kTW; ++w) {
            lo = imin(lo, sh.red[w]);
            hi = imax(hi, sh.red[kTW + w]);
        }
        const int64_t range = P > 0 ? (int64_t)hi - (int64_t)lo : 0;
        shift = range > 0 ? 64u - (uint32_t)__clzll((unsigned long long)range) : 0u;
        flat = ((int64_t)nr << shift) < (1ll << 31);
        xmin = lo;
        if (flat) {
            if (q0 < P) {
                // combined row index k of element q0: sh.cS[k] <= q0 < sh.cS[k + 1]
                int k = 0, kh = 2 * nr + 2;
                while (kh - k > 1) {
                    const int mid = (k + kh) >> 1;
                    if (sh.cS[mid] <= q0) k = mid;
                    else kh = mid;
                }
                int32_t nxt = sh.cS[k + 1];
                uint32_t off = ((uint32_t)(k < nr ? k : k - nr - 1) << shift) - (uint32_t)xmin;
#pragma unroll
                for (int u = 0; u < VT; ++u) {
                    const int32_t q = q0 + u;
                    if (q < P) {
                        if (q >= nxt) {
                            do {
                                ++k;
                                nxt = sh.cS[k + 1];
                            } while (q >= nxt);
                            off = ((uint32_t)(k < nr ? k : k - nr - 1) << shift) - (uint32_t)xmin;
                        }
                        const int32_t slot = q < nA ? a0 + q : b0 + (q - nA);
                        s_k[sk(slot)] = make_int2((int32_t)((uint32_t)xb[u] + off), (int32_t)((uint32_t)xe[u] + off));
                    }
                }
            }
            if (tid <= VT) {
                s_k[sk(a0 + nA + tid)] = make_int2(INT32_MAX, kHiY);
                s_k[sk(b0 + nB + tid)] = make_int2(INT32_MAX, kHiY);
            }
            if (tid == 0) {
                s_k[sk(a0 - 1)] = make_int2(INT32_MIN, INT32_MIN);
                s_k[sk(b0 - 1)] = make_int2(INT32_MIN, INT32_MIN);
            }
        }
    }
    __syncthreads();  // staged keys (or the fallback's start)

    int32_t tile_total = 0;
    if (flat) {
        const int32_t p0 = tid * VT;
        const uint32_t valid = p0 >= P ? 0u : (P - p0 >= VT ? (1u << VT) - 1u : (1u << (P - p0)) - 1u);
        int32_t rb[VT], re[VT];
        uint32_t emit_mask = 0, alive_mask = (1u << VT) - 1u;
        int32_t agg = INT32_MIN;
        if (p0 < P) {
            // co-rank of p0 over the whole tile
            int32_t lo = p0 - nB > 0 ? p0 - nB : 0, hi = p0 < nA ? p0 : nA;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                const int2 ka = s_k[sk(a0 + mid)];
                const int2 kb = s_k[sk(b0 + (p0 - 1 - mid))];
                if (key_a_first<OP>(ka.x, kb.x, kb.y)) lo = mid + 1;
                else hi = mid;
            }
            int32_t ia = a0 + lo, jb = b0 + (p0 - lo);  // slots of the current A and B elements
            int32_t cAx, cAy, cBx, cBy, pAx = 0, pAy = 0, pBy = 0;
            {
                const int2 t0 = s_k[sk(ia)], t1 = s_k[sk(jb)];
                cAx = t0.x, cAy = t0.y, cBx = t1.x, cBy = t1.y;
                if (OP != 0) {
                    const int2 t2 = s_k[sk(ia - 1)], t3 = s_k[sk(jb - 1)];
                    pAx = t2.x, pAy = t2.y, pBy = t3.y;
                }
            }
#pragma unroll
            for (int v = 0; v < VT; ++v) {
                const bool isA = key_a_first<OP>(cAx, cBx, cBy);
                const int32_t nidx = (isA ? ia : jb) + 1;
                const int2 nraw = s_k[sk(nidx)];
                if (OP == 0) {
                    rb[v] = isA ? cAx : cBx;
                    re[v] = isA ? cAy : cBy;
                    if (XF) {
                        const bool clip = isA ? A.xf.mode == XF_CLIP : B.xf.mode == XF_CLIP;
                        if (clip && rb[v] >= re[v]) {
                            alive_mask &= ~(1u << v);
                            re[v] = INT32_MIN;
                        }
                    }
                    agg = imax(agg, re[v]);
                } else if (OP == 1) {
                    // A element: clipped by the B element covering its begin
                    const int32_t ceA = pBy > cAx ? pBy : (cBx == cAx ? cBy : INT32_MIN);
                    const bool emA = ceA != INT32_MIN;
                    // B element: inside the previous A element
                    const bool emB = pAx < cBx && cBx < pAy;
                    emit_mask |= ((isA ? emA : emB) ? 1u : 0u) << v;
                    rb[v] = isA ? cAx : cBx;
                    re[v] = isA ? imin(cAy, ceA) : imin(cBy, pAy);
                } else {
                    // A element not covered at its begin: runs to the next B begin
                    const bool emA = cBx > cAx;
                    const int32_t peA = cBx < cAy ? cBx : cAy;
                    // B element ending inside the previous A element, not
                    // followed by a B element that begins at its end
                    const bool emB = cBy < pAy && nraw.x > cBy;
                    const int32_t peB = nraw.x < pAy ? nraw.x : pAy;
                    emit_mask |= ((isA ? emA : emB) ? 1u : 0u) << v;
                    rb[v] = isA ? cAx : cBy;
                    re[v] = isA ? peA : peB;
                }
                if (OP != 0) {
                    pAx = isA ? cAx : pAx;
                    pAy = isA ? cAy : pAy;
                    pBy = isA ? pBy : cBy;
                }
                cAx = isA ? nraw.x : cAx;
                cAy = isA ? nraw.y : cAy;
                cBx = isA ? cBx : nraw.x;
                cBy = isA ? cBy : nraw.y;
                ia = isA ? nidx : ia;
                jb = isA ? jb : nidx;
            }
        }
        int32_t carry = INT32_MIN;
        if (OP == 0) {
            // positions past the tile's end hold sentinels: re = INT32_MIN
            carry = block_excl_max<kTW>(agg, sh.wmax);
            int32_t pm = carry;
            if (p0 < P) {
#pragma unroll
                for (int v = 0; v < VT; ++v) {
                    if (rb[v] > pm && (alive_mask & (1u << v))) emit_mask |= 1u << v;
                    pm = imax(pm, re[v]);
                }
            }
        }
        emit_mask &= valid;
        const int32_t my_base = block_excl_sum<kTW>(__popc(emit_mask), sh.wsum, &tile_total);
        // every walk is done: the staged inputs are free
        if (tid == 0) lookback_publish(lb, tile, (uint32_t)tile_total, epoch);
        int32_t w = my_base;
        if (OP == 0) {
            if (p0 < P) {
                // a start writes its begin and closes the run before it
                int32_t pm = carry;
#pragma unroll
                for (int v = 0; v < VT; ++v) {
                    if (emit_mask & (1u << v)) {
                        s_k[sk(w)].x = rb[v];
                        if (w > 0) s_k[sk(w - 1)].y = pm;
                        ++w;
                    }
                    pm = imax(pm, re[v]);
                }
                // the owner of the tile's last position closes the last run
                if (P - p0 <= VT && tile_total > 0) s_k[sk(tile_total - 1)].y = pm;
            }
        } else {
#pragma unroll
            for (int v = 0; v < VT; ++v) {
                if (emit_mask & (1u << v)) {
                    s_k[sk(w)] = make_int2(rb[v], re[v]);
                    ++w;
                }
            }
        }
    } else {
        // fallback: one wavefront per row, counts then their exclusive scan
        int32_t* s_cnt = (int32_t*)s_k;
        for (int r = wave; r < nr; r += kTW) {
            int32_t ra, na, rbb, nb;
            row_extent(A, row0 + r, ra, na);
            row_extent(B, row0 + r, rbb, nb);
            const int32_t c = setop_row<64, OP, false>(A, ra, na, B, rbb, nb, nullptr, nullptr, 0);
            if (lane == 0) s_cnt[r] = c;
        }
        __syncthreads();
        const int32_t c = tid < nr ? s_cnt[tid] : 0;  // nr <= kMaxRPT <= TB
        const int32_t ex = block_excl_sum<kTW>(c, sh.wsum, &tile_total);
        if (tid < nr) s_cnt[tid] = ex;
        if (tid == 0) lookback_publish(lb, tile, (uint32_t)tile_total, epoch);
    }
    return FlatTile{row0, nr, tile_total, xmin, shift, flat};
}

// Writes a produced tile at output offset `excl`: its intervals, its rows'
// offsets, and for the last tile the total.  Called by every thread.
template <int OP, int TB>
__device__ __forceinline__ void flat_store(const FlatTile& ft, const Operand& A, const Operand& B, int64_t U, int64_t excl,
                                           const int2* __restrict__ s_k, int32_t* __restrict__ out_off,
                                           int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t capacity,
                                           int32_t* __restrict__ d_total, uint32_t* status) {
    constexpr int kTW = TB / STX_WAVE;
    const int tid = threadIdx.x, wave = wave_id();
    const int64_t row0 = ft.row0;
    const int nr = ft.nr;
    const int32_t tile_total = ft.total;
    const uint32_t shift = ft.shift;
    const int32_t xmin = ft.xmin;
    const bool flat = ft.flat;
    const bool over = excl + tile_total > capacity;
    if (over && tid == 0 && tile_total > 0) atomicOr(status, STX_FLAG_CAPACITY);
    if (flat) {
        for (int32_t k = tid; k < tile_total; k += TB) {
            const int2 o = s_k[sk(k)];
            const int32_t r = o.x >> shift;
            if (!over) {
                const uint32_t boff = ((uint32_t)r << shift) - (uint32_t)xmin;
                ob[excl + k] = (int32_t)((uint32_t)o.x - boff);
                oe[excl + k] = (int32_t)((uint32_t)o.y - boff);
            }
            // rows (row of output k - 1, row of output k] start at output k
            // (the bound keeps every store inside the tile's rows)
            const int32_t rp = k > 0 ? (s_k[sk(k - 1)].x >> shift) : -1;
            const int32_t rt = r < nr ? r : nr - 1;
            for (int32_t rr = rp + 1; rr <= rt; ++rr) out_off[row0 + rr] = (int32_t)(excl + k);
        }
        // rows after the last output's row
        const int32_t rl = tile_total > 0 ? (s_k[sk(tile_total - 1)].x >> shift) : -1;
        for (int32_t rr = rl + 1 + tid; rr < nr; rr += TB) out_off[row0 + rr] = (int32_t)(excl + tile_total);
    } else {
        const int32_t* s_cnt = (const int32_t*)s_k;
        for (int r = tid; r < nr; r += TB) out_off[row0 + r] = (int32_t)(excl + s_cnt[r]);
        if (!over) {
            for (int r = wave; r < nr; r += kTW) {
                int32_t ra, na, rbb, nb;
                row_extent(A, row0 + r, ra, na);
                row_extent(B, row0 + r, rbb, nb);
                setop_row<64, OP, true>(A, ra, na, B, rbb, nb, ob, oe, excl + s_cnt[r]);
            }
        }
    }
    if (row0 + nr == U && tid == 0) {
        out_off[U] = (int32_t)(excl + tile_total);
        if (d_total) *d_total = (int32_t)(excl + tile_total);
    }
}


template <int OP, bool XF, int TB, int VT>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(tile_waves_per_eu<VT>()))) void setop_flat_kernel(
    Operand A, Operand B, int64_t U, int32_t RPT, uint32_t ntiles, int32_t* __restrict__ out_off,
    int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t capacity, int32_t* __restrict__ d_total,
    uint32_t* status, void* scratch, uint32_t epoch, int64_t spin_limit) {
    __shared__ __align__(16) int2 s_k[sk_size(TB * VT)];
    __shared__ FlatShared<TB> sh;
    __shared__ uint32_t s_excl;
    LookbackScratch lb = lookback_scratch(scratch, ntiles);
    const uint32_t tile = blockIdx.x;
    const int64_t row0 = (int64_t)tile * RPT;
    const int nr = (int)(U - row0 < RPT ? (U - row0 > 0 ? U - row0 : 0) : RPT);
    const FlatTile ft = flat_produce<OP, XF, TB, VT>(A, B, row0, nr, tile, s_k, sh, lb, epoch);
    if (wave_id() == 0) {
        const uint32_t e =
            lookback_exclusive<1, true, kFlatSleep, true>(lb, tile, (uint32_t)ft.total, epoch, status, spin_limit);
        if (lane_id() == 0) s_excl = e;
    }
    __syncthreads();
    flat_store<OP, TB>(ft, A, B, U, (int64_t)s_excl, s_k, out_off, ob, oe, capacity, d_total, status);
}

// ---------------------------------------------------------------------------
// Flat tile kernel, register-output form (identity row maps, operands read as
// stored, two operands).  The same keyed merge as setop_flat_kernel, with
// less issue work per interval:
//   * staging: thread t loads 8 consecutive intervals of ONE operand (A
//     chunks first, then B chunks), so its 8 LDS slots are contiguous (A and
//     B start on slots that are multiples of 8) and one row lookup serves the
//     chunk;
//   * the walk keeps each position's begin key and, for union, the largest
//     end key before it (lm); a position starts an output run when its begin
//     is above both lm and the block carry;
//   * the block scans combine the per-wave values lane-parallel (one LDS
//     read and a 16-lane DPP scan per thread);
//   * once every walk is done the staged keys are dead: each emitted
//     position writes its decoded interval straight into the compacted
//     output (slot = the position's rank among the tile's emitted
//     positions), and after the look-back one coalesced copy moves the
//     tile's outputs to HBM;
//   * row offsets: row r of the tile starts at merged position M[r] = (A
//     intervals before r) + (B intervals before r), and its offset is the
//     number of outputs emitted before M[r]: the exclusive count of the
//     thread owning M[r] plus the emitted positions below M[r] in its mask.
// A tile that this layout cannot hold goes through setop_flat_kernel's tile
// code (flat or one wavefront per row) in the same launch.
constexpr int kFlat2Base = 8;  // first A slot; A's lower sentinel sits at slot 7

// Runs one tile through setop_flat_kernel's tile code (the flat layout of
// that kernel, or one wavefront per row), look-back and store included.
template <int OP, int TB, int VT>
__device__ __forceinline__ void flat_tile_whole(const Operand& A, const Operand& B, int64_t U, int64_t row0, int nr,
                                                uint32_t tile, int2* s_k, FlatShared<TB>& sh, LookbackScratch lb,
                                                uint32_t epoch, int32_t* out_off, int32_t* ob, int32_t* oe,
                                                int64_t capacity, int32_t* d_total, uint32_t* status,
                                                int64_t spin_limit, uint32_t* s_excl) {
    const FlatTile ft = flat_produce<OP, false, TB, VT>(A, B, row0, nr, tile, s_k, sh, lb, epoch);
    if (wave_id() == 0) {
        const uint32_t e =
            lookback_exclusive<1, true, kFlatSleep, true>(lb, tile, (uint32_t)ft.total, epoch, status, spin_limit);
        if (lane_id() == 0) *s_excl = e;
    }
    __syncthreads();
    flat_store<OP, TB>(ft, A, B, U, (int64_t)*s_excl, s_k, out_off, ob, oe, capacity, d_total, status);
}

template <int OP, int TB, int VT>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(tile_waves_per_eu<VT>()))) void setop_flat2_kernel(
    Operand A, Operand B, int64_t U, int32_t RPT, const int2* __restrict__ tab, uint32_t ntiles,
    int32_t* __restrict__ out_off, int32_t* __restrict__ ob, int32_t* __restrict__ oe, int64_t capacity,
    int32_t* __restrict__ d_total, uint32_t* status, void* scratch, uint32_t epoch, int64_t spin_limit) {
    static_assert(VT == 8, "one 8-interval chunk per thread");
    constexpr int kTW = TB / STX_WAVE;
    static_assert(kTW <= 16, "the range and scan combines read one 16-lane row");
    constexpr int kPCAP = TB * VT;
    constexpr int kMaxRows = flat_rows_per_tile<TB>();
    constexpr int32_t kHiY = OP == 2 ? INT32_MAX : INT32_MIN;
    __shared__ __align__(16) int2 s_k[sk_size(kPCAP)];
    // tile-local row starts of A at [0, nr] and of B at [kMaxRows + 2, ...];
    // the same memory is setop_flat_kernel's FlatShared for tiles it takes
    __shared__ __align__(16) int32_t s_aux[sizeof(FlatShared<TB>) / 4 > 2 * kMaxRows + 4 ? sizeof(FlatShared<TB>) / 4
                                                                                         : 2 * kMaxRows + 4];
    __shared__ int32_t s_red[2 * kTW];
    __shared__ int32_t s_wmax[kTW];
    __shared__ int32_t s_wsum[kTW];
    __shared__ uint32_t s_excl;
    LookbackScratch lb = lookback_scratch(scratch, ntiles);
    const uint32_t tile = blockIdx.x;
    const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
    // rows of the tile: RPT per tile, or the tile table's count-balanced
    // starts (tiles past the last row start at U, and no tile waits on them)
    int64_t row0;
    int nr;
    if (tab) {
        row0 = tab[tile].x;
        if (row0 >= U && tile > 0) return;
        const int64_t n = (int64_t)tab[tile + 1].x - row0;
        nr = (int)(n < kMaxRows ? n : kMaxRows);
        if (n > kMaxRows && tid == 0) atomicOr(status, STX_FLAG_CAPACITY);  // a table not made for these counts
    } else {
        row0 = (int64_t)tile * RPT;
        nr = (int)(U - row0 < RPT ? (U - row0 > 0 ? U - row0 : 0) : RPT);
    }
    const int32_t a_base = A.off[row0];
    const int32_t b_base = B.off[row0];
    const int32_t nA = A.off[row0 + nr] - a_base;
    const int32_t nB = B.off[row0 + nr] - b_base;
    const int32_t P = nA + nB;
    constexpr int32_t a0 = kFlat2Base;
    const int32_t b0 = (a0 + nA + VT + 2 + 7) & ~7;  // past A's VT + 1 upper sentinels and B's lower one
    if (!(b0 + nB + VT + 1 <= kPCAP && nr <= kMaxRows)) {  // uniform: the tile does not fit this layout
        flat_tile_whole<OP, TB, VT>(A, B, U, row0, nr, tile, s_k, *reinterpret_cast<FlatShared<TB>*>(s_aux), lb,
                                    epoch, out_off, ob, oe, capacity, d_total, status, spin_limit, &s_excl);
        return;
    }
    int32_t* s_ra = s_aux;
    int32_t* s_rb = s_aux + kMaxRows + 2;
    for (int r = tid; r <= nr; r += TB) {
        s_ra[r] = A.off[row0 + r] - a_base;
        s_rb[r] = B.off[row0 + r] - b_base;
    }
    if (tid == 0) {
        s_ra[nr + 1] = INT32_MAX;
        s_rb[nr + 1] = INT32_MAX;
    }
    // this thread's chunk: 8 consecutive intervals of A (chunks [0, ca)) or B
    const int32_t ca = (nA + 7) >> 3;
    const bool inA = tid < ca;
    const int32_t c = inA ? tid : tid - ca;
    const int32_t q0 = c << 3;
    const int32_t nq = inA ? nA : nB;
    const int32_t n_mine = nq - q0 < 0 ? 0 : (nq - q0 > 8 ? 8 : nq - q0);  // 0 past the B chunks
    const int32_t* pb = inA ? A.beg + a_base + q0 : B.beg + b_base + q0;
    const int32_t* pe = inA ? A.end + a_base + q0 : B.end + b_base + q0;
    int32_t xb[VT], xe[VT];
    if (n_mine == 8) {
        int4 v0, v1, w0, w1;
        __builtin_memcpy(&v0, __builtin_assume_aligned(pb, 4), 16);
        __builtin_memcpy(&v1, __builtin_assume_aligned(pb + 4, 4), 16);
        __builtin_memcpy(&w0, __builtin_assume_aligned(pe, 4), 16);
        __builtin_memcpy(&w1, __builtin_assume_aligned(pe + 4, 4), 16);
        xb[0] = v0.x, xb[1] = v0.y, xb[2] = v0.z, xb[3] = v0.w, xb[4] = v1.x, xb[5] = v1.y, xb[6] = v1.z, xb[7] = v1.w;
        xe[0] = w0.x, xe[1] = w0.y, xe[2] = w0.z, xe[3] = w0.w, xe[4] = w1.x, xe[5] = w1.y, xe[6] = w1.z, xe[7] = w1.w;
    } else {
#pragma unroll
        for (int u = 0; u < VT; ++u) {
            xb[u] = u < n_mine ? pb[u] : INT32_MAX;
            xe[u] = u < n_mine ? pe[u] : INT32_MIN;
        }
    }
    // range of every stored value (begins and ends alike, so that any stored
    // interval lands inside its row's key band)
    int32_t lo = INT32_MAX, hi = INT32_MIN;
#pragma unroll
    for (int u = 0; u < VT; ++u) {
        const bool ok = u < n_mine;
        lo = imin(lo, ok ? imin(xb[u], xe[u]) : INT32_MAX);
        hi = imax(hi, ok ? imax(xb[u], xe[u]) : INT32_MIN);
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
    const int64_t range = P > 0 ? (int64_t)hi - (int64_t)lo : 0;
    const uint32_t shift = range > 0 ? 64u - (uint32_t)__clzll((unsigned long long)range) : 0u;
    if (((int64_t)nr << shift) >= (1ll << 31)) {  // uniform: rows x key band do not fit 31 bits
        __syncthreads();  // every thread has read s_red / s_aux before the tile code reuses s_aux
        flat_tile_whole<OP, TB, VT>(A, B, U, row0, nr, tile, s_k, *reinterpret_cast<FlatShared<TB>*>(s_aux), lb,
                                    epoch, out_off, ob, oe, capacity, d_total, status, spin_limit, &s_excl);
        return;
    }
    const int32_t xmin = lo;
    // keys (tile row << shift) + (x - xmin); a chunk's slots past its
    // operand's end hold the operand's upper sentinel
    if (n_mine > 0) {
        const int32_t* rs = inA ? s_ra : s_rb;
        int k = 0, kh = nr;  // row of element q0: rs[k] <= q0 < rs[k + 1]
        while (kh - k > 1) {
            const int mid = (k + kh) >> 1;
            if (rs[mid] <= q0) k = mid;
            else kh = mid;
        }
        int32_t nxt = rs[k + 1];
        uint32_t off = ((uint32_t)k << shift) - (uint32_t)xmin;
        const int32_t slot = (inA ? a0 : b0) + q0;  // a multiple of 8: the chunk's 8 slots are contiguous
        int2* dst = s_k + sk(slot);
#pragma unroll
        for (int u = 0; u < VT; ++u) {
            const int32_t q = q0 + u;
            if (q >= nxt && u < n_mine) {
                do {
                    ++k;
                    nxt = rs[k + 1];
                } while (q >= nxt);
                off = ((uint32_t)k << shift) - (uint32_t)xmin;
            }
            dst[u] = u < n_mine ? make_int2((int32_t)((uint32_t)xb[u] + off), (int32_t)((uint32_t)xe[u] + off))
                                : make_int2(INT32_MAX, kHiY);
        }
    }
    if (tid <= VT) {
        s_k[sk(a0 + nA + tid)] = make_int2(INT32_MAX, kHiY);
        s_k[sk(b0 + nB + tid)] = make_int2(INT32_MAX, kHiY);
    }
    if (tid == 0) {
        s_k[sk(a0 - 1)] = make_int2(INT32_MIN, INT32_MIN);
        s_k[sk(b0 - 1)] = make_int2(INT32_MIN, INT32_MIN);
    }
    __syncthreads();  // staged keys

    // merge-path walk of positions [p0, p0 + VT)
    const int32_t p0 = tid * VT;
    const uint32_t valid = p0 >= P ? 0u : (P - p0 >= VT ? (1u << VT) - 1u : (1u << (P - p0)) - 1u);
    int32_t rb[VT], re[VT];
    uint32_t emit = 0;
    int32_t lm = INT32_MIN;  // union: largest end key before the current position (this thread's part)
    {
        int32_t lo2 = p0 - nB > 0 ? p0 - nB : 0, hi2 = p0 < nA ? p0 : nA;
        if (p0 >= P) lo2 = hi2 = nA;  // positions past the end: every walk reads sentinels
        while (lo2 < hi2) {
            const int32_t mid = (lo2 + hi2) >> 1;
            const int2 ka = s_k[sk(a0 + mid)];
            const int2 kb = s_k[sk(b0 + (p0 - 1 - mid))];
            if (key_a_first<OP>(ka.x, kb.x, kb.y)) lo2 = mid + 1;
            else hi2 = mid;
        }
        int32_t ia = a0 + lo2, jb = b0 + (p0 >= P ? nB : p0 - lo2);
        int32_t cAx, cAy, cBx, cBy, pAx = 0, pAy = 0, pBy = 0;
        {
            const int2 t0 = s_k[sk(ia)], t1 = s_k[sk(jb)];
            cAx = t0.x, cAy = t0.y, cBx = t1.x, cBy = t1.y;
            if (OP != 0) {
                const int2 t2 = s_k[sk(ia - 1)], t3 = s_k[sk(jb - 1)];
                pAx = t2.x, pAy = t2.y, pBy = t3.y;
            }
        }
#pragma unroll
        for (int v = 0; v < VT; ++v) {
            const bool isA = key_a_first<OP>(cAx, cBx, cBy);
            const int32_t nidx = (isA ? ia : jb) + 1;
            const int2 nraw = s_k[sk(nidx)];
            if (OP == 0) {
                const int32_t x = isA ? cAx : cBx;
                const int32_t y = isA ? cAy : cBy;
                rb[v] = x;
                re[v] = lm;  // the largest end key before position v
                emit |= (x > lm ? 1u : 0u) << v;
                lm = imax(lm, y);
            } else if (OP == 1) {
                const int32_t ceA = pBy > cAx ? pBy : (cBx == cAx ? cBy : INT32_MIN);
                const bool emA = ceA != INT32_MIN;
                const bool emB = pAx < cBx && cBx < pAy;
                emit |= ((isA ? emA : emB) ? 1u : 0u) << v;
                rb[v] = isA ? cAx : cBx;
                re[v] = isA ? imin(cAy, ceA) : imin(cBy, pAy);
            } else {
                const bool emA = cBx > cAx;
                const int32_t peA = cBx < cAy ? cBx : cAy;
                const bool emB = cBy < pAy && nraw.x > cBy;
                const int32_t peB = nraw.x < pAy ? nraw.x : pAy;
                emit |= ((isA ? emA : emB) ? 1u : 0u) << v;
                rb[v] = isA ? cAx : cBy;
                re[v] = isA ? peA : peB;
            }
            if (OP != 0) {
                pAx = isA ? cAx : pAx;
                pAy = isA ? cAy : pAy;
                pBy = isA ? pBy : cBy;
            }
            cAx = isA ? nraw.x : cAx;
            cAy = isA ? nraw.y : cAy;
            cBx = isA ? cBx : nraw.x;
            cBy = isA ? cBy : nraw.y;
            ia = isA ? nidx : ia;
            jb = isA ? jb : nidx;
        }
    }
    int32_t carry = INT32_MIN;
    if (OP == 0) {
        // union: a position starts a run when its begin is above every end
        // before it; begins grow along the walk, so those above the carry
        // form a suffix of the thread's positions
        carry = block_excl_max<kTW>(lm, s_wmax);
        int32_t below = 0;
#pragma unroll
        for (int v = 0; v < VT; ++v) below += rb[v] <= carry ? 1 : 0;
        emit &= ~((1u << below) - 1u);
    }
    emit &= valid;
    int32_t tile_total = 0;
    const int32_t my_base = block_excl_sum<kTW>(__popc(emit), s_wsum, &tile_total);
    if (tid == 0) lookback_publish(lb, tile, (uint32_t)tile_total, epoch);
    // every walk is done (barrier in the scans): s_k now holds the tile's
    // outputs, decoded and compacted (pairs [0, tile_total)), and past them
    // each thread's exclusive count | emit mask << 16 for the row offsets
    const uint32_t kmask = shift ? (1u << shift) - 1u : 0u;
    int32_t* s_out = reinterpret_cast<int32_t*>(s_k);
    uint32_t* s_pk = reinterpret_cast<uint32_t*>(s_k + kPCAP);
    s_pk[tid] = (uint32_t)my_base | (emit << 16);
    {
        int32_t w = my_base;
#pragma unroll
        for (int v = 0; v < VT; ++v) {
            if (emit & (1u << v)) {
                s_out[2 * w] = (int32_t)(((uint32_t)rb[v] & kmask) + (uint32_t)xmin);
                if (OP == 0) {
                    if (w > 0) s_out[2 * w - 1] = (int32_t)(((uint32_t)imax(carry, re[v]) & kmask) + (uint32_t)xmin);
                } else {
                    s_out[2 * w + 1] = (int32_t)(((uint32_t)re[v] & kmask) + (uint32_t)xmin);
                }
                ++w;
            }
        }
        // union: the owner of the tile's last position closes the last run
        if (OP == 0 && tile_total > 0 && p0 < P && P - p0 <= VT)
            s_out[2 * tile_total - 1] = (int32_t)(((uint32_t)imax(carry, lm) & kmask) + (uint32_t)xmin);
    }
    if (wave == 0) {
        const uint32_t e =
            lookback_exclusive<1, true, kFlatSleep, true>(lb, tile, (uint32_t)tile_total, epoch, status, spin_limit);
        if (lane == 0) s_excl = e;
    }
    __syncthreads();
    const int64_t excl = (int64_t)s_excl;
    const bool over = excl + tile_total > capacity;
    if (over && tid == 0 && tile_total > 0) atomicOr(status, STX_FLAG_CAPACITY);
    if (!over) {
        int32_t* obt = ob + excl;
        int32_t* oet = oe + excl;
        for (int32_t k = tid; k < tile_total; k += TB) {
            const int2 o = s_k[k];
            obt[k] = o.x;
            oet[k] = o.y;
        }
    }
    for (int r = tid; r < nr; r += TB) {  // row r starts at merged position m
        const int32_t m = s_ra[r] + s_rb[r];
        const uint32_t pk = s_pk[m >> 3];
        const uint32_t below = (pk >> 16) & ((1u << (m & 7)) - 1u);
        out_off[row0 + r] = (int32_t)(excl + (int64_t)(pk & 0xffffu) + __popc(below));
    }
    if (row0 + nr == U && tid == 0) {
        out_off[U] = (int32_t)(excl + tile_total);
        if (d_total) *d_total = (int32_t)(excl + tile_total);
    }
}

// One block: aggregates (flag 1) of tiles [0, ntiles) -> inclusive prefixes
// (flag 2), same epoch.
constexpr int kPrefixThreads = 1024;

__global__ __launch_bounds__(kPrefixThreads) void prefix_states_kernel(void* scratch, uint32_t ntiles, uint32_t epoch) {
    __shared__ uint32_t s_scan[kPrefixThreads / 64 + 1];
    LookbackScratch lb = lookback_scratch(scratch, ntiles);
    const uint32_t per = (ntiles + kPrefixThreads - 1) / kPrefixThreads;
    const uint32_t t0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (t0 + k < ntiles) sum += (uint32_t)state_load(&lb.state[t0 + k]);
    // block exclusive scan of the per-thread sums
    const int lane = lane_id(), wave = wave_id();
    uint32_t incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_scan[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < kPrefixThreads / 64; ++w) {
            const uint32_t v = s_scan[w];
            s_scan[w] = acc;
            acc += v;
        }
    }
    __syncthreads();
    uint32_t run = s_scan[wave] + incl - sum;
    for (uint32_t k = 0; k < per; ++k) {
        if (t0 + k < ntiles) {
            run += (uint32_t)state_load(&lb.state[t0 + k]);
            state_store(&lb.state[t0 + k], epoch, 2u, run);
        }
    }
}

// ---------------------------------------------------------------------------
// Generic exclusive scan of int32 counts (single pass, decoupled look-back).
// out[k] = sum(in[0..k)), out[n] = total; *d_total = total.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__global__ __launch_bounds__(kBlock) void scan_i32_kernel(const int32_t* __restrict__ in, const int32_t* d_n, int64_t n_bound,
                                                          uint32_t ntiles, int32_t* __restrict__ out,
                                                          int32_t* __restrict__ d_total, uint32_t* status, void* scratch,
                                                          uint32_t epoch, int64_t spin_limit) {
    __shared__ int32_t s_scan[17];
    __shared__ uint32_t s_excl;
    LookbackScratch lb = lookback_scratch(scratch, ntiles);
    const uint32_t tile = blockIdx.x;
    const int64_t n = d_n ? (int64_t)*d_n : n_bound;
    const int64_t t0 = (int64_t)tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t vals[kScanItems];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t idx = t0 + k;
        vals[k] = idx < n ? in[idx] : 0;
        sum += vals[k];
    }
    int32_t total = 0;
    int32_t ex = block_exclusive_scan<int32_t>(sum, s_scan, &total);
    if (wave_id() == 0) {
        uint32_t e = lookback_exclusive(lb, tile, (uint32_t)total, epoch, status, spin_limit);
        if (lane_id() == 0) s_excl = e;
    }
    __syncthreads();
    int32_t run = (int32_t)s_excl + ex;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t idx = t0 + k;
        if (idx <= n) out[idx] = run;
        if (idx == n && d_total) *d_total = run;
        run += vals[k];
    }
    if (threadIdx.x == kBlock - 1 && t0 + kScanItems == n) {
        out[n] = run;
        if (d_total) *d_total = run;
    }
}

// ---------------------------------------------------------------------------
// Row-id alignment: the sorted union of two strictly increasing row lists
// (each optionally shifted) and, per union row, the source row in each list
// (-1 when absent).  Replaces _align_interval_sets / _reindex_interval_set
// (expressions.py:58-141): the intervals themselves are never copied, the
// set-op kernel reads them through the maps.
__device__ __forceinline__ int64_t lower_bound_shift(const int32_t* a, int64_t n, int32_t shift, int32_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] + shift < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ void rows_newflag_kernel(const int32_t* __restrict__ ra, int64_t na, int32_t sa, const int32_t* __restrict__ rb,
                                    int64_t nb, int32_t sb, int32_t* __restrict__ flag) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    int32_t v = rb[j] + sb;
    int64_t p = lower_bound_shift(ra, na, sa, v);
    flag[j] = (p < na && ra[p] + sa == v) ? 0 : 1;
}

__global__ void rows_scatter_kernel(const int32_t* __restrict__ ra, int64_t na, int32_t sa, const int32_t* __restrict__ rb,
                                    int64_t nb, int32_t sb, const int32_t* __restrict__ flag,
                                    const int32_t* __restrict__ newpos, int32_t* __restrict__ rows_u,
                                    int32_t* __restrict__ map_a, int32_t* __restrict__ map_b) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < na) {
        int32_t v = ra[t] + sa;
        int64_t jl = lower_bound_shift(rb, nb, sb, v);
        int64_t u = t + newpos[jl];
        rows_u[u] = v;
        map_a[u] = (int32_t)t;
        map_b[u] = (jl < nb && rb[jl] + sb == v) ? (int32_t)jl : -1;
    } else if (t < na + nb) {
        int64_t j = t - na;
        if (!flag[j]) return;
        int32_t v = rb[j] + sb;
        int64_t u = newpos[j] + lower_bound_shift(ra, na, sa, v);
        rows_u[u] = v;
        map_a[u] = -1;
        map_b[u] = (int32_t)j;
    }
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v, int32_t step) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v + (int32_t)k * step;
}

__global__ void count_plus_kernel(int32_t* dst, const int32_t* src, int32_t add) { *dst = *src + add; }

__global__ void add_i32_kernel(int32_t* p, int64_t n, int32_t v) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] += v;
}

}  // namespace

// ===========================================================================
// Host side (C-ABI).
#include "stx_host.h"

// Rows per tile for the LDS tile kernel: kTileFill of the tile capacity at
// the mean row length (the row-bucket schedule of the tile path); a tile
// whose rows exceed the capacity takes the in-kernel per-row fallback.
constexpr double kTileFill = 0.9;
template <int TB, int VT>
static int32_t setop_tile_rows(int64_t elems, int64_t rows, double fill = kTileFill, int max_rows = max_rows_per_tile<TB>()) {
    double avg = rows > 0 ? (double)elems / (double)rows : 1.0;
    if (avg < 1.0) avg = 1.0;
    int64_t r = (int64_t)(fill * TB * VT / avg);
    if (r < 1) r = 1;
    if (r > max_rows) r = max_rows;
    return (int32_t)r;
}

constexpr uint32_t kTwoPassTiles = 0xffffffffu;  // measured slower than one pass at every size tried
// Group 0 takes the 1024 x 8 flat tiles filled to kFlatFill: at 128 stacked
// 4096 x 65536 pairs, union 178.7 us against 197.7 (512 x 8 flat tiles,
// filled to 90 %) and 201.0 (row-aware tiles); intersection 165.1 / 175.1 /
// 176.9; difference 170.3 / 185.4 / 188.5.  All tie on one and on 16 pairs.
constexpr bool kFlatDefault = true;
constexpr double kFlatFill = 0.95;

template <int OP, int MODE, int TB, int VT>
static void launch_tile_op(uint32_t ntiles, hipStream_t stream, const Operand& A, const Operand& B, int64_t U, int32_t rpt,
                           int32_t* out_off, int32_t* ob, int32_t* oe, int64_t capacity, int32_t* d_total,
                           uint32_t* status, void* scratch, uint32_t epoch, int64_t spin) {
    // operands read as stored (no transform on load): the plain kernel
    const bool xf = A.xf.mode != XF_NONE || (B.off != nullptr && B.xf.mode != XF_NONE);
    if (xf)
        hipLaunchKernelGGL((setop_tile_kernel<OP, MODE, true, TB, VT>), dim3(ntiles), dim3(TB), 0, stream, A, B, U, rpt,
                           ntiles, out_off, ob, oe, capacity, d_total, status, scratch, epoch, spin);
    else
        hipLaunchKernelGGL((setop_tile_kernel<OP, MODE, false, TB, VT>), dim3(ntiles), dim3(TB), 0, stream, A, B, U,
                           rpt, ntiles, out_off, ob, oe, capacity, d_total, status, scratch, epoch, spin);
}

template <int MODE, int TB = 512, int VT = 8>
static void launch_tile(int32_t op, uint32_t ntiles, hipStream_t stream, const Operand& A, const Operand& B, int64_t U,
                        int32_t rpt, int32_t* out_off, int32_t* ob, int32_t* oe, int64_t capacity, int32_t* d_total,
                        uint32_t* status, void* scratch, uint32_t epoch, int64_t spin) {
    if (op == 0)
        launch_tile_op<0, MODE, TB, VT>(ntiles, stream, A, B, U, rpt, out_off, ob, oe, capacity, d_total, status,
                                        scratch, epoch, spin);
    else if (op == 1)
        launch_tile_op<1, MODE, TB, VT>(ntiles, stream, A, B, U, rpt, out_off, ob, oe, capacity, d_total, status,
                                        scratch, epoch, spin);
    else
        launch_tile_op<2, MODE, TB, VT>(ntiles, stream, A, B, U, rpt, out_off, ob, oe, capacity, d_total, status,
                                        scratch, epoch, spin);
}

template <int OP, int TB, int VT>
static void launch_flat_op(uint32_t ntiles, hipStream_t stream, const Operand& A, const Operand& B, int64_t U, int32_t rpt,
                           int32_t* out_off, int32_t* ob, int32_t* oe, int64_t capacity, int32_t* d_total,
                           uint32_t* status, void* scratch, uint32_t epoch, int64_t spin) {
    const bool xf = A.xf.mode != XF_NONE || (B.off != nullptr && B.xf.mode != XF_NONE);
    if (xf)
        hipLaunchKernelGGL((setop_flat_kernel<OP, true, TB, VT>), dim3(ntiles), dim3(TB), 0, stream, A, B, U, rpt, ntiles,
                           out_off, ob, oe, capacity, d_total, status, scratch, epoch, spin);
    else
        hipLaunchKernelGGL((setop_flat_kernel<OP, false, TB, VT>), dim3(ntiles), dim3(TB), 0, stream, A, B, U, rpt,
                           ntiles, out_off, ob, oe, capacity, d_total, status, scratch, epoch, spin);
}

template <int TB, int VT>
static void launch_flat(int32_t op, uint32_t ntiles, hipStream_t stream, const Operand& A, const Operand& B, int64_t U,
                        int32_t rpt, int32_t* out_off, int32_t* ob, int32_t* oe, int64_t capacity, int32_t* d_total,
                        uint32_t* status, void* scratch, uint32_t epoch, int64_t spin) {
    if (op == 0)
        launch_flat_op<0, TB, VT>(ntiles, stream, A, B, U, rpt, out_off, ob, oe, capacity, d_total, status, scratch,
                                  epoch, spin);
    else if (op == 1)
        launch_flat_op<1, TB, VT>(ntiles, stream, A, B, U, rpt, out_off, ob, oe, capacity, d_total, status, scratch,
                                  epoch, spin);
    else
        launch_flat_op<2, TB, VT>(ntiles, stream, A, B, U, rpt, out_off, ob, oe, capacity, d_total, status, scratch,
                                  epoch, spin);
}

template <int TB, int VT>
static void launch_flat2(int32_t op, uint32_t ntiles, hipStream_t stream, const Operand& A, const Operand& B, int64_t U,
                         int32_t rpt, int32_t* out_off, int32_t* ob, int32_t* oe, int64_t capacity, int32_t* d_total,
                         uint32_t* status, void* scratch, uint32_t epoch, int64_t spin,
                         const int2* tab = nullptr) {
    if (op == 0)
        hipLaunchKernelGGL((setop_flat2_kernel<0, TB, VT>), dim3(ntiles), dim3(TB), 0, stream, A, B, U, rpt, tab,
                           ntiles, out_off, ob, oe, capacity, d_total, status, scratch, epoch, spin);
    else if (op == 1)
        hipLaunchKernelGGL((setop_flat2_kernel<1, TB, VT>), dim3(ntiles), dim3(TB), 0, stream, A, B, U, rpt, tab,
                           ntiles, out_off, ob, oe, capacity, d_total, status, scratch, epoch, spin);
    else
        hipLaunchKernelGGL((setop_flat2_kernel<2, TB, VT>), dim3(ntiles), dim3(TB), 0, stream, A, B, U, rpt, tab,
                           ntiles, out_off, ob, oe, capacity, d_total, status, scratch, epoch, spin);
}

// Count-balanced tiles (schedule -8): balanced_table_entry with c(r) the
// intervals of both operands in rows below r, so a tile holds fewer than T
// intervals plus its last row, and at most M rows.  One thread per row and
// per table entry past h(U).
constexpr int64_t kBalancedRoom = 1024 * 8 - 34;  // intervals a 1024 x 8 register-output tile holds
constexpr int64_t kBalancedTarget = kBalancedRoom - 256;  // room left for the tile's last row

__global__ void tile_table_kernel(const int32_t* __restrict__ a_off, const int32_t* __restrict__ b_off, int64_t U,
                                  int64_t T, int64_t M, int64_t ntiles, int2* __restrict__ tab, uint32_t* status) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t a0 = a_off[0], b0 = b_off[0];
    auto h = [&](int64_t r) { return ((int64_t)a_off[r] - a0 + (int64_t)b_off[r] - b0) / T + r / M; };
    balanced_table_entry(i, U, ntiles, h, tab, status);
}

inline int64_t balanced_tiles(int64_t elems, int64_t U) { return elems / kBalancedTarget + U / 1024 + 1; }

inline int64_t balanced_scratch_bytes(int64_t elems, int64_t U) {
    return balanced_table_scratch_bytes(balanced_tiles(elems, U));
}

template <int G>
static int32_t launch_setop_g(int32_t op, const Operand& A, const Operand& B, const int32_t* d_nrows, int64_t nrows_bound,
                              int32_t* out_off, int32_t* ob, int32_t* oe, int64_t capacity, int32_t* d_total,
                              uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    FusedLaunch L{d_nrows, nrows_bound, out_off, ob, oe, capacity, d_total, status, scratch, scratch_bytes, epoch, stream};
    if (op == 0) return launch_fused_rows<G>(SetopRow<0>{A, B}, L, "setop_union");
    if (op == 1) return launch_fused_rows<G>(SetopRow<1>{A, B}, L, "setop_intersection");
    return launch_fused_rows<G>(SetopRow<2>{A, B}, L, "setop_difference");
}

// ---- allocation-free variant -------------------------------------------------
// For callers that cannot allocate (code under graph capture, tight serving
// loops): the look-back state of stx_setop_noalloc lives in this file-scope
// device array instead of a caller workspace.  Each GPU holds its own copy
// (zero at module load) until the process ends; its address on a device is
// taken once, with that device current, through hipGetSymbolAddress and
// kept per device.  Launches are told apart by a per-device epoch counter
// (as the caller-scratch path does), so nothing is cleared between calls;
// at epoch wrap the array is zeroed by a fill kernel.  Calls that overlap on
// one GPU (other streams, other threads) share the array and race.
constexpr int64_t kNoAllocScratchBytes = 1 << 20;  // 64-byte header + 131,064 tile states
constexpr int kNoAllocMaxDevices = 64;
constexpr uint32_t kEpochLimit = 1u << 30;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

struct NoAllocState {
    void* ptr = nullptr;  // device address of `scratch` on this device
    uint32_t epoch = 0;   // last epoch handed out
};
static NoAllocState g_noalloc[kNoAllocMaxDevices];
static std::mutex g_noalloc_mu;

// Device address of the static array on the current device and the epoch
// for the next launch (zeroing the array on `stream` first when the epoch
// wraps).  Returns a status code.
static int32_t noalloc_acquire(hipStream_t stream, void** ptr, uint32_t* epoch) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices)
        return stx_fail(STX_EHIP, "setop_noalloc: no current HIP device (or device id above 63)");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocState& st = g_noalloc[dev];
    if (st.ptr == nullptr) {
        void* p = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || p == nullptr)
            return stx_fail(STX_EHIP, "setop_noalloc: hipGetSymbolAddress of the static scratch failed");
        st.ptr = p;
    }
    if (st.epoch + 1 >= kEpochLimit) {
        hipLaunchKernelGGL(zero_u64_kernel, dim3(256), dim3(256), 0, stream, (unsigned long long*)st.ptr,
                           kNoAllocScratchBytes / 8);
        st.epoch = 0;
    }
    *ptr = st.ptr;
    *epoch = ++st.epoch;
    return STX_OK;
}

extern "C" {

int32_t stx_scan_i32(const int32_t* in, const int32_t* d_n, int64_t n_bound, int32_t* out, int32_t* d_total,
                     uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    if (n_bound < 0) return STX_EINVAL;
    uint32_t ntiles = (uint32_t)((n_bound + 1 + kScanTile - 1) / kScanTile);
    if (ntiles == 0) ntiles = 1;
    if (lookback_bytes(ntiles) > scratch_bytes) return stx_fail(STX_EINVAL, "scan: scratch too small");
    hipLaunchKernelGGL(scan_i32_kernel, dim3(ntiles), dim3(kBlock), 0, stream, in, d_n, n_bound, ntiles, out, d_total,
                       status, scratch, epoch, stx_spin_limit());
    return stx_launch_status("scan_i32_kernel");
}

int64_t stx_scan_scratch_bytes(int64_t n_bound) {
    int64_t ntiles = (n_bound + 1 + kScanTile - 1) / kScanTile;
    return lookback_bytes(ntiles < 1 ? 1 : ntiles);
}

int64_t stx_setop_scratch_bytes(int64_t nrows_bound, int32_t group) {
    // the tile paths never use more tiles than the 4-lane row-group path
    (void)group;
    return fused_rows_scratch_bytes(nrows_bound, 4) > lookback_bytes(nrows_bound + 1)
               ? fused_rows_scratch_bytes(nrows_bound, 4)
               : lookback_bytes(nrows_bound + 1);
}

int64_t stx_setop_scratch_bytes_n(int64_t nrows_bound, int64_t elems, int32_t group) {
    const int64_t base = stx_setop_scratch_bytes(nrows_bound, group);
    if (group != -8) return base;
    const int64_t b = balanced_scratch_bytes(elems > 0 ? elems : 0, nrows_bound);
    return b > base ? b : base;
}

int32_t stx_setop(int32_t op, const int32_t* a_beg, const int32_t* a_end, const int32_t* a_off, const int32_t* a_map,
                  int32_t a_xf_mode, int32_t a_xf_param, int64_t a_n, const int32_t* b_beg, const int32_t* b_end,
                  const int32_t* b_off, const int32_t* b_map, int32_t b_xf_mode, int32_t b_xf_param, int64_t b_n,
                  const int32_t* d_nrows, int64_t nrows_bound, int32_t group, int32_t* out_off, int32_t* out_beg,
                  int32_t* out_end, int64_t capacity, int32_t* d_total, uint32_t* status, void* scratch,
                  int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    if (op < 0 || op > 2 || nrows_bound < 0 || a_off == nullptr) return stx_fail(STX_EINVAL, "setop: bad arguments");
    if (op != 0 && b_off == nullptr) return stx_fail(STX_EINVAL, "setop: intersection/difference need two operands");
    Operand A{a_beg, a_end, a_off, a_map, XForm{a_xf_mode, a_xf_param}};
    Operand B{b_beg, b_end, b_off, b_map, XForm{b_xf_mode, b_xf_param}};
    if ((a_xf_mode == XF_COARSEN && a_xf_param < 1) || (b_xf_mode == XF_COARSEN && b_xf_param < 1))
        return stx_fail(STX_EINVAL, "setop: coarsening ratio must be >= 1");
    const bool tile_path = a_map == nullptr && b_map == nullptr && d_nrows == nullptr && group <= 0;
    if (tile_path) {
        const int64_t elems = a_n + (b_off ? b_n : 0);
        const int64_t spin = stx_spin_limit();
        if (group == -5) {
            // flat tiles (rows merged as one keyed sequence), 512 x 8
            const int32_t rpt = setop_tile_rows<512, 8>(elems, nrows_bound, kTileFill, flat_rows_per_tile<512>());
            const uint32_t ntiles = (uint32_t)((nrows_bound + rpt) / rpt);
            if (lookback_bytes(ntiles) > scratch_bytes) return stx_fail(STX_EINVAL, "setop: scratch too small");
            launch_flat<512, 8>(op, ntiles, stream, A, B, nrows_bound, rpt, out_off, out_beg, out_end, capacity, d_total,
                                status, scratch, epoch, spin);
            return stx_launch_status("setop_flat_kernel");
        }
        const bool plain2 = a_xf_mode == XF_NONE && b_off != nullptr && b_xf_mode == XF_NONE;
        if (group == -8 && plain2 && nrows_bound < INT32_MAX) {
            // register-output tiles with count-balanced rows: a tile table
            // in the scratch after the look-back states
            const int64_t nt = balanced_tiles(elems, nrows_bound);
            if (balanced_scratch_bytes(elems, nrows_bound) > scratch_bytes)
                return stx_fail(STX_EINVAL, "setop: scratch too small (stx_setop_scratch_bytes_n)");
            int2* tab = (int2*)((char*)scratch + balanced_table_offset(nt));
            const int64_t nthreads = (nrows_bound > nt ? nrows_bound : nt) + 1;
            hipLaunchKernelGGL(tile_table_kernel, dim3((unsigned)((nthreads + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               stream, a_off, b_off, nrows_bound, kBalancedTarget, (int64_t)flat_rows_per_tile<1024>(),
                               nt, tab, status);
            launch_flat2<1024, 8>(op, (uint32_t)nt, stream, A, B, nrows_bound, 1, out_off, out_beg, out_end, capacity,
                                  d_total, status, scratch, epoch, spin, tab);
            return stx_launch_status("setop_flat2_kernel(balanced)");
        }
        if ((group == -7 || (group == 0 && kFlatDefault)) && plain2) {
            // flat tiles in the register-output form, 1024 x 8 (two per CU),
            // filled to 95 %; other operands take the flat tiles below
            const int32_t rpt = setop_tile_rows<1024, 8>(elems, nrows_bound, kFlatFill, flat_rows_per_tile<1024>());
            const uint32_t ntiles = (uint32_t)((nrows_bound + rpt) / rpt);
            if (lookback_bytes(ntiles) > scratch_bytes) return stx_fail(STX_EINVAL, "setop: scratch too small");
            launch_flat2<1024, 8>(op, ntiles, stream, A, B, nrows_bound, rpt, out_off, out_beg, out_end, capacity,
                                  d_total, status, scratch, epoch, spin);
            return stx_launch_status("setop_flat2_kernel");
        }
        if (group == -6 || group == -7 || group == -8 || (group == 0 && kFlatDefault)) {
            // flat tiles, 1024 x 8 (two per CU), filled to 95 %: the mean
            // tile holds ~100 rows or more, so its size varies little
            const int32_t rpt = setop_tile_rows<1024, 8>(elems, nrows_bound, kFlatFill, flat_rows_per_tile<1024>());
            const uint32_t ntiles = (uint32_t)((nrows_bound + rpt) / rpt);
            if (lookback_bytes(ntiles) > scratch_bytes) return stx_fail(STX_EINVAL, "setop: scratch too small");
            launch_flat<1024, 8>(op, ntiles, stream, A, B, nrows_bound, rpt, out_off, out_beg, out_end, capacity,
                                 d_total, status, scratch, epoch, spin);
            return stx_launch_status("setop_flat_kernel");
        }
        if (group == -3 || group == -4) {
            // narrower tile shapes, single pass: -3 = 128 x 16, -4 = 64 x 32
            const int32_t rpt = group == -3 ? setop_tile_rows<128, 16>(elems, nrows_bound)
                                            : setop_tile_rows<64, 32>(elems, nrows_bound);
            const uint32_t ntiles = (uint32_t)((nrows_bound + rpt) / rpt);
            if (lookback_bytes(ntiles) > scratch_bytes) return stx_fail(STX_EINVAL, "setop: scratch too small");
            if (group == -3)
                launch_tile<kFused, 128, 16>(op, ntiles, stream, A, B, nrows_bound, rpt, out_off, out_beg, out_end,
                                             capacity, d_total, status, scratch, epoch, spin);
            else
                launch_tile<kFused, 64, 32>(op, ntiles, stream, A, B, nrows_bound, rpt, out_off, out_beg, out_end,
                                            capacity, d_total, status, scratch, epoch, spin);
            return stx_launch_status("setop_tile_kernel");
        }
        const int32_t rpt = setop_tile_rows<512, 8>(elems, nrows_bound);
        const uint32_t ntiles = (uint32_t)((nrows_bound + rpt) / rpt);
        // row-aware tiles: -1 single pass; -2 two passes
        const bool two_pass = group == -2 || (group == 0 && ntiles >= kTwoPassTiles);
        if (lookback_bytes(ntiles) > scratch_bytes) return stx_fail(STX_EINVAL, "setop: scratch too small");
        if (!two_pass) {
            launch_tile<kFused>(op, ntiles, stream, A, B, nrows_bound, rpt, out_off, out_beg, out_end, capacity, d_total,
                                status, scratch, epoch, spin);
            return stx_launch_status("setop_tile_kernel");
        }
        launch_tile<kCount>(op, ntiles, stream, A, B, nrows_bound, rpt, out_off, out_beg, out_end, capacity, d_total,
                            status, scratch, epoch, spin);
        hipLaunchKernelGGL(prefix_states_kernel, dim3(1), dim3(kPrefixThreads), 0, stream, scratch, ntiles, epoch);
        launch_tile<kWrite>(op, ntiles, stream, A, B, nrows_bound, rpt, out_off, out_beg, out_end, capacity, d_total,
                            status, scratch, epoch, spin);
        return stx_launch_status("setop_tile_kernel(two-pass)");
    }
    const int32_t g = group > 0 ? group : 64;
    if (g >= 64)
        return launch_setop_g<64>(op, A, B, d_nrows, nrows_bound, out_off, out_beg, out_end, capacity, d_total, status,
                                  scratch, scratch_bytes, epoch, stream);
    if (g >= 16)
        return launch_setop_g<16>(op, A, B, d_nrows, nrows_bound, out_off, out_beg, out_end, capacity, d_total, status,
                                  scratch, scratch_bytes, epoch, stream);
    return launch_setop_g<4>(op, A, B, d_nrows, nrows_bound, out_off, out_beg, out_end, capacity, d_total, status,
                             scratch, scratch_bytes, epoch, stream);
}

int32_t stx_rows_union(const int32_t* rows_a, int64_t na, int32_t shift_a, const int32_t* rows_b, int64_t nb,
                       int32_t shift_b, int32_t* rows_u, int32_t* map_a, int32_t* map_b, int32_t* d_nu, int32_t* tmp,
                       uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream) {
    // tmp: 2*(nb+1) int32 (flags, positions)
    if (na < 0 || nb < 0) return stx_fail(STX_EINVAL, "rows_union: bad sizes");
    int32_t* flag = tmp;
    int32_t* pos = tmp + (nb + 1);
    if (nb > 0)
        hipLaunchKernelGGL(rows_newflag_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, stream, rows_a, na, shift_a,
                           rows_b, nb, shift_b, flag);
    int32_t rc = stx_scan_i32(flag, nullptr, nb, pos, nullptr, status, scratch, scratch_bytes, epoch, stream);
    if (rc) return rc;
    if (na + nb > 0)
        hipLaunchKernelGGL(rows_scatter_kernel, dim3((unsigned)((na + nb + 255) / 256)), dim3(256), 0, stream, rows_a, na,
                           shift_a, rows_b, nb, shift_b, flag, pos, rows_u, map_a, map_b);
    hipLaunchKernelGGL(count_plus_kernel, dim3(1), dim3(1), 0, stream, d_nu, pos + nb, (int32_t)na);
    return stx_launch_status("rows_union");
}


int64_t stx_setop_noalloc_max_scratch_bytes(void) { return kNoAllocScratchBytes; }

int32_t stx_setop_noalloc(int32_t op, const int32_t* a_beg, const int32_t* a_end, const int32_t* a_off, int64_t a_n,
                          const int32_t* b_beg, const int32_t* b_end, const int32_t* b_off, int64_t b_n, int64_t nrows,
                          int32_t* out_off, int32_t* out_beg, int32_t* out_end, int64_t capacity, int32_t* d_total,
                          uint32_t* status, hipStream_t stream) {
    if (op < 0 || op > 2 || nrows < 0 || a_off == nullptr || a_n < 0 || b_n < 0)
        return stx_fail(STX_EINVAL, "setop_noalloc: bad arguments");
    if (op != 0 && b_off == nullptr)
        return stx_fail(STX_EINVAL, "setop_noalloc: intersection/difference need two operands");
    const Operand A{a_beg, a_end, a_off, nullptr, XForm{XF_NONE, 0}};
    const Operand B{b_beg, b_end, b_off, nullptr, XForm{XF_NONE, 0}};
    const int64_t elems = a_n + (b_off ? b_n : 0);
    // the flat 1024 x 8 tiles of stx_setop's default path (register-output
    // form for two operands)
    const int32_t rpt = setop_tile_rows<1024, 8>(elems, nrows, kFlatFill, flat_rows_per_tile<1024>());
    const int64_t ntiles = (nrows + rpt) / rpt;
    if (lookback_bytes(ntiles) > kNoAllocScratchBytes)
        return stx_fail(STX_EINVAL, "setop_noalloc: the operands need more look-back tiles than the 1 MiB static "
                                    "scratch holds (131,064 tiles of up to 8,192 intervals); use stx_setop with a "
                                    "caller scratch");
    void* sp = nullptr;
    uint32_t epoch = 0;
    const int32_t rc = noalloc_acquire(stream, &sp, &epoch);
    if (rc) return rc;
    if (b_off != nullptr) {
        launch_flat2<1024, 8>(op, (uint32_t)ntiles, stream, A, B, nrows, rpt, out_off, out_beg, out_end, capacity,
                              d_total, status, sp, epoch, stx_spin_limit());
        return stx_launch_status("setop_flat2_kernel(noalloc)");
    }
    launch_flat<1024, 8>(op, (uint32_t)ntiles, stream, A, B, nrows, rpt, out_off, out_beg, out_end, capacity, d_total,
                         status, sp, epoch, stx_spin_limit());
    return stx_launch_status("setop_flat_kernel(noalloc)");
}

}  // extern "C"
