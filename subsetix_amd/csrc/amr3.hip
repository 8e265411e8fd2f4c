// Dense-level batched three-level AMR: regrid and step kernels whose launch
// shapes do not depend on the data, so a whole regrid and the steps after it
// run with no host synchronisation and can be captured once as one graph.
//
// Layout (per level L = 0, 1, 2; instance i has side s = n_i 2^L):
//   * values: dense row-major per instance, cell (y, x) at cell0_i + y p + x
//     (int64; p the row pitch).  Level 0 is the full domain, stored with a
//     one-cell ring of zeros around each instance (p = s + 2; the ring is
//     never written, so a neighbour outside the instance reads the zero
//     exterior); levels 1 and 2 (p = s) hold meaningful values only on their
//     level sets;
//   * masks (levels 0 and 1): one 64-bit word per 64-cell row chunk, word
//     (y, m) at word0_i + y cw_i + m, bit b = cell x = 64 m + b; bits past the
//     row end are 0;
//   * work lists: 64-bit chunk descriptors (i << 48) | (y << 24) | m with a
//     device-side count; consumers are persistent grid-stride kernels (one
//     wavefront, or one lane, per chunk).
// Level tables: int64[5 n] per level: side, first cell (cell (0, 0)), first
// word, words per row, row pitch.
//
// The level sets follow from two refine masks: level 1 = prolong(R0),
// level 2 = prolong(R1) (R0 at level-0 resolution, R1 at level 1).  Ghost
// cells (the one-cell ring of a level inside its domain) are never stored:
// a ghost's value is its parent's current value, read where the stencil
// needs it.  Every other value, tag and mask follows the same arithmetic,
// neighbour rules and set algebra as the interval-set engine, cell for cell.
#include "stx_common.h"
#include "stx_host.h"

using namespace stx;

namespace {

constexpr int kT = 256;
constexpr int64_t kMaxBlocks = 4096;

struct Lvl {
    const int64_t* t;
    __device__ __forceinline__ int32_t side(int i) const { return (int32_t)t[5 * i]; }
    __device__ __forceinline__ int64_t cell0(int i) const { return t[5 * i + 1]; }
    __device__ __forceinline__ int64_t word0(int i) const { return t[5 * i + 2]; }
    __device__ __forceinline__ int32_t cw(int i) const { return (int32_t)t[5 * i + 3]; }
    __device__ __forceinline__ int32_t pitch(int i) const { return (int32_t)t[5 * i + 4]; }
    __device__ __forceinline__ int64_t word(int i, int y, int m) const { return word0(i) + (int64_t)y * cw(i) + m; }
};

struct Chunk {
    int32_t i, y, m;
};

__device__ __forceinline__ Chunk decode(uint64_t d) {
    return Chunk{(int32_t)(d >> 48), (int32_t)((d >> 24) & 0xffffffu), (int32_t)(d & 0xffffffu)};
}

__device__ __forceinline__ uint64_t encode(int32_t i, int32_t y, int32_t m) {
    return ((uint64_t)i << 48) | ((uint64_t)y << 24) | (uint64_t)m;
}

__device__ __forceinline__ uint64_t valid_bits(int32_t side, int32_t m) {
    const int32_t rem = side - 64 * m;
    return rem >= 64 ? ~0ull : (rem <= 0 ? 0ull : ((1ull << rem) - 1ull));
}

// mask word (y, m) of instance i, 0 outside the instance
__device__ __forceinline__ uint64_t mword(const uint64_t* M, const Lvl& L, int i, int y, int m) {
    if (y < 0 || y >= L.side(i) || m < 0 || m >= L.cw(i)) return 0ull;
    return M[L.word(i, y, m)];
}

// each of the 32 bits of h doubled (bit b -> bits 2b, 2b + 1)
__device__ __forceinline__ uint64_t spread2(uint32_t h) {
    uint64_t x = h;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x | (x << 1);
}

// word (Y, M) of the ratio-2 prolongation of a parent-level mask (0 outside
// the child instance)
__device__ __forceinline__ uint64_t child_word(const uint64_t* Mp, const Lvl& Lp, const Lvl& Lc, int i, int Y, int M) {
    const int32_t s = Lc.side(i);
    if (Y < 0 || Y >= s || M < 0 || M >= Lc.cw(i)) return 0ull;
    const uint64_t pw = Mp[Lp.word(i, Y >> 1, M >> 1)];
    return spread2((uint32_t)(pw >> ((M & 1) * 32))) & valid_bits(s, M);
}

// one row's contribution to a Moore (8-neighbour) dilation of word c, with
// the words left / right of it
__device__ __forceinline__ uint64_t row_dil(uint64_t l, uint64_t c, uint64_t r) {
    return c | (c << 1) | (l >> 63) | (c >> 1) | (r << 63);
}

__device__ __forceinline__ int64_t nitems(const int32_t* d_n, int64_t bound) {
    return d_n ? (int64_t)*d_n : bound;
}

inline unsigned grid_for(int64_t items, int per_block) {
    int64_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    return (unsigned)(b < kMaxBlocks ? b : kMaxBlocks);
}

template <typename T>
__device__ __forceinline__ bool grad_tag(T c, T l, T r, T d, T u, bool hl, bool hr, bool hd, bool hu, T smin) {
    T gx = (T)0, gy = (T)0;
    if (hl && hr) gx = (r - l) * (T)0.5;
    else if (hr) gx = r - c;
    else if (hl) gx = c - l;
    if (hd && hu) gy = (u - d) * (T)0.5;
    else if (hu) gy = u - c;
    else if (hd) gy = c - d;
    const T gx2 = gx * gx;
    const T gy2 = gy * gy;
    return gx2 + gy2 >= smin;
}

// Range of |a| in which div_rn's product form is exact (its residuals stay
// normal numbers); outside it, and for 0, inf and NaN, it divides.
template <typename T>
struct DivRange;
template <>
struct DivRange<float> {
    static constexpr float lo = 0x1p-64f, hi = 0x1p64f;
};
template <>
struct DivRange<double> {
    static constexpr double lo = 0x1p-800, hi = 0x1p800;
};

// a / d correctly rounded, for d > 0 and y = RN(1 / d) (a per-instance
// coefficient).  q0 = a y lies within 1.5 ulp of a / d; one correction,
// q1 = q0 + (a - q0 d) y, leaves it within 1/2 ulp + 2^-22 ulp, a faithful
// quotient; with q1 faithful and y correctly rounded, the residual
// a - q1 d is exact and q1 + (a - q1 d) y rounds to RN(a / d) (Markstein's
// theorem).  Checked exhaustively in FP32 for every float a in a binade and
// every d = 1/n, n <= 16384, at both ends of the range.
__device__ __forceinline__ float fma_rn(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_rn(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ __forceinline__ T div_rn(T a, T d, T y) {
    const T aa = a < (T)0 ? -a : a;
    if (aa >= DivRange<T>::lo && aa <= DivRange<T>::hi) {
        const T q0 = a * y;
        const T r0 = fma_rn(-q0, d, a);
        const T q1 = fma_rn(r0, y, q0);
        const T r1 = fma_rn(-q1, d, a);
        return fma_rn(r1, y, q1);
    }
    return a == (T)0 ? a : a / d;
}

// Coefficients per instance: a, b, dt, dx, dy, RN(1/dx), RN(1/dy).
constexpr int kCoef = 7;

template <typename T>
__device__ __forceinline__ T upwind(T c, T l, T r, T d, T u, const T* k) {
    const T a = k[0], b = k[1], dt = k[2], dx = k[3], dy = k[4];
    const T nx = (a >= (T)0) ? c - l : r - c;
    const T ny = (b >= (T)0) ? c - d : u - c;
    const T dudx = div_rn<T>(nx, dx, k[5]);
    const T dudy = div_rn<T>(ny, dy, k[6]);
    const T t1 = a * dudx;
    const T t2 = b * dudy;
    return c - dt * (t1 + t2);
}

#define WAVE_LOOP(n)                                                                  \
    const int lane = lane_id();                                                       \
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;      \
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;                    \
    for (int64_t t = wave0; t < (n); t += nwaves)

#define LANE_LOOP(n)                                                                  \
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (n); t += (int64_t)gridDim.x * blockDim.x)

// ---- regrid ------------------------------------------------------------------


// out = Moore dilation of `in` by one cell, clipped to the instance (one lane
// per listed word).
__global__ __launch_bounds__(kT) void dilate_kernel(Lvl L, const uint64_t* __restrict__ list, const int32_t* d_n,
                                                    int64_t bound, const uint64_t* __restrict__ in,
                                                    uint64_t* __restrict__ out) {
    const int64_t n = nitems(d_n, bound);
    LANE_LOOP(n) {
        const Chunk ch = decode(list[t]);
        uint64_t acc = 0;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = ch.y + dy;
            acc |= row_dil(mword(in, L, ch.i, y, ch.m - 1), mword(in, L, ch.i, y, ch.m), mword(in, L, ch.i, y, ch.m + 1));
        }
        out[L.word(ch.i, ch.y, ch.m)] = acc & valid_bits(L.side(ch.i), ch.m);
    }
}

// R = D & erode(prolong(Rp)): the cells of D whose 8 neighbours inside the
// instance all belong to the level set prolong(Rp) (one lane per word).
__global__ __launch_bounds__(kT) void refine_kernel(Lvl Lp, Lvl Lc, const uint64_t* __restrict__ list, const int32_t* d_n,
                                                    int64_t bound, const uint64_t* __restrict__ Rp,
                                                    const uint64_t* __restrict__ D, uint64_t* __restrict__ R) {
    const int64_t n = nitems(d_n, bound);
    LANE_LOOP(n) {
        const Chunk ch = decode(list[t]);
        const int32_t s = Lc.side(ch.i);
        uint64_t outside = 0;  // dilation of (instance cells not in the level set)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = ch.y + dy;
            uint64_t w[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int m = ch.m - 1 + k;
                const bool in_dom = y >= 0 && y < s && m >= 0 && m < Lc.cw(ch.i);
                w[k] = in_dom ? (~child_word(Rp, Lp, Lc, ch.i, y, m) & valid_bits(s, m)) : 0ull;
            }
            outside |= row_dil(w[0], w[1], w[2]);
        }
        const uint64_t l = child_word(Rp, Lp, Lc, ch.i, ch.y, ch.m);
        const int64_t wd = Lc.word(ch.i, ch.y, ch.m);
        R[wd] = D[wd] & l & ~outside;
    }
}



// Work-list construction.  mode 0: the listed parent chunk itself when its
// E word is non-zero; mode 1: the child chunks (rows 2y, 2y + 1; chunks 2m,
// 2m + 1) under the non-zero halves of the E word.
__device__ __forceinline__ int32_t emit_count(int mode, const Lvl& Lp, const Lvl& Lc, const Chunk& ch, uint64_t e) {
    if (mode == 0) return e != 0 ? 1 : 0;
    int32_t c = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h)
        if (2 * ch.m + h < Lc.cw(ch.i) && (uint32_t)(e >> (32 * h)) != 0u) c += 2;
    return c;
}

__global__ __launch_bounds__(kT) void list_count_kernel(int mode, Lvl Lp, Lvl Lc, const uint64_t* __restrict__ list,
                                                        const int32_t* d_n, int64_t bound, const uint64_t* __restrict__ E,
                                                        int32_t* __restrict__ counts) {
    const int64_t n = nitems(d_n, bound);
    LANE_LOOP(n) {
        const Chunk ch = decode(list[t]);
        counts[t] = emit_count(mode, Lp, Lc, ch, E[Lp.word(ch.i, ch.y, ch.m)]);
    }
}

__global__ __launch_bounds__(kT) void list_emit_kernel(int mode, Lvl Lp, Lvl Lc, const uint64_t* __restrict__ list,
                                                       const int32_t* d_n, int64_t bound, const uint64_t* __restrict__ E,
                                                       const int32_t* __restrict__ offsets, uint64_t* __restrict__ out) {
    const int64_t n = nitems(d_n, bound);
    LANE_LOOP(n) {
        const uint64_t d = list[t];
        const Chunk ch = decode(d);
        const uint64_t e = E[Lp.word(ch.i, ch.y, ch.m)];
        int64_t o = offsets[t];
        if (mode == 0) {
            if (e != 0) out[o] = d;
            continue;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (2 * ch.m + h < Lc.cw(ch.i) && (uint32_t)(e >> (32 * h)) != 0u) {
                out[o++] = encode(ch.i, 2 * ch.y, 2 * ch.m + h);
                out[o++] = encode(ch.i, 2 * ch.y + 1, 2 * ch.m + h);
            }
        }
    }
}

// Zero up to three masks over the listed words.
__global__ __launch_bounds__(kT) void zero_kernel(Lvl L, const uint64_t* __restrict__ list, const int32_t* d_n,
                                                  int64_t bound, uint64_t* m0, uint64_t* m1, uint64_t* m2) {
    const int64_t n = nitems(d_n, bound);
    LANE_LOOP(n) {
        const Chunk ch = decode(list[t]);
        const int64_t w = L.word(ch.i, ch.y, ch.m);
        if (m0) m0[w] = 0ull;
        if (m1) m1[w] = 0ull;
        if (m2) m2[w] = 0ull;
    }
}

// Set bits of a mask over the listed words, added to *out (one atomic per
// block).
__global__ __launch_bounds__(kT) void popc_kernel(Lvl L, const uint64_t* __restrict__ list, const int32_t* d_n,
                                                  int64_t bound, const uint64_t* __restrict__ M,
                                                  unsigned long long* out) {
    __shared__ unsigned long long s[kT / 64];
    const int64_t n = nitems(d_n, bound);
    unsigned long long acc = 0;
    LANE_LOOP(n) {
        const Chunk ch = decode(list[t]);
        acc += (unsigned long long)__popcll(M[L.word(ch.i, ch.y, ch.m)]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane_id() == 0) s[wave_id()] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = 0;
        for (int w = 0; w < kT / 64; ++w) b += s[w];
        if (b) atomicAdd(out, b);
    }
}

// ---- step ----------------------------------------------------------------------




// ---- work records ----------------------------------------------------------
// Built once per regrid so that a step's kernels start from one coalesced
// load per record (no table or mask look-ups):
//   stencil record (8 words), for a chunk holding cells of the level set:
//     desc, first cell of the chunk, parent cell under its first cell,
//     side | parent pitch << 32, the level-set words of the chunk, of the rows
//     below and above, and the edge bits (bit 0: the cell left of the chunk,
//     bit 1: the cell right of it);
//   restriction record (4 words), for a parent chunk holding refined cells:
//     first parent cell, first child cell (row 2y), child pitch, refine word;
//   level-0 record (4 words, static, one per chunk): first cell, mask word,
//     side | pitch << 32, chunk descriptor.
constexpr int kSW = 8;
constexpr int kQW = 4;

__device__ __forceinline__ uint64_t bcast(uint64_t w, int src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)w, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(w >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// Records of one listed chunk.  kind 0: the listed chunk is a parent
// (level Lp) chunk; a record for each of its child chunks (rows 2y, 2y + 1,
// chunks 2m + h) under a non-zero half h of its refine word Rp, i.e. the
// child chunks holding cells of the level set prolong(Rp).  kind 1: one
// record for the listed chunk when its refine word Rp is non-zero.
__device__ __forceinline__ int rec_count(int kind, const Lvl& Lc, const Chunk& ch, uint64_t rw) {
    if (kind == 1) return rw != 0 ? 1 : 0;
    int c = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h)
        if (2 * ch.m + h < Lc.cw(ch.i) && (uint32_t)(rw >> (32 * h)) != 0u) c += 2;
    return c;
}

__global__ __launch_bounds__(kT) void rec_count_kernel(int kind, Lvl Lp, Lvl Lc, const uint64_t* __restrict__ list,
                                                       const int32_t* d_n, int64_t bound, const uint64_t* __restrict__ Rp,
                                                       int32_t* __restrict__ counts) {
    const int64_t n = nitems(d_n, bound);
    LANE_LOOP(n) {
        const Chunk ch = decode(list[t]);
        counts[t] = rec_count(kind, Lc, ch, Rp[Lp.word(ch.i, ch.y, ch.m)]);
    }
}

__global__ __launch_bounds__(kT) void rec_emit_kernel(int kind, Lvl Lp, Lvl Lc, const uint64_t* __restrict__ list,
                                                      const int32_t* d_n, int64_t bound, const uint64_t* __restrict__ Rp,
                                                      const int32_t* __restrict__ offsets, uint64_t* __restrict__ out) {
    const int64_t n = nitems(d_n, bound);
    LANE_LOOP(n) {
        const Chunk ch = decode(list[t]);
        const int i = ch.i;
        const uint64_t rw = Rp[Lp.word(i, ch.y, ch.m)];
        if (rec_count(kind, Lc, ch, rw) == 0) continue;
        int64_t o = offsets[t];
        if (kind == 1) {
            const int32_t pq = Lp.pitch(i), pc = Lc.pitch(i);
            uint64_t* r = out + kQW * o;
            r[0] = (uint64_t)(Lp.cell0(i) + (int64_t)ch.y * pq + 64 * ch.m);
            r[1] = (uint64_t)(Lc.cell0(i) + (int64_t)(2 * ch.y) * pc + 128 * ch.m);
            r[2] = (uint64_t)pc;
            r[3] = rw;
            continue;
        }
        // own level unpadded (pitch = side); sp: the parent level's pitch
        const int32_t s = Lc.side(i), sp = Lp.pitch(i);
        for (int h = 0; h < 2; ++h) {
            const int M = 2 * ch.m + h;
            if (M >= Lc.cw(i) || (uint32_t)(rw >> (32 * h)) == 0u) continue;
            for (int dy = 0; dy < 2; ++dy, ++o) {
                const int Y = 2 * ch.y + dy;
                uint64_t* r = out + kSW * o;
                const uint64_t wl = child_word(Rp, Lp, Lc, i, Y, M - 1), wr = child_word(Rp, Lp, Lc, i, Y, M + 1);
                r[0] = encode(i, Y, M);
                r[1] = (uint64_t)(Lc.cell0(i) + (int64_t)Y * s + 64 * M);
                r[2] = (uint64_t)(Lp.cell0(i) + (int64_t)ch.y * sp + 32 * M);
                r[3] = (uint64_t)(uint32_t)s | ((uint64_t)(uint32_t)sp << 32);
                r[4] = child_word(Rp, Lp, Lc, i, Y, M);
                r[5] = child_word(Rp, Lp, Lc, i, Y - 1, M);
                r[6] = child_word(Rp, Lp, Lc, i, Y + 1, M);
                r[7] = ((wl >> 63) & 1ull) | ((wr & 1ull) << 1);
            }
        }
    }
}

// Level-0 records of every chunk (built once).
__global__ __launch_bounds__(kT) void tag0_rec_build_kernel(Lvl L, const uint64_t* __restrict__ list, int64_t n,
                                                            uint64_t* __restrict__ out) {
    LANE_LOOP(n) {
        const Chunk ch = decode(list[t]);
        const int32_t s = L.side(ch.i), p = L.pitch(ch.i);
        uint64_t* r = out + kQW * t;
        r[0] = (uint64_t)(L.cell0(ch.i) + (int64_t)ch.y * p + 64 * ch.m);
        r[1] = (uint64_t)L.word(ch.i, ch.y, ch.m);
        r[2] = (uint64_t)(uint32_t)s | ((uint64_t)(uint32_t)p << 32);
        r[3] = list[t];
    }
}

constexpr int kG = 8;  // records whose loads a wavefront issues together

// Level-0 record kernels run one block (4 wavefronts x 16 records) per tile
// of 64 records, tiles handed to the 8 XCDs in chunks of 32 consecutive
// tiles, so the rows above and below a chunk's rows mostly sit in the same
// XCD's L2.
constexpr int kXcd = 8, kXcdChunk = 32, kRecTile = 64;

__device__ __forceinline__ int64_t xcd_tile(int64_t block) {
    const int64_t xcd = block % kXcd, nth = block / kXcd;
    return ((nth / kXcdChunk) * kXcd + xcd) * kXcdChunk + nth % kXcdChunk;
}

inline unsigned xcd_grid(int64_t n) {
    const int64_t tiles = (n + kRecTile - 1) / kRecTile;
    const int64_t per = (int64_t)kXcd * kXcdChunk;
    return (unsigned)((tiles + per - 1) / per * per);
}

// Persistent record kernels whose count is known only on the device: the
// records split into one contiguous slab per XCD (a multiple of `per_wave`
// records each), walked by that XCD's wavefronts (blocks b with b % 8 the
// XCD) in steps of `per_wave` records; the grid is a multiple of 8 blocks.
struct Slab {
    int64_t begin, end, step;
};

__device__ __forceinline__ Slab xcd_slab(int64_t n, int per_wave) {
    const int64_t xcd = blockIdx.x % kXcd, j = blockIdx.x / kXcd;
    const int64_t wpb = blockDim.x >> 6;
    const int64_t waves = (int64_t)(gridDim.x / kXcd) * wpb;
    const int64_t per = ((n + kXcd - 1) / kXcd + per_wave - 1) / per_wave * per_wave;
    const int64_t end = (xcd + 1) * per;
    return Slab{xcd * per + (j * wpb + wave_id()) * per_wave, end < n ? end : n, waves * per_wave};
}

inline unsigned slab_grid(int64_t items, int per_block) {
    const unsigned g = grid_for(items, per_block);
    return (g + kXcd - 1) / kXcd * kXcd;
}

// Level-0 gradient tags from the static records, 16 records per wavefront
// iteration in groups of kG.  The zero ring around each instance (and the
// slack after the last one) keeps every neighbour load in bounds, so every
// load of a group is issued unconditionally before any is used; slots past
// n repeat the last record (their tags are not stored).
template <typename T>
__global__ __launch_bounds__(kT) void tag0_rec_kernel(const uint64_t* __restrict__ rec, int64_t n, const T* __restrict__ V,
                                                      T smin, uint64_t* __restrict__ out) {
    constexpr int U = 64 / kQW;
    const int lane = lane_id();
    {
        const int64_t base = xcd_tile(blockIdx.x) * kRecTile + (int64_t)wave_id() * U;
        if (base >= n) return;
        const int64_t tr = base + (lane / kQW);
        const uint64_t w = rec[(tr < n ? tr : n - 1) * kQW + (lane % kQW)];
#pragma unroll
        for (int g = 0; g < U; g += kG) {
            T cv[kG], l[kG], r[kG], d[kG], up[kG];
#pragma unroll
            for (int u = 0; u < kG; ++u) {
                const int k = g + u;
                const int64_t c = (int64_t)bcast(w, kQW * k) + lane;
                const int64_t p = (int64_t)(bcast(w, kQW * k + 2) >> 32);
                cv[u] = V[c];
                l[u] = V[c - 1];
                r[u] = V[c + 1];
                d[u] = V[c - p];
                up[u] = V[c + p];
            }
#pragma unroll
            for (int u = 0; u < kG; ++u) {
                const int k = g + u;
                const int64_t widx = (int64_t)bcast(w, kQW * k + 1);
                const int32_t s = (int32_t)(uint32_t)bcast(w, kQW * k + 2);
                const Chunk ch = decode(bcast(w, kQW * k + 3));
                const int32_t y = ch.y, x = 64 * ch.m + lane;
                const bool tag = x < s && grad_tag<T>(cv[u], l[u], r[u], d[u], up[u], x > 0, x + 1 < s, y > 0,
                                                      y + 1 < s, smin);
                const uint64_t bits = __ballot(tag);
                if (lane == 0 && base + k < n) out[widx] = bits;
            }
        }
    }
}

// Level-0 upwind over the padded full domain from the same records: the
// ring supplies the zero exterior, so the update needs no neighbour test.
template <typename T>
__global__ __launch_bounds__(kT) void upwind0_rec_kernel(const uint64_t* __restrict__ rec, int64_t n,
                                                         const T* __restrict__ V, T* __restrict__ Vn,
                                                         const T* __restrict__ coef) {
    constexpr int U = 64 / kQW;
    const int lane = lane_id();
    {
        const int64_t base = xcd_tile(blockIdx.x) * kRecTile + (int64_t)wave_id() * U;
        if (base >= n) return;
        const int64_t tr = base + (lane / kQW);
        const uint64_t w = rec[(tr < n ? tr : n - 1) * kQW + (lane % kQW)];
#pragma unroll
        for (int g = 0; g < U; g += kG) {
            T cv[kG], l[kG], r[kG], d[kG], up[kG];
#pragma unroll
            for (int u = 0; u < kG; ++u) {
                const int k = g + u;
                const int64_t c = (int64_t)bcast(w, kQW * k) + lane;
                const int64_t p = (int64_t)(bcast(w, kQW * k + 2) >> 32);
                cv[u] = V[c];
                l[u] = V[c - 1];
                r[u] = V[c + 1];
                d[u] = V[c - p];
                up[u] = V[c + p];
            }
#pragma unroll
            for (int u = 0; u < kG; ++u) {
                const int k = g + u;
                const int32_t s = (int32_t)(uint32_t)bcast(w, kQW * k + 2);
                const Chunk ch = decode(bcast(w, kQW * k + 3));
                if (base + k >= n || 64 * ch.m + lane >= s) continue;
                const int64_t c = (int64_t)bcast(w, kQW * k) + lane;
                Vn[c] = upwind<T>(cv[u], l[u], r[u], d[u], up[u], coef + kCoef * ch.i);
            }
        }
    }
}

// Fields of stencil record u of a wavefront's batch.
struct SRec {
    int32_t i, y, x0, s, sp;
    int64_t c0, prow;
    uint64_t wc, wd, wu, edges;
};

__device__ __forceinline__ SRec srec(uint64_t w, int u) {
    SRec r;
    const Chunk ch = decode(bcast(w, kSW * u));
    r.i = ch.i, r.y = ch.y, r.x0 = 64 * ch.m;
    r.c0 = (int64_t)bcast(w, kSW * u + 1);
    r.prow = (int64_t)bcast(w, kSW * u + 2);
    const uint64_t sides = bcast(w, kSW * u + 3);
    r.s = (int32_t)(uint32_t)sides, r.sp = (int32_t)(sides >> 32);
    r.wc = bcast(w, kSW * u + 4), r.wd = bcast(w, kSW * u + 5), r.wu = bcast(w, kSW * u + 6);
    r.edges = bcast(w, kSW * u + 7);
    return r;
}

// Upwind over one list of stencil records (levels 1 or 2): a neighbour in
// the instance but not in the level set is a ghost, whose value is its
// parent's current value.  Branch-free loads as in tag0_rec_kernel: each
// neighbour's address is selected (own level, parent, or the cell itself
// when outside the instance) and all loads of the 8 records are issued
// before any is used.
template <typename T>
__device__ __forceinline__ void stencil_records(const uint64_t* __restrict__ rec, int64_t n, const T* __restrict__ Vp,
                                                const T* __restrict__ Vc, T* __restrict__ Vn, const T* __restrict__ coef,
                                                int lane) {
    constexpr int U = 64 / kSW;
    const Slab sl = xcd_slab(n, U);
    for (int64_t base = sl.begin; base < sl.end; base += sl.step) {
        const int64_t tr = base + (lane / kSW);
        // words of records past n read as 0: no cell of the level set
        const uint64_t w = tr < n ? rec[tr * kSW + (lane % kSW)] : 0ull;
        T cv[U], l[U], rr[U], d[U], up[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const SRec r = srec(w, u);
            const bool act = (r.wc >> lane) & 1ull;
            const int32_t x = r.x0 + lane;
            const int64_t c = act ? r.c0 + lane : r.c0;
            const bool inl = lane > 0 ? ((r.wc >> (lane - 1)) & 1ull) : (r.edges & 1ull);
            const bool inr = lane < 63 ? ((r.wc >> (lane + 1)) & 1ull) : ((r.edges >> 1) & 1ull);
            const bool ind = (r.wd >> lane) & 1ull, inu = (r.wu >> lane) & 1ull;
            // parent of (y, x +- 1): the parent row, column (x +- 1) / 2; of
            // (y -+ 1, x): one parent row down / up when y is even / odd
            const int64_t pc = r.prow + (lane >> 1);
            const T* al = !act || x == 0 ? Vc + c : (inl ? Vc + (c - 1) : Vp + (r.prow + ((lane - 1) >> 1)));
            const T* ar = !act || x + 1 >= r.s ? Vc + c : (inr ? Vc + (c + 1) : Vp + (r.prow + ((lane + 1) >> 1)));
            const T* ad = !act || r.y == 0 ? Vc + c : (ind ? Vc + (c - r.s) : Vp + ((r.y & 1) ? pc : pc - r.sp));
            const T* au = !act || r.y + 1 >= r.s ? Vc + c : (inu ? Vc + (c + r.s) : Vp + ((r.y & 1) ? pc + r.sp : pc));
            cv[u] = Vc[c];
            l[u] = *al;
            rr[u] = *ar;
            d[u] = *ad;
            up[u] = *au;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const SRec r = srec(w, u);
            if (!((r.wc >> lane) & 1ull)) continue;
            const int32_t x = r.x0 + lane;
            const T lv = x == 0 ? (T)0 : l[u];
            const T rv = x + 1 >= r.s ? (T)0 : rr[u];
            const T dv = r.y == 0 ? (T)0 : d[u];
            const T uv = r.y + 1 >= r.s ? (T)0 : up[u];
            Vn[r.c0 + lane] = upwind<T>(cv[u], lv, rv, dv, uv, coef + kCoef * r.i);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kT) void stencil_rec_kernel(const uint64_t* __restrict__ rec1, const int32_t* d_n1,
                                                         const uint64_t* __restrict__ rec2, const int32_t* d_n2,
                                                         const T* __restrict__ V0, const T* __restrict__ V1,
                                                         const T* __restrict__ V2, T* __restrict__ V1n,
                                                         T* __restrict__ V2n, const T* __restrict__ coef1,
                                                         const T* __restrict__ coef2) {
    const int lane = lane_id();
    stencil_records<T>(rec1, *d_n1, V0, V1, V1n, coef1, lane);
    stencil_records<T>(rec2, *d_n2, V1, V2, V2n, coef2, lane);
}

// Level-set values after a regrid from stencil records (new level set):
// cells not in the old level set take their parent's value.
template <typename T>
__global__ __launch_bounds__(kT) void lupdate_rec_kernel(Lvl Lp, Lvl Lc, const uint64_t* __restrict__ rec, const int32_t* d_n,
                                                         const uint64_t* __restrict__ Rold, const T* __restrict__ Vp,
                                                         T* __restrict__ Vc) {
    constexpr int U = 64 / kSW;
    const int64_t n = *d_n;
    const int lane = lane_id();
    const Slab sl = xcd_slab(n, U);
    for (int64_t base = sl.begin; base < sl.end; base += sl.step) {
        const int64_t tr = base + (lane / kSW);
        const uint64_t w = tr < n ? rec[tr * kSW + (lane % kSW)] : 0ull;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (base + u >= n) continue;
            const SRec r = srec(w, u);
            const uint64_t upd = r.wc & ~child_word(Rold, Lp, Lc, r.i, r.y, r.x0 >> 6);
            if ((upd >> lane) & 1ull) Vc[r.c0 + lane] = Vp[r.prow + (lane >> 1)];
        }
    }
}

// Gradient tags of a refined level from its stencil records.
template <typename T>
__global__ __launch_bounds__(kT) void tagl_rec_kernel(Lvl Lc, const uint64_t* __restrict__ rec, const int32_t* d_n,
                                                      const T* __restrict__ V, T smin, uint64_t* __restrict__ out) {
    constexpr int U = 64 / kSW;
    const int64_t n = *d_n;
    const int lane = lane_id();
    const Slab sl = xcd_slab(n, U);
    for (int64_t base = sl.begin; base < sl.end; base += sl.step) {
        const int64_t tr = base + (lane / kSW);
        const uint64_t w = tr < n ? rec[tr * kSW + (lane % kSW)] : 0ull;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (base + u >= n) continue;
            const SRec r = srec(w, u);
            bool tag = false;
            if ((r.wc >> lane) & 1ull) {
                const bool hl = lane > 0 ? ((r.wc >> (lane - 1)) & 1ull) : (r.edges & 1ull);
                const bool hr = lane < 63 ? ((r.wc >> (lane + 1)) & 1ull) : ((r.edges >> 1) & 1ull);
                const bool hd = (r.wd >> lane) & 1ull, hu = (r.wu >> lane) & 1ull;
                const int64_t c = r.c0 + lane;
                const T cv = V[c];
                const T l = hl ? V[c - 1] : (T)0, rr = hr ? V[c + 1] : (T)0;
                const T d = hd ? V[c - r.s] : (T)0, up = hu ? V[c + r.s] : (T)0;
                tag = grad_tag<T>(cv, l, rr, d, up, hl, hr, hd, hu, smin);
            }
            const uint64_t bits = __ballot(tag);
            if (lane == 0) out[Lc.word(r.i, r.y, r.x0 >> 6)] = bits;
        }
    }
}

// Restriction from records, 16 per wavefront iteration in groups of kG
// (branch-free loads): a refined parent cell takes the mean of its four
// children, summed dy-major in Acc.  The four children of a chunk's first
// cell exist whenever the chunk holds a refined cell; records past n read as
// 0 and load cells 0 and 1.
template <typename T, typename Acc>
__global__ __launch_bounds__(kT) void restrict_rec_kernel(const uint64_t* __restrict__ rec, const int32_t* d_n,
                                                          const T* __restrict__ Vc, T* __restrict__ Vp, Acc norm) {
    constexpr int U = 64 / kQW;
    const int64_t n = *d_n;
    const int lane = lane_id();
    const Slab sl = xcd_slab(n, U);
    for (int64_t base = sl.begin; base < sl.end; base += sl.step) {
        const int64_t tr = base + (lane / kQW);
        const uint64_t w = tr < n ? rec[tr * kQW + (lane % kQW)] : 0ull;
#pragma unroll
        for (int g = 0; g < U; g += kG) {
            T a[kG], b[kG], c[kG], e[kG];
#pragma unroll
            for (int u = 0; u < kG; ++u) {
                const int k = g + u;
                const uint64_t r = bcast(w, kQW * k + 3);
                const int64_t cc = (int64_t)bcast(w, kQW * k + 1);
                const int64_t sc = (int64_t)bcast(w, kQW * k + 2);
                const int64_t ci = ((r >> lane) & 1ull) ? cc + 2 * lane : cc;
                a[u] = Vc[ci];
                b[u] = Vc[ci + 1];
                c[u] = Vc[ci + sc];
                e[u] = Vc[ci + sc + 1];
            }
#pragma unroll
            for (int u = 0; u < kG; ++u) {
                const int k = g + u;
                const uint64_t r = bcast(w, kQW * k + 3);
                if (!((r >> lane) & 1ull)) continue;
                const int64_t pc = (int64_t)bcast(w, kQW * k);
                Acc acc = (Acc)0;
                acc = acc + (Acc)a[u];
                acc = acc + (Acc)b[u];
                acc = acc + (Acc)c[u];
                acc = acc + (Acc)e[u];
                Vp[pc + lane] = (T)(acc * norm);
            }
        }
    }
}

// cells_total += steps * (base + 4 * (bits[0] + bits[1]))  (level-1 and
// level-2 cells are four per refined parent cell)
__global__ void accumulate_kernel(unsigned long long* total, const unsigned long long* bits, int64_t base, int64_t steps) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *total += (unsigned long long)steps * ((unsigned long long)base + 4ull * (bits[0] + bits[1]));
}

}  // namespace

extern "C" {

int32_t stx_amr3_dilate(const int64_t* lt, const uint64_t* list, const int32_t* d_n, int64_t bound, const uint64_t* in,
                        uint64_t* out, hipStream_t stream) {
    if (bound <= 0) return STX_OK;
    hipLaunchKernelGGL(dilate_kernel, dim3(grid_for(bound, kT)), dim3(kT), 0, stream, Lvl{lt}, list, d_n, bound, in, out);
    return stx_launch_status("amr3_dilate");
}

int32_t stx_amr3_refine(const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n, int64_t bound,
                        const uint64_t* rp, const uint64_t* d, uint64_t* r, hipStream_t stream) {
    if (bound <= 0) return STX_OK;
    hipLaunchKernelGGL(refine_kernel, dim3(grid_for(bound, kT)), dim3(kT), 0, stream, Lvl{ltp}, Lvl{ltc}, list, d_n, bound,
                       rp, d, r);
    return stx_launch_status("amr3_refine");
}

int32_t stx_amr3_list_count(int32_t mode, const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n,
                            int64_t bound, const uint64_t* e, int32_t* counts, hipStream_t stream) {
    if (mode != 0 && mode != 1) return stx_fail(STX_EINVAL, "amr3_list_count: mode must be 0 or 1");
    if (bound <= 0) return STX_OK;
    hipLaunchKernelGGL(list_count_kernel, dim3(grid_for(bound, kT)), dim3(kT), 0, stream, (int)mode, Lvl{ltp}, Lvl{ltc},
                       list, d_n, bound, e, counts);
    return stx_launch_status("amr3_list_count");
}

int32_t stx_amr3_list_emit(int32_t mode, const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n,
                           int64_t bound, const uint64_t* e, const int32_t* offsets, uint64_t* out, hipStream_t stream) {
    if (mode != 0 && mode != 1) return stx_fail(STX_EINVAL, "amr3_list_emit: mode must be 0 or 1");
    if (bound <= 0) return STX_OK;
    hipLaunchKernelGGL(list_emit_kernel, dim3(grid_for(bound, kT)), dim3(kT), 0, stream, (int)mode, Lvl{ltp}, Lvl{ltc},
                       list, d_n, bound, e, offsets, out);
    return stx_launch_status("amr3_list_emit");
}

int32_t stx_amr3_zero(const int64_t* lt, const uint64_t* list, const int32_t* d_n, int64_t bound, uint64_t* m0,
                      uint64_t* m1, uint64_t* m2, hipStream_t stream) {
    if (bound <= 0) return STX_OK;
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(bound, kT)), dim3(kT), 0, stream, Lvl{lt}, list, d_n, bound, m0, m1, m2);
    return stx_launch_status("amr3_zero");
}

int32_t stx_amr3_popc(const int64_t* lt, const uint64_t* list, const int32_t* d_n, int64_t bound, const uint64_t* m,
                      uint64_t* out, hipStream_t stream) {
    if (bound <= 0) return STX_OK;
    hipLaunchKernelGGL(popc_kernel, dim3(grid_for(bound, kT)), dim3(kT), 0, stream, Lvl{lt}, list, d_n, bound, m,
                       (unsigned long long*)out);
    return stx_launch_status("amr3_popc");
}

int32_t stx_amr3_accumulate(uint64_t* total, const uint64_t* bits, int64_t base, int64_t steps, hipStream_t stream) {
    hipLaunchKernelGGL(accumulate_kernel, dim3(1), dim3(64), 0, stream, (unsigned long long*)total,
                       (const unsigned long long*)bits, base, steps);
    return stx_launch_status("amr3_accumulate");
}


int32_t stx_amr3_rec_count(int32_t kind, const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n,
                           int64_t bound, const uint64_t* rp, int32_t* counts, hipStream_t stream) {
    if (kind != 0 && kind != 1) return stx_fail(STX_EINVAL, "amr3_rec_count: kind must be 0 or 1");
    if (bound <= 0) return STX_OK;
    hipLaunchKernelGGL(rec_count_kernel, dim3(grid_for(bound, kT)), dim3(kT), 0, stream, (int)kind, Lvl{ltp}, Lvl{ltc},
                       list, d_n, bound, rp, counts);
    return stx_launch_status("amr3_rec_count");
}

int32_t stx_amr3_rec_emit(int32_t kind, const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n,
                          int64_t bound, const uint64_t* rp, const int32_t* offsets, uint64_t* out, hipStream_t stream) {
    if (kind != 0 && kind != 1) return stx_fail(STX_EINVAL, "amr3_rec_emit: kind must be 0 or 1");
    if (bound <= 0) return STX_OK;
    hipLaunchKernelGGL(rec_emit_kernel, dim3(grid_for(bound, kT)), dim3(kT), 0, stream, (int)kind, Lvl{ltp}, Lvl{ltc},
                       list, d_n, bound, rp, offsets, out);
    return stx_launch_status("amr3_rec_emit");
}

int32_t stx_amr3_tag0_records(const int64_t* lt0, const uint64_t* list, int64_t n, uint64_t* out, hipStream_t stream) {
    if (n <= 0) return STX_OK;
    hipLaunchKernelGGL(tag0_rec_build_kernel, dim3(grid_for(n, kT)), dim3(kT), 0, stream, Lvl{lt0}, list, n, out);
    return stx_launch_status("amr3_tag0_records");
}

int32_t stx_amr3_tag0_rec(int32_t dtype, const uint64_t* rec, int64_t n, const void* v, double t, uint64_t* out,
                          hipStream_t stream) {
    if (n <= 0) return STX_OK;
    const unsigned g = xcd_grid(n);
    if (dtype == 4)
        hipLaunchKernelGGL(tag0_rec_kernel<float>, dim3(g), dim3(kT), 0, stream, rec, n, (const float*)v,
                           sqrt_threshold<float, uint32_t>((float)t), out);
    else if (dtype == 8)
        hipLaunchKernelGGL(tag0_rec_kernel<double>, dim3(g), dim3(kT), 0, stream, rec, n, (const double*)v,
                           sqrt_threshold<double, uint64_t>(t), out);
    else
        return stx_fail(STX_EINVAL, "amr3_tag0_rec: dtype must be 4 or 8");
    return stx_launch_status("amr3_tag0_rec");
}

int32_t stx_amr3_stencil_rec(int32_t dtype, const uint64_t* rec1, const int32_t* d_n1, const uint64_t* rec2,
                             const int32_t* d_n2, int64_t bound, const void* v0, const void* v1, const void* v2, void* v1n,
                             void* v2n, const void* coef1, const void* coef2, hipStream_t stream) {
    if (bound <= 0) return STX_OK;
    const unsigned g = slab_grid((bound + 7) / 8, kT / 64);
    if (dtype == 4)
        hipLaunchKernelGGL(stencil_rec_kernel<float>, dim3(g), dim3(kT), 0, stream, rec1, d_n1, rec2, d_n2,
                           (const float*)v0, (const float*)v1, (const float*)v2, (float*)v1n, (float*)v2n,
                           (const float*)coef1, (const float*)coef2);
    else if (dtype == 8)
        hipLaunchKernelGGL(stencil_rec_kernel<double>, dim3(g), dim3(kT), 0, stream, rec1, d_n1, rec2, d_n2,
                           (const double*)v0, (const double*)v1, (const double*)v2, (double*)v1n, (double*)v2n,
                           (const double*)coef1, (const double*)coef2);
    else
        return stx_fail(STX_EINVAL, "amr3_stencil_rec: dtype must be 4 or 8");
    return stx_launch_status("amr3_stencil_rec");
}

int32_t stx_amr3_lupdate_rec(int32_t dtype, const int64_t* ltp, const int64_t* ltc, const uint64_t* rec, const int32_t* d_n,
                             int64_t bound, const uint64_t* rold, const void* vp, void* vc, hipStream_t stream) {
    if (bound <= 0) return STX_OK;
    const unsigned g = slab_grid((bound + 7) / 8, kT / 64);
    if (dtype == 4)
        hipLaunchKernelGGL(lupdate_rec_kernel<float>, dim3(g), dim3(kT), 0, stream, Lvl{ltp}, Lvl{ltc}, rec, d_n, rold,
                           (const float*)vp, (float*)vc);
    else if (dtype == 8)
        hipLaunchKernelGGL(lupdate_rec_kernel<double>, dim3(g), dim3(kT), 0, stream, Lvl{ltp}, Lvl{ltc}, rec, d_n, rold,
                           (const double*)vp, (double*)vc);
    else
        return stx_fail(STX_EINVAL, "amr3_lupdate_rec: dtype must be 4 or 8");
    return stx_launch_status("amr3_lupdate_rec");
}

int32_t stx_amr3_tagl_rec(int32_t dtype, const int64_t* ltc, const uint64_t* rec, const int32_t* d_n, int64_t bound,
                          const void* v, double t, uint64_t* out, hipStream_t stream) {
    if (bound <= 0) return STX_OK;
    const unsigned g = slab_grid((bound + 7) / 8, kT / 64);
    if (dtype == 4)
        hipLaunchKernelGGL(tagl_rec_kernel<float>, dim3(g), dim3(kT), 0, stream, Lvl{ltc}, rec, d_n, (const float*)v,
                           sqrt_threshold<float, uint32_t>((float)t), out);
    else if (dtype == 8)
        hipLaunchKernelGGL(tagl_rec_kernel<double>, dim3(g), dim3(kT), 0, stream, Lvl{ltc}, rec, d_n, (const double*)v,
                           sqrt_threshold<double, uint64_t>(t), out);
    else
        return stx_fail(STX_EINVAL, "amr3_tagl_rec: dtype must be 4 or 8");
    return stx_launch_status("amr3_tagl_rec");
}

int32_t stx_amr3_restrict_rec(int32_t dtype, int32_t acc64, const uint64_t* rec, const int32_t* d_n, int64_t bound,
                              const void* vc, void* vp, hipStream_t stream) {
    if (bound <= 0) return STX_OK;
    const unsigned g = slab_grid((bound + 15) / 16, kT / 64);
    if (dtype == 4 && acc64)
        hipLaunchKernelGGL((restrict_rec_kernel<float, double>), dim3(g), dim3(kT), 0, stream, rec, d_n, (const float*)vc,
                           (float*)vp, 0.25);
    else if (dtype == 4)
        hipLaunchKernelGGL((restrict_rec_kernel<float, float>), dim3(g), dim3(kT), 0, stream, rec, d_n, (const float*)vc,
                           (float*)vp, 0.25f);
    else if (dtype == 8)
        hipLaunchKernelGGL((restrict_rec_kernel<double, double>), dim3(g), dim3(kT), 0, stream, rec, d_n,
                           (const double*)vc, (double*)vp, 0.25);
    else
        return stx_fail(STX_EINVAL, "amr3_restrict_rec: dtype must be 4 or 8");
    return stx_launch_status("amr3_restrict_rec");
}

int32_t stx_amr3_upwind0_rec(int32_t dtype, const uint64_t* rec, int64_t n, const void* v, void* vn, const void* coef,
                             hipStream_t stream) {
    if (n <= 0) return STX_OK;
    const unsigned g = xcd_grid(n);
    if (dtype == 4)
        hipLaunchKernelGGL(upwind0_rec_kernel<float>, dim3(g), dim3(kT), 0, stream, rec, n, (const float*)v, (float*)vn,
                           (const float*)coef);
    else if (dtype == 8)
        hipLaunchKernelGGL(upwind0_rec_kernel<double>, dim3(g), dim3(kT), 0, stream, rec, n, (const double*)v,
                           (double*)vn, (const double*)coef);
    else
        return stx_fail(STX_EINVAL, "amr3_upwind0_rec: dtype must be 4 or 8");
    return stx_launch_status("amr3_upwind0_rec");
}

}  // extern "C"
