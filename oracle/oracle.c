/*
 * oracle.c -- CPU restatement of the subsetix interval-set AMR hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (subsetix_amd) never calls it.
 *
 * Every function restates the reference algorithm it cites (paths relative to
 * the sbstndb/subsetix repository).  The reference is CuPy + NVRTC kernels;
 * the kernels are restated loop by loop, the CuPy glue (lexsort, unique,
 * searchsorted, cumsum) is restated with plain C sorting and scans.
 *
 * Floating point: built with -ffp-contract=off so every product/sum is
 * rounded on its own, which is the order the HIP kernels also use.
 *
 * Conventions: a set is CSR over "rows" (strictly increasing int32 row ids),
 * row offsets off[R+1], half-open [beg, end) int32 intervals.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int64_t nrows;
    int64_t n;
    int32_t* rows;
    int32_t* off;
    int32_t* beg;
    int32_t* end;
} ox_set;

enum { OX_OK = 0, OX_EINVAL = -1, OX_EALIGN = -2, OX_ENOMEM = -3 };

static void* xalloc(size_t n) {
    void* p = malloc(n ? n : 4);
    return p;
}

static void set_alloc(ox_set* s, int64_t nrows, int64_t n) {
    s->nrows = nrows;
    s->n = n;
    s->rows = (int32_t*)xalloc(sizeof(int32_t) * (size_t)nrows);
    s->off = (int32_t*)xalloc(sizeof(int32_t) * (size_t)(nrows + 1));
    s->beg = (int32_t*)xalloc(sizeof(int32_t) * (size_t)n);
    s->end = (int32_t*)xalloc(sizeof(int32_t) * (size_t)n);
    s->off[0] = 0;
}

void ox_free(ox_set* s) {
    if (!s) return;
    free(s->rows);
    free(s->off);
    free(s->beg);
    free(s->end);
    s->rows = s->off = s->beg = s->end = NULL;
    s->nrows = s->n = 0;
}

void ox_free_buf(void* p) { free(p); }

static void set_copy(const ox_set* in, ox_set* out) {
    set_alloc(out, in->nrows, in->n);
    memcpy(out->rows, in->rows, sizeof(int32_t) * (size_t)in->nrows);
    memcpy(out->off, in->off, sizeof(int32_t) * (size_t)(in->nrows + 1));
    memcpy(out->beg, in->beg, sizeof(int32_t) * (size_t)in->n);
    memcpy(out->end, in->end, sizeof(int32_t) * (size_t)in->n);
}

static void set_empty_rows(ox_set* out, const int32_t* rows, int64_t nrows) {
    set_alloc(out, nrows, 0);
    if (nrows) memcpy(out->rows, rows, sizeof(int32_t) * (size_t)nrows);
    memset(out->off, 0, sizeof(int32_t) * (size_t)(nrows + 1));
}

static void set_empty_dense(ox_set* out, int64_t h) {
    set_alloc(out, h, 0);
    for (int64_t r = 0; r < h; ++r) out->rows[r] = (int32_t)r;
    memset(out->off, 0, sizeof(int32_t) * (size_t)(h + 1));
}

/* Row id of every interval: expressions.py:236-249 (searchsorted over
 * row_offsets[1:]). */
static int32_t* interval_rows(const ox_set* s) {
    int32_t* r = (int32_t*)xalloc(sizeof(int32_t) * (size_t)s->n);
    for (int64_t row = 0; row < s->nrows; ++row)
        for (int32_t k = s->off[row]; k < s->off[row + 1]; ++k) r[k] = s->rows[row];
    return r;
}

static int64_t lower_bound_i32(const int32_t* a, int64_t n, int32_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* Sorted-unique union of two sorted-unique int arrays (cupy.union1d).
 * Returns the count; out must hold na+nb. */
static int64_t union1d(const int32_t* a, int64_t na, const int32_t* b, int64_t nb, int32_t* out) {
    int64_t i = 0, j = 0, k = 0;
    while (i < na || j < nb) {
        int32_t v;
        if (j >= nb || (i < na && a[i] < b[j])) v = a[i++];
        else if (i >= na || b[j] < a[i]) v = b[j++];
        else { v = a[i]; ++i; ++j; }
        if (k == 0 || out[k - 1] != v) out[k++] = v;
    }
    return k;
}

/* _reindex_interval_set: expressions.py:58-128.  Intervals keep their order;
 * rows of rows_out absent from the input become empty rows. */
static void reindex(const ox_set* in, const int32_t* rows_out, int64_t nout, ox_set* out) {
    int64_t total = 0;
    int32_t* cnt = (int32_t*)xalloc(sizeof(int32_t) * (size_t)nout);
    for (int64_t u = 0; u < nout; ++u) {
        int64_t p = lower_bound_i32(in->rows, in->nrows, rows_out[u]);
        cnt[u] = (p < in->nrows && in->rows[p] == rows_out[u]) ? in->off[p + 1] - in->off[p] : 0;
        total += cnt[u];
    }
    set_alloc(out, nout, total);
    memcpy(out->rows, rows_out, sizeof(int32_t) * (size_t)nout);
    int64_t w = 0;
    for (int64_t u = 0; u < nout; ++u) {
        int64_t p = lower_bound_i32(in->rows, in->nrows, rows_out[u]);
        if (cnt[u]) {
            for (int32_t k = in->off[p]; k < in->off[p + 1]; ++k) {
                out->beg[w] = in->beg[k];
                out->end[w] = in->end[k];
                ++w;
            }
        }
        out->off[u + 1] = (int32_t)w;
    }
    free(cnt);
}

/* ---------------------------------------------------------------------- */
/* Binary set operations: kernels.py:11-236 (count_segments/write_segments),
 * driven as in expressions.py:446-528.  op 0 = union, 1 = intersection,
 * 2 = difference (expressions.py:529).  One walk function serves both the
 * count (out == NULL) and the write pass, as the two reference kernels share
 * their control flow line for line. */
static int32_t walk_row(int op, const int32_t* lb, const int32_t* le, int32_t ls, int32_t lt,
                        const int32_t* rb, const int32_t* re, int32_t rs, int32_t rt,
                        int32_t* ob, int32_t* oe) {
    int32_t i = ls, j = rs, c = 0;
#define EMIT(B, E) do { if (ob) { ob[c] = (B); oe[c] = (E); } ++c; } while (0)
    if (op == 0) {
        int has = 0, as = 0, ae = 0;
        while (i < lt || j < rt) {
            int take_l = (j >= rt) || (i < lt && lb[i] <= rb[j]);
            int s = take_l ? lb[i] : rb[j];
            int e = take_l ? le[i] : re[j];
            if (!has) { has = 1; as = s; ae = e; }
            else if (s <= ae) { if (e > ae) ae = e; }
            else { EMIT(as, ae); as = s; ae = e; }
            if (take_l) ++i; else ++j;
        }
        if (has) EMIT(as, ae);
    } else if (op == 1) {
        while (i < lt && j < rt) {
            int as = lb[i], ae = le[i], bs = rb[j], be = re[j];
            int s = as > bs ? as : bs;
            int e = ae < be ? ae : be;
            if (s < e) EMIT(s, e);
            if (ae <= be) ++i; else ++j;
        }
    } else {
        while (i < lt) {
            int as = lb[i], ae = le[i], cur = as;
            while (cur < ae) {
                while (j < rt && re[j] <= cur) ++j;
                if (j >= rt || rb[j] >= ae) { if (cur < ae) EMIT(cur, ae); break; }
                if (rb[j] > cur) {
                    int se = rb[j] < ae ? rb[j] : ae;
                    if (se > cur) EMIT(cur, se);
                }
                if (re[j] <= cur) ++j; else cur = re[j];
            }
            ++i;
        }
    }
#undef EMIT
    return c;
}

int ox_setop(int op, const ox_set* lhs, const ox_set* rhs, ox_set* out) {
    if (op < 0 || op > 2) return OX_EINVAL;
    /* _align_interval_sets: expressions.py:131-141 */
    ox_set la, ra;
    int same = lhs->nrows == rhs->nrows &&
               (lhs->nrows == 0 || memcmp(lhs->rows, rhs->rows, sizeof(int32_t) * (size_t)lhs->nrows) == 0);
    if (same) {
        la = *lhs;
        ra = *rhs;
    } else {
        int32_t* u = (int32_t*)xalloc(sizeof(int32_t) * (size_t)(lhs->nrows + rhs->nrows));
        int64_t nu = union1d(lhs->rows, lhs->nrows, rhs->rows, rhs->nrows, u);
        reindex(lhs, u, nu, &la);
        reindex(rhs, u, nu, &ra);
        free(u);
    }
    int64_t R = la.nrows;
    int32_t* cnt = (int32_t*)xalloc(sizeof(int32_t) * (size_t)R);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < R; ++r)
        cnt[r] = walk_row(op, la.beg, la.end, la.off[r], la.off[r + 1], ra.beg, ra.end, ra.off[r], ra.off[r + 1], NULL, NULL);
    int64_t total = 0;
    for (int64_t r = 0; r < R; ++r) total += cnt[r];
    set_alloc(out, R, total);
    if (R) memcpy(out->rows, la.rows, sizeof(int32_t) * (size_t)R);
    for (int64_t r = 0; r < R; ++r) out->off[r + 1] = out->off[r] + cnt[r];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < R; ++r)
        walk_row(op, la.beg, la.end, la.off[r], la.off[r + 1], ra.beg, ra.end, ra.off[r], ra.off[r + 1],
                 out->beg + out->off[r], out->end + out->off[r]);
    free(cnt);
    if (!same) { ox_free(&la); ox_free(&ra); }
    return OX_OK;
}

/* ---------------------------------------------------------------------- */
/* _build_interval_set_from_rows: morphology.py:21-107.  Stable lexsort by
 * (row, begin), rows = unique(rows) U unique(rows_hint), then the per-row
 * merge of kernels.py:292-381 (count/write_merged_segments). */
typedef struct { int32_t row, b, e; int64_t idx; } seg_t;

static int seg_cmp(const void* pa, const void* pb) {
    const seg_t* a = (const seg_t*)pa;
    const seg_t* b = (const seg_t*)pb;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->b != b->b) return a->b < b->b ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

static int i32_cmp(const void* pa, const void* pb) {
    int32_t a = *(const int32_t*)pa, b = *(const int32_t*)pb;
    return a < b ? -1 : (a > b);
}

static int64_t sort_unique(int32_t* a, int64_t n) {
    if (n == 0) return 0;
    qsort(a, (size_t)n, sizeof(int32_t), i32_cmp);
    int64_t k = 1;
    for (int64_t i = 1; i < n; ++i)
        if (a[i] != a[k - 1]) a[k++] = a[i];
    return k;
}

int ox_rebuild(const int32_t* rows, const int32_t* beg, const int32_t* end, int64_t n,
               const int32_t* hint, int64_t nhint, int has_hint, ox_set* out) {
    int32_t* hu = NULL;
    int64_t nhu = 0;
    if (has_hint) {
        hu = (int32_t*)xalloc(sizeof(int32_t) * (size_t)nhint);
        if (nhint) memcpy(hu, hint, sizeof(int32_t) * (size_t)nhint);
        nhu = sort_unique(hu, nhint);
    }
    if (n == 0) {
        set_empty_rows(out, hu, nhu);
        free(hu);
        return OX_OK;
    }
    seg_t* s = (seg_t*)xalloc(sizeof(seg_t) * (size_t)n);
    for (int64_t k = 0; k < n; ++k) { s[k].row = rows[k]; s[k].b = beg[k]; s[k].e = end[k]; s[k].idx = k; }
    qsort(s, (size_t)n, sizeof(seg_t), seg_cmp);
    int32_t* ur = (int32_t*)xalloc(sizeof(int32_t) * (size_t)n);
    int64_t nur = 0;
    for (int64_t k = 0; k < n; ++k)
        if (nur == 0 || ur[nur - 1] != s[k].row) ur[nur++] = s[k].row;
    int32_t* ro = (int32_t*)xalloc(sizeof(int32_t) * (size_t)(nur + nhu));
    int64_t nro = has_hint ? union1d(hu, nhu, ur, nur, ro) : nur;
    if (!has_hint) memcpy(ro, ur, sizeof(int32_t) * (size_t)nur);
    /* raw offsets per output row, then merge */
    int32_t* raw = (int32_t*)xalloc(sizeof(int32_t) * (size_t)(nro + 1));
    raw[0] = 0;
    {
        int64_t k = 0;
        for (int64_t r = 0; r < nro; ++r) {
            while (k < n && s[k].row == ro[r]) ++k;
            raw[r + 1] = (int32_t)k;
        }
    }
    int32_t* cnt = (int32_t*)xalloc(sizeof(int32_t) * (size_t)nro);
    int64_t total = 0;
    for (int64_t r = 0; r < nro; ++r) {
        int32_t st = raw[r], sp = raw[r + 1];
        int32_t c = 0;
        if (st < sp) {
            int32_t ae = s[st].e;
            c = 1;
            for (int32_t k = st + 1; k < sp; ++k) {
                if (s[k].b <= ae) { if (s[k].e > ae) ae = s[k].e; }
                else { ++c; ae = s[k].e; }
            }
        }
        cnt[r] = c;
        total += c;
    }
    set_alloc(out, nro, total);
    memcpy(out->rows, ro, sizeof(int32_t) * (size_t)nro);
    for (int64_t r = 0; r < nro; ++r) out->off[r + 1] = out->off[r] + cnt[r];
    for (int64_t r = 0; r < nro; ++r) {
        int32_t st = raw[r], sp = raw[r + 1], w = out->off[r];
        if (st >= sp) continue;
        int32_t ab = s[st].b, ae = s[st].e;
        for (int32_t k = st + 1; k < sp; ++k) {
            if (s[k].b <= ae) { if (s[k].e > ae) ae = s[k].e; }
            else { out->beg[w] = ab; out->end[w] = ae; ++w; ab = s[k].b; ae = s[k].e; }
        }
        out->beg[w] = ab; out->end[w] = ae;
    }
    free(s); free(ur); free(ro); free(raw); free(cnt); free(hu);
    return OX_OK;
}

/* ---------------------------------------------------------------------- */
/* Morphology: morphology.py:115-140 (translate), :211-265 (dilate),
 * :268-319 (clip), :344-358 (full). */
int ox_translate(const ox_set* in, int32_t dx, int32_t dy, ox_set* out) {
    set_copy(in, out);
    for (int64_t k = 0; k < out->n; ++k) { out->beg[k] += dx; out->end[k] += dx; }
    for (int64_t r = 0; r < out->nrows; ++r) out->rows[r] += dy;
    return OX_OK;
}

int ox_dilate(const ox_set* in, int32_t hx, int32_t hy, ox_set* out) {
    if (hx < 0 || hy < 0) return OX_EINVAL;
    if (hx == 0 && hy == 0) { set_copy(in, out); return OX_OK; }
    if (in->n == 0) { set_empty_rows(out, in->rows, in->nrows); return OX_OK; }
    int32_t* ir = interval_rows(in);
    int32_t* b = (int32_t*)xalloc(sizeof(int32_t) * (size_t)in->n);
    int32_t* e = (int32_t*)xalloc(sizeof(int32_t) * (size_t)in->n);
    for (int64_t k = 0; k < in->n; ++k) {
        b[k] = hx > 0 ? in->beg[k] - hx : in->beg[k];
        e[k] = hx > 0 ? in->end[k] + hx : in->end[k];
    }
    int rc;
    if (hy == 0) {
        rc = ox_rebuild(ir, b, e, in->n, in->rows, in->nrows, 1, out);
    } else {
        int64_t m = 2 * (int64_t)hy + 1, tot = in->n * m;
        int32_t* rr = (int32_t*)xalloc(sizeof(int32_t) * (size_t)tot);
        int32_t* bb = (int32_t*)xalloc(sizeof(int32_t) * (size_t)tot);
        int32_t* ee = (int32_t*)xalloc(sizeof(int32_t) * (size_t)tot);
        for (int64_t k = 0; k < in->n; ++k)
            for (int64_t s = 0; s < m; ++s) {
                rr[k * m + s] = ir[k] + (int32_t)(s - hy);
                bb[k * m + s] = b[k];
                ee[k * m + s] = e[k];
            }
        rc = ox_rebuild(rr, bb, ee, tot, NULL, 0, 0, out);
        free(rr); free(bb); free(ee);
    }
    free(ir); free(b); free(e);
    return rc;
}

int ox_full(int32_t w, int32_t h, ox_set* out) {
    if (h <= 0 || w <= 0) { set_alloc(out, 0, 0); return OX_OK; }
    set_alloc(out, h, h);
    for (int32_t r = 0; r < h; ++r) {
        out->rows[r] = r; out->beg[r] = 0; out->end[r] = w; out->off[r + 1] = r + 1;
    }
    return OX_OK;
}

int ox_clip(const ox_set* in, int32_t w, int32_t h, ox_set* out) {
    if (w <= 0 || h < 0) return OX_EINVAL;
    ox_set dom, cl;
    ox_full(w, h, &dom);
    ox_setop(1, &dom, in, &cl);
    ox_free(&dom);
    if (cl.n == 0) { ox_free(&cl); set_empty_dense(out, h); return OX_OK; }
    int32_t* ir = interval_rows(&cl);
    int64_t m = 0;
    for (int64_t k = 0; k < cl.n; ++k)
        if (ir[k] >= 0 && ir[k] < h) { ir[m] = ir[k]; cl.beg[m] = cl.beg[k]; cl.end[m] = cl.end[k]; ++m; }
    int rc;
    if (m == 0) { set_empty_dense(out, h); rc = OX_OK; }
    else {
        int32_t* hint = (int32_t*)xalloc(sizeof(int32_t) * (size_t)h);
        for (int32_t r = 0; r < h; ++r) hint[r] = r;
        rc = ox_rebuild(ir, cl.beg, cl.end, m, hint, h, 1, out);
        free(hint);
    }
    free(ir);
    ox_free(&cl);
    return rc;
}

/* ---------------------------------------------------------------------- */
/* Multilevel sets: multilevel.py:213-264 (_prolong_set_impl), :272-362
 * (restrict_set), :169-210 (_coarse_to_fine_interval_indices). */
int ox_prolong_set(const ox_set* in, int32_t ratio, ox_set* out, int32_t** base_out) {
    if (ratio < 1) return OX_EINVAL;
    if (ratio == 1) {
        set_copy(in, out);
        if (base_out) {
            *base_out = (int32_t*)xalloc(sizeof(int32_t) * (size_t)in->n);
            for (int64_t k = 0; k < in->n; ++k) (*base_out)[k] = (int32_t)k;
        }
        return OX_OK;
    }
    if (in->nrows == 0 || in->n == 0) {
        int64_t nr = in->nrows * ratio;
        int32_t* fr = (int32_t*)xalloc(sizeof(int32_t) * (size_t)nr);
        for (int64_t r = 0; r < in->nrows; ++r)
            for (int32_t k = 0; k < ratio; ++k) fr[r * ratio + k] = in->rows[r] * ratio + k;
        qsort(fr, (size_t)nr, sizeof(int32_t), i32_cmp);
        set_empty_rows(out, fr, nr);
        free(fr);
        if (base_out) *base_out = (int32_t*)xalloc(4);
        return OX_OK;
    }
    int64_t tot = in->n * ratio;
    int32_t* ir = interval_rows(in);
    seg_t* s = (seg_t*)xalloc(sizeof(seg_t) * (size_t)tot);
    for (int64_t k = 0; k < in->n; ++k)
        for (int32_t d = 0; d < ratio; ++d) {
            seg_t* t = &s[k * ratio + d];
            t->row = ir[k] * ratio + d;
            t->b = in->beg[k] * ratio;
            t->e = in->end[k] * ratio;
            t->idx = k * ratio + d;
        }
    qsort(s, (size_t)tot, sizeof(seg_t), seg_cmp);
    int64_t nr = 0;
    for (int64_t k = 0; k < tot; ++k)
        if (k == 0 || s[k].row != s[k - 1].row) ++nr;
    set_alloc(out, nr, tot);
    int64_t r = -1;
    for (int64_t k = 0; k < tot; ++k) {
        if (k == 0 || s[k].row != s[k - 1].row) { ++r; out->rows[r] = s[k].row; out->off[r] = (int32_t)k; }
        out->beg[k] = s[k].b;
        out->end[k] = s[k].e;
    }
    out->off[nr] = (int32_t)tot;
    if (base_out) {
        *base_out = (int32_t*)xalloc(sizeof(int32_t) * (size_t)tot);
        for (int64_t k = 0; k < tot; ++k) (*base_out)[k] = (int32_t)(s[k].idx / ratio);
    }
    free(s); free(ir);
    return OX_OK;
}

static int32_t floordiv(int32_t a, int32_t b) {
    int32_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

int ox_restrict_set(const ox_set* in, int32_t ratio, ox_set* out) {
    if (ratio < 1) return OX_EINVAL;
    if (in->nrows == 0) { set_alloc(out, 0, 0); return OX_OK; }
    if (in->n == 0) {
        int32_t* cr = (int32_t*)xalloc(sizeof(int32_t) * (size_t)in->nrows);
        for (int64_t r = 0; r < in->nrows; ++r) cr[r] = floordiv(in->rows[r], ratio);
        int64_t nc = sort_unique(cr, in->nrows);
        set_empty_rows(out, cr, nc);
        free(cr);
        return OX_OK;
    }
    int32_t* ir = interval_rows(in);
    int32_t* cr = (int32_t*)xalloc(sizeof(int32_t) * (size_t)in->n);
    int32_t* cb = (int32_t*)xalloc(sizeof(int32_t) * (size_t)in->n);
    int32_t* ce = (int32_t*)xalloc(sizeof(int32_t) * (size_t)in->n);
    int64_t m = 0;
    for (int64_t k = 0; k < in->n; ++k) {
        cr[k] = floordiv(ir[k], ratio);
        cb[k] = floordiv(in->beg[k], ratio);
        ce[k] = floordiv(in->end[k] + ratio - 1, ratio);
        if (ce[k] > cb[k]) ++m;
    }
    int rc;
    if (m == 0) {
        int64_t nc = sort_unique(cr, in->n);
        set_empty_rows(out, cr, nc);
        rc = OX_OK;
    } else {
        int64_t w = 0;
        for (int64_t k = 0; k < in->n; ++k)
            if (ce[k] > cb[k]) { cr[w] = cr[k]; cb[w] = cb[k]; ce[w] = ce[k]; ++w; }
        rc = ox_rebuild(cr, cb, ce, m, NULL, 0, 0, out);
    }
    free(ir); free(cr); free(cb); free(ce);
    return rc;
}

/* Returns OX_OK and *idx (size nc*ratio) or OX_EALIGN when the fine rows are
 * not exactly {r*ratio + k} for every coarse row r. */
int ox_c2f_indices(const ox_set* c, const ox_set* f, int32_t ratio, int32_t** idx_out, int64_t* n_out) {
    *idx_out = NULL;
    *n_out = 0;
    if (ratio < 1) return OX_EINVAL;
    if (c->nrows == 0) { *idx_out = (int32_t*)xalloc(4); return OX_OK; }
    if (f->nrows != c->nrows * ratio) return OX_EALIGN;
    for (int64_t r = 0; r < c->nrows; ++r)
        for (int32_t k = 0; k < ratio; ++k)
            if (f->rows[r * ratio + k] != c->rows[r] * ratio + k) return OX_EALIGN;
    int32_t* idx = (int32_t*)xalloc(sizeof(int32_t) * (size_t)(c->n * ratio));
    for (int64_t r = 0; r < c->nrows; ++r)
        for (int32_t i = c->off[r]; i < c->off[r + 1]; ++i) {
            int32_t ord = i - c->off[r];
            for (int32_t k = 0; k < ratio; ++k) idx[(int64_t)i * ratio + k] = f->off[r * ratio + k] + ord;
        }
    *idx_out = idx;
    *n_out = c->n * ratio;
    return OX_OK;
}

static int32_t* cell_offsets(const ox_set* s) {
    int32_t* o = (int32_t*)xalloc(sizeof(int32_t) * (size_t)(s->n + 1));
    o[0] = 0;
    for (int64_t k = 0; k < s->n; ++k) o[k + 1] = o[k] + (s->end[k] - s->beg[k]);
    return o;
}

/* prolong_field kernels.py:384-455 over the map above.  dtype 4 = float,
 * 8 = double (values are copied as raw elements). */
int ox_prolong_field(const ox_set* c, const void* cvals, const ox_set* f, void* fvals, int32_t ratio, int esize) {
    int32_t* idx; int64_t nidx;
    int rc = ox_c2f_indices(c, f, ratio, &idx, &nidx);
    if (rc) return rc;
    int32_t* co = cell_offsets(c);
    int32_t* fo = cell_offsets(f);
    for (int64_t i = 0; i < c->n; ++i) {
        if (co[i + 1] - co[i] <= 0) continue;
        for (int32_t d = 0; d < ratio; ++d) {
            int32_t fi = idx[i * ratio + d];
            int32_t fw = fo[fi + 1] - fo[fi];
            for (int32_t k = 0; k < fw; ++k) {
                int32_t src = co[i] + k / ratio;
                memcpy((char*)fvals + (size_t)(fo[fi] + k) * esize, (const char*)cvals + (size_t)src * esize, (size_t)esize);
            }
        }
    }
    free(idx); free(co); free(fo);
    return OX_OK;
}

/* restrict_field kernels.py:458-598 (reducer 0 mean, 1 sum, 2 min, 3 max).
 * in_kind: 0 = float in / float accumulate (the reference FP32 kernel),
 *          1 = float in / double accumulate / float out,
 *          2 = double in / double accumulate / double out. */
int ox_restrict_field(const ox_set* c, void* cvals, const ox_set* f, const void* fvals, int32_t ratio,
                      int reducer, int in_kind) {
    int32_t* idx; int64_t nidx;
    int rc = ox_c2f_indices(c, f, ratio, &idx, &nidx);
    if (rc) return rc;
    /* multilevel.py:443-451 alignment check */
    for (int64_t i = 0; i < c->n; ++i)
        for (int32_t d = 0; d < ratio; ++d) {
            int32_t fi = idx[i * ratio + d];
            if (fi < 0 || fi >= f->n || f->beg[fi] != c->beg[i] * ratio || f->end[fi] != c->end[i] * ratio) {
                free(idx);
                return OX_EALIGN;
            }
        }
    int32_t* co = cell_offsets(c);
    int32_t* fo = cell_offsets(f);
    int32_t samples = ratio * ratio;
    float normf = (float)(reducer == 0 ? 1.0 / samples : 1.0);
    double normd = reducer == 0 ? 1.0 / samples : 1.0;
    for (int64_t i = 0; i < c->n; ++i) {
        int32_t width = co[i + 1] - co[i];
        for (int32_t k = 0; k < width; ++k) {
            float accf = 0.0f, mnf = 0.0f, mxf = 0.0f;
            double accd = 0.0, mnd = 0.0, mxd = 0.0;
            int have = 0;
            for (int32_t dy = 0; dy < ratio; ++dy) {
                int32_t fi = idx[i * ratio + dy];
                int32_t st = fo[fi] + k * ratio, fe = fo[fi + 1];
                for (int32_t dx = 0; dx < ratio; ++dx) {
                    int32_t id = st + dx;
                    if (id >= fe) break;
                    if (in_kind == 0) {
                        float v = ((const float*)fvals)[id];
                        if (reducer == 2) { if (!have || v < mnf) mnf = v; }
                        else if (reducer == 3) { if (!have || v > mxf) mxf = v; }
                        else accf += v;
                    } else {
                        double v = in_kind == 1 ? (double)((const float*)fvals)[id] : ((const double*)fvals)[id];
                        if (reducer == 2) { if (!have || v < mnd) mnd = v; }
                        else if (reducer == 3) { if (!have || v > mxd) mxd = v; }
                        else accd += v;
                    }
                    have = 1;
                }
            }
            int64_t dst = co[i] + k;
            if (in_kind == 0) {
                float r = reducer == 0 ? accf * normf : reducer == 1 ? accf : reducer == 2 ? mnf : mxf;
                ((float*)cvals)[dst] = r;
            } else {
                double r = reducer == 0 ? accd * normd : reducer == 1 ? accd : reducer == 2 ? mnd : mxd;
                if (in_kind == 1) ((float*)cvals)[dst] = (float)r;
                else ((double*)cvals)[dst] = r;
            }
        }
    }
    free(idx); free(co); free(fo);
    return OX_OK;
}

/* ---------------------------------------------------------------------- */
/* Upwind stencil: interval_stencil.py:16-113 (sample_interval_cell +
 * interval_upwind), host checks :143-191.  Zero exterior, per-cell binary
 * search of the neighbour rows. */
#define UPWIND_BODY(T)                                                                              \
    int32_t* ir = interval_rows(s);                                                                 \
    int32_t* co = cell_offsets(s);                                                                  \
    int32_t* lookup = (int32_t*)xalloc(sizeof(int32_t) * (size_t)height);                           \
    for (int32_t r = 0; r < height; ++r) lookup[r] = -1;                                            \
    for (int64_t r = 0; r < s->nrows; ++r) {                                                        \
        if (s->rows[r] < 0 || s->rows[r] >= height) { free(ir); free(co); free(lookup); return OX_EINVAL; } \
        lookup[s->rows[r]] = (int32_t)r;                                                            \
    }                                                                                               \
    _Pragma("omp parallel for schedule(dynamic, 16)")                                               \
    for (int64_t i = 0; i < s->n; ++i) {                                                            \
        int32_t row = ir[i], b0 = s->beg[i], e0 = s->end[i];                                        \
        if (e0 <= b0) continue;                                                                     \
        for (int32_t k = 0; k < e0 - b0; ++k) {                                                     \
            int32_t x = b0 + k;                                                                     \
            T c = u[co[i] + k];                                                                     \
            T nb[4];                                                                                \
            int32_t qr[4] = {row, row, row - 1, row + 1}, qx[4] = {x - 1, x + 1, x, x};             \
            for (int q = 0; q < 4; ++q) {                                                           \
                T v = (T)0;                                                                         \
                if (qr[q] >= 0 && qr[q] < height && qx[q] >= 0 && qx[q] < width) {                 \
                    int32_t ord = lookup[qr[q]];                                                    \
                    if (ord >= 0) {                                                                 \
                        int32_t lo = s->off[ord], hi = s->off[ord + 1] - 1;                         \
                        while (lo <= hi) {                                                          \
                            int32_t mid = (lo + hi) >> 1;                                           \
                            if (qx[q] < s->beg[mid]) hi = mid - 1;                                  \
                            else if (qx[q] >= s->end[mid]) lo = mid + 1;                            \
                            else { v = u[co[mid] + (qx[q] - s->beg[mid])]; break; }                 \
                        }                                                                           \
                    }                                                                               \
                }                                                                                   \
                nb[q] = v;                                                                          \
            }                                                                                       \
            T dudx = (a >= (T)0) ? (c - nb[0]) / dx : (nb[1] - c) / dx;                             \
            T dudy = (b >= (T)0) ? (c - nb[2]) / dy : (nb[3] - c) / dy;                             \
            T t1 = a * dudx;                                                                        \
            T t2 = b * dudy;                                                                        \
            T sm = t1 + t2;                                                                         \
            T m = dt * sm;                                                                          \
            out[co[i] + k] = c - m;                                                                 \
        }                                                                                           \
    }                                                                                               \
    free(ir); free(co); free(lookup);                                                               \
    return OX_OK;

int ox_upwind_f32(const ox_set* s, const float* u, float* out, int32_t width, int32_t height,
                  float a, float b, float dt, float dx, float dy) {
    if (width <= 0 || height <= 0) return OX_EINVAL;
    UPWIND_BODY(float)
}

int ox_upwind_f64(const ox_set* s, const double* u, double* out, int32_t width, int32_t height,
                  double a, double b, double dt, double dx, double dy) {
    if (width <= 0 || height <= 0) return OX_EINVAL;
    UPWIND_BODY(double)
}

/* FP64 mass: sum of u * (dx*dy) over all cells.  Not in the reference; the
 * order is a plain sequential sum, tests compare with a relative tolerance. */
double ox_mass_f32(const float* u, int64_t n, double cell_area) {
    double acc = 0.0;
    for (int64_t k = 0; k < n; ++k) acc += (double)u[k];
    return acc * cell_area;
}

double ox_mass_f64(const double* u, int64_t n, double cell_area) {
    double acc = 0.0;
    for (int64_t k = 0; k < n; ++k) acc += u[k];
    return acc * cell_area;
}

/* ---------------------------------------------------------------------- */
/* Regrid: gradient magnitude regrid.py:181-218 (dense W x H, central
 * differences x0.5 inside, one-sided at the edges), threshold runs
 * regrid.py:20-124 (maximal runs of v >= t per row; the intended result,
 * without the width % 32 double write of the reference write kernel). */
void ox_gradient_f32(const float* d, int32_t w, int32_t h, float* out) {
#pragma omp parallel for
    for (int64_t idx = 0; idx < (int64_t)w * h; ++idx) {
        int32_t x = (int32_t)(idx % w), y = (int32_t)(idx / w);
        float c = d[idx];
        float gx = 0.0f, gy = 0.0f;
        if (w > 1) {
            if (x == 0) gx = d[idx + 1] - c;
            else if (x == w - 1) gx = c - d[idx - 1];
            else gx = (d[idx + 1] - d[idx - 1]) * 0.5f;
        }
        if (h > 1) {
            if (y == 0) gy = d[idx + w] - c;
            else if (y == h - 1) gy = c - d[idx - w];
            else gy = (d[idx + w] - d[idx - w]) * 0.5f;
        }
        float gx2 = gx * gx;
        float gy2 = gy * gy;
        out[idx] = sqrtf(gx2 + gy2);
    }
}

int ox_threshold(const float* d, int32_t w, int32_t h, float t, ox_set* out) {
    if (w <= 0 || h <= 0) return OX_EINVAL;
    int32_t* cnt = (int32_t*)xalloc(sizeof(int32_t) * (size_t)h);
    int64_t total = 0;
    for (int32_t y = 0; y < h; ++y) {
        int32_t c = 0, prev = 0;
        for (int32_t x = 0; x < w; ++x) {
            int on = d[(int64_t)y * w + x] >= t;
            if (on && !prev) ++c;
            prev = on;
        }
        cnt[y] = c;
        total += c;
    }
    set_alloc(out, h, total);
    for (int32_t y = 0; y < h; ++y) { out->rows[y] = y; out->off[y + 1] = out->off[y] + cnt[y]; }
    for (int32_t y = 0; y < h; ++y) {
        int32_t wp = out->off[y], prev = 0;
        for (int32_t x = 0; x < w; ++x) {
            int on = d[(int64_t)y * w + x] >= t;
            if (on && !prev) out->beg[wp] = x;
            if (!on && prev) out->end[wp++] = x;
            prev = on;
        }
        if (prev) out->end[wp++] = w;
    }
    free(cnt);
    return OX_OK;
}

/* ---------------------------------------------------------------------- */
/* Interval subset gather/scatter: fields.py:127-294 (per subset interval,
 * locate the covering super interval, then copy run by run), with the row
 * containment check of fields.py:310-318.  mode 0 = gather (super ->
 * subset), 1 = scatter (subset -> super). */
int ox_subset_copy(int mode, const ox_set* sup, void* sup_vals, const ox_set* sub, void* sub_vals, int esize) {
    int32_t* so = cell_offsets(sup);
    int32_t* bo = cell_offsets(sub);
    int32_t* ir = interval_rows(sub);
    for (int64_t i = 0; i < sub->n; ++i) {
        int64_t rp = lower_bound_i32(sup->rows, sup->nrows, ir[i]);
        if (rp >= sup->nrows || sup->rows[rp] != ir[i]) { free(so); free(bo); free(ir); return OX_EINVAL; }
    }
    for (int64_t i = 0; i < sub->n; ++i) {
        int32_t row = (int32_t)lower_bound_i32(sup->rows, sup->nrows, ir[i]);
        int32_t start = sub->beg[i], stop = sub->end[i];
        if (stop <= start) continue;
        int32_t base = bo[i], off = 0;
        int32_t rs = sup->off[row], rt = sup->off[row + 1];
        if (rs >= rt) continue;
        int32_t lo = rs, hi = rt - 1, idx = rs;
        while (lo <= hi) {
            int32_t mid = (lo + hi) >> 1;
            if (start < sup->beg[mid]) hi = mid - 1;
            else if (start >= sup->end[mid]) lo = mid + 1;
            else { idx = mid; break; }
        }
        if (lo > hi) idx = lo;
        if (idx < rs) idx = rs;
        int32_t cur = start;
        while (cur < stop && idx < rt) {
            int32_t b = sup->beg[idx], e = sup->end[idx];
            if (cur < b) { cur = b; continue; }
            int32_t ce = stop < e ? stop : e;
            int32_t len = ce - cur;
            if (len > 0) {
                int32_t sb = so[idx] + (cur - b);
                for (int32_t t = 0; t < len; ++t) {
                    char* sp = (char*)sup_vals + (size_t)(sb + t) * esize;
                    char* bp = (char*)sub_vals + (size_t)(base + off + t) * esize;
                    if (mode == 0) memcpy(bp, sp, (size_t)esize); else memcpy(sp, bp, (size_t)esize);
                }
                off += len;
                cur = ce;
            }
            if (cur >= e) ++idx;
            if (len == 0 && cur < stop) ++cur;
        }
    }
    free(so); free(bo); free(ir);
    return OX_OK;
}


/* ---------------------------------------------------------------------- */
/* Level transfers onto a target subset, by cell lookup.  The reference's
 * synchronisation (fields.py:570-612) restricts the whole fine field with
 * restrict_field (multilevel.py:410-479), gathers the refine cells and
 * scatters them into the coarse field, then gathers the coarse-only cells,
 * prolongs them (multilevel.py:365-407) and scatters them into the fine
 * field.  Cell for cell that is:
 *   prolong onto T:  dst(y, x) = src(floor(y/R), floor(x/R))
 *   restrict onto T: dst(y, x) = (sum over dy < R, then dx < R, of
 *                    src(yR+dy, xR+dx)) * (1/R^2)
 * for every cell (y, x) of T.  This restatement looks every cell up by row
 * (binary search over the row ids) and interval (binary search in the row),
 * with no segment tables and no cursors, so it shares no indexing with the
 * HIP kernels.  Cells of T missing from dst are skipped; a missing prolong
 * parent leaves dst as it is; missing restrict children add nothing.
 * kind: 0 = float values, float sums, norm = (float)(1/R^2) (the reference
 * FP32 kernel, multilevel.py:462-479); 1 = float values, double sums, then
 * (float)(sum * (1/R^2)); 2 = double throughout.  mode 0 = prolong,
 * 1 = restrict; ratio 1 prolong is a plain copy of the common cells. */
typedef struct {
    const ox_set* s;
    int32_t* co;
} ox_cells;

static void cells_init(ox_cells* c, const ox_set* s) {
    c->s = s;
    c->co = cell_offsets(s);
}

static int64_t cell_find(const ox_cells* c, int32_t y, int32_t x) {
    const ox_set* s = c->s;
    int64_t r = lower_bound_i32(s->rows, s->nrows, y);
    if (r >= s->nrows || s->rows[r] != y) return -1;
    int32_t lo = s->off[r], hi = s->off[r + 1] - 1;
    while (lo <= hi) {
        int32_t mid = lo + ((hi - lo) >> 1);
        if (x < s->beg[mid]) hi = mid - 1;
        else if (x >= s->end[mid]) lo = mid + 1;
        else return (int64_t)c->co[mid] + (x - s->beg[mid]);
    }
    return -1;
}

int ox_transfer(int mode, int kind, int32_t ratio, const ox_set* t, const ox_set* d, void* dvals, const ox_set* s,
                const void* svals) {
    if (ratio < 1 || ratio > 64 || (mode != 0 && mode != 1) || kind < 0 || kind > 2) return OX_EINVAL;
    ox_cells dc, sc;
    cells_init(&dc, d);
    cells_init(&sc, s);
    int32_t* tir = interval_rows(t);
    const float normf = (float)(1.0 / (double)(ratio * ratio));
    const double normd = 1.0 / (double)(ratio * ratio);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < t->n; ++i) {
        const int32_t y = tir[i];
        for (int32_t x = t->beg[i]; x < t->end[i]; ++x) {
            const int64_t di = cell_find(&dc, y, x);
            if (di < 0) continue;
            if (mode == 0) {
                const int64_t si = cell_find(&sc, floordiv(y, ratio), floordiv(x, ratio));
                if (si < 0) continue;
                if (kind == 2) ((double*)dvals)[di] = ((const double*)svals)[si];
                else ((float*)dvals)[di] = ((const float*)svals)[si];
                continue;
            }
            float accf = 0.0f;
            double accd = 0.0;
            for (int32_t dy = 0; dy < ratio; ++dy)
                for (int32_t dx = 0; dx < ratio; ++dx) {
                    const int64_t si = cell_find(&sc, y * ratio + dy, x * ratio + dx);
                    if (si < 0) continue;
                    if (kind == 0) accf = accf + ((const float*)svals)[si];
                    else if (kind == 1) accd = accd + (double)((const float*)svals)[si];
                    else accd = accd + ((const double*)svals)[si];
                }
            if (kind == 0) ((float*)dvals)[di] = accf * normf;
            else if (kind == 1) ((float*)dvals)[di] = (float)(accd * normd);
            else ((double*)dvals)[di] = accd * normd;
        }
    }
    free(tir); free(dc.co); free(sc.co);
    return OX_OK;
}

/* Gradient tag on a sparse level: the central-difference gradient of
 * regrid.py:181-218, restated for an interval field whose neighbours may be
 * absent (a neighbour is a cell of the same set; one-sided differences where
 * only one exists, 0 where neither does), then maximal runs of
 * sqrt(gx^2 + gy^2) >= t per row (regrid.py:20-124), the rows being the
 * set's rows.  esize 4: float arithmetic, t rounded to float; 8: double.
 * Runs are maximal over adjacent cells, also across touching intervals. */
#define GRAD_TAG_BODY(T, SQRT)                                                                       \
    const T* u = (const T*)vals;                                                                     \
    const T tt = (T)t;                                                                               \
    unsigned char* tag = (unsigned char*)xalloc((size_t)(nc > 0 ? nc : 1));                          \
    _Pragma("omp parallel for schedule(dynamic, 64)")                                                \
    for (int64_t i = 0; i < s->n; ++i) {                                                             \
        const int32_t y = ir[i];                                                                     \
        for (int32_t x = s->beg[i]; x < s->end[i]; ++x) {                                            \
            const int64_t c = (int64_t)cs.co[i] + (x - s->beg[i]);                                   \
            const int64_t l = cell_find(&cs, y, x - 1), r = cell_find(&cs, y, x + 1);                \
            const int64_t dn = cell_find(&cs, y - 1, x), up = cell_find(&cs, y + 1, x);              \
            T gx = (T)0, gy = (T)0;                                                                  \
            if (l >= 0 && r >= 0) gx = (u[r] - u[l]) * (T)0.5;                                       \
            else if (r >= 0) gx = u[r] - u[c];                                                       \
            else if (l >= 0) gx = u[c] - u[l];                                                       \
            if (dn >= 0 && up >= 0) gy = (u[up] - u[dn]) * (T)0.5;                                   \
            else if (up >= 0) gy = u[up] - u[c];                                                     \
            else if (dn >= 0) gy = u[c] - u[dn];                                                     \
            const T gx2 = gx * gx;                                                                   \
            const T gy2 = gy * gy;                                                                   \
            tag[c] = SQRT(gx2 + gy2) >= tt;                                                          \
        }                                                                                            \
    }

int ox_grad_tag(const ox_set* s, const void* vals, int esize, double t, ox_set* out) {
    if (esize != 4 && esize != 8) return OX_EINVAL;
    ox_cells cs;
    cells_init(&cs, s);
    const int64_t nc = cs.co[s->n];
    int32_t* ir = interval_rows(s);
    unsigned char* tagp;
    {
        if (esize == 4) {
            GRAD_TAG_BODY(float, sqrtf)
            tagp = tag;
        } else {
            GRAD_TAG_BODY(double, sqrt)
            tagp = tag;
        }
    }
    /* runs: a tagged cell starts a run unless the previous cell of the row
     * (x - 1, in the set) is tagged */
    int32_t* cnt = (int32_t*)xalloc(sizeof(int32_t) * (size_t)(s->nrows > 0 ? s->nrows : 1));
    int64_t total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            set_alloc(out, s->nrows, total);
            for (int64_t r = 0; r < s->nrows; ++r) { out->rows[r] = s->rows[r]; out->off[r + 1] = out->off[r] + cnt[r]; }
        }
        for (int64_t r = 0; r < s->nrows; ++r) {
            int32_t k = 0, prev_x = 0, wp = pass ? out->off[r] : 0;
            int prev_on = 0;
            for (int32_t i = s->off[r]; i < s->off[r + 1]; ++i)
                for (int32_t x = s->beg[i]; x < s->end[i]; ++x) {
                    const int on = tagp[cs.co[i] + (x - s->beg[i])];
                    const int joined = prev_on && prev_x == x - 1;
                    if (prev_on && !(on && joined)) { if (pass) out->end[wp++] = prev_x + 1; }
                    if (on && !joined) { if (pass) out->beg[wp] = x; ++k; }
                    prev_on = on;
                    prev_x = x;
                }
            if (prev_on && pass) out->end[wp++] = prev_x + 1;
            if (!pass) { cnt[r] = k; total += k; }
        }
    }
    free(cnt); free(tagp); free(ir); free(cs.co);
    return OX_OK;
}

int ox_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ox_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
