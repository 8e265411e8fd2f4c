/*
 * subsetix_hip.h -- C-ABI of libsubsetix_hip.so (gfx950 / MI355X).
 *
 * Drop-in boundary for the subsetix interval-set AMR hot path.  The reference
 * (sbstndb/subsetix) drives CUDA-C kernels embedded in Python strings through
 * cupy.RawKernel(src, name)(grid, block, args) with raw device pointers and
 * numpy scalars; each entry point below replaces one of those kernels (or the
 * CuPy glue around it) and cites the reference code it stands for.
 *
 * Conventions
 *   - Interval sets are CSR: row_offsets[R+1] (int32, starts at 0), begin/end
 *     [N] int32 half-open, rows[R] strictly increasing int32 row ids.
 *   - All pointers are device pointers (HIP); sizes are int64_t.
 *   - Every call is asynchronous on `stream` and returns 0 or a negative
 *     status (STX_E*); stx_last_error() describes the last failure.
 *   - Device-detected conditions (capacity overflow, misalignment, rows out of
 *     range, look-back wait bound) are OR-ed into the caller's `status` word.
 *   - The library never allocates device memory: outputs, temporaries and the
 *     look-back `scratch` are caller buffers.  The scratch is zeroed once by
 *     the caller; every launch that uses it passes a new `epoch` (1, 2, ...,
 *     < 2^30; re-zero the scratch before wrapping), so nothing is cleared
 *     between launches.  Calls sharing a scratch buffer must be ordered (same
 *     stream).
 */
#ifndef SUBSETIX_HIP_H
#define SUBSETIX_HIP_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    STX_OK = 0,
    STX_EINVAL = -1,
    STX_ECAPACITY = -2,
    STX_EALIGN = -3,
    STX_EHIP = -4,
    STX_ESPIN = -5
};

/* status word bits */
enum { STX_FLAG_CAPACITY = 1, STX_FLAG_SPIN = 2, STX_FLAG_ALIGN = 4, STX_FLAG_ROWS = 8, STX_FLAG_DOMAIN = 16 };

/* operand transforms applied on load (set algebra / merges) */
enum { STX_XF_NONE = 0, STX_XF_WIDEN = 1, STX_XF_COARSEN = 2, STX_XF_CLIP = 3, STX_XF_SCALE = 4 };

/* ---- housekeeping ------------------------------------------------------- */
int32_t stx_abi_version(void);
const char* stx_last_error(void);
void stx_set_spin_limit(int64_t polls);

/* ---- scans ---------------------------------------------------------------
 * Replaces cp.cumsum over counts (expressions.py:502, morphology.py:60,84,
 * multilevel.py:261,316,340, interval_field.py:73, regrid.py:314,
 * fields.py:94).  out[k] = sum(in[0..k)), out[n] = total (n = *d_n when
 * d_n != NULL, else n_bound). */
int32_t stx_scan_i32(const int32_t* in, const int32_t* d_n, int64_t n_bound, int32_t* out, int32_t* d_total,
                     uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
int64_t stx_scan_scratch_bytes(int64_t n_bound);

/* ---- set algebra ---------------------------------------------------------
 * Replaces count_segments / write_segments (kernels.py:11-236) driven by
 * _apply_binary (expressions.py:446-528), and the merge rebuild
 * count/write_merged_segments (kernels.py:292-381) when b_off == NULL
 * (op must then be 0).  op: 0 union, 1 intersection, 2 difference
 * (expressions.py:529).  One launch: count, scan and write (single pass,
 * decoupled look-back).  Operand X is read through an optional row map
 * x_map[output row] -> source row (-1 = empty; replaces the re-indexing of
 * _align_interval_sets, expressions.py:58-141) and an optional transform
 * (STX_XF_*).  Output: out_off[U+1], out_beg/out_end (capacity entries),
 * *d_total.  U = *d_nrows when non-NULL, else nrows_bound.
 * group <= 0: LDS tile path when no maps are given: 0 and -7: flat tiles in
 * the register-output form for two operands read as stored (else as -6);
 * -8: the same with count-balanced tiles (a tile table built on the device
 * from the offsets; for inputs whose density varies between row ranges,
 * where tiles of equal row counts would overflow); -6: flat tiles (a
 * tile's rows merged as one keyed sequence, single pass), 1024 threads x
 * 8 positions; -5: flat tiles, 512 x 8; -1: row-aware tiles, 512 x 8, single
 * pass; -2: the same in two passes; -3: row-aware 128 x 16; -4: row-aware
 * 64 x 32.  With maps, the per-row path.  4 / 16 / 64: lane-group size of
 * the per-row path. */
int32_t stx_setop(int32_t op, const int32_t* a_beg, const int32_t* a_end, const int32_t* a_off, const int32_t* a_map,
                  int32_t a_xf_mode, int32_t a_xf_param, int64_t a_n, const int32_t* b_beg, const int32_t* b_end,
                  const int32_t* b_off, const int32_t* b_map, int32_t b_xf_mode, int32_t b_xf_param, int64_t b_n,
                  const int32_t* d_nrows, int64_t nrows_bound, int32_t group, int32_t* out_off, int32_t* out_beg,
                  int32_t* out_end, int64_t capacity, int32_t* d_total, uint32_t* status, void* scratch,
                  int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
int64_t stx_setop_scratch_bytes(int64_t nrows_bound, int32_t group);
/* Scratch for a call with `elems` input intervals (a_n + b_n): schedule -8
 * keeps its tile table in the scratch as well. */
int64_t stx_setop_scratch_bytes_n(int64_t nrows_bound, int64_t elems, int32_t group);

/* Allocation-free set operation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * stx_setop with no maps, no transforms, d_nrows = NULL and group 0 (the
 * flat-tile kernel), but the look-back state lives in a static 1 MiB device
 * array of the library (one per GPU, held until the process ends) instead of
 * a caller scratch; the caller passes only inputs and outputs (status is
 * OR-ed as for stx_setop).  Operands needing more than the array's 131,064
 * look-back tiles are rejected with STX_EINVAL.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams or threads)
 * race.  Allocates nothing (no hipMalloc, hipFree or allocating library call).
 */
int32_t stx_setop_noalloc(int32_t op, const int32_t* a_beg, const int32_t* a_end, const int32_t* a_off, int64_t a_n,
                          const int32_t* b_beg, const int32_t* b_end, const int32_t* b_off, int64_t b_n, int64_t nrows,
                          int32_t* out_off, int32_t* out_beg, int32_t* out_end, int64_t capacity, int32_t* d_total,
                          uint32_t* status, hipStream_t stream);
/* Size in bytes of that static array (the operand limit above). */
int64_t stx_setop_noalloc_max_scratch_bytes(void);

/* Sorted union of two row-id lists (shifted) with source maps: replaces the
 * cp.union1d + searchsorted of _align_interval_sets / _reindex_interval_set
 * (expressions.py:58-141, :138).  tmp: 2*(nb+1) int32; outputs sized na+nb;
 * *d_nu = union size. */
int32_t stx_rows_union(const int32_t* rows_a, int64_t na, int32_t shift_a, const int32_t* rows_b, int64_t nb,
                       int32_t shift_b, int32_t* rows_u, int32_t* map_a, int32_t* map_b, int32_t* d_nu, int32_t* tmp,
                       uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);

/* ---- row structure -------------------------------------------------------
 * CSR checks of IntervalSet.__post_init__ (expressions.py:166-227): flag
 * bits 1 off[0]!=0, 2 off[R]!=N, 4 rows not increasing, 8 offsets
 * decreasing, 16 rows != arange(R). */
int32_t stx_validate_csr(const int32_t* off, int64_t R, const int32_t* rows, int64_t N, uint32_t* flags,
                         hipStream_t stream);
/* IntervalSet.interval_rows (expressions.py:236-249): row id and/or row
 * ordinal of every interval. */
int32_t stx_interval_rows(const int32_t* off, int64_t R, const int32_t* rows, int64_t N, int32_t* out_row,
                          int32_t* out_ord, hipStream_t stream);
/* Dense row map for clip_interval_set (morphology.py:268-319): map[y] =
 * index of row y0+y in rows, or -1. */
int32_t stx_rows_lookup(const int32_t* rows, int64_t R, int32_t y0, int64_t H, int32_t* map, hipStream_t stream);
/* Dense offsets of a set over the rows y0 .. y0 + n - 1 (n + 1 entries):
 * out[j] = intervals in rows below y0 + j, counted from off[0]; rows the set
 * does not hold are empty.  Row-shifted operands of the vertical dilation
 * (morphology.py:211-248 unions 2 hy + 1 row translates) are views of such
 * offsets. */
int32_t stx_dense_offsets(const int32_t* rows, const int32_t* off, int64_t R, int64_t y0, int64_t n, int32_t* out,
                          hipStream_t stream);
/* Drop rows without intervals (the unique(rows) of the rebuild,
 * morphology.py:43-48).  tmp: 2*(R+1) int32. */
int32_t stx_compact_rows(const int32_t* rows, const int32_t* off, int64_t R, int32_t shift, int32_t* rows_out,
                         int32_t* off_out, int32_t* d_count, int32_t* tmp, uint32_t* status, void* scratch,
                         int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
/* Point location: replaces the locate_cells kernel of
 * locate_interval_cells (interval_field.py:159-297) and the host searches of
 * _locate_interval / get_cell / set_cell (interval_field.py:92-156).
 * out_interval[q]: interval holding (qrow[q], qcol[q]) or -1; out_cell[q]
 * (optional; needs the cell offsets co): value index, -1 for an inactive
 * cell of an existing row, -2 for an absent row.  Either output may be NULL. */
int32_t stx_locate_cells(const int32_t* rows, const int32_t* off, const int32_t* beg, const int32_t* end,
                         const int32_t* co, int64_t R, const int32_t* qrow, const int32_t* qcol, int64_t nq,
                         int32_t* out_interval, int32_t* out_cell, hipStream_t stream);
/* Coarse rows floor(row/ratio) of restrict_set (multilevel.py:284-316) and
 * per-phase fine source rows maps[ph*cap + c]. */
int32_t stx_coarse_rows(const int32_t* rows, const int32_t* off, int64_t R, int32_t ratio, int32_t all_rows,
                        int32_t* crow, int32_t* maps, int64_t cap, int32_t* d_count, int32_t* tmp, uint32_t* status,
                        void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
/* _prolong_set_impl (multilevel.py:213-264) of a set whose rows all hold
 * intervals: K*ratio rows, N*ratio intervals; fbase (optional) maps each fine
 * interval to its coarse interval. */
int32_t stx_prolong_set(const int32_t* rows, const int32_t* off, int64_t K, const int32_t* beg, const int32_t* end,
                        int64_t N, int32_t ratio, int32_t* frows, int32_t* foff, int32_t* fbeg, int32_t* fend,
                        int32_t* fbase, hipStream_t stream);
/* Rows of the prolongation of an empty set (multilevel.py:225-232). */
int32_t stx_prolong_rows(const int32_t* rows, int64_t K, int32_t ratio, int32_t* frows, hipStream_t stream);
int32_t stx_fill_i32(int32_t* p, int64_t n, int32_t value, int32_t step, hipStream_t stream);
/* Box dilation by [-hx, hx] x [-hy, hy] (dilate_interval_set,
 * morphology.py:211-265: the union of the translated copies), for
 * 1 <= hy <= stx_dilate_max_halo(): rows holding intervals are compacted,
 * the output rows are those within hy of one of them, and each output row
 * is the union of its (<= 2 hy + 1) input rows widened by hx.  Outputs:
 * rows_out / off_out (row_cap >= R (2 hy + 1), off_out row_cap + 1 entries),
 * beg_out / end_out (capacity >= (2 hy + 1) N), *d_ny rows, *d_total
 * intervals.  tmp: stx_dilate_tmp_ints(R, hy) int32.  Three scans on the
 * scratch use epochs epoch, epoch + 1, epoch + 2. */
int32_t stx_dilate_max_halo(void);
/* The same dilation over the dense output rows row_lo - hy .. row_hi + hy
 * (row_lo / row_hi: the set's first and last row id; nt = row_hi - row_lo +
 * 1 + 2 hy): tiles of consecutive output rows, balanced by interval count
 * (a table in the scratch), each merge their 2 hy + 1 widened input rows in
 * LDS, then the rows without intervals are dropped.
 * Outputs as stx_dilate (row_cap >= nt, capacity >= (2 hy + 1) N); row ids
 * other than the set's first / last set STX_FLAG_ROWS.  tmp:
 * stx_dilate_dense_tmp_ints(nt, hy) int32; scratch:
 * stx_dilate_dense_scratch_bytes(nt, N) with epochs epoch and epoch + 1. */
/* Erosion (erode_interval_set, morphology.py:361-386: U - clip(dilate(U -
 * clip(S))), U = [0, W) x [0, H), the exterior counting as inside) in one
 * launch of row tiles balanced by interval count (a table in the scratch),
 * for 0 <= hy <= stx_dilate_max_halo(): each tile takes
 * the complement inside [0, W) of its rows' clipped input window (rows
 * outside [0, H) give none), its gaps widened by hx, the 2 hy + 1 row union
 * and the complement of that.  Outputs: off_out (H + 1 entries; the rows are
 * 0 .. H - 1), beg_out / end_out (capacity >= (2 hy + 1)(N + H) + H),
 * *d_total.  tmp: stx_erode_dense_tmp_ints(H, hy) int32; scratch:
 * stx_erode_dense_scratch_bytes(H, N), one epoch. */
int64_t stx_erode_dense_tmp_ints(int64_t H, int32_t hy);
int64_t stx_erode_dense_scratch_bytes(int64_t H, int64_t N);
int32_t stx_erode_dense(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg, const int32_t* end,
                        int64_t N, int32_t hx, int32_t hy, int32_t W, int64_t H, int32_t* off_out, int32_t* beg_out,
                        int32_t* end_out, int64_t capacity, int32_t* d_total, int32_t* tmp, int64_t tmp_ints,
                        uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
int64_t stx_dilate_dense_tmp_ints(int64_t nt, int32_t hy);
int64_t stx_dilate_dense_scratch_bytes(int64_t nt, int64_t N);
int32_t stx_dilate_dense(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg, const int32_t* end,
                         int64_t N, int32_t hx, int32_t hy, int32_t row_lo, int32_t row_hi, int32_t* rows_out,
                         int32_t* off_out, int32_t* beg_out, int32_t* end_out, int64_t row_cap, int64_t capacity,
                         int32_t* d_ny, int32_t* d_total, int32_t* tmp, int64_t tmp_ints, uint32_t* status,
                         void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
int64_t stx_dilate_tmp_ints(int64_t R, int32_t hy);
int32_t stx_dilate(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg, const int32_t* end,
                   int32_t hx, int32_t hy, int32_t* rows_out, int32_t* off_out, int32_t* beg_out, int32_t* end_out,
                   int64_t row_cap, int64_t capacity, int32_t* d_ny, int32_t* d_total, int32_t* tmp, int64_t tmp_ints,
                   uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
/* translate_interval_set (morphology.py:115-140): dst = src + v. */
int32_t stx_add_i32(const int32_t* src, int32_t* dst, int64_t n, int32_t v, hipStream_t stream);

/* ---- fields -------------------------------------------------------------- */
/* Interval lengths for create_interval_field's cell offsets
 * (interval_field.py:64-75; scan them with stx_scan_i32). */
int32_t stx_interval_lengths(const int32_t* beg, const int32_t* end, int64_t n, int32_t* len, hipStream_t stream);
/* esize 1/4/8 fill of n elements with the bit pattern `bits`. */
int32_t stx_fill_value(int32_t esize, void* p, int64_t n, int64_t bits, hipStream_t stream);
/* prolong_field_f32 / _f64 (kernels.py:384-455; multilevel.py:365-407) for
 * a coarse field without empty rows: K rows (c_off), cell offsets c_co.
 * dtype 4 float32, 8 float64.  group: lane-group size (4/16/64). */
int32_t stx_prolong_field(int32_t dtype, int32_t group, const int32_t* c_off, int64_t K, const int32_t* c_co,
                          const void* cv, int32_t ratio, void* fv, hipStream_t stream);
/* Alignment checks of restrict_field (multilevel.py:169-210, :433-451);
 * sets STX_FLAG_ALIGN. */
int32_t stx_restrict_align(const int32_t* c_rows, const int32_t* c_off, int64_t K, const int32_t* c_beg,
                           const int32_t* c_end, const int32_t* f_rows, const int32_t* f_off, const int32_t* f_beg,
                           const int32_t* f_end, int32_t ratio, uint32_t* status, hipStream_t stream);
/* restrict_field_f32 / _f64 (kernels.py:458-598).  kind 0: float in, float
 * accumulate (the reference kernel); 1: float in, double accumulate, float
 * out; 2: double.  reducer 0 mean, 1 sum, 2 min, 3 max (multilevel.py:44). */
int32_t stx_restrict_field(int32_t kind, int32_t group, const int32_t* c_off, int64_t K, const int32_t* c_co,
                           const int32_t* f_off, const int32_t* f_co, const void* fv, int32_t ratio, int32_t reducer,
                           void* cv, hipStream_t stream);
/* Row ids within [0, height) (interval_stencil.py:183-191); STX_FLAG_ROWS. */
int32_t stx_rows_in_range(const int32_t* rows, int64_t R, int32_t height, uint32_t* status, hipStream_t stream);
/* interval_upwind + sample_interval_cell (interval_stencil.py:16-113):
 * zero-exterior first-order upwind.  dtype 4: float arithmetic (scalars
 * rounded to float), 8: double. */
int32_t stx_upwind(int32_t dtype, int32_t group, const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg,
                   const int32_t* end, const int32_t* co, const void* u, void* out, int32_t width, int32_t height,
                   double a, double b, double dt, double dx, double dy, hipStream_t stream);
/* Segment form of the same update (interval_stencil.py:16-113).  A
 * per-geometry table, built once and reused for every step on that geometry
 * and domain width, cuts every interval into segments over which the
 * intervals of the rows above and below do not change, also cut at columns 0
 * and width (8 int32 per segment).  Build:
 * stx_stencil_seg_count (per-interval counts, their scan into seg_off[N+1],
 * total in *d_total), then stx_stencil_seg_build (records; row_inst / inst_w
 * optional, for a stack of instances; plan = 2 int32 per tile of
 * stx_stencil_tile_cells() cells, stx_stencil_plan_tiles(C) tiles).  co: N+1
 * cell offsets; ord: row index of every interval.  Row ids must lie in
 * [0, height).  stx_upwind_seg: inst_coef == NULL: one instance (width, a, b,
 * dt, dx, dy as given; dtype 4 rounds them to float); otherwise the batched
 * form of stx_upwind_batched with a table built with row_inst / inst_w
 * (instances whose dx is 0 are skipped). */
int64_t stx_stencil_tile_cells(void);
int64_t stx_stencil_plan_tiles(int64_t cells);
int32_t stx_stencil_seg_count(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg,
                              const int32_t* end, const int32_t* co, const int32_t* ord, int64_t N, int32_t width,
                              const int32_t* row_inst, const int32_t* inst_w, int32_t* counts, int32_t* seg_off,
                              int32_t* d_total, uint32_t* status, void* scratch, int64_t scratch_bytes,
                              uint32_t epoch, hipStream_t stream);
int32_t stx_stencil_seg_build(const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg,
                              const int32_t* end, const int32_t* co, const int32_t* ord, int64_t N, int32_t width,
                              const int32_t* row_inst, const int32_t* inst_w, const int32_t* seg_off, int64_t nseg,
                              int32_t* segs, int64_t C, int32_t* plan, int64_t ntiles, hipStream_t stream);
int32_t stx_upwind_seg(int32_t dtype, const int32_t* plan, int64_t ntiles, int64_t C, const int32_t* segs,
                       const void* u, void* out, double a, double b, double dt, double dx, double dy,
                       const void* inst_coef, hipStream_t stream);
/* Gradient tags of an interval field through its segment table (the tag of
 * the regrid step, regrid.py:40-120 / K13, as stx_grad_tag): runs of cells
 * with |grad u| >= t, per row.  Every cell must lie in [0, w) of its row for
 * the widths the table was built with (else STX_FLAG_DOMAIN).  words: 6 * W + 1
 * uint32 of scratch, W = stx_grad_tag_seg_words(C); off / co / R: the set's
 * row offsets, cell offsets and row count; outputs as stx_grad_tag. */
int64_t stx_grad_tag_seg_words(int64_t cells);
int32_t stx_grad_tag_seg(int32_t dtype, const int32_t* plan, int64_t ntiles, int64_t C, const int32_t* segs,
                         const int32_t* off, const int32_t* co, int64_t R, const void* u, double t, uint32_t* words,
                         int64_t W, int32_t* out_off, int32_t* out_beg, int32_t* out_end, int64_t capacity,
                         int32_t* d_total, uint32_t* status, void* scratch, int64_t scratch_bytes, uint32_t epoch,
                         hipStream_t stream);
/* FP64 mass sum(u)*scale with a fixed reduction order (no reference
 * counterpart; conservation check of the AMR step).  part: stx_mass_parts(n)
 * doubles. */
int64_t stx_mass_parts(int64_t n);
int32_t stx_mass(int32_t dtype, const void* u, int64_t n, double scale, double* part, double* out, hipStream_t stream);
/* Row location for gather/scatter (fields.py:310-318): per subset interval,
 * the index of its row in the super rows (-1 -> STX_FLAG_ROWS). */
int32_t stx_locate_rows(const int32_t* sub_rows, const int32_t* sub_off, int64_t Rs, const int32_t* sup_rows, int64_t Rp,
                        int32_t* srcrow_per_iv, uint32_t* status, hipStream_t stream);
/* subsetix_interval_subset_gather / _scatter (fields.py:127-294).  mode 0
 * gather (super -> subset), 1 scatter; esize 1, 4 or 8. */
int32_t stx_subset_copy(int32_t mode, int32_t esize, int32_t group, const int32_t* sub_beg, const int32_t* sub_end,
                        const int32_t* sub_co, const int32_t* sub_srcrow, int64_t nsub, const int32_t* sup_off,
                        const int32_t* sup_beg, const int32_t* sup_end, const int32_t* sup_co, void* sup_v, void* sub_v,
                        hipStream_t stream);
/* _interval_field_value_intervals (fields.py:41-100): runs of cells equal to
 * `target_bits` (esize 1 or 4), never across an interval. */
int32_t stx_value_runs(int32_t esize, int32_t group, const int32_t* off, int64_t R, const int32_t* beg, const int32_t* co,
                       const void* values, int64_t target_bits, int32_t* out_off, int32_t* out_beg, int32_t* out_end,
                       int64_t capacity, int32_t* d_total, uint32_t* status, void* scratch, int64_t scratch_bytes,
                       uint32_t epoch, hipStream_t stream);

/* Batched upwind over a stack of independent instances: row k belongs to
 * instance row_inst[k] (width inst_w[i], coefficients inst_coef[5i..5i+4] =
 * a, b, dt, dx, dy in the field type); instance row ids are separated by
 * gaps.  An instance whose dx is 0 is retired: its output cells are left
 * unwritten.  No reference counterpart (the reference advances one instance
 * per call, simulation.py:330-354). */
int32_t stx_upwind_batched(int32_t dtype, int32_t group, const int32_t* rows, const int32_t* off, int64_t R,
                           const int32_t* beg, const int32_t* end, const int32_t* co, const void* u, void* out,
                           const int32_t* row_inst, const int32_t* inst_w, const void* inst_coef, hipStream_t stream);
/* Prolong (mode 0) / restrict (mode 1) onto the cells of a target set T that
 * lie in destination geometry D, reading source geometry S: the restrict /
 * gather / scatter and the gather / prolong / scatter chains of
 * synchronize_interval_fields (fields.py:570-612) in one pass each.
 * Geometries: rows, off, beg, end, cell offsets, row count. */
int32_t stx_transfer(int32_t mode, int32_t dtype, int32_t acc64, int32_t group, int32_t ratio, const int32_t* t_rows,
                     const int32_t* t_off, const int32_t* t_beg, const int32_t* t_end, const int32_t* t_co, int64_t t_R,
                     const int32_t* d_rows, const int32_t* d_off, const int32_t* d_beg, const int32_t* d_end,
                     const int32_t* d_co, int64_t d_R, const int32_t* s_rows, const int32_t* s_off, const int32_t* s_beg,
                     const int32_t* s_end, const int32_t* s_co, int64_t s_R, const void* src, void* dst,
                     hipStream_t stream);
/* Segment schedule of stx_transfer (same chains, fields.py:570-612): mode 0
 * ratio-2 prolong, 1 ratio-2 restrict, 2 ratio-1 copy (the ratio-1 prolong of
 * stx_transfer).  stx_transfer_seg_count (per target interval counts,
 * their scan into seg_off[tN+1], total in *d_total), stx_transfer_seg_build
 * (records, plan = 2 int32 per tile of stx_transfer_tile_cells() target
 * cells), then stx_transfer_seg per step.  t_ord: row index of every target
 * interval; CT: target cell count; src_cells: source field size (cells the
 * source geometry does not hold read as 0, which equals skipping them). */
int64_t stx_transfer_tile_cells(void);
int32_t stx_transfer_seg_count(int32_t mode, const int32_t* t_rows, const int32_t* t_off, const int32_t* t_beg,
                               const int32_t* t_end, const int32_t* t_co, int64_t t_R, const int32_t* t_ord,
                               int64_t tN, const int32_t* d_rows, const int32_t* d_off, const int32_t* d_beg,
                               const int32_t* d_end, const int32_t* d_co, int64_t d_R, const int32_t* s_rows,
                               const int32_t* s_off, const int32_t* s_beg, const int32_t* s_end, const int32_t* s_co,
                               int64_t s_R, int32_t* counts, int32_t* seg_off, int32_t* d_total, uint32_t* status,
                               void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
int32_t stx_transfer_seg_build(int32_t mode, const int32_t* t_rows, const int32_t* t_off, const int32_t* t_beg,
                               const int32_t* t_end, const int32_t* t_co, int64_t t_R, const int32_t* t_ord,
                               int64_t tN, const int32_t* d_rows, const int32_t* d_off, const int32_t* d_beg,
                               const int32_t* d_end, const int32_t* d_co, int64_t d_R, const int32_t* s_rows,
                               const int32_t* s_off, const int32_t* s_beg, const int32_t* s_end, const int32_t* s_co,
                               int64_t s_R, const int32_t* seg_off, int64_t nseg, int32_t* segs, int64_t CT,
                               int32_t* plan, int64_t ntiles, hipStream_t stream);
int32_t stx_transfer_seg(int32_t mode, int32_t dtype, int32_t acc64, const int32_t* plan, int64_t ntiles, int64_t CT,
                         const int32_t* segs, const void* src, int64_t src_cells, void* dst, hipStream_t stream);

/* Instance of every row of a stack (last i with bases[i] <= rows[k]). */
int32_t stx_row_instance(const int32_t* rows, int64_t R, const int32_t* bases, int32_t ninst, int32_t* out,
                         hipStream_t stream);
/* FP64 mass per instance: out[i] = scale[i] * sum of u over cells
 * [cstart[i], cstart[i+1]) with a fixed reduction order that depends only on
 * the instance's cell count (the same mass in any stack). */
int32_t stx_instance_mass(int32_t dtype, const void* u, const int64_t* cstart, int32_t ninst, const double* scale,
                          double* out, hipStream_t stream);

/* ---- regrid -------------------------------------------------------------- */
/* Gradient tagging on an interval field (generalises regrid.py:175-223 +
 * 20-124 to sparse levels: one-sided differences where a neighbour is
 * missing); runs of |grad u| >= t, per row of the field. */
int32_t stx_grad_tag(int32_t dtype, int32_t group, const int32_t* rows, const int32_t* off, int64_t R, const int32_t* beg,
                     const int32_t* end, const int32_t* co, const void* u, double t, int32_t* out_off, int32_t* out_beg,
                     int32_t* out_end, int64_t capacity, int32_t* d_total, uint32_t* status, void* scratch,
                     int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
/* subsetix_count/write_threshold_intervals (regrid.py:20-124) and, with
 * gradient != 0, subsetix_gradient_magnitude_flat (regrid.py:175-223) fused:
 * maximal runs of v >= t per row of a dense width x height float field. */
int32_t stx_threshold(int32_t gradient, const float* data, int32_t width, int32_t height, float t, int32_t* out_off,
                      int32_t* out_beg, int32_t* out_end, int64_t capacity, int32_t* d_total, uint32_t* status,
                      void* scratch, int64_t scratch_bytes, uint32_t epoch, hipStream_t stream);
int64_t stx_threshold_scratch_bytes(int32_t height);
/* subsetix_gradient_magnitude_flat (regrid.py:175-223) as a separate pass. */
int32_t stx_gradient(const float* data, int32_t width, int32_t height, float* out, hipStream_t stream);


/* ---- batched three-level AMR with a device-only regrid ---------------------
 * Replaces the host-driven regrid of simulation.py:291-328 / regrid.py:404-442
 * (tag, grade, regrid) and the per-level step (simulation.py:335-354) for a
 * stack of independent instances, with launch shapes fixed by the instance
 * geometry so that a regrid and the steps after it are captured once as one
 * graph.  Levels are dense row-major per instance; the refine sets are bit
 * masks (one uint64 per 64-cell row chunk); work lists are chunk descriptors
 * (i << 48 | y << 24 | m) with device-side counts (d_n; NULL = bound).
 * lt* : int64[5 n] per level (side, first cell, first mask word, words per
 * row, row pitch; level 0 carries a one-cell zero ring: pitch = side + 2).  t: tag threshold (|grad u| >= t). */
int32_t stx_amr3_dilate(const int64_t* lt, const uint64_t* list, const int32_t* d_n, int64_t bound, const uint64_t* in,
                        uint64_t* out, hipStream_t stream);
int32_t stx_amr3_refine(const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n, int64_t bound,
                        const uint64_t* rp, const uint64_t* d, uint64_t* r, hipStream_t stream);
int32_t stx_amr3_list_count(int32_t mode, const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n,
                            int64_t bound, const uint64_t* e, int32_t* counts, hipStream_t stream);
int32_t stx_amr3_list_emit(int32_t mode, const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n,
                           int64_t bound, const uint64_t* e, const int32_t* offsets, uint64_t* out, hipStream_t stream);
int32_t stx_amr3_zero(const int64_t* lt, const uint64_t* list, const int32_t* d_n, int64_t bound, uint64_t* m0,
                      uint64_t* m1, uint64_t* m2, hipStream_t stream);
int32_t stx_amr3_popc(const int64_t* lt, const uint64_t* list, const int32_t* d_n, int64_t bound, const uint64_t* m,
                      uint64_t* out, hipStream_t stream);
/* Work records (built once per regrid; see amr3.hip): kind 0 = stencil
 * records of the listed chunks of level ltc holding cells of prolong(rp),
 * kind 1 = restriction records of the listed chunks of level ltp whose refine
 * word (rp) is non-zero.  Record-driven step and regrid kernels read one
 * coalesced record per chunk. */
int32_t stx_amr3_rec_count(int32_t kind, const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n,
                           int64_t bound, const uint64_t* rp, int32_t* counts, hipStream_t stream);
int32_t stx_amr3_rec_emit(int32_t kind, const int64_t* ltp, const int64_t* ltc, const uint64_t* list, const int32_t* d_n,
                          int64_t bound, const uint64_t* rp, const int32_t* offsets, uint64_t* out, hipStream_t stream);
int32_t stx_amr3_tag0_records(const int64_t* lt0, const uint64_t* list, int64_t n, uint64_t* out, hipStream_t stream);
int32_t stx_amr3_tag0_rec(int32_t dtype, const uint64_t* rec, int64_t n, const void* v, double t, uint64_t* out,
                          hipStream_t stream);
/* Level-0 upwind over the padded full domain of every instance (records of
 * stx_amr3_tag0_records). */
int32_t stx_amr3_upwind0_rec(int32_t dtype, const uint64_t* rec, int64_t n, const void* v, void* vn, const void* coef,
                             hipStream_t stream);
int32_t stx_amr3_stencil_rec(int32_t dtype, const uint64_t* rec1, const int32_t* d_n1, const uint64_t* rec2,
                             const int32_t* d_n2, int64_t bound, const void* v0, const void* v1, const void* v2, void* v1n,
                             void* v2n, const void* coef1, const void* coef2, hipStream_t stream);
int32_t stx_amr3_lupdate_rec(int32_t dtype, const int64_t* ltp, const int64_t* ltc, const uint64_t* rec, const int32_t* d_n,
                             int64_t bound, const uint64_t* rold, const void* vp, void* vc, hipStream_t stream);
int32_t stx_amr3_tagl_rec(int32_t dtype, const int64_t* ltc, const uint64_t* rec, const int32_t* d_n, int64_t bound,
                          const void* v, double t, uint64_t* out, hipStream_t stream);
int32_t stx_amr3_restrict_rec(int32_t dtype, int32_t acc64, const uint64_t* rec, const int32_t* d_n, int64_t bound,
                              const void* vc, void* vp, hipStream_t stream);
int32_t stx_amr3_accumulate(uint64_t* total, const uint64_t* bits, int64_t base, int64_t steps, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SUBSETIX_HIP_H */
