"""Batched three-level AMR whose regrid runs entirely on the device: no host
synchronisation inside a regrid or a step, so ``regrid + k steps`` is
captured once as one hipGraph and replayed for the whole run.

Same algorithm as :class:`subsetix_amd.batch.BatchedAMR`, cell for cell: gradient tag on level 0,
grading dilation, a one-cell level-1 ring forced onto its parents, level-1
gradient tag gated by the eroded level 1, prolonged level sets, new cells
prolonged from their parents, kept values elsewhere; a step prolongs ghost
cells, advances every level with the zero-exterior upwind update and
restricts finest to coarsest with FP64 sums.

What changes is the representation, chosen so that no launch depends on a
size known only after a kernel has run:

* level values are dense row-major per instance (level L of an instance
  with coarse size n is (n 2^L)^2 cells); only the level sets' cells carry
  values on levels 1 and 2;
* the refine sets are bit masks, one 64-bit word per 64-cell row chunk, at
  level-0 (R0) and level-1 (R1) resolution; level set L = prolong(R_{L-1});
  set algebra (dilation, erosion, prolongation) is word arithmetic;
* ghost cells are not stored: the stencil reads a ghost neighbour as its
  parent's current value, which is what the prolongation of ghosts would
  have written;
* the work of levels 1 and 2 is driven by chunk lists built on the device
  (count, exclusive scan, emit; counts stay on the device); each regrid then
  turns the level sets' chunks into work records that carry everything a
  step needs (cell bases, row sides, the level-set words of the chunk and of
  its neighbours), so a step's kernels start from one coalesced load per
  record; every grid and buffer is sized from the instance geometry alone
  (capacity bounds), so nothing is ever re-captured.

Level 0 (the full domain of every instance) is stored with a one-cell ring
of zeros around each instance, so its update and gradient tag read every
neighbour without a test (the ring is the zero exterior); both run from
static records, one per 64-cell row chunk.  Every record kernel is a
persistent grid-stride loop in which a wavefront loads 8 or 16 records at
once and issues all value loads of a group of records before using any, so
it keeps dozens of loads in flight.  A step is four launches (level-0
upwind, level-1/2 upwind, two restrictions); a regrid about thirty, all
device-side.
"""

from __future__ import annotations

import math
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib, autotune
from ._lib import _F64, _I32, _I64, _P, check, context, ptr
from .batch import BatchConfig

_lib.register("stx_amr3_dilate", _I32, [_P, _P, _P, _I64, _P, _P, _P])
_lib.register("stx_amr3_refine", _I32, [_P, _P, _P, _P, _I64, _P, _P, _P, _P])
_lib.register("stx_amr3_list_count", _I32, [_I32, _P, _P, _P, _P, _I64, _P, _P, _P])
_lib.register("stx_amr3_list_emit", _I32, [_I32, _P, _P, _P, _P, _I64, _P, _P, _P, _P])
_lib.register("stx_amr3_zero", _I32, [_P, _P, _P, _I64, _P, _P, _P, _P])
_lib.register("stx_amr3_popc", _I32, [_P, _P, _P, _I64, _P, _P, _P])
_lib.register("stx_amr3_accumulate", _I32, [_P, _P, _I64, _I64, _P])
_lib.register("stx_amr3_rec_count", _I32, [_I32, _P, _P, _P, _P, _I64, _P, _P, _P])
_lib.register("stx_amr3_rec_emit", _I32, [_I32, _P, _P, _P, _P, _I64, _P, _P, _P, _P])
_lib.register("stx_amr3_tag0_records", _I32, [_P, _P, _I64, _P, _P])
_lib.register("stx_amr3_tag0_rec", _I32, [_I32, _P, _I64, _P, _F64, _P, _P])
_lib.register("stx_amr3_upwind0_rec", _I32, [_I32, _P, _I64, _P, _P, _P, _P])
_lib.register("stx_amr3_stencil_rec", _I32, [_I32, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P])
_lib.register("stx_amr3_lupdate_rec", _I32, [_I32, _P, _P, _P, _P, _I64, _P, _P, _P, _P])
_lib.register("stx_amr3_tagl_rec", _I32, [_I32, _P, _P, _P, _I64, _P, _F64, _P, _P])
_lib.register("stx_amr3_restrict_rec", _I32, [_I32, _I32, _P, _P, _I64, _P, _P, _P])

_SW, _QW = 8, 4  # words per stencil / restriction record


def _encode(i, y, m):
    return (np.int64(i) << 48) | (y.astype(np.int64) << 24) | m.astype(np.int64)


class _Level:
    """Host and device tables of one level: side, cell (0, 0), first mask
    word, words per row, row pitch (int64[n, 5]).  ``pad``: width of the ring
    of zeros around each instance's block (level 0: 1), so a block is
    (side + 2 pad)^2 cells; ``start`` holds the block boundaries and
    ``ncells`` the allocation (plus 64 cells of slack after a padded level,
    which whole-chunk loads past the last row may touch)."""

    def __init__(self, sizes: np.ndarray, level: int, dev, pad: int = 0):
        self.side = (sizes << level).astype(np.int64)
        self.pad = int(pad)
        self.pitch = self.side + 2 * self.pad
        self.cells = self.side * self.side
        self.cw = (self.side + 63) // 64
        self.words = self.side * self.cw
        self.start = np.concatenate([[0], np.cumsum(self.pitch * self.pitch)]).astype(np.int64)
        self.origin = self.start[:-1] + self.pad * (self.pitch + 1)
        self.word0 = np.concatenate([[0], np.cumsum(self.words)]).astype(np.int64)
        tab = np.stack([self.side, self.origin, self.word0[:-1], self.cw, self.pitch], axis=1).astype(np.int64)
        self.table = torch.from_numpy(np.ascontiguousarray(tab.reshape(-1))).to(dev)
        self.ncells = int(self.start[-1]) + (64 if self.pad else 0)
        self.interior = int(self.cells.sum())
        self.nwords = int(self.word0[-1])

    def all_chunks(self) -> np.ndarray:
        out = []
        for i, (s, cw) in enumerate(zip(self.side.tolist(), self.cw.tolist())):
            ys = np.repeat(np.arange(s, dtype=np.int64), cw)
            ms = np.tile(np.arange(cw, dtype=np.int64), s)
            out.append(_encode(i, ys, ms))
        return np.concatenate(out)


class DeviceBatchedAMR:
    """Many independent three-level advection instances (see the module
    docstring).  ``instances``: dicts with ``n`` (coarse size), ``bumps``
    ((cx, cy, sigma, amplitude) Gaussian bumps) and ``velocity``."""

    def __init__(self, instances: List[dict], *, dtype=torch.float32, config: Optional[BatchConfig] = None,
                 device=None):
        self.cfg = config or BatchConfig()
        if self.cfg.levels != 3 or self.cfg.ratio != 2:
            raise ValueError("the device-regrid engine runs three levels at ratio 2")
        self.dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        self.dcode = 4 if dtype == torch.float32 else 8
        self.instances = instances
        self.n = len(instances)
        if self.n >= (1 << 15):
            raise ValueError("at most 32767 instances")
        sizes = np.array([int(i["n"]) for i in instances], dtype=np.int64)
        if sizes.min() < 1 or (sizes.max() << 2) >= (1 << 24):
            raise ValueError("instance sizes must be in [1, 2^22)")
        self.sizes = sizes
        dev = self.dev
        self.L = [_Level(sizes, lv, dev, pad=1 if lv == 0 else 0) for lv in range(3)]
        npdt = np.float32 if dtype == torch.float32 else np.float64
        # per-level coefficients (a, b, dt, dx, dy) as the interval-set engine,
        vel = np.array([i["velocity"] for i in instances], dtype=np.float64)
        finest = sizes * 4
        denom = np.abs(vel[:, 0]) * finest + np.abs(vel[:, 1]) * finest
        dts = np.where(denom > 0, self.cfg.cfl / np.maximum(denom, 1e-300), 0.0)
        # and the correctly rounded reciprocals of dx, dy that the update's
        # division uses (1 / d in FP64 rounds correctly to FP32 as well)
        self.coef = []
        for lv in range(3):
            nl = sizes * (1 << lv)
            c = np.stack([vel[:, 0], vel[:, 1], dts, 1.0 / nl, 1.0 / nl], axis=1).astype(npdt)
            rec = (1.0 / c[:, 3:5].astype(np.float64)).astype(npdt)
            self.coef.append(torch.from_numpy(np.ascontiguousarray(np.concatenate([c, rec], axis=1).reshape(-1)))
                             .to(dev))
        # dense level values, two buffers each
        self.V = [[torch.zeros(lv.ncells, dtype=dtype, device=dev) for _ in range(2)] for lv in self.L]
        self.cur = 0
        # masks: level 0 (tag, two temps, refine x2, E) and level 1 (tag, temps, refine x2, E)
        z = lambda k: torch.zeros(k, dtype=torch.int64, device=dev)  # noqa: E731
        n0w, n1w = self.L[0].nwords, self.L[1].nwords
        self.T0, self.A0, self.B0, self.E0 = z(n0w), z(n0w), z(n0w), z(n0w)
        self.R0 = [z(n0w), z(n0w)]
        self.T1, self.D1, self.E1 = z(n1w), z(n1w), z(n1w)
        self.R1 = [z(n1w), z(n1w)]
        # work lists (capacity bounds from the geometry) and their device counts
        self.l0all = torch.from_numpy(self.L[0].all_chunks()).to(dev)
        self.n0 = int(self.l0all.numel())
        self.cap1 = 4 * self.n0
        self.cap2 = 4 * self.cap1
        self.l0r = z(self.n0)
        self.a1 = [z(self.cap1), z(self.cap1)]
        # work records: stencil records of levels 1 and 2, restriction records
        # 2 -> 1 and 1 -> 0, level-0 tag records (static)
        self.S1, self.S2 = z(_SW * self.cap1), z(_SW * self.cap2)
        self.Q1, self.Q0 = z(_QW * self.cap1), z(_QW * self.n0)
        self.rec0 = z(_QW * self.n0)
        # [n0r, n1[0], n1[1], nS1, nS2, nQ1, nQ0, -]
        self.counts = torch.zeros(8, dtype=torch.int32, device=dev)
        self.tmp_cnt = torch.empty(self.cap1 + 1, dtype=torch.int32, device=dev)
        self.tmp_off = torch.empty(self.cap1 + 1, dtype=torch.int32, device=dev)
        self.scan_bytes = int(_lib.lib().stx_scan_scratch_bytes(self.cap1))
        self.scan_scratch = torch.zeros(self.scan_bytes // 4 + 1, dtype=torch.int32, device=dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.bits = torch.zeros(2, dtype=torch.int64, device=dev)  # popcounts of R0, R1 (last regrid)
        self.cell_total = torch.zeros(1, dtype=torch.int64, device=dev)
        self.parity = 1  # the first regrid writes R[0], A1[0]
        check(_lib.lib().stx_amr3_tag0_records(ptr(self.L[0].table), ptr(self.l0all), self.n0, ptr(self.rec0),
                                               context(dev).stream), "tag0_records")
        # level-0 block boundaries (the zero ring adds nothing to a mass)
        self.cstart0 = torch.from_numpy(self.L[0].start).to(dev)
        self.cell_area0 = torch.from_numpy((1.0 / (sizes * sizes)).astype(np.float64)).to(dev)
        self._init_values()
        self.step_count = 0
        self._last_regrid = 0
        self._graphs: Dict[tuple, torch.cuda.CUDAGraph] = {}
        self._gstream = torch.cuda.Stream(dev)
        self.regrids = 0
        self.regrid()
        if self.cfg.autotune:
            # one step's kernels under the schedule autotuner; the state does
            # not move (the step reads buffer cur and writes the other one)
            with autotune.tuning():
                self._step(self.cur)

    # ------------------------------------------------------------------ setup
    def _init_values(self) -> None:
        v0 = self.V[0][0]
        for k, inst in enumerate(self.instances):
            n = int(inst["n"])
            xs = (torch.arange(n, device=self.dev, dtype=torch.float64) + 0.5) / n
            Y, X = torch.meshgrid(xs, xs, indexing="ij")
            u = torch.zeros((n, n), dtype=torch.float64, device=self.dev)
            for cx, cy, sg, amp in inst["bumps"]:
                u += amp * torch.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sg * sg))
            self._block(v0, 0, k).copy_(u.to(self.dtype))

    def _block(self, v: torch.Tensor, level: int, i: int) -> torch.Tensor:
        """(side, side) view of instance i's cells in a level buffer."""
        lv = self.L[level]
        p, pad, s = int(lv.pitch[i]), lv.pad, int(lv.side[i])
        blk = v[int(lv.start[i]):int(lv.start[i + 1])].view(p, p)
        return blk[pad:pad + s, pad:pad + s]

    # ------------------------------------------------------------------ lists
    def _scan(self, counts: torch.Tensor, d_n, bound: int, d_total) -> None:
        """Exclusive scan of counts into tmp_off; total into d_total (the
        scratch is cleared first: a captured replay reuses one epoch)."""
        ctx = context(self.dev)
        self.scan_scratch.zero_()
        check(_lib.lib().stx_scan_i32(ptr(counts), ptr(d_n), int(bound), ptr(self.tmp_off), ptr(d_total),
                                      ptr(self.status), ptr(self.scan_scratch), self.scan_bytes, 1, ctx.stream), "scan")

    def _build_list(self, mode: int, lp: int, plist, d_pn, pbound: int, E, out, d_out) -> None:
        lib, st = _lib.lib(), context(self.dev).stream
        Lp, Lc = self.L[lp].table, self.L[min(lp + 1, 2)].table
        check(lib.stx_amr3_list_count(mode, ptr(Lp), ptr(Lc), ptr(plist), ptr(d_pn), pbound, ptr(E),
                                      ptr(self.tmp_cnt), st), "list_count")
        self._scan(self.tmp_cnt, d_pn, pbound, d_out)
        check(lib.stx_amr3_list_emit(mode, ptr(Lp), ptr(Lc), ptr(plist), ptr(d_pn), pbound, ptr(E), ptr(self.tmp_off),
                                     ptr(out), st), "list_emit")

    def _build_records(self, kind: int, lp: int, plist, d_pn, pbound: int, R, out, d_out) -> None:
        lib, st = _lib.lib(), context(self.dev).stream
        Lp, Lc = self.L[lp].table, self.L[lp + 1].table
        check(lib.stx_amr3_rec_count(kind, ptr(Lp), ptr(Lc), ptr(plist), ptr(d_pn), pbound, ptr(R),
                                     ptr(self.tmp_cnt), st), "rec_count")
        self._scan(self.tmp_cnt, d_pn, pbound, d_out)
        check(lib.stx_amr3_rec_emit(kind, ptr(Lp), ptr(Lc), ptr(plist), ptr(d_pn), pbound, ptr(R), ptr(self.tmp_off),
                                    ptr(out), st), "rec_emit")

    # ------------------------------------------------------------------ regrid
    def regrid(self) -> None:
        """Device-only regrid (no host synchronisation)."""
        lib, st = _lib.lib(), context(self.dev).stream
        p = 1 - self.parity
        self.parity = p
        self.regrids += 1
        L0, L1, L2 = (lv.table for lv in self.L)
        cur = self.cur
        V0, V1, V2 = self.V[0][cur], self.V[1][cur], self.V[2][cur]
        thr = self.cfg.threshold
        t0, t1 = float(thr[0]), float(thr[min(1, len(thr) - 1)])
        n0 = self.n0
        cnt = self.counts
        d_n0r, d_n1, d_n1old = cnt[0:1], cnt[1 + p:2 + p], cnt[2 - p:3 - p]
        d_s1, d_s2, d_q1, d_q0 = cnt[3:4], cnt[4:5], cnt[5:6], cnt[6:7]
        a1, a1old = self.a1[p], self.a1[1 - p]
        R0n, R0o, R1n, R1o = self.R0[p], self.R0[1 - p], self.R1[p], self.R1[1 - p]
        # level 0: tags, grading dilation(s), the ring dilation, the level-1 extent
        check(lib.stx_amr3_tag0_rec(self.dcode, ptr(self.rec0), n0, ptr(V0), t0, ptr(self.T0), st), "tag0")
        src = self.T0
        for k in range(int(self.cfg.grading)):
            dst = self.A0 if src is not self.A0 else self.B0
            check(lib.stx_amr3_dilate(ptr(L0), ptr(self.l0all), None, n0, ptr(src), ptr(dst), st), "dilate")
            src = dst
        check(lib.stx_amr3_dilate(ptr(L0), ptr(self.l0all), None, n0, ptr(src), ptr(R0n), st), "dilate")
        check(lib.stx_amr3_dilate(ptr(L0), ptr(self.l0all), None, n0, ptr(R0n), ptr(self.E0), st), "dilate")
        # level-1 masks: clear what the last two regrids wrote, then the new chunk lists
        check(lib.stx_amr3_zero(ptr(L1), ptr(a1), ptr(d_n1), self.cap1, ptr(R1n), None, None, st), "zero")
        check(lib.stx_amr3_zero(ptr(L1), ptr(a1old), ptr(d_n1old), self.cap1, ptr(self.T1), ptr(self.D1),
                                ptr(self.E1), st), "zero")
        self._build_list(0, 0, self.l0all, None, n0, self.E0, self.l0r, d_n0r)
        self._build_list(1, 0, self.l0all, None, n0, self.E0, a1, d_n1)
        self._build_records(0, 0, self.l0r, d_n0r, n0, R0n, self.S1, d_s1)
        # level 1: values, tags, grading, erosion
        check(lib.stx_amr3_lupdate_rec(self.dcode, ptr(L0), ptr(L1), ptr(self.S1), ptr(d_s1), self.cap1, ptr(R0o),
                                       ptr(V0), ptr(V1), st), "lupdate")
        check(lib.stx_amr3_tagl_rec(self.dcode, ptr(L1), ptr(self.S1), ptr(d_s1), self.cap1, ptr(V1), t1,
                                    ptr(self.T1), st), "tagl")
        src = self.T1
        for k in range(int(self.cfg.grading)):
            dst = self.D1 if src is not self.D1 else self.E1
            check(lib.stx_amr3_dilate(ptr(L1), ptr(a1), ptr(d_n1), self.cap1, ptr(src), ptr(dst), st), "dilate")
            src = dst
        check(lib.stx_amr3_refine(ptr(L0), ptr(L1), ptr(a1), ptr(d_n1), self.cap1, ptr(R0n), ptr(src), ptr(R1n), st),
              "refine")
        # level 2, and the restriction records of the step
        self._build_records(0, 1, a1, d_n1, self.cap1, R1n, self.S2, d_s2)
        check(lib.stx_amr3_lupdate_rec(self.dcode, ptr(L1), ptr(L2), ptr(self.S2), ptr(d_s2), self.cap2, ptr(R1o),
                                       ptr(V1), ptr(V2), st), "lupdate")
        self._build_records(1, 1, a1, d_n1, self.cap1, R1n, self.Q1, d_q1)
        self._build_records(1, 0, self.l0r, d_n0r, n0, R0n, self.Q0, d_q0)
        # refined cell counts (level sets = 4 cells per refined parent)
        self.bits.zero_()
        check(lib.stx_amr3_popc(ptr(L0), ptr(self.l0all), None, n0, ptr(R0n), ptr(self.bits[0:1]), st), "popc")
        check(lib.stx_amr3_popc(ptr(L1), ptr(a1), ptr(d_n1), self.cap1, ptr(R1n), ptr(self.bits[1:2]), st), "popc")

    # ------------------------------------------------------------------ step
    def _step(self, cur: int) -> None:
        lib, st = _lib.lib(), context(self.dev).stream
        nxt = 1 - cur
        V = self.V
        check(lib.stx_amr3_upwind0_rec(self.dcode, ptr(self.rec0), self.n0, ptr(V[0][cur]), ptr(V[0][nxt]),
                                       ptr(self.coef[0]), st), "upwind0")
        cnt = self.counts
        check(lib.stx_amr3_stencil_rec(self.dcode, ptr(self.S1), ptr(cnt[3:4]), ptr(self.S2), ptr(cnt[4:5]),
                                       self.cap2 + self.cap1, ptr(V[0][cur]), ptr(V[1][cur]), ptr(V[2][cur]),
                                       ptr(V[1][nxt]), ptr(V[2][nxt]), ptr(self.coef[1]), ptr(self.coef[2]), st),
              "amr3_stencil")
        acc = 1 if self.cfg.accumulate == "f64" else 0
        check(lib.stx_amr3_restrict_rec(self.dcode, acc, ptr(self.Q1), ptr(cnt[5:6]), self.cap1, ptr(V[2][nxt]),
                                        ptr(V[1][nxt]), st), "restrict")
        check(lib.stx_amr3_restrict_rec(self.dcode, acc, ptr(self.Q0), ptr(cnt[6:7]), self.n0, ptr(V[1][nxt]),
                                        ptr(V[0][nxt]), st), "restrict")

    def _count(self, steps: int) -> None:
        check(_lib.lib().stx_amr3_accumulate(ptr(self.cell_total), ptr(self.bits), self.L[0].interior, int(steps),
                                             context(self.dev).stream), "accumulate")

    def _steps(self, n: int) -> None:
        for _ in range(n):
            self._step(self.cur)
            self.cur = 1 - self.cur
        self._count(n)

    def _cycle_graph(self, parity_after: int, cur: int, k: int) -> torch.cuda.CUDAGraph:
        """Graph of one regrid (writing the masks of ``parity_after``) and k
        steps from buffer ``cur``; captured once per (parity, buffer)."""
        key = (parity_after, cur, k)
        g = self._graphs.get(key)
        if g is None:
            stream = self._gstream
            stream.wait_stream(torch.cuda.current_stream(self.dev))
            keep = (self.parity, self.cur, self.regrids)
            with torch.cuda.stream(stream):
                context(self.dev)  # per-stream context allocated outside the capture
                g = torch.cuda.CUDAGraph()
                self.parity, self.cur = 1 - parity_after, cur
                g.capture_begin()
                try:
                    self.regrid()
                    self._steps(k)
                finally:
                    g.capture_end()
            self.parity, self.cur, self.regrids = keep
            torch.cuda.current_stream(self.dev).wait_stream(stream)
            self._graphs[key] = g
        return g

    def run(self, steps: int) -> None:
        """Advance every instance by ``steps`` steps, regridding every k steps
        (no host synchronisation; with ``use_graph`` each full regrid cycle is
        one graph replay)."""
        k = int(self.cfg.regrid_every)
        remaining = int(steps)
        while remaining > 0:
            if self.step_count % k == 0 and self._last_regrid != self.step_count:
                if self.cfg.use_graph and remaining >= k and k % 2 == 0:
                    g = self._cycle_graph(1 - self.parity, self.cur, k)
                    with torch.cuda.stream(self._gstream):
                        self._gstream.wait_stream(torch.cuda.current_stream(self.dev))
                        g.replay()
                    torch.cuda.current_stream(self.dev).wait_stream(self._gstream)
                    self.parity = 1 - self.parity
                    self.regrids += 1
                    self._last_regrid = self.step_count
                    self.step_count += k
                    remaining -= k
                    continue
                self.regrid()
                self._last_regrid = self.step_count
            n = min(k - self.step_count % k, remaining)
            self._steps(n)
            self.step_count += n
            remaining -= n

    # ------------------------------------------------------------------ results
    def cell_updates(self) -> int:
        """Level-set cells advanced since construction (one host read)."""
        return int(self.cell_total.item())

    def cells(self) -> Dict[str, int]:
        b = self.bits.tolist()
        return {"level0": self.L[0].interior, "level1": 4 * b[0], "level2": 4 * b[1]}

    def check_status(self) -> None:
        if int(self.status.item()):
            raise _lib.StxError(f"device status flags {int(self.status.item())}")

    def masses(self) -> torch.Tensor:
        out = torch.empty(self.n, dtype=torch.float64, device=self.dev)
        check(_lib.lib().stx_instance_mass(self.dcode, ptr(self.V[0][self.cur]), ptr(self.cstart0), self.n,
                                           ptr(self.cell_area0), ptr(out), context(self.dev).stream), "instance_mass")
        return out

    def level_mask(self, level: int, i: int) -> np.ndarray:
        """Dense bool (s, s) mask of instance i's level set (host copy)."""
        lv = self.L[level]
        s = int(lv.side[i])
        if level == 0:
            return np.ones((s, s), bool)
        lp = self.L[level - 1]
        R = (self.R0 if level == 1 else self.R1)[self.parity]
        sp, cw = int(lp.side[i]), int(lp.cw[i])
        words = np.ascontiguousarray(R[int(lp.word0[i]):int(lp.word0[i + 1])].cpu().numpy())
        # bit b of word (y, m) is cell x = 64 m + b (little-endian bytes)
        bits = np.unpackbits(words.view(np.uint8), bitorder="little").reshape(sp, cw * 64)[:, :sp].astype(bool)
        return np.kron(bits, np.ones((2, 2), bool))

    def level_values(self, level: int, i: int) -> np.ndarray:
        """Dense (s, s) values of instance i at a level (host copy; only the
        level set's cells are meaningful on levels 1 and 2)."""
        return self._block(self.V[level][self.cur], level, i).cpu().numpy()
