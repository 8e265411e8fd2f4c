"""GPU: the allocation-free set operation (stx_setop_noalloc, static device
scratch) against stx_setop with a caller scratch, through the C-ABI: two
calls in a row, the second on a smaller input, plus the size limit."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(rows, seed):
    from subsetix_amd import workloads

    return workloads.setalgebra_pair(rows=rows, width=65536, seed_a=seed, seed_b=seed + 1)


def _dev(a, torch):
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    return t(a.begin), t(a.end), t(a.row_offsets)


def _noalloc(op, A, B, nrows, cap, torch):
    from subsetix_amd import _lib
    from subsetix_amd._lib import ptr

    out_off = torch.empty(nrows + 1, dtype=torch.int32, device="cuda")
    ob = torch.empty(cap, dtype=torch.int32, device="cuda")
    oe = torch.empty(cap, dtype=torch.int32, device="cuda")
    total = torch.zeros(1, dtype=torch.int32, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = _lib.lib().stx_setop_noalloc(op, ptr(A[0]), ptr(A[1]), ptr(A[2]), A[0].numel(), ptr(B[0]), ptr(B[1]),
                                      ptr(B[2]), B[0].numel(), nrows, ptr(out_off), ptr(ob), ptr(oe), cap, ptr(total),
                                      ptr(status), stream)
    return rc, out_off, ob, oe, total, status


def test_noalloc_setop_equals_setop():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import subsetix_amd  # noqa: F401  (registers the bindings)
    from subsetix_amd import _core, _lib

    assert _lib.lib().stx_setop_noalloc_max_scratch_bytes() == 1 << 20
    for rows, seed in ((4096, 1), (700, 5)):  # the second call is smaller
        a, b = _pair(rows, seed)
        A, B = _dev(a, torch), _dev(b, torch)
        cap = a.n + b.n
        for op in (0, 1, 2):
            print("noalloc", rows, op, flush=True)
            rc, off, ob, oe, total, status = _noalloc(op, A, B, rows, cap, torch)
            assert rc == 0, _lib.lib().stx_last_error()
            e_off, e_ob, e_oe, ctx = _core.setop_raw(op, A, B, nrows=rows, capacity=cap, group=-6)
            (n,) = _core.read_counts(ctx, ctx.words[1])
            torch.cuda.synchronize()
            assert int(status.item()) == 0
            assert int(total.item()) == n
            np.testing.assert_array_equal(off.cpu().numpy(), e_off.cpu().numpy())
            np.testing.assert_array_equal(ob[:n].cpu().numpy(), e_ob[:n].cpu().numpy())
            np.testing.assert_array_equal(oe[:n].cpu().numpy(), e_oe[:n].cpu().numpy())


def test_noalloc_rejects_inputs_beyond_the_static_scratch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import subsetix_amd  # noqa: F401
    from subsetix_amd import _lib

    # 2^27 rows with no intervals need far more than 131,064 tiles; the
    # check runs on the host before any launch, so the tiny buffers are
    # never touched
    z = torch.zeros(2, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    p = z.data_ptr()
    rc = _lib.lib().stx_setop_noalloc(0, p, p, p, 0, p, p, p, 0, 1 << 27, p, p, p, 0, p, p, stream)
    assert rc != 0
    assert b"static" in _lib.lib().stx_last_error()
