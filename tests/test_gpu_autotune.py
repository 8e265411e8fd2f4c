"""GPU: the schedule autotuner.

* every candidate schedule a family offers, forced through the tuner, gives
  the oracle's result bit-exactly (set operations with aligned and with
  sparse rows, upwind, level transfers);
* a tuned batched AMR run (real timings) equals the untuned run bit-exactly,
  and its table holds the step kernels' families;
* the composite driver's tuning pass leaves the state unchanged.
"""

import numpy as np
import pytest

from oracle import oracle as ox
from subsetix_amd import autotune
from tests.golden_util import random_canonical
from tests.gpu_util import to_device, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


class Forced(autotune.Autotuner):
    """Runs every candidate once and then picks the k-th (the last one when
    a family has fewer)."""

    def __init__(self, k):
        super().__init__()
        self.k = k
        self.enabled = True
        self.ran = {}

    def tune(self, key, factory):
        cands = factory()
        names = list(cands)
        for fn in cands.values():
            fn()
        choice = names[min(self.k, len(names) - 1)]
        self.table[key] = autotune.Entry(choice, {})
        self.ran[key] = names
        return choice


@pytest.fixture
def install():
    old = autotune._DEFAULT

    def put(t):
        autotune._DEFAULT = t
        return t

    yield put
    autotune._DEFAULT = old


@pytest.mark.parametrize("k", range(11))
def test_setop_candidates_vs_oracle(torch_mod, install, k):
    from subsetix_amd.expressions import binary_op

    t = install(Forced(k))
    rng = np.random.default_rng(100 + k)
    a = random_canonical(rng, 300, 4000, mean_n=20)
    b = random_canonical(rng, 300, 4000, mean_n=20, touching=True)
    # sparse rows: different row sets (row maps; lane-group candidates only)
    sa = random_canonical(rng, 80, 500, mean_n=6, rows=np.sort(rng.choice(200, 80, replace=False)))
    sb = random_canonical(rng, 90, 500, mean_n=6, rows=np.sort(rng.choice(200, 90, replace=False)))
    for x, y in ((a, b), (sa, sb)):
        dx, dy = to_device(x), to_device(y)
        for name in ("union", "intersection", "difference"):
            print("setop candidate", k, name, x.n, y.n, flush=True)
            got = to_host(binary_op(dx, dy, name))
            exp = ox.setop(name, x, y)
            for f in ("rows", "off", "beg", "end"):
                np.testing.assert_array_equal(getattr(got, f), getattr(exp, f), err_msg=f"{name} {f}")
    fams = {key[0]: names for key, names in t.ran.items()}
    assert "setop" in fams
    tile_keys = [names for key, names in t.ran.items() if key[0] == "setop" and key[2]]
    map_keys = [names for key, names in t.ran.items() if key[0] == "setop" and not key[2]]
    assert tile_keys and all(n == [-7, -8, -6, -5, -1, -2, -3, -4, 4, 16, 64] for n in tile_keys)
    assert map_keys and all(n == [4, 16, 64] for n in map_keys)


@pytest.mark.parametrize("k", range(4))
def test_upwind_candidates_vs_oracle(torch_mod, install, k):
    torch = torch_mod
    from subsetix_amd import interval_stencil as st
    from subsetix_amd.interval_field import create_interval_field

    t = install(Forced(k))
    rng = np.random.default_rng(7 + k)
    cases = [(random_canonical(rng, 40, 900, mean_n=9, touching=True), 900, 40),
             (ox.make(list(range(33)), [0] * 32, [1500] * 32), 1500, 32)]
    for s, w, h in cases:
        ncell = int(ox.cell_offsets(s)[-1])
        u = rng.random(ncell).astype(np.float32)
        field = create_interval_field(to_device(s), 0.0, dtype=torch.float32)
        field.values.copy_(torch.from_numpy(u))
        kw = dict(width=w, height=h, a=0.6, b=-0.35, dt=0.01, dx=1.0 / w, dy=1.0 / h)
        print("upwind candidate", k, ncell, flush=True)
        got = st.step_upwind_interval(field, **kw).cpu().numpy()
        np.testing.assert_array_equal(got, ox.upwind(s, u, **kw))
    names = [n for key, n in t.ran.items() if key[0] == "upwind"]
    assert names and all(n == ["seg", 4, 16, 64] for n in names)


@pytest.mark.parametrize("k", range(4))
def test_transfer_candidates_vs_model(torch_mod, install, k):
    import subsetix_amd as sx
    from tests.test_gpu_levels import _prolong_restrict_case

    t = install(Forced(k))
    for dt in ("f32", "f64"):
        _prolong_restrict_case(sx, dt, touching=bool(k % 2))
    names = [n for key, n in t.ran.items() if key[0] == "transfer"]
    assert len(names) >= 4 and all(n == ["seg", 4, 16, 64] for n in names)


def _level_state(eng):
    out = []
    for lv in range(eng.cfg.levels):
        s = to_host(eng.ext[lv])
        out.append((s.rows.copy(), s.off.copy(), s.beg.copy(), s.end.copy(), eng.buf[lv][eng.cur].cpu().numpy()))
    return out


@pytest.mark.parametrize("k", [None, 1, 3])
def test_batched_tuned_equals_untuned(torch_mod, install, k):
    from subsetix_amd import workloads
    from subsetix_amd.batch import BatchConfig, BatchedAMR

    t = install(autotune.Autotuner(reps=1) if k is None else Forced(k))
    t.enabled = False  # the engine enables tuning around its own regrids
    insts = workloads.batched_instances(count=4, seed=21, nmin=16, nmax=64)
    base = dict(levels=3, regrid_every=4, threshold=(0.02, 0.02), grading=1)
    print("batched tuned", k, flush=True)
    tuned = BatchedAMR(insts, config=BatchConfig(autotune=True, **base))
    tuned.run(10)
    assert not autotune.enabled()  # tuning is scoped to the engine's regrids
    t.enabled = False
    autotune._DEFAULT = autotune.Autotuner()  # fixed rules for the reference run
    plain = BatchedAMR(insts, config=BatchConfig(autotune=False, **base))
    plain.run(10)
    for x, y in zip(_level_state(tuned), _level_state(plain)):
        for p, q in zip(x, y):
            np.testing.assert_array_equal(p, q)
    np.testing.assert_array_equal(tuned.masses().cpu().numpy(), plain.masses().cpu().numpy())
    fams = {key[0] for key in t.table}
    assert {"setop", "grad_tag", "upwind", "transfer"} <= fams, fams


def test_composite_tuning_pass_keeps_state(torch_mod, install):
    torch = torch_mod
    import subsetix_amd as sx
    from subsetix_amd.amr2.composite import CompositeTwoLevel

    install(autotune.Autotuner(reps=1))
    rng = np.random.default_rng(3)
    refine = to_device(random_canonical(rng, 48, 64, mean_n=2))
    runs = []
    for tune in (True, False):
        comp = CompositeTwoLevel(64, 64, refine, dtype=torch.float32, velocity=(0.8, -0.4))
        g = torch.Generator(device="cuda").manual_seed(5)
        cv = torch.rand(64 * 64, generator=g, device="cuda")
        fv = torch.rand(comp.fine_field.values.numel(), generator=g, device="cuda")
        comp.set_values(cv, fv)
        comp.prepare(tune=tune)
        for _ in range(5):
            comp.step()
        runs.append((comp.coarse.values.cpu().numpy(), comp.fine_field.values.cpu().numpy()))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert any(key[0] == "upwind" for key in autotune.default().table)
