"""GPU: stacked two-level composites (the serving cohorts) against each
instance run alone with CompositeTwoLevel: coarse and extended fine fields
and FP64 mass bit-exact after each instance's own step count (instances
retired once copied out), eager steps and captured-graph replays, float32
and float64."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _alone(spec, dtype, steps):
    import torch

    from subsetix_amd.amr2.composite import CompositeTwoLevel
    from subsetix_amd.multilevel import restrict_set
    from subsetix_amd.serve import _device_set

    dev = torch.device("cuda", 0)
    nc = spec["n_fine"] // 2
    comp = CompositeTwoLevel(nc, nc, restrict_set(_device_set(spec["discs"], dev), 2), dtype=dtype,
                             velocity=spec["velocity"])
    gen = torch.Generator(device=dev).manual_seed(spec["seed"])
    cv = torch.rand(nc * nc, generator=gen, device=dev, dtype=torch.float32)
    fv = torch.rand(comp.fine_field.values.numel(), generator=gen, device=dev, dtype=torch.float32)
    comp.set_values(cv, fv)
    for _ in range(steps):
        comp.step()
    return comp.coarse.values.cpu().numpy(), comp.fine_field.values.cpu().numpy(), comp.mass()


def _specs():
    from subsetix_amd import workloads
    from subsetix_amd.serve import _host_set

    out = []
    for k, (n, vel) in enumerate(((64, (1.0, 0.5)), (98, (-0.7, 0.3)), (40, (0.2, -1.0)), (128, (0.0, 0.0)))):
        discs, _, _ = workloads.stencil_discs_config(n=n, seed=30 + k)
        out.append({"discs": _host_set(discs), "n_fine": n, "velocity": vel, "seed": 100 + k})
    return out


@pytest.mark.parametrize("dt", ["f32", "f64"])
@pytest.mark.parametrize("graphs", [False, True])
def test_stacked_equals_alone(dt, graphs):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from subsetix_amd.amr2.stacked import StackedComposite

    dtype = torch.float32 if dt == "f32" else torch.float64
    dev = torch.device("cuda", 0)
    specs = _specs()
    steps = [5, 2, 7, 5]
    st = StackedComposite(specs, dtype=dtype, device=dev)
    for j, sp in enumerate(specs):
        gen = torch.Generator(device=dev).manual_seed(sp["seed"])
        nc = sp["n_fine"] // 2
        cv = torch.rand(nc * nc, generator=gen, device=dev, dtype=torch.float32)
        fv = torch.rand(st.ext_cells(j), generator=gen, device=dev, dtype=torch.float32)
        st.set_instance_values(j, cv, fv)
    st.finish_init()
    stream = torch.cuda.Stream(dev) if graphs else None
    snaps, done = {}, 0
    for target in sorted(set(steps)):
        st.advance(target - done, stream, use_graph=graphs)
        done = target
        for j, k in enumerate(steps):
            if k == target:
                with torch.cuda.stream(stream) if stream is not None else torch.cuda.stream(
                        torch.cuda.current_stream(dev)):
                    snaps[j] = st.snapshot(j)
                    st.retire(j)  # later steps skip it; the others must not change
    torch.cuda.synchronize()
    for j, sp in enumerate(specs):
        print("stacked instance", j, dt, "graphs", graphs, flush=True)
        c, f, m = _alone(sp, dtype, steps[j])
        np.testing.assert_array_equal(snaps[j]["coarse"].cpu().numpy(), c)
        np.testing.assert_array_equal(snaps[j]["fine"].cpu().numpy(), f)
        assert float(snaps[j]["mass"].item()) == m
        assert st.cells_per_step(j) == c.size + f.size
