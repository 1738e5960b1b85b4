"""Stream ordering and failure cleanup of the C ABI on the GPU.

* A library call made right after torch queued work on its default stream
  that writes the call's input (no host synchronisation in between) sees that
  input complete: the context's own stream is ordered with the null stream.
* A search plan created on one stream and finalized after the context moved
  to another stream returns the same maps (the plan follows the context's
  stream and orders its queued work before the new stream's).
* A plan whose mask has no positive voxel is refused with -1 and leaves no
  device memory behind.
Tolerances: bit-exact (the same kernels run in the same order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_call_after_default_stream_work_sees_its_input():
    import torch

    from tom_amd import api
    say("stream order: rotate after default-stream producer")
    dev = torch.device("cuda:0")
    n = 64
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    src = torch.randn((n, n, n), generator=g, device=dev)
    euler = np.array([30.0, 45.0, 60.0], np.float32)
    torch.cuda.synchronize()
    ref = api.tom_rotate(src, euler).clone()
    torch.cuda.synchronize()
    for trial in range(3):
        # a long chain of default-stream kernels, then the producer of `x`
        a = torch.randn((4096, 4096), generator=g, device=dev)
        for _ in range(12):
            a = torch.tanh(a @ a * (1.0 / 64.0))
        x = torch.empty_like(src)
        x.copy_(src + 0.0 * a[0, 0])  # depends on the whole chain
        out = api.tom_rotate(x, euler)  # no torch.cuda.synchronize() before it
        torch.cuda.synchronize()
        assert torch.equal(out, ref), trial


def test_plan_follows_a_stream_change():
    import ctypes

    import torch

    from tom_amd import api
    say("stream order: plan created on one stream, finalized on another")
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(3)
    N, n = (64, 64, 64), 16
    V = torch.from_numpy(rng.standard_normal(N[::-1]).astype(np.float32)).to(dev)
    T = rng.standard_normal((n, n, n)).astype(np.float32)
    M = api.tom_spheremask_gauss((n, n, n), 6.0, 0.0)
    tn, mu, sg = api.prepare_template(T, M, None)
    angles = np.array([[0, 0, 0], [30, 60, 90], [90, 30, 45], [180, 0, 120], [10, 20, 30]], np.float32)
    ctx = api.Context(0)
    try:
        with api.SearchPlan(V, (n, n, n), M, batch=2, ctx=ctx) as p:
            p.set_template(tn, None, mu, sg)
            p.run(angles, 0)
            s0, a0 = p.finalize()
            torch.cuda.synchronize()
            s0, a0 = s0.clone(), a0.clone()
            p.set_template(tn, None, mu, sg)
            p.run(angles, 0)  # queued on the context's own stream
            side = torch.cuda.Stream(device=dev)
            ctx.check(ctx.lib.tomhip_set_stream(ctx.h, ctypes.c_void_p(side.cuda_stream)))
            s1, a1 = p.finalize()
            torch.cuda.synchronize()
            assert torch.equal(s0, s1) and torch.equal(a0, a1)
            ctx.check(ctx.lib.tomhip_set_stream(ctx.h, None))
    finally:
        ctx.close()


def test_zero_mask_refused_without_leak():
    import torch

    from tom_amd import api
    say("zero mask: refused, nothing left allocated")
    dev = torch.device("cuda:0")
    V = torch.zeros((256, 512, 512), dtype=torch.float32, device=dev)
    Mz = np.zeros((48, 48, 48), np.float32)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(4):
        with pytest.raises(api.TomHipError) as e:
            api.SearchPlan(V, (48, 48, 48), Mz, batch=16)
        assert e.value.code == -1
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    # one plan of this size holds more than 6 GB; four leaked ones would show
    assert free0 - free1 < (1 << 30), (free0, free1)
