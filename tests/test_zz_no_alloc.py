"""GPU: mcgpu_simulate_device_noalloc (no context, no allocation; scratch in
a fixed device array) against mcgpu_simulate_device on the same inputs,
through the C ABI: every history's final state, exit code, event and
collision counts and the tally bit-identical, on two calls in a row (the
second with fewer particles and cells, so stale scratch of the first call
would show), plus the argument limits.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run_noalloc(bank, sigs, ar, w, wl, result, events, collisions, options):
    import torch
    from mcgpu import _bank_struct, _ptr, lib
    from mcgpu._lib import Stats
    st = Stats()
    b = _bank_struct(bank)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib().mcgpu_simulate_device_noalloc(
        bank["x"].numel(), C.byref(b), _ptr(sigs), _ptr(ar), _ptr(w), wl.min_index,
        wl.max_index, wl.dx, wl.min_weight, _ptr(result), _ptr(events), _ptr(collisions),
        C.byref(options), C.byref(st), stream)
    torch.cuda.synchronize()
    return rc, st.as_dict()


def _case(ctx, wl, w0):
    """Both entry points on the workload wl, weights starting at w0."""
    import torch
    import mcgpu
    init = mcgpu.bank_to_device(wl.bank)
    sigs = torch.from_numpy(wl.sigs).cuda()
    ar = torch.from_numpy(wl.absorption_rates).cuda()
    n = len(wl.bank)
    out = {}
    for name in ("context", "noalloc"):
        bank = {k: v.clone() for k, v in init.items()}
        w = torch.from_numpy(w0.copy()).cuda()
        res = torch.full((n,), 7, dtype=torch.int8, device="cuda")
        ev = torch.zeros(n, dtype=torch.int64, device="cuda")
        co = torch.zeros(n, dtype=torch.int64, device="cuda")
        opt = mcgpu.make_options("history")
        if name == "context":
            st, rc = ctx.simulate_device(bank, sigs, ar, w, wl.min_index, wl.max_index, wl.dx,
                                         wl.min_weight, result=res, events=ev, collisions=co,
                                         options=opt)
        else:
            rc, st = _run_noalloc(bank, sigs, ar, w, wl, res, ev, co, opt)
        assert rc == 0, (name, rc)
        out[name] = (mcgpu.bank_from_device(bank), w.cpu().numpy(), res.cpu().numpy(),
                     ev.cpu().numpy(), co.cpu().numpy(), st)
    a, b = out["context"], out["noalloc"]
    for f in ("seed", "x", "mu", "wmc", "index"):
        x, y = a[0][f], b[0][f]
        assert np.array_equal(x.view(np.uint64 if x.itemsize == 8 else np.uint32),
                              y.view(np.uint64 if y.itemsize == 8 else np.uint32)), f
    assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64)), "tally"
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
    for k in ("events", "collisions", "n_left", "n_right", "n_killed"):
        assert a[5][k] == b[5][k], k
    assert a[5]["events"] == int(a[3].sum()) > 0
    return a


def test_noalloc_matches_context_entry_point_twice(ctx):
    from mcgpu import configs
    first = configs.slab10k(30_000)          # 10,000 cells: tally in LDS
    rng = np.random.default_rng(11)
    _say("noalloc: 10,000 cells, 30,000 particles")
    _case(ctx, first, rng.random(first.m))
    second = configs.slab100(5_000)         # then fewer cells and particles
    _say("noalloc: 100 cells, 5,000 particles")
    _case(ctx, second, np.zeros(second.m))


def test_noalloc_largest_layer_and_limits(ctx):
    import torch
    import mcgpu
    from mcgpu import configs
    from mcgpu._lib import NOALLOC_MAX_CELLS
    base = configs.slab100(2_000)
    m = NOALLOC_MAX_CELLS                  # tally too large for LDS: global words
    reps = m // base.m + 1
    wl = configs.Workload("max_cells", 0.0, 1.0, 0, m, 1.0 / m, base.min_weight,
                          np.tile(base.sigs, reps)[:m] * 20.0,
                          np.tile(base.absorption_rates, reps)[:m], base.bank, True, True)
    _say(f"noalloc: {m} cells, 2,000 particles")
    _case(ctx, wl, np.zeros(m))

    bank = mcgpu.bank_to_device(base.bank)
    big = torch.zeros(m + 1, dtype=torch.float64, device="cuda")
    wl2 = configs.Workload("too_many_cells", 0.0, 1.0, 0, m + 1, 1.0 / (m + 1),
                           base.min_weight, None, None, base.bank, True, True)
    rc, _ = _run_noalloc(bank, big, big, big, wl2, None, None, None, mcgpu.make_options())
    assert rc == mcgpu.ERR_ARG
    sigs = torch.from_numpy(base.sigs).cuda()
    ar = torch.from_numpy(base.absorption_rates).cuda()
    w = torch.zeros(base.m, dtype=torch.float64, device="cuda")
    rc, _ = _run_noalloc(bank, sigs, ar, w, base, None, None, None, mcgpu.make_options("event"))
    assert rc == mcgpu.ERR_ARG
    assert float(w.abs().sum()) == 0.0      # rejected before anything ran
