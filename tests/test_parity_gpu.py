"""GPU parity: the HIP transport against the CPU oracle, through the C ABI.

Bar: every history's final particle (seed, x, mu, wmc, index), exit code,
event and collision count bit-identical to the FP64 oracle; per-cell tallies
bit-identical to the oracle's exact fixed-point sum of the same terms and
within 1e-12 relative of the oracle's sequential FP64 sum (the reference's
summation, src/layer.cpp:179,323-329).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEQ_RTOL = 1e-12  # tallies vs the reference-order FP64 sum (summation order only)

KERNELS = ["history", "event"]


def _say(msg):
    print(msg, flush=True)


def _workload(name, n):
    from mcgpu import configs
    return {"slab100": configs.slab100, "slab10k": configs.slab10k, "layer_rank2of5": configs.layer_rank2of5,
            "adjoint1000": configs.adjoint1000}[name](n)


def _oracle_run(oracle, wl, bank=None, weights=None):
    return oracle.simulate_bank(64, wl.index_start, wl.m, wl.dx, wl.min_weight, wl.sigs,
                                wl.absorption_rates, wl.bank if bank is None else bank,
                                nthread=8, weights=weights)


def _assert_bank_equal(got, want):
    for f in ("seed", "x", "mu", "wmc", "index"):
        a, b = np.ascontiguousarray(got[f]), np.ascontiguousarray(want[f])
        eq = (a.view(np.uint64 if a.itemsize == 8 else np.uint32) ==
              b.view(np.uint64 if b.itemsize == 8 else np.uint32))
        if not eq.all():
            i = int(np.flatnonzero(~eq)[0])
            raise AssertionError(f"field {f} differs at particle {i}: {a[i]!r} vs {b[i]!r} "
                                 f"({(~eq).sum()} particles)")


def _assert_tally(w_gpu, o, w0=None):
    exact = o["tally_exact"] if w0 is None else w0 + o["tally_exact"]
    assert np.array_equal(w_gpu.view(np.uint64), exact.view(np.uint64))
    seq = o["weights"]
    denom = np.maximum(np.abs(seq), 1e-300)
    assert np.max(np.abs(w_gpu - seq) / denom) <= SEQ_RTOL


# ------------------------------------------------------------- math / RNG

def test_device_log_exp_correctly_rounded(ctx, oracle):
    import torch
    rng = np.random.default_rng(1)
    s = rng.integers(1, 2 ** 63, 1_000_000, dtype=np.uint64)
    s >>= rng.integers(0, 60, len(s)).astype(np.uint64)
    h = s.astype(np.float64) * 2.0 ** -63
    k = np.arange(1, 50_001, dtype=np.float64)
    h = np.concatenate([h, 1.0 - k * 2.0 ** -53, [1.0, 2.0 ** -63, 5e-324, 0.0]])
    x = np.concatenate([-rng.random(1_000_000) * 10.0 ** rng.uniform(-18, 2.9, 1_000_000),
                        k * 2.0 ** -53, -k * 2.0 ** -53, -708.0 - rng.random(20_000) * 38,
                        [0.0, -np.inf, np.nan, -745.14]])
    for which, v in (("log", h), ("exp", x)):
        dv = torch.from_numpy(v).cuda()
        out = torch.empty_like(dv)
        ctx.math_eval(which, dv, out)
        got = out.cpu().numpy()
        want = oracle.math_eval(which, v)
        same = (got.view(np.uint64) == want.view(np.uint64)) | (np.isnan(got) & np.isnan(want))
        assert same.all(), (which, v[~same][:5], got[~same][:5], want[~same][:5])


def test_device_rng_known_answers(ctx, oracle):
    """src/test_curandom.cu recipe: the rnd_seed chain from 30061994, then one
    rnd_real step per element, bit-identical to the host (FP64 here)."""
    import torch
    n = 1_000_000
    chain = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx.rng_seed_chain(30061994, chain)
    want_chain = oracle.rnd_seed_chain(30061994, n)
    assert np.array_equal(chain.cpu().numpy().view(np.uint64), want_chain)
    reals = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.rng_real_step(chain, reals)
    s_after, r = oracle.rnd_real_step(want_chain, 64)
    assert np.array_equal(chain.cpu().numpy().view(np.uint64), s_after)
    assert np.array_equal(reals.cpu().numpy().view(np.uint64), r.view(np.uint64))


def test_device_uniform_generator_and_sqrt():
    """The benchmark builds source directions on the device with
    mcgpu.configs.uniform_torch and torch.sqrt: both must equal the host
    generator bit for bit."""
    import torch
    from mcgpu import configs
    for seed, offset in ((2003, 0), (4004, 300_000_000)):
        u = configs.uniform_torch(seed, 500_000, offset)
        want = configs.uniform(seed, 500_000, offset)
        assert np.array_equal(u.cpu().numpy().view(np.uint64), want.view(np.uint64))
        assert np.array_equal(torch.sqrt(u).cpu().numpy().view(np.uint64),
                              np.sqrt(want).view(np.uint64))


def test_device_create_particles(ctx, oracle):
    """mcgpu_create_particles against the oracle's Layer: decompose_domain with
    the reference's config.yaml source (x_ini = 0.7071..., 1000 cells, seed
    root 5127801, weight 1/N) creates its first MAX_PARTICLES_VECTOR = 1e6
    particles eagerly (src/layer.cpp:71-82); a later lazy chunk (particles
    1e6 .. 1.3e6) against the oracle's seed chain + one rnd_real step."""
    import torch
    import mcgpu
    x_ini, n_total = 0.7071067690849304, 2_500_000
    L = oracle.OracleLayer.decompose_domain(64, 0.0, 1.0, x_ini, 1, 0, 1000, n_total, 1e-12)
    want = L.particles.copy()
    assert len(want) == 1_000_000

    def dev_bank(n):
        return {"seed": torch.empty(n, dtype=torch.int64, device="cuda"),
                "x": torch.empty(n, dtype=torch.float64, device="cuda"),
                "mu": torch.empty(n, dtype=torch.float64, device="cuda"),
                "wmc": torch.empty(n, dtype=torch.float64, device="cuda"),
                "index": torch.empty(n, dtype=torch.int32, device="cuda")}
    b = dev_bank(len(want))
    ctx.create_particles(5127801, 0, b, x_ini, 1.0 / n_total, L.dx)
    _assert_bank_equal(mcgpu.bank_from_device(b), want)
    first, n = 1_000_000, 300_000
    b = dev_bank(n)
    ctx.create_particles(5127801, first, b, x_ini, 1.0 / n_total, L.dx)
    got = mcgpu.bank_from_device(b)
    seeds, r = oracle.rnd_real_step(oracle.rnd_seed_chain(5127801, first + n)[first:])
    assert np.array_equal(got["seed"], seeds)
    assert np.array_equal(got["mu"].view(np.uint64), (2 * r - 1).view(np.uint64))
    assert np.all(got["index"] == want["index"][0]) and np.all(got["x"] == x_ini)


# ------------------------------------------------------------- transport

# the 100-cell slab at its full 1e5 particles; the others at prefixes of
# their banks (full sizes below, through size-independent properties)
CASES = [("slab100", 100_000), ("layer_rank2of5", 20_000), ("adjoint1000", 3_000), ("slab10k", 400)]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name,n", CASES)
def test_host_api_bit_exact(ctx, oracle, name, n, kernel):
    import mcgpu
    wl = _workload(name, n)
    _say(f"case {name} n={n} kernel={kernel} host api")
    o = _oracle_run(oracle, wl)
    bank = wl.bank.copy()
    w = np.zeros(wl.m)
    result, st, rc = ctx.simulate(bank, wl.sigs, wl.absorption_rates, w, wl.min_index,
                                  wl.max_index, wl.dx, wl.min_weight,
                                  options=mcgpu.make_options(kernel))
    _assert_bank_equal(bank, o["bank"])
    assert np.array_equal(result, o["result"])
    assert st["events"] == int(o["events"].sum())
    assert st["collisions"] == int(o["collisions"].sum())
    assert st["n_left"] == int((o["result"] == -1).sum())
    assert st["n_right"] == int((o["result"] == 1).sum())
    assert st["n_killed"] == int((o["result"] == 0).sum())
    _assert_tally(w, o)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name,n", CASES)
def test_device_api_bit_exact(ctx, oracle, name, n, kernel):
    import torch
    import mcgpu
    wl = _workload(name, n)
    _say(f"case {name} n={n} kernel={kernel} device api")
    o = _oracle_run(oracle, wl)
    bank = mcgpu.bank_to_device(wl.bank)
    w = torch.zeros(wl.m, dtype=torch.float64, device="cuda")
    res = torch.empty(n, dtype=torch.int8, device="cuda")
    ev = torch.empty(n, dtype=torch.int64, device="cuda")
    co = torch.empty(n, dtype=torch.int64, device="cuda")
    sigs = torch.from_numpy(wl.sigs).cuda()
    ar = torch.from_numpy(wl.absorption_rates).cuda()
    ctx.simulate_device(bank, sigs, ar, w, wl.min_index, wl.max_index, wl.dx, wl.min_weight,
                        result=res, events=ev, collisions=co, options=mcgpu.make_options(kernel))
    _assert_bank_equal(mcgpu.bank_from_device(bank), o["bank"])
    assert np.array_equal(res.cpu().numpy().astype(np.int32), o["result"])
    assert np.array_equal(ev.cpu().numpy().view(np.uint64), o["events"])
    assert np.array_equal(co.cpu().numpy().view(np.uint64), o["collisions"])
    _assert_tally(w.cpu().numpy(), o)


LAUNCH_OPTS = [
    dict(kernel="history", block_size=64), dict(kernel="history", block_size=1024),
    dict(kernel="history", events_per_refill=1), dict(kernel="history", events_per_refill=64),
    dict(kernel="history", grid_blocks=1), dict(kernel="history", grid_blocks=3, block_size=128),
    dict(kernel="event", block_size=64), dict(kernel="event", block_size=1024),
    dict(kernel="event", particles_per_lane=1), dict(kernel="event", particles_per_lane=8),
    dict(kernel="event", compaction_interval=1), dict(kernel="event", compaction_interval=3),
    dict(kernel="event", compaction_interval=16), dict(kernel="event", compaction_interval=64),
    dict(kernel="event", grid_blocks=1), dict(kernel="event", grid_blocks=5, block_size=128,
                                              particles_per_lane=2, compaction_interval=2)]


@pytest.mark.parametrize("opts", LAUNCH_OPTS)
def test_launch_shape_does_not_change_results(ctx, oracle, opts):
    import mcgpu
    wl = _workload("slab100", 5_000)
    _say(f"launch options {opts}")
    o = _oracle_run(oracle, wl)
    bank = wl.bank.copy()
    w = np.zeros(wl.m)
    result, st, rc = ctx.simulate(bank, wl.sigs, wl.absorption_rates, w, wl.min_index,
                                  wl.max_index, wl.dx, wl.min_weight,
                                  options=mcgpu.make_options(**opts))
    _assert_bank_equal(bank, o["bank"])
    assert np.array_equal(result, o["result"])
    _assert_tally(w, o)


@pytest.mark.parametrize("block", [1088, 2048, 96 + 1])
def test_event_block_size_limit(ctx, block):
    """The event kernel takes at most 1024 threads per block, in multiples of
    64; a larger or ragged block is an argument error, with nothing run."""
    import mcgpu
    wl = _workload("slab100", 1_000)
    bank = wl.bank.copy()
    w = np.zeros(wl.m)
    _, _, rc = ctx.simulate(bank, wl.sigs, wl.absorption_rates, w, wl.min_index, wl.max_index,
                            wl.dx, wl.min_weight,
                            options=mcgpu.make_options("event", block_size=block),
                            allow=(mcgpu.ERR_ARG,))
    assert rc == mcgpu.ERR_ARG
    assert "1024" in ctx.error()
    assert np.array_equal(bank, wl.bank)
    assert not w.any()


def _layer_case(oracle, m, n, seed=0, start=0, x_mode="face"):
    rng = np.random.default_rng(seed)
    from mcgpu import configs
    sigs = rng.lognormal(0.3, 0.6, m)
    ar = rng.uniform(0.05, 0.95, m)
    dx = 1.0 / m
    bank = np.zeros(n, dtype=oracle.P64)
    bank["seed"] = configs.seed_chain(n)
    bank["x"] = start * dx
    bank["index"] = start
    bank["mu"] = np.sqrt(rng.random(n))
    bank["wmc"] = 1.0
    return sigs, ar, dx, bank


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("m", [12_000, 10_230, 10_100, 5_000, 64])
def test_tally_paths_lds_and_global(ctx, oracle, m, kernel):
    """m = 12000 exceeds one block's LDS (global 128-bit tally), 5000 keeps
    the tally in LDS with the cross sections in L2, 64 keeps both in LDS.
    10,100 and 10,230 sit either side of the event kernel's limit, a 160 KiB
    block less its own 2 KiB (163,648 and 163,680 bytes of tally + 2,048)."""
    _say(f"tally path m={m}")
    sigs, ar, dx, bank = _layer_case(oracle, m, 300 if m > 1000 else 3000, seed=m)
    o = oracle.simulate_bank(64, 0, m, dx, 1e-3, sigs, ar, bank, nthread=8)
    b = bank.copy()
    w = np.zeros(m)
    import mcgpu
    result, st, rc = ctx.simulate(b, sigs, ar, w, 0, m, dx, 1e-3,
                                  options=mcgpu.make_options(kernel))
    _assert_bank_equal(b, o["bank"])
    assert np.array_equal(result, o["result"])
    _assert_tally(w, o)


@pytest.mark.parametrize("kernel", KERNELS)
def test_edge_particles(ctx, oracle, kernel):
    """Particles the loop never steps (outside the layer, below the minimum
    weight, NaN weight) get the reference's exit codes; mu = 0 goes right;
    interior layer with nonzero index_start; weights accumulate (+=)."""
    m, start = 50, 1000
    rng = np.random.default_rng(3)
    sigs = rng.lognormal(0.0, 0.5, m)
    ar = rng.uniform(0.1, 0.9, m)
    dx = 0.2 / m
    from mcgpu import configs
    n = 12
    bank = np.zeros(n, dtype=oracle.P64)
    bank["seed"] = configs.seed_chain(n)
    bank["wmc"] = 1.0
    bank["index"] = [start - 1, start + m, start + 3, start + 7, start, start + m - 1,
                     start + 10, start + 20, start + 25, start + 30, start + 40, start + 49]
    bank["x"] = (bank["index"] + 0.5) * dx
    bank["mu"] = [0.5, -0.5, 0.0, 1e-5, 1.0, -1.0, 0.3, -0.9, 1e-4, -1e-4, 0.7, 0.2]
    bank["wmc"][3] = np.nan
    bank["wmc"][6] = 1e-4  # below min weight at entry
    w0 = rng.random(m)
    o = oracle.simulate_bank(64, start, m, dx, 1e-3, sigs, ar, bank, nthread=1, weights=w0.copy())
    b = bank.copy()
    w = w0.copy()
    import mcgpu
    result, st, rc = ctx.simulate(b, sigs, ar, w, start, start + m, dx, 1e-3,
                                  options=mcgpu.make_options(kernel))
    _assert_bank_equal(b, o["bank"])
    assert np.array_equal(result, o["result"])
    assert result[0] == -1 and result[1] == 1 and result[6] == 0 and result[3] == -1
    assert np.array_equal(w.view(np.uint64), (w0 + o["tally_exact"]).view(np.uint64))


def test_empty_bank(ctx):
    import torch
    import mcgpu
    m = 10
    w = np.full(m, 0.5)
    b = np.zeros(0, dtype=mcgpu.PARTICLE)
    result, st, rc = ctx.simulate(b, np.ones(m), np.full(m, 0.5), w, 0, m, 0.1, 1e-3)
    assert rc == 0 and st["events"] == 0 and np.all(w == 0.5)
    bank = mcgpu.bank_to_device(b)
    wd = torch.full((m,), 0.5, dtype=torch.float64, device="cuda")
    ones = torch.ones(m, dtype=torch.float64, device="cuda")
    st, rc = ctx.simulate_device(bank, ones, ones * 0.5, wd, 0, m, 0.1, 1e-3)
    assert rc == 0 and torch.all(wd == 0.5)


@pytest.mark.parametrize("kernel", KERNELS)
def test_event_cap_reports_error(ctx, oracle, kernel):
    import mcgpu
    wl = _workload("slab100", 500)
    b = wl.bank.copy()
    w = np.zeros(wl.m)
    result, st, rc = ctx.simulate(b, wl.sigs, wl.absorption_rates, w, wl.min_index, wl.max_index,
                                  wl.dx, wl.min_weight,
                                  options=mcgpu.make_options(kernel, max_events_per_history=5),
                                  allow=(mcgpu.ERR_EVENT_CAP,))
    assert rc == mcgpu.ERR_EVENT_CAP
    assert np.any(result == 2)


@pytest.mark.parametrize("cap,ci,refill", [(5, 0, 0), (37, 0, 0), (37, 1, 0), (200, 16, 0),
                                           (250, 0, 64)])
def test_event_cap_same_in_both_kernels(ctx, oracle, cap, ci, refill):
    """A binding cap on events per history: both kernels stop each longer
    history after `cap` events, with exit code 2, the same state and the
    same tally, bit for bit; every history that ends sooner is the
    oracle's.  The event kernel reads a slot's count origin only while the
    cap can bind, and moves it at compactions (ci: every pass, every 16th,
    the default 64th).  The history kernel drops the per-event cap test
    while its wave's counts are an interval below the cap: with 64 events
    between refills and a cap of 250 it runs three intervals that way, then
    tests every event."""
    import torch
    import mcgpu
    wl = _workload("slab100", 3_000)
    _say(f"event cap {cap}, compaction interval {ci}")
    o = _oracle_run(oracle, wl)
    n = len(wl.bank)
    got = {}
    for kernel in KERNELS:
        bank = mcgpu.bank_to_device(wl.bank)
        w = torch.zeros(wl.m, dtype=torch.float64, device="cuda")
        res = torch.empty(n, dtype=torch.int8, device="cuda")
        ev = torch.empty(n, dtype=torch.int64, device="cuda")
        co = torch.empty(n, dtype=torch.int64, device="cuda")
        opts = mcgpu.make_options(kernel, max_events_per_history=cap,
                                  compaction_interval=ci if kernel == "event" else 0,
                                  events_per_refill=refill if kernel == "history" else 0)
        _, rc = ctx.simulate_device(bank, torch.from_numpy(wl.sigs).cuda(),
                                    torch.from_numpy(wl.absorption_rates).cuda(), w,
                                    wl.min_index, wl.max_index, wl.dx, wl.min_weight,
                                    result=res, events=ev, collisions=co, options=opts,
                                    allow=(mcgpu.ERR_EVENT_CAP,))
        got[kernel] = (mcgpu.bank_from_device(bank), res.cpu().numpy().astype(np.int32),
                       ev.cpu().numpy().view(np.uint64), co.cpu().numpy().view(np.uint64),
                       w.cpu().numpy(), rc)
    hb, hr, he, hc, hw, hrc = got["history"]
    eb, er, ee, ec, ew, erc = got["event"]
    _assert_bank_equal(eb, hb)
    assert np.array_equal(er, hr) and np.array_equal(ee, he) and np.array_equal(ec, hc)
    assert np.array_equal(ew.view(np.uint64), hw.view(np.uint64))
    longer = o["events"] > cap
    assert longer.any() and hrc == erc == mcgpu.ERR_EVENT_CAP
    assert np.all(he[longer] == cap) and np.all(hr[longer] == 2)
    short = ~longer
    assert np.array_equal(hr[short], o["result"][short])
    assert np.array_equal(he[short], o["events"][short])
    assert np.array_equal(hc[short], o["collisions"][short])
    _assert_bank_equal(hb[short], o["bank"][short])


def test_out_of_range_tally_term(ctx, oracle):
    """A term of 2^39 or more cannot be summed exactly: the cell's tally
    becomes NaN (as the oracle's) and the call reports MCGPU_ERR_RANGE."""
    import mcgpu
    m = 4
    sigs, ar, dx = np.full(m, 50.0), np.full(m, 0.9), 0.25
    bank = np.zeros(3, dtype=mcgpu.PARTICLE)
    bank["seed"] = [11, 22, 33]
    bank["mu"] = 1.0
    bank["wmc"] = [1.0, 1e13, 1.0]
    o = oracle.simulate_bank(64, 0, m, dx, 1e-3, sigs, ar, bank, nthread=1)
    b = bank.copy()
    w = np.zeros(m)
    result, st, rc = ctx.simulate(b, sigs, ar, w, 0, m, dx, 1e-3, allow=(mcgpu.ERR_RANGE,))
    assert rc == mcgpu.ERR_RANGE
    assert np.isnan(w[0]) and np.isnan(o["tally_exact"][0])
    _assert_bank_equal(b, o["bank"])


def test_repeated_calls_accumulate(ctx, oracle):
    """Two calls on two halves equal one call on the whole (per history), and
    the tally adds the two exact partial sums."""
    wl = _workload("adjoint1000", 2000)
    o1 = _oracle_run(oracle, wl, bank=wl.bank[:1000])
    o2 = _oracle_run(oracle, wl, bank=wl.bank[1000:])
    w = np.zeros(wl.m)
    b1, b2 = wl.bank[:1000].copy(), wl.bank[1000:].copy()
    ctx.simulate(b1, wl.sigs, wl.absorption_rates, w, wl.min_index, wl.max_index, wl.dx,
                 wl.min_weight)
    ctx.simulate(b2, wl.sigs, wl.absorption_rates, w, wl.min_index, wl.max_index, wl.dx,
                 wl.min_weight)
    _assert_bank_equal(b1, o1["bank"])
    _assert_bank_equal(b2, o2["bank"])
    assert np.array_equal(w, o1["tally_exact"] + o2["tally_exact"])


# ---------------------------------------------------- the reference's GPU recipe

_RECIPE = {}


def _reference_recipe(oracle):
    """The recipe's layer, bank and oracle run (computed once for both kernels)."""
    if not _RECIPE:
        import mcgpu
        x_ini = np.sqrt(2.0) / 2.0
        L = oracle.OracleLayer.decompose_domain(64, 0.0, 1.0, x_ini, 1, 0, 1000, 1_000_000, 0.0)
        # copies: the layer's arrays are views of memory freed with the layer
        bank0 = np.array(L.particles, dtype=mcgpu.PARTICLE)
        sigs, ar = np.array(L.sigs), np.array(L.absorption_rates)
        o = oracle.simulate_bank(64, 0, 1000, L.dx, 0.0, sigs, ar, bank0, nthread=16)
        _RECIPE.update(bank0=bank0, sigs=sigs, ar=ar, dx=L.dx, o=o)
    return _RECIPE


@pytest.mark.parametrize("kernel", KERNELS)
def test_reference_gpu_recipe_at_its_size(ctx, oracle, kernel):
    """The reference's own CPU-against-GPU recipe (src/test_culayer.cu:27-87):
    decompose_domain(0, 1, sqrt(2)/2, 1 rank, 1000 cells, 1e6 particles,
    minimum weight 0) -- the created source (seed chain from 5127801,
    isotropic mu, weight 1e-6) -- transported to the end.  The reference
    compares the tallies to 1e-4; here every history's final state, exit
    code, event and collision counts and the exact tally against the FP64
    oracle, bit for bit."""
    import torch
    import mcgpu
    _say(f"reference GPU recipe: 1e6 particles, kernel {kernel}")
    rc = _reference_recipe(oracle)
    bank0, sigs, ar, dx, o = rc["bank0"], rc["sigs"], rc["ar"], rc["dx"], rc["o"]
    n = len(bank0)
    assert n == 1_000_000 and np.all(bank0["wmc"] == 1e-6)
    bank = mcgpu.bank_to_device(bank0)
    w = torch.zeros(1000, dtype=torch.float64, device="cuda")
    res = torch.empty(n, dtype=torch.int8, device="cuda")
    ev = torch.empty(n, dtype=torch.int64, device="cuda")
    co = torch.empty(n, dtype=torch.int64, device="cuda")
    st, _ = ctx.simulate_device(bank, torch.from_numpy(sigs).cuda(), torch.from_numpy(ar).cuda(),
                                w, 0, 1000, dx, 0.0, result=res, events=ev, collisions=co,
                                options=mcgpu.make_options(kernel))
    _say(f"  events={st['events']} kernel_ms={st['kernel_ms']:.1f}")
    _assert_bank_equal(mcgpu.bank_from_device(bank), o["bank"])
    assert np.array_equal(res.cpu().numpy().astype(np.int32), o["result"])
    assert np.array_equal(ev.cpu().numpy().view(np.uint64), o["events"])
    assert np.array_equal(co.cpu().numpy().view(np.uint64), o["collisions"])
    assert st["n_killed"] == 0          # minimum weight 0: every history escapes
    _assert_tally(w.cpu().numpy(), o)


# ---------------------------------------------------- full-size properties

@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("name,n", [("layer_rank2of5", 20_000_000), ("slab10k", 1_000_000)])
def test_full_size_properties_and_sampled_histories(ctx, oracle, name, n, kernel):
    """At full bank sizes (2e7 inflow particles on the rank-2 layer of the
    5,000-cell slab; 1e6 on the 10,000-cell slab): counts and weight balance
    close exactly, every exit state is consistent, and a random sample of
    histories, re-run alone on the oracle from their initial states, matches
    bit for bit (histories are independent, so the sample is a valid
    per-history check at scale)."""
    import torch
    import mcgpu
    wl = _workload(name, n)
    _say(f"full size {name} n={n} kernel={kernel}")
    bank = mcgpu.bank_to_device(wl.bank)
    w = torch.zeros(wl.m, dtype=torch.float64, device="cuda")
    res = torch.empty(n, dtype=torch.int8, device="cuda")
    sigs = torch.from_numpy(wl.sigs).cuda()
    ar = torch.from_numpy(wl.absorption_rates).cuda()
    st, rc = ctx.simulate_device(bank, sigs, ar, w, wl.min_index, wl.max_index, wl.dx,
                                 wl.min_weight, result=res, options=mcgpu.make_options(kernel))
    _say(f"  events={st['events']} kernel_ms={st['kernel_ms']:.1f}")
    r = res.cpu().numpy()
    fin = mcgpu.bank_from_device(bank)
    assert st["n_left"] + st["n_right"] + st["n_killed"] == n
    assert st["n_left"] == (r == -1).sum() and st["n_right"] == (r == 1).sum()
    assert np.all(fin["index"][r == -1] == wl.min_index - 1)
    assert np.all(fin["index"][r == 1] == wl.max_index)
    assert np.all(fin["wmc"][r == 0] < wl.min_weight)
    total_in = wl.bank["wmc"].sum()
    balance = w.sum().item() + fin["wmc"].sum()
    assert abs(balance - total_in) <= 1e-9 * total_in
    idx = np.sort(np.random.default_rng(0).choice(n, 2000, replace=False))
    o = _oracle_run(oracle, wl, bank=wl.bank[idx])
    _assert_bank_equal(fin[idx], o["bank"])
    assert np.array_equal(r[idx].astype(np.int32), o["result"])


@pytest.mark.parametrize("kernel", KERNELS)
def test_crossing_decision_and_edge_cases(ctx, oracle, kernel):
    """The step decides most crossings from a plain-double log with a margin
    and divides by a stored reciprocal of mu; both must reproduce the
    reference's branch and quotients exactly.  Cells without interaction
    (sigma_i <= EPS), collision-dominated cells (mean free path ~ dx/50),
    tiny cross sections, directions at and below EPS, particles exactly on
    the edge they move to (zero numerator) and just beyond it (negative
    edge distance)."""
    _say(f"crossing decision edge cases, {kernel}")
    m = 240
    rng = np.random.default_rng(11)
    dx = 1.0 / m
    kind = np.arange(m) % 4
    sigs = np.where(kind == 0, rng.lognormal(0.0, 0.5, m),
                    np.where(kind == 1, 50.0 / dx * rng.uniform(0.5, 1.5, m),
                             np.where(kind == 2, 1e-6, rng.lognormal(1.0, 1.0, m))))
    ar = np.where(kind == 0, 1.0, rng.uniform(0.05, 0.95, m))   # kind 0: sigma_i = 0
    from mcgpu import configs
    n = 6000
    bank = np.zeros(n, dtype=oracle.P64)
    bank["seed"] = configs.seed_chain(n)
    bank["wmc"] = 1.0
    idx = rng.integers(0, m, n)
    bank["index"] = idx
    frac = rng.random(n)
    sel = rng.integers(0, 6, n)
    mu = rng.uniform(-1.0, 1.0, n)
    mu = np.where(sel == 1, rng.choice([1e-4, -1e-4, 1e-5, 0.0, 2e-4, -2e-4], n), mu)
    x = (idx + frac) * dx
    x = np.where(sel == 2, idx * dx, x)                       # on the left edge
    x = np.where(sel == 3, (idx + 1) * dx, x)                 # on the right edge
    x = np.where(sel == 4, np.nextafter((idx + 1) * dx, 2.0), x)  # just beyond it
    x = np.where(sel == 5, np.nextafter(idx * dx, -1.0), x)   # just before the left edge
    mu = np.where(sel == 2, -np.abs(mu) - 1e-3, mu)           # moving to the left edge
    mu = np.where(sel == 3, np.abs(mu) + 1e-3, mu)            # moving to the right edge
    bank["x"] = x
    bank["mu"] = mu
    o = oracle.simulate_bank(64, 0, m, dx, 1e-3, sigs, ar, bank, nthread=8)
    b = bank.copy()
    w = np.zeros(m)
    import mcgpu
    result, st, rc = ctx.simulate(b, sigs, ar, w, 0, m, dx, 1e-3,
                                  options=mcgpu.make_options(kernel))
    _assert_bank_equal(b, o["bank"])
    assert np.array_equal(result, o["result"])
    _assert_tally(w, o)
    assert st["collisions"] > 0 and st["events"] > st["collisions"]
