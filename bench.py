#!/usr/bin/env python3
"""Benchmark: particle-events/s of the MI355X transport on one GPU.

Headline (`value`): one slab of 10,000 cells on [0,1] with
per-cell sigma_a ~ LogNormal(0, 0.5), sigma_s ~ LogNormal(0.5, 0.5), 1e8
particles entering at x = 0 with mu = sqrt(U), weight 1, minimum weight 1e-3,
FP64.  One step = one mcgpu_simulate_device call transporting the whole bank
(resident in HBM before the timed region) to termination, including the
restore of the bank from a pristine device copy.

Also reported, in "variants":
  event     -- the event-based kernel on the same bank (one step);
  host_api  -- mcgpu_simulate with a host bank (copies over PCIe included) on
               the first 1e7 particles of the 10,000-cell slab;
  cycle     -- the reference's headline job (its scripts' config.yaml: 1000
               cells, 1e8 particles, 1e5 per cycle, sync) with 1 and 8 ranks,
               every rank on this GPU (mcgpu_run --ranks R, device-resident
               cycle loop), next to the reference's published wall times;
  batched   -- 4096 independent slabs (5.8e8 particles) in one
               persistent launch replayed from a hipGraph, launch shape chosen
               by the autotuner on a quarter of the slabs; and the same
               quarter of the slabs run one mcgpu_simulate_device call per
               slab ("per-problem launches") next to its batched time;
  adjoint   -- one slab of 1,000 cells, 1e7 particles: forward with
               per-history scores + seed-replay adjoint, against the plain
               forward, for d(absorbed weight in cells 800-899)/d(sigma);
  adjoint_batched -- the same across 64 such slabs (6.4e8 particles) with
               the batched forward and adjoint launches.
Each variant's events are device-counted.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mc-mpi_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_EVENT = 72  # algorithmic: the 36-byte particle state read + written per event

# Per-event figures of the 10,000-cell slab measured with rocprofv3 on the
# build this file ships with (1e7 particles; the summaries are
# profiles/history_slab10k_counters.md and profiles/event_slab10k_hbm_counters.md):
#  - bytes between L2 and the memory fabric (FETCH_SIZE / WRITE_SIZE), scaled by
#    the copy-kernel calibration (read x2.0, write x1.0).  These are L2 misses:
#    an upper bound on HBM bytes, since the Infinity Cache (MALL) in front of
#    HBM can serve them;
#  - VALU wave-instructions (SQ_INSTS_VALU).
MEASURED_TRAFFIC_PER_EVENT = {"history": 0.0616, "event": 19.89}
# (the event kernel's VALU figure is of a 1e8-particle run, the size bench.py
# times: at 1e7 the end of the bank, with its partly empty chunks, weighs more)
MEASURED_VALU_PER_EVENT = {"history": 1.620, "event": 2.918}
# gfx950 issues one wave64 VALU instruction per SIMD every 4 cycles:
# 256 CUs x 4 SIMDs x 2.4 GHz / 4 = 6.144e11 wave-instructions/s.
VALU_PEAK_WAVE_INSTR = 256 * 4 * 2.4e9 / 4


def valu_rate(kernel, events, seconds):
    w = MEASURED_VALU_PER_EVENT.get(kernel)
    if w is None:
        return None
    a = w * events / seconds
    return {"wave_instructions_per_event": w, "achieved": a, "peak": VALU_PEAK_WAVE_INSTR,
            "unit": "wave-instructions/s", "frac": a / VALU_PEAK_WAVE_INSTR}


def roofline(kernel, events, seconds):
    """The roofline block of one plain-kernel launch (events over seconds).

    Both kernels are reported against VALU issue: `frac` is the issue rate
    (measured wave-instructions per event x events/s) over gfx950's 6.144e11
    wave-instructions/s, and the bytes are a side field (`hbm`).
    history kernel: the state stays in registers (about 0.06 B/event of
      traffic); bound by VALU issue.
    event kernel: every event reloads and stores its particle, but the
      fields a pass touches (about 3.7 MB per XCD for the four-slot pools)
      stay in L2: about 3.4 B/event read and 16.5 B/event written reach the
      memory fabric, against the 72 B/event algorithmic yardstick of the
      event formulation (the 36-byte state read and written).  Neither
      VALU issue (about 74 %) nor HBM (about 40 %) is saturated: the
      limit is the vector memory pipeline, whose texture address unit is
      busy about 73 % of the cycles on average and 84 % on the busiest CU
      (pool fields, the cell and exp-table gathers)."""
    measured_b = MEASURED_TRAFFIC_PER_EVENT.get(kernel)
    traffic = None if measured_b is None else measured_b * events
    alg_gbs = events * BYTES_PER_EVENT / seconds / 1e9
    hbm = {"algorithmic_bytes_per_event": BYTES_PER_EVENT,
           "algorithmic_GBs": alg_gbs, "algorithmic_frac": alg_gbs / HBM_PEAK_GBS,
           "measured_bytes_per_event": measured_b,
           "measured_GBs": None if traffic is None else traffic / seconds / 1e9,
           "measured_frac": None if traffic is None else traffic / seconds / 1e9 / HBM_PEAK_GBS,
           "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    valu = valu_rate(kernel, events, seconds)
    if kernel == "history":
        note = ("history kernel: one history per lane, state in registers; bound by FP64 VALU "
                "issue (frac = measured wave-instructions/event x events/s / 6.144e11); ")
    else:
        note = ("event kernel: one event per live particle per pass, state reloaded each event "
                "from per-wave pools whose touched fields stay in L2; frac = measured "
                "wave-instructions/event x events/s / 6.144e11, about 0.74: neither VALU issue "
                "nor HBM (hbm.measured_frac) is saturated, the limit is the vector memory "
                "pipeline (texture address unit busy ~73 % of cycles: pool fields, cell and "
                "exp-table gathers); ")
    return {"bound": "valu", "achieved": valu["achieved"], "peak": valu["peak"],
            "unit": valu["unit"], "frac": valu["frac"], "traffic": traffic,
            "wave_instructions_per_event": valu["wave_instructions_per_event"],
            "hbm": hbm,
            "note": note + "traffic = measured L2-miss bytes per launch"}


def host_cores():
    """(threads to use, nproc, CPU quota of this cgroup or None): all CPUs this
    process may run on (os.sched_getaffinity, what nproc prints), capped by
    the cgroup's CPU quota when one is set below that."""
    nproc = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    threads = nproc if quota is None else min(nproc, quota)
    return threads, nproc, quota


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def device_bank(ctx, n, mu_seed, torch):
    """A source bank built on the device: the reference seed chain by
    jump-ahead (mcgpu_rng_seed_chain), x = 0, index = 0, weight 1, and
    directions mu = sqrt(uniform(mu_seed, .)) from the splitmix64 generator of
    mcgpu.configs evaluated with torch on the device (a prefix is checked
    bit for bit against the host generator)."""
    from mcgpu import configs
    seed = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx.rng_seed_chain(configs.ROOT_SEED, seed)
    mu = torch.empty(n, dtype=torch.float64, device="cuda")
    chunk = 50_000_000
    for o in range(0, n, chunk):
        k = min(chunk, n - o)
        mu[o:o + k] = torch.sqrt(configs.uniform_torch(mu_seed, k, o))
    k = min(n, 100_000)
    want = np.sqrt(configs.uniform(mu_seed, k))
    if not np.array_equal(mu[:k].cpu().numpy().view(np.uint64), want.view(np.uint64)):
        raise RuntimeError("device-generated directions differ from the host generator")
    return {"seed": seed, "x": torch.zeros(n, dtype=torch.float64, device="cuda"), "mu": mu,
            "wmc": torch.ones(n, dtype=torch.float64, device="cuda"),
            "index": torch.zeros(n, dtype=torch.int32, device="cuda")}


def restore(dst, src):
    for k in dst:
        dst[k].copy_(src[k])


def cpu_baselines(n_ref, n_port, n_ref1):
    """CPU legs on prefixes of the 10,000-cell slab's bank, each about 10 s
    of work, on all host cores this process may use (host_cores()):
      reference -- the reference's own Layer::simulate (src/layer.cpp:220-361,
                   FP32 as the reference is built, OpenMP) from oracle/_ref on
                   the first n_ref particles, and on one thread on the first
                   n_ref1; events counted by the FP32 oracle on the same bank
                   (its per-history results equal the reference's);
      port      -- the FP64 oracle (the same arithmetic as the GPU path) on the
                   first n_port particles.
    Returns (reference or None when oracle/_ref is not built, port)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from mcgpu import configs
    wl = configs.slab10k(max(n_ref, n_port, n_ref1))
    threads, nproc, quota = host_cores()
    hostinfo = {"nproc": nproc, "cgroup_cpu_quota": quota}

    t0 = time.perf_counter()
    o = pyoracle.simulate_bank(64, wl.index_start, wl.m, wl.dx, wl.min_weight, wl.sigs,
                               wl.absorption_rates, wl.bank[:n_port], nthread=threads,
                               exact=False)
    dt = time.perf_counter() - t0
    ev = int(o["events"].sum())
    port = {"value": ev / dt, "unit": "particle-events/s", "cores": threads, "kind": "port",
            **hostinfo,
            "sample": f"first {n_port} particles of the 10,000-cell slab's bank ({ev} events, "
                      f"{dt:.1f} s on {threads} threads; FP64 oracle restating "
                      "src/layer.cpp with correctly rounded log/exp)"}
    log(f"cpu port: {port['value']:.3e} events/s")
    if not os.path.exists(pyoracle.REF_PATH):
        return None, port

    sigs32 = wl.sigs.astype(np.float32)
    ar32 = wl.absorption_rates.astype(np.float32)

    def ref_leg(n, nthread):
        b32 = np.zeros(n, dtype=pyoracle.P32)
        for f in ("seed", "x", "mu", "wmc", "index"):
            b32[f] = wl.bank[f][:n]
        R = pyoracle.RefLayer.new(0.0, 1.0, 0, wl.m, wl.min_weight)
        R.sigs[:] = sigs32
        R.absorption_rates[:] = ar32
        o32 = pyoracle.simulate_bank(32, 0, wl.m, np.float32(R.dx), np.float32(wl.min_weight),
                                     sigs32, ar32, b32, nthread=threads)
        ev32 = int(o32["events"].sum())
        R.push(b32)
        t0 = time.perf_counter()
        R.simulate(-1, nthread)
        dt = time.perf_counter() - t0
        assert R.nb_active == 0
        return ev32, dt

    ev32, dt = ref_leg(n_ref, threads)
    ev1, dt1 = ref_leg(n_ref1, 1)
    ref = {"value": ev32 / dt, "unit": "particle-events/s", "cores": threads,
           "kind": "reference", **hostinfo,
           "value_1_thread": ev1 / dt1,
           "sample": f"first {n_ref} particles of the 10,000-cell slab's bank in FP32 ({ev32} "
                     f"events, {dt:.1f} s): the reference's Layer::simulate(-1, {threads}) "
                     "built from src/layer.cpp + src/random.cpp (-O3 -fopenmp), one rank, no "
                     f"MPI; value_1_thread: the first {n_ref1} particles ({ev1} events, "
                     f"{dt1:.1f} s) with Layer::simulate(-1, 1)"}
    log(f"cpu reference: {ref['value']:.3e} events/s on {threads} threads, "
        f"{ref['value_1_thread']:.3e} on 1 (nproc {nproc}, quota {quota})")
    return ref, port


def cpu_reference_job(n):
    """The reference's own Layer (oracle/_ref, FP32 as the reference builds,
    OpenMP) on the first n particles of its headline job, one rank:
    decompose_domain + Layer::simulate(-1, threads) does the same histories
    as the one-rank cycle loop.  Wall seconds, extrapolated to the job's
    1e8 particles.  None when oracle/_ref is not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if not os.path.exists(pyoracle.REF_PATH):
        return None
    threads = host_cores()[0]
    R = pyoracle.RefLayer.decompose_domain(0.0, 1.0, float(np.float32(0.7071067690849304)), 1, 0,
                                           1000, n, float(np.float32(1e-12)))
    t0 = time.perf_counter()
    R.simulate(-1, threads)
    dt = time.perf_counter() - t0
    assert R.nb_active == 0
    res = {"particles": n, "seconds": dt, "threads": threads,
           "extrapolated_wall_s_1e8": dt * 1e8 / n,
           "note": "the reference's Layer::simulate(-1, threads) built from src/layer.cpp + "
                   "src/random.cpp (-O3 -fopenmp, FP32), one rank, this host"}
    log(f"cpu reference job: {res}")
    return res


def run_plain(ctx, args, torch, mcgpu, configs):
    n = args.particles
    wl = configs.slab10k(n, with_bank=False)
    log(f"10,000-cell slab: building the {n}-particle bank on the device")
    bank0 = device_bank(ctx, n, 2 * 1000 + 3, torch)  # slab10k_mu
    bank = {k: v.clone() for k, v in bank0.items()}
    sigs = torch.from_numpy(wl.sigs).cuda()
    ar = torch.from_numpy(wl.absorption_rates).cuda()
    w = torch.zeros(wl.m, dtype=torch.float64, device="cuda")

    def step(opt):
        restore(bank, bank0)
        w.zero_()
        st, _ = ctx.simulate_device(bank, sigs, ar, w, wl.min_index, wl.max_index, wl.dx,
                                    wl.min_weight, options=opt)
        return st

    opt = mcgpu.make_options(args.kernel, block_size=args.block, particles_per_lane=args.ppl,
                             compaction_interval=args.ci, events_per_refill=args.refill)
    for i in range(args.warmup):
        st = step(opt)
        log(f"warmup {i}: {st['events']} events, kernel {st['kernel_ms']:.1f} ms")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms, events = [], []
    for i in range(args.steps):
        st = step(opt)
        kernel_ms.append(st["kernel_ms"])
        events.append(st["events"])
        log(f"step {i}: {st['events']} events, kernel {st['kernel_ms']:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    res = {"events_per_step": events[-1], "elapsed_s": elapsed, "kernel_ms": kernel_ms,
           "value": sum(events) / elapsed}
    other = None
    if not args.plain_only:
        # Both kernels on the same 1e8-particle bank, outside the timed region,
        # with per-history exit codes and event / collision counts; every
        # history's final state and counts and the tally must be bit-identical.
        other_kernel = "event" if args.kernel == "history" else "history"
        outs = {}
        for kern in (args.kernel, other_kernel):
            res8 = torch.empty(n, dtype=torch.int8, device="cuda")
            evs = torch.empty(n, dtype=torch.int64, device="cuda")
            cols = torch.empty(n, dtype=torch.int64, device="cuda")
            restore(bank, bank0)
            w.zero_()
            st, _ = ctx.simulate_device(bank, sigs, ar, w, wl.min_index, wl.max_index, wl.dx,
                                        wl.min_weight, result=res8, events=evs, collisions=cols,
                                        options=mcgpu.make_options(kern))
            torch.cuda.synchronize()
            log(f"{kern} kernel: {st['events']} events, kernel {st['kernel_ms']:.1f} ms")
            outs[kern] = (st, {k: v.clone() for k, v in bank.items()}, w.clone(), res8, evs, cols)
        a, b = outs[args.kernel], outs[other_kernel]
        mism = {f: int((a[1][f].view(torch.int64 if a[1][f].element_size() == 8 else torch.int32)
                        != b[1][f].view(torch.int64 if b[1][f].element_size() == 8 else torch.int32)).sum())
                for f in bank}
        mism["exit_code"] = int((a[3] != b[3]).sum())
        mism["events"] = int((a[4] != b[4]).sum())
        mism["collisions"] = int((a[5] != b[5]).sum())
        tally_cells = int((a[2].view(torch.int64) != b[2].view(torch.int64)).sum())
        same = all(v == 0 for v in mism.values()) and tally_cells == 0 and \
            a[0]["events"] == b[0]["events"] and a[0]["collisions"] == b[0]["collisions"]
        parity = {"kernels": [args.kernel, other_kernel], "histories": n,
                  "bit_identical": same, "mismatching_histories_per_field": mism,
                  "mismatching_tally_cells": tally_cells,
                  "events": [a[0]["events"], b[0]["events"]],
                  "collisions": [a[0]["collisions"], b[0]["collisions"]]}
        log(f"history vs event kernel on the full bank: {parity}")
        if not same:
            log("ERROR: the two kernels disagree on the full bank")
        del outs, a, b
        # the other kernel timed on its own, without the per-history outputs
        st = step(mcgpu.make_options(other_kernel))
        torch.cuda.synchronize()
        log(f"{other_kernel} kernel (timed): {st['events']} events, "
            f"kernel {st['kernel_ms']:.1f} ms")
        other = {"kernel": other_kernel, "events": st["events"], "kernel_ms": st["kernel_ms"],
                 "events_per_s": st["events"] / st["kernel_ms"] * 1e3,
                 "full_bank_parity": parity}
    del bank, bank0
    torch.cuda.empty_cache()
    return res, other


def run_host_api(ctx, args, torch, mcgpu, configs):
    """The drop-in path with host buffers (mcgpu_simulate: host AoS bank in,
    copied to the device and back inside the call), timed on the wall clock
    around the call, on the first 1e7 particles of the 10,000-cell slab."""
    n = args.host_particles
    wl = configs.slab10k(n)
    bank = wl.bank.copy()
    w = np.zeros(wl.m)
    t0 = time.perf_counter()
    _, st, _ = ctx.simulate(bank, wl.sigs, wl.absorption_rates, w, wl.min_index, wl.max_index,
                            wl.dx, wl.min_weight)
    dt = time.perf_counter() - t0
    out = {"workload": f"single slab of 10000 cells, first {n} particles, host AoS bank (40 B per particle)",
           "events": st["events"], "wall_s": dt, "kernel_ms": st["kernel_ms"],
           "events_per_s_pcie_inclusive": st["events"] / dt,
           "events_per_s_kernel": st["events"] / st["kernel_ms"] * 1e3}
    log(f"host api: {out}")
    return out


# The reference's published whole-program wall times for its headline job
# (scripts/batch.py:55-79: 1000 cells, x_ini = 0.70710677, min weight 1e-12,
# 1e8 particles, 1e5 per cycle, `mpirun -n R ./main config.yaml sync`, FP32,
# nthread 1), as recorded in py_data.pkl (si36 cluster, 8 ranks per node).
REFERENCE_JOB_WALL_S = {1: 2895.78, 2: 1612.93, 4: 952.40, 8: 646.36}
REFERENCE_JOB_YAML = ("x_min: 0.000000000000000000e+00\nx_max: 1.000000000000000000e+00\n"
                      "x_ini: 7.071067690849304199e-01\n"
                      "particle_min_weight: 9.999999960041972003e-13\nnb_cells: 1000\n"
                      "nb_particles: {n}\ncycle_time: 1.e-05\n"
                      "nb_particles_per_cycle: 100000\nnthread: 1\nstatistics_cycle_time: 1.e-1\n")


def run_cycle(ctx, args, torch, mcgpu, configs):
    """The reference's headline job, `mpirun -n R ./main config.yaml sync`
    with its scripts' config (1000 cells, 1e8 particles, 1e5 per cycle), run
    by `mcgpu_run config.yaml --ranks R`: every rank's layer on this GPU,
    the whole cycle loop device-resident.  Wall seconds of the loop as
    mcgpu_run reports them (src/main.cpp:90-91 prints the same quantity),
    next to the reference's published wall time for the same job."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "mc-mpi_amd", "mcgpu_run")
    n = args.cycle_particles
    out = {"workload": f"the reference's headline job: config.yaml of scripts/batch.py, 1000 "
                       f"cells, {n} particles from x_ini = 0.70710677, min weight 1e-12, 1e5 "
                       "per cycle, sync; mcgpu_run --ranks R (all ranks on this GPU)",
           "runs": {}}
    runs = []
    for R in (int(r) for r in args.cycle_ranks.split(",")):
        runs.append((str(R), R, []))
        if R == 1:  # the same job with its cycles kept one after the other
            runs.append(("1_cycle_by_cycle", 1, ["--cycle-by-cycle"]))
    for key, R, extra in runs:
        with tempfile.TemporaryDirectory() as d:
            cfg = os.path.join(d, "config.yaml")
            with open(cfg, "w") as f:
                f.write(REFERENCE_JOB_YAML.format(n=n))
            r = subprocess.run([exe, cfg, "--ranks", str(R)] + extra, cwd=d, capture_output=True,
                               text=True, timeout=600)
            with open(os.path.join(d, "out", "weights.csv")) as f:
                weights_csv = f.read()
        if r.returncode != 0:
            raise RuntimeError(f"mcgpu_run failed: {r.stderr[-500:]}")
        secs = float(r.stdout.split()[0])
        # the summary line; a profiler attached to the child adds lines of its own
        summary = [ln for ln in r.stderr.splitlines() if ln.startswith("events ")]
        if not summary:
            raise RuntimeError(f"mcgpu_run printed no summary: {r.stderr[-500:]}")
        fields = dict(kv.split(" ", 1) for kv in summary[-1].split(", "))
        events = int(fields["events"])
        run = {"events": events, "cycles": int(fields["cycles"]), "wall_s": secs,
               "kernel_s": float(fields["kernel_ms"]) / 1e3, "events_per_s": events / secs,
               "weights_csv_sha1": __import__("hashlib").sha1(weights_csv.encode()).hexdigest()}
        if extra:
            run["mode"] = "cycle by cycle (--cycle-by-cycle)"
        elif R == 1:
            run["mode"] = ("one rank: the cycles' particle sets follow from the counts, so all "
                           "cycles are transported concurrently, tallies added per cycle in order")
        if n == 100_000_000 and R in REFERENCE_JOB_WALL_S:
            ref = REFERENCE_JOB_WALL_S[R]
            run["reference_published_wall_s"] = ref
            run["speedup_vs_published"] = ref / secs
        out["runs"][key] = run
        log(f"cycle, {key}: {run}")
    if "1" in out["runs"] and "1_cycle_by_cycle" in out["runs"]:
        a, b = out["runs"]["1"], out["runs"]["1_cycle_by_cycle"]
        out["one_rank_modes_identical"] = (a["weights_csv_sha1"] == b["weights_csv_sha1"] and
                                           a["events"] == b["events"] and
                                           a["cycles"] == b["cycles"])
    return out


def run_batched(ctx, args, torch, mcgpu, configs):
    log("4096-slab batch: layout and bank")
    slabs, sigs, ar, total = configs.slabs4096_layout(particle_scale=args.batch_scale)
    bank0 = device_bank(ctx, total, 4 * 1000 + 4, torch)  # slabs4096_mu
    bank = {k: v.clone() for k, v in bank0.items()}
    d_sigs = torch.from_numpy(sigs).cuda()
    d_ar = torch.from_numpy(ar).cuda()
    w = torch.zeros(len(sigs), dtype=torch.float64, device="cuda")
    sub = slabs[::4]

    # autotune on the quarter of the slabs (inputs restored by the tuner)
    plan = mcgpu.Batch(ctx, sub, d_sigs, d_ar, w, bank)
    times, chosen = plan.autotune()
    plan.close()
    log(f"autotune (quarter): {['%.0f' % t for t in times]} ms -> {chosen}")
    opt = mcgpu.make_options(block_size=chosen["block_size"],
                             particles_per_lane=chosen["particles_per_lane"],
                             events_per_refill=chosen["events_per_refill"])

    # batched vs per-problem launches on the same quarter; every slab's bank,
    # exit codes and tally compared bit for bit between the two
    restore(bank, bank0)
    w.zero_()
    res_b = torch.zeros(total, dtype=torch.int8, device="cuda")
    plan = mcgpu.Batch(ctx, sub, d_sigs, d_ar, w, bank, result=res_b, options=opt)
    st_sub = plan.run()
    plan.close()
    torch.cuda.synchronize()
    out_b = ({k: v.clone() for k, v in bank.items()}, w.clone())
    restore(bank, bank0)
    w.zero_()
    res_p = torch.zeros(total, dtype=torch.int8, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_pp = 0
    for (lo, hi, dx, mw, coff, poff, n) in sub:
        b = {k: v[poff:poff + n] for k, v in bank.items()}
        st, _ = ctx.simulate_device(b, d_sigs[coff:coff + hi - lo], d_ar[coff:coff + hi - lo],
                                    w[coff:coff + hi - lo], lo, hi, dx, mw,
                                    result=res_p[poff:poff + n])
        ev_pp += st["events"]
    torch.cuda.synchronize()
    t_pp = time.perf_counter() - t0
    log(f"quarter: batched {st_sub['kernel_ms']:.0f} ms, per-problem {t_pp * 1e3:.0f} ms")
    mism = {f: int((out_b[0][f].view(torch.int64 if bank[f].element_size() == 8 else torch.int32)
                    != bank[f].view(torch.int64 if bank[f].element_size() == 8 else torch.int32)).sum())
            for f in bank}
    mism["exit_code"] = int((res_b != res_p).sum())
    tally_cells = int((out_b[1].view(torch.int64) != w.view(torch.int64)).sum())
    n_sub = sum(s_[6] for s_ in sub)
    same = all(v == 0 for v in mism.values()) and tally_cells == 0 and ev_pp == st_sub["events"]
    quarter_parity = {"slabs": len(sub), "histories": n_sub, "bit_identical": same,
                      "mismatching_histories_per_field": mism,
                      "mismatching_tally_cells": tally_cells}
    log(f"quarter, batched vs per-problem: {quarter_parity}")
    if not same:
        log("ERROR: batched and per-problem runs disagree")
    del out_b, res_b, res_p

    # the full batch
    restore(bank, bank0)
    w.zero_()
    torch.cuda.synchronize()
    plan = mcgpu.Batch(ctx, slabs, d_sigs, d_ar, w, bank, options=opt)
    st = plan.run()
    plan.close()
    log(f"full batch: {st['events']} events, {st['kernel_ms']:.0f} ms")
    out = {"workload": f"{len(slabs)} independent slabs, {total} particles, {len(sigs)} cells",
           "events": st["events"], "graph_ms": st["kernel_ms"],
           "events_per_s": st["events"] / st["kernel_ms"] * 1e3, "autotuned": chosen,
           "autotune_ms": times,
           "quarter": {"slabs": len(sub), "events": st_sub["events"],
                       "batched_ms": st_sub["kernel_ms"], "per_problem_ms": t_pp * 1e3,
                       "per_problem_events": ev_pp,
                       "batched_events_per_s": st_sub["events"] / st_sub["kernel_ms"] * 1e3,
                       "per_problem_events_per_s": ev_pp / t_pp,
                       "speedup": (st_sub["events"] / st_sub["kernel_ms"] * 1e3) / (ev_pp / t_pp),
                       "parity": quarter_parity}}
    del bank, bank0
    torch.cuda.empty_cache()
    return out


def run_adjoint(ctx, args, torch, mcgpu, configs):
    wl = configs.adjoint1000(args.adjoint_particles)
    init = mcgpu.bank_to_device(wl.bank)
    sa = torch.from_numpy(wl.sigma_a).cuda()
    ss = torch.from_numpy(wl.sigma_s).cuda()
    g = torch.from_numpy(wl.target).cuda()
    n = len(wl.bank)
    res = {}
    for rep in range(2):  # first repetition warms up
        work = {k: v.clone() for k, v in init.items()}
        w = torch.zeros(wl.m, dtype=torch.float64, device="cuda")
        plain = ctx.forward_cells(work, sa, ss, w, 0, wl.m, wl.dx, wl.min_weight)
        work = {k: v.clone() for k, v in init.items()}
        w.zero_()
        score = torch.empty(n, dtype=torch.float64, device="cuda")
        fwd = ctx.forward_cells(work, sa, ss, w, 0, wl.m, wl.dx, wl.min_weight, g=g,
                                history_score=score)
        ga = torch.zeros(wl.m, dtype=torch.float64, device="cuda")
        gi = torch.zeros(wl.m, dtype=torch.float64, device="cuda")
        adj = ctx.adjoint_cells(init, sa, ss, 0, wl.m, wl.dx, wl.min_weight, g, ga, gi,
                                history_score=score)
        adj2 = ctx.adjoint_cells(init, sa, ss, 0, wl.m, wl.dx, wl.min_weight, g, ga, gi)
        res = {"workload": f"single slab, 1000 cells, {n} particles, target cells 800-899",
               "forward_ms": plain["kernel_ms"], "forward_events": plain["events"],
               "forward_scored_ms": fwd["kernel_ms"], "adjoint_ms": adj["kernel_ms"],
               "adjoint_without_scores_ms": adj2["kernel_ms"],
               "fwd_plus_adjoint_over_forward": (fwd["kernel_ms"] + adj["kernel_ms"]) /
               plain["kernel_ms"],
               "events_per_s_fwd_plus_adjoint": plain["events"] /
               (fwd["kernel_ms"] + adj["kernel_ms"]) * 1e3}
    log(f"adjoint: {res}")
    return res


def run_adjoint_batched(ctx, args, torch, mcgpu, configs):
    """The 1,000-cell adjoint workload across 64 slabs: batched forward with scores +
    batched adjoint against the batched plain forward."""
    slabs, sa, ss, g, total = configs.adjoint1000_batch(64, args.adjoint_batch_particles)
    log(f"1,000-cell slabs batched: {len(slabs)} slabs, {total} particles")
    init = device_bank(ctx, total, 5004, torch)  # adjoint1000_batch_mu
    d_sa, d_ss, d_g = (torch.from_numpy(v).cuda() for v in (sa, ss, g))
    w = torch.zeros(len(sa), dtype=torch.float64, device="cuda")
    work = {k: v.clone() for k, v in init.items()}
    plain = ctx.forward_cells_batch(slabs, work, d_sa, d_ss, w)
    for k in work:
        work[k].copy_(init[k])
    w.zero_()
    score = torch.empty(total, dtype=torch.float64, device="cuda")
    fwd = ctx.forward_cells_batch(slabs, work, d_sa, d_ss, w, g=d_g, history_score=score)
    del work
    ga = torch.zeros(len(sa), dtype=torch.float64, device="cuda")
    gi = torch.zeros(len(sa), dtype=torch.float64, device="cuda")
    adj = ctx.adjoint_cells_batch(slabs, init, d_sa, d_ss, d_g, ga, gi, history_score=score)
    out = {"workload": f"64 slabs of 1000 cells, {total} particles, "
                       "target cells 800-899 of every slab",
           "forward_ms": plain["kernel_ms"], "forward_events": plain["events"],
           "forward_scored_ms": fwd["kernel_ms"], "adjoint_ms": adj["kernel_ms"],
           "fwd_plus_adjoint_over_forward": (fwd["kernel_ms"] + adj["kernel_ms"]) /
           plain["kernel_ms"],
           "events_per_s_forward": plain["events"] / plain["kernel_ms"] * 1e3,
           "events_per_s_fwd_plus_adjoint": plain["events"] /
           (fwd["kernel_ms"] + adj["kernel_ms"]) * 1e3}
    log(f"adjoint batched: {out}")
    del init, score
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=100_000_000)
    ap.add_argument("--kernel", default="history")
    ap.add_argument("--block", type=int, default=0, help="threads per block (0 = auto)")
    ap.add_argument("--ppl", type=int, default=0, help="event kernel: pool slots per lane")
    ap.add_argument("--ci", type=int, default=0, help="event kernel: compaction interval")
    ap.add_argument("--refill", type=int, default=0, help="history kernel: events per refill")
    ap.add_argument("--cpu-ref-sample", type=int, default=2_000_000,
                    help="particles for the reference CPU leg")
    ap.add_argument("--cpu-ref1-sample", type=int, default=100_000,
                    help="particles for the one-thread reference CPU leg")
    ap.add_argument("--cpu-port-sample", type=int, default=200_000,
                    help="particles for the FP64 oracle CPU leg")
    ap.add_argument("--batch-scale", type=float, default=1.0, help="particle scale of the 4096-slab batch")
    ap.add_argument("--adjoint-particles", type=int, default=10_000_000)
    ap.add_argument("--adjoint-batch-particles", type=int, default=10_000_000,
                    help="particles per slab of the batched 1,000-cell slabs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true",
                    help="only the 10,000-cell slab steps (for profiling runs)")
    ap.add_argument("--skip", default="",
                    help="comma-separated variants to leave out (host_api, cycle, batched, "
                         "adjoint, adjoint_batched)")
    ap.add_argument("--cycle-particles", type=int, default=100_000_000,
                    help="particles of the reference-job cycle variant")
    ap.add_argument("--cycle-ranks", default="1,8",
                    help="comma-separated rank counts of the cycle variant")
    ap.add_argument("--cpu-job-sample", type=int, default=4_000_000,
                    help="particles of the reference CPU leg of the cycle variant")
    ap.add_argument("--host-particles", type=int, default=10_000_000,
                    help="particles of the host-buffer (PCIe-inclusive) variant")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (untimed), the "
                         "known-size reference kernel for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import mcgpu
    from mcgpu import configs

    ctx = mcgpu.Context(0)
    if args.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    plain, other = run_plain(ctx, args, torch, mcgpu, configs)
    ev_step = plain["events_per_step"]
    k_avg_s = sum(plain["kernel_ms"]) / len(plain["kernel_ms"]) / 1e3
    out = {
        "metric": "particle-events/s on 1 MI355X (plain vs batched vs fwd+adjoint); % HBM roofline",
        "value": plain["value"],
        "unit": "particle-events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["elapsed_s"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"single slab, 10000 cells, {args.particles} particles "
                               "from x=0, mu=sqrt(U), min weight 1e-3, FP64",
                   "kernel": args.kernel, "events_per_step": ev_step,
                   "kernel_ms_avg": k_avg_s * 1e3},
        "roofline": roofline(args.kernel, ev_step, k_avg_s),
        "cpu_baseline": None,
        "variants": {},
    }
    if other is not None:
        other["roofline"] = roofline(other["kernel"], other["events"], other["kernel_ms"] / 1e3)
        out["variants"][other["kernel"]] = other
    if not args.plain_only:
        runs = {"host_api": run_host_api, "cycle": run_cycle, "batched": run_batched,
                "adjoint": run_adjoint, "adjoint_batched": run_adjoint_batched}
        for name, fn in runs.items():
            if args.skip and name in args.skip.split(","):
                continue
            out["variants"][name] = fn(ctx, args, torch, mcgpu, configs)
    if not args.no_cpu:
        log("timing the CPU baselines")
        ref, port = cpu_baselines(args.cpu_ref_sample, args.cpu_port_sample,
                                  args.cpu_ref1_sample)
        out["cpu_baseline"] = ref if ref is not None else port
        out["variants"]["cpu_fp64_port"] = port
        if "cycle" in out["variants"]:
            job = cpu_reference_job(args.cpu_job_sample)
            if job is not None:
                out["variants"]["cycle"]["reference_cpu_this_host"] = job
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
