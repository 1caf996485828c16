#!/usr/bin/env python3
"""Benchmark: particle-events/s of the MI355X transport on one GPU.

Workload (config 2): one slab of 10 000 cells on [0,1] with per-cell
sigma_a ~ LogNormal(0, 0.5), sigma_s ~ LogNormal(0.5, 0.5), 1e8 particles
entering at x = 0 with mu = sqrt(U), weight 1, minimum weight 1e-3, FP64.
One step = one mcgpu_simulate_device call transporting the whole 1e8-particle
bank (resident in HBM before timing) to termination, including the restore
of the bank from a pristine copy.

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this
build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mc-mpi_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_EVENT = 72  # read + write of the 36-byte particle state per event


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def device_bank(ctx, n, torch):
    """config 2's source bank built on the device: the reference seed chain by
    jump-ahead (mcgpu_rng_seed_chain), mu = sqrt(U) from the same generator
    as mcgpu.configs.config2, x = 0, index = 0, weight 1."""
    from mcgpu import configs
    seed = torch.empty(n, dtype=torch.int64, device="cuda")
    ctx.rng_seed_chain(configs.ROOT_SEED, seed)
    mu = torch.empty(n, dtype=torch.float64, device="cuda")
    chunk = 10_000_000
    for o in range(0, n, chunk):
        k = min(chunk, n - o)
        mu[o:o + k] = torch.from_numpy(configs.config2_mu(k, offset=o)).cuda()
    return {"seed": seed, "x": torch.zeros(n, dtype=torch.float64, device="cuda"), "mu": mu,
            "wmc": torch.ones(n, dtype=torch.float64, device="cuda"),
            "index": torch.zeros(n, dtype=torch.int32, device="cuda")}


def cpu_baseline(wl_small, seconds_hint=15.0):
    """The FP64 oracle (port of src/layer.cpp) on a bounded prefix of the same
    bank, all host cores up to 16, correctly rounded log/exp."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    o = pyoracle.simulate_bank(64, wl_small.index_start, wl_small.m, wl_small.dx,
                               wl_small.min_weight, wl_small.sigs, wl_small.absorption_rates,
                               wl_small.bank, nthread=threads, exact=False)
    dt = time.perf_counter() - t0
    ev = int(o["events"].sum())
    return {"value": ev / dt, "unit": "particle-events/s", "cores": threads, "kind": "port",
            "sample": f"first {len(wl_small.bank)} particles of the config-2 bank "
                      f"({ev} events, {dt:.1f} s, FP64 oracle, correctly rounded log/exp)"}


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
    ap.add_argument("--cpu-sample", type=int, default=40_000)
    ap.add_argument("--no-cpu", action="store_true")
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

    n = args.particles
    wl = configs.config2(n, with_bank=False)
    ctx = mcgpu.Context(0)
    if args.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log(f"building the {n}-particle bank on the device")
    bank0 = device_bank(ctx, n, torch)
    bank = {k: v.clone() for k, v in bank0.items()}
    sigs = torch.from_numpy(wl.sigs).cuda()
    ar = torch.from_numpy(wl.absorption_rates).cuda()
    w = torch.zeros(wl.m, dtype=torch.float64, device="cuda")
    opt = mcgpu.make_options(args.kernel, block_size=args.block, particles_per_lane=args.ppl,
                             compaction_interval=args.ci, events_per_refill=args.refill)

    def step():
        for k in bank:
            bank[k].copy_(bank0[k])
        w.zero_()
        st, _ = ctx.simulate_device(bank, sigs, ar, w, wl.min_index, wl.max_index, wl.dx,
                                    wl.min_weight, options=opt)
        return st

    for i in range(args.warmup):
        st = step()
        log(f"warmup {i}: {st['events']} events, kernel {st['kernel_ms']:.1f} ms")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms, events = [], []
    for i in range(args.steps):
        st = step()
        kernel_ms.append(st["kernel_ms"])
        events.append(st["events"])
        log(f"step {i}: {st['events']} events, kernel {st['kernel_ms']:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    ev_step = events[-1]
    value = sum(events) / elapsed
    k_avg_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    achieved = ev_step * BYTES_PER_EVENT / k_avg_s / 1e9
    out = {
        "metric": "particle-events/s on 1 MI355X (plain vs batched vs fwd+adjoint); % HBM roofline",
        "value": value,
        "unit": "particle-events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"config2: single slab, 10000 cells, {n} particles from x=0, "
                               "mu=sqrt(U), min weight 1e-3, FP64",
                   "kernel": args.kernel, "events_per_step": ev_step},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("timing the CPU baseline")
        small = configs.config2(args.cpu_sample)
        out["cpu_baseline"] = cpu_baseline(small)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
