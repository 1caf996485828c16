"""Batched transport (mcgpu_batch_*): many independent slabs in one
persistent launch replayed from a captured hipGraph.

Bar: every slab's final bank, exit codes and weights bit-identical to
transporting that slab alone with mcgpu_simulate_device, and a sample of
slabs bit-identical to the CPU oracle; graph replays (with eager work and
synchronizations between them) each reproduce the eager result; the
autotuner leaves inputs unchanged and its choice gives the same results.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _batch_inputs(n_slabs, scale, extra_slabs=()):
    from mcgpu import configs
    slabs, sigs, ar, total = configs.slabs4096_layout(n_slabs=n_slabs, particle_scale=scale)
    slabs = list(slabs)
    sigs, ar = list(sigs), list(ar)
    coff = len(sigs)
    for (m, n) in extra_slabs:  # extra slabs: (cells, particles); n may be 0
        rng = np.random.default_rng(m + n)
        sigs += list(rng.lognormal(0, 1, m))
        ar += list(1 - rng.uniform(0.1, 0.99, m))
        slabs.append((0, m, 1.0 / m, 1e-3, coff, total, n))
        coff += m
        total += n
    bank = configs.slabs4096_bank(total)
    return slabs, np.array(sigs), np.array(ar), bank


def _device(slabs, sigs, ar, bank):
    import torch
    import mcgpu
    return (torch.from_numpy(sigs).cuda(), torch.from_numpy(ar).cuda(),
            torch.zeros(len(sigs), dtype=torch.float64, device="cuda"),
            mcgpu.bank_to_device(bank),
            torch.full((len(bank),), 9, dtype=torch.int8, device="cuda"))


def _unbatched(ctx, slabs, sigs, ar, bank):
    """Each slab alone through mcgpu_simulate_device."""
    import torch
    import mcgpu
    w_all = np.zeros(len(sigs))
    out_bank = bank.copy()
    res_all = np.zeros(len(bank), np.int8)
    for (lo, hi, dx, mw, coff, poff, n) in slabs:
        m = hi - lo
        if n == 0:
            continue
        b = mcgpu.bank_to_device(bank[poff:poff + n])
        w = torch.zeros(m, dtype=torch.float64, device="cuda")
        r = torch.empty(n, dtype=torch.int8, device="cuda")
        ctx.simulate_device(b, torch.from_numpy(sigs[coff:coff + m]).cuda(),
                            torch.from_numpy(ar[coff:coff + m]).cuda(), w, lo, hi, dx, mw,
                            result=r)
        out_bank[poff:poff + n] = mcgpu.bank_from_device(b)
        res_all[poff:poff + n] = r.cpu().numpy()
        w_all[coff:coff + m] = w.cpu().numpy()
    return out_bank, res_all, w_all


def _same_bank(a, b):
    for f in ("seed", "x", "mu", "wmc", "index"):
        x, y = np.ascontiguousarray(a[f]), np.ascontiguousarray(b[f])
        if not np.array_equal(x.view(np.uint64 if x.itemsize == 8 else np.uint32),
                              y.view(np.uint64 if y.itemsize == 8 else np.uint32)):
            return False
    return True


# cells in L2 / LDS, empty slab, and a tally larger than LDS (added in global memory)
EXTRA = ((8192, 3000), (5200, 2000), (40, 0), (33, 5), (12000, 1500))


def test_batched_equals_unbatched_and_oracle(ctx, oracle):
    import mcgpu
    slabs, sigs, ar, bank = _batch_inputs(48, 0.01, EXTRA)
    _say(f"batch of {len(slabs)} slabs, {len(bank)} particles, {len(sigs)} cells")
    ref_bank, ref_res, ref_w = _unbatched(ctx, slabs, sigs, ar, bank)
    d_sigs, d_ar, d_w, d_bank, d_res = _device(slabs, sigs, ar, bank)
    batch = mcgpu.Batch(ctx, slabs, d_sigs, d_ar, d_w, d_bank, d_res)
    st = batch.run()
    got = mcgpu.bank_from_device(d_bank)
    assert _same_bank(got, ref_bank)
    assert np.array_equal(d_res.cpu().numpy(), ref_res)
    assert np.array_equal(d_w.cpu().numpy().view(np.uint64), ref_w.view(np.uint64))
    assert st["n_left"] + st["n_right"] + st["n_killed"] == len(bank)
    # a sample of slabs against the oracle, whole slab
    for s in (0, 7, len(slabs) - 4, len(slabs) - 3, len(slabs) - 1):
        lo, hi, dx, mw, coff, poff, n = slabs[s]
        o = oracle.simulate_bank(64, lo, hi - lo, dx, mw, sigs[coff:coff + hi - lo],
                                 ar[coff:coff + hi - lo], bank[poff:poff + n], nthread=8)
        assert _same_bank(got[poff:poff + n], o["bank"]), s
        assert np.array_equal(ref_res[poff:poff + n].astype(np.int32), o["result"])
        assert np.array_equal(d_w.cpu().numpy()[coff:coff + hi - lo], o["tally_exact"])
    batch.close()


def test_batched_layout_prefix_at_a_tenth_of_its_particles(ctx, oracle):
    """The first 256 slabs of the 4096-slab layout (32 to 8192 cells, a tenth
    of their particles: 3.9e6 histories): every slab's bank, exit codes and
    tally bit-identical to its own mcgpu_simulate_device call, and six slabs
    (the largest and smallest by cells and by particles among them) to the
    oracle."""
    import mcgpu
    slabs, sigs, ar, bank = _batch_inputs(256, 0.1)
    _say(f"layout prefix: {len(slabs)} slabs, {len(bank)} particles, {len(sigs)} cells")
    ref_bank, ref_res, ref_w = _unbatched(ctx, slabs, sigs, ar, bank)
    d_sigs, d_ar, d_w, d_bank, d_res = _device(slabs, sigs, ar, bank)
    batch = mcgpu.Batch(ctx, slabs, d_sigs, d_ar, d_w, d_bank, d_res)
    st = batch.run()
    batch.close()
    got = mcgpu.bank_from_device(d_bank)
    res = d_res.cpu().numpy()
    w = d_w.cpu().numpy()
    assert _same_bank(got, ref_bank)
    assert np.array_equal(res, ref_res)
    assert np.array_equal(w.view(np.uint64), ref_w.view(np.uint64))
    assert st["n_left"] + st["n_right"] + st["n_killed"] == len(bank)
    m = np.array([s[1] - s[0] for s in slabs])
    n = np.array([s[6] for s in slabs])
    pick = {int(np.argmax(m)), int(np.argmin(m)), int(np.argmin(n))}
    pick |= set(int(i) for i in np.argsort(n)[-3:] if m[i] < 3000)  # large banks, short slabs
    for s in sorted(pick):
        lo, hi, dx, mw, coff, poff, k = slabs[s]
        _say(f"  oracle: slab {s}, {hi - lo} cells, {k} particles")
        o = oracle.simulate_bank(64, lo, hi - lo, dx, mw, sigs[coff:coff + hi - lo],
                                 ar[coff:coff + hi - lo], bank[poff:poff + k], nthread=16)
        assert _same_bank(got[poff:poff + k], o["bank"]), s
        assert np.array_equal(res[poff:poff + k].astype(np.int32), o["result"]), s
        assert np.array_equal(w[coff:coff + hi - lo].view(np.uint64),
                              o["tally_exact"].view(np.uint64)), s


def test_graph_replays_match_eager(ctx):
    """Replay the captured graph three times, restoring the inputs with eager
    torch copies and synchronizing between replays; every replay must equal
    the eager (unbatched) result."""
    import torch
    import mcgpu
    slabs, sigs, ar, bank = _batch_inputs(32, 0.02)
    ref_bank, ref_res, ref_w = _unbatched(ctx, slabs, sigs, ar, bank)
    d_sigs, d_ar, d_w, d_bank, d_res = _device(slabs, sigs, ar, bank)
    pristine = {k: v.clone() for k, v in d_bank.items()}
    batch = mcgpu.Batch(ctx, slabs, d_sigs, d_ar, d_w, d_bank, d_res)
    for rep in range(3):
        for k in d_bank:
            d_bank[k].copy_(pristine[k])
        d_w.zero_()
        d_res.fill_(9)
        torch.cuda.synchronize()
        batch.run()
        torch.cuda.synchronize()
        _say(f"replay {rep}")
        assert _same_bank(mcgpu.bank_from_device(d_bank), ref_bank), rep
        assert np.array_equal(d_res.cpu().numpy(), ref_res), rep
        assert np.array_equal(d_w.cpu().numpy().view(np.uint64), ref_w.view(np.uint64)), rep
    batch.close()


def test_autotune_restores_inputs_and_keeps_results(ctx):
    import torch
    import mcgpu
    slabs, sigs, ar, bank = _batch_inputs(24, 0.02, EXTRA[:2])
    ref_bank, ref_res, ref_w = _unbatched(ctx, slabs, sigs, ar, bank)
    d_sigs, d_ar, d_w, d_bank, d_res = _device(slabs, sigs, ar, bank)
    batch = mcgpu.Batch(ctx, slabs, d_sigs, d_ar, d_w, d_bank, d_res)
    times, chosen = batch.autotune()
    _say(f"autotune times {times} chosen {chosen}")
    torch.cuda.synchronize()
    assert _same_bank(mcgpu.bank_from_device(d_bank), bank)   # inputs restored
    assert np.all(d_w.cpu().numpy() == 0)
    assert all(t > 0 for t in times)
    assert chosen == batch.options()
    batch.run()
    assert _same_bank(mcgpu.bank_from_device(d_bank), ref_bank)
    assert np.array_equal(d_w.cpu().numpy().view(np.uint64), ref_w.view(np.uint64))
    batch.close()
