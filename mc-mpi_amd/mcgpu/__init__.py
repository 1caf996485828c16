"""mcgpu: Python access to the MI355X particle-transport library (libmcgpu.so).

Thin ctypes wrapper over the C ABI of include/mcgpu.h; every call goes to the
HIP kernels.  Device memory for the device-resident entry points is plain
torch tensors (their data pointers are passed through the C ABI).
"""
import ctypes as C

import numpy as np

from ._lib import (CYCLE_WINDOW_FN, ERR_ARG, ERR_EVENT_CAP, ERR_HIP, ERR_NO_DEVICE, ERR_NOMEM,
                   ERR_RANGE, KERNEL_EVENT, KERNEL_HISTORY, OK, PARTICLE, Bank, CycleRank,
                   CycleStats, McgpuError, Options, Slab, Stats, lib, status_string)

__all__ = ["Context", "Batch", "PARTICLE", "McgpuError", "KERNEL_HISTORY", "KERNEL_EVENT", "lib",
           "bank_to_device", "bank_from_device", "make_options", "options_dict", "check_tensor",
           "check_bank"]

_KERNELS = {"history": KERNEL_HISTORY, "event": KERNEL_EVENT}


def make_options(kernel="history", block_size=0, grid_blocks=0, events_per_refill=0,
                 compaction_interval=0, particles_per_lane=0, max_events_per_history=0):
    o = Options()
    lib().mcgpu_default_options(C.byref(o))
    o.kernel = _KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    o.block_size = block_size
    o.grid_blocks = grid_blocks
    o.events_per_refill = events_per_refill
    o.compaction_interval = compaction_interval
    o.particles_per_lane = particles_per_lane
    o.max_events_per_history = max_events_per_history
    return o


def options_dict(o):
    return {k: getattr(o, k) for k, _ in o._fields_}


def _stream(stream):
    """The stream a call is ordered on: the given one, else torch's current
    stream (so the call is ordered after the torch ops that filled its
    inputs); 0 / None is HIP's default stream."""
    if stream is not None:
        return stream
    try:
        import torch
        if torch.cuda.is_initialized():
            return torch.cuda.current_stream().cuda_stream
    except ImportError:
        pass
    return None


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


_BANK_DTYPES = None


def _bank_dtypes():
    global _BANK_DTYPES
    if _BANK_DTYPES is None:
        import torch
        _BANK_DTYPES = {"seed": (torch.int64, torch.uint64), "x": (torch.float64,),
                        "mu": (torch.float64,), "wmc": (torch.float64,),
                        "index": (torch.int32,)}
    return _BANK_DTYPES


def check_tensor(t, name, dtypes, n=None, device=None, optional=False):
    """Validate a tensor handed to the C ABI, before any launch: a torch
    tensor of one of `dtypes`, contiguous, with n elements (if given), in
    device memory of HIP device `device` (if given).  The library reads and
    writes exactly the element counts it is told, so a wrong dtype, length or
    a host tensor would be out-of-bounds device memory access.  Raises
    TypeError (not a tensor, wrong dtype) or ValueError (length, layout,
    device)."""
    import torch
    if t is None:
        if optional:
            return
        raise ValueError(f"{name}: required")
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    if not isinstance(dtypes, tuple):
        dtypes = (dtypes,)
    if t.dtype not in dtypes:
        raise TypeError(f"{name}: dtype {t.dtype}, expected "
                        + " or ".join(str(d) for d in dtypes))
    if n is not None and t.numel() != n:
        raise ValueError(f"{name}: {t.numel()} elements, expected {n}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if device is not None and (t.device.type != "cuda" or t.device.index != device):
        raise ValueError(f"{name}: on {t.device}, expected cuda:{device}")


def check_bank(bank, name, n=None, device=None):
    """Validate a device SoA bank (dict seed/x/mu/wmc/index); returns its length."""
    if not isinstance(bank, dict) or set(bank) != set(_bank_dtypes()):
        raise TypeError(f"{name}: expected a dict with keys seed, x, mu, wmc, index")
    if n is None:
        n = bank["x"].numel() if hasattr(bank["x"], "numel") else None
    for f, dts in _bank_dtypes().items():
        check_tensor(bank[f], f"{name}[{f!r}]", dts, n, device)
    return n


def _check_slabs(slabs, n_particles):
    """Check a slab table (tuples min_index, max_index, dx, min_weight,
    cell_offset, particle_offset, n) against a bank of n_particles; returns
    the number of cells the table spans (max of cell_offset + cells)."""
    cells = 0
    for k, s in enumerate(slabs):
        lo, hi, _dx, _mw, coff, poff, n = s
        if hi <= lo or coff < 0 or poff < 0 or n < 0 or poff + n > n_particles:
            raise ValueError(f"slab {k}: {s} does not fit a bank of {n_particles} particles")
        cells = max(cells, coff + hi - lo)
    return cells


def _bank_struct(bank):
    return Bank(_ptr(bank["seed"]), _ptr(bank["x"]), _ptr(bank["mu"]), _ptr(bank["wmc"]),
                _ptr(bank["index"]))


def bank_to_device(particles, device="cuda"):
    """numpy PARTICLE array -> dict of torch tensors (SoA bank) on `device`."""
    import torch
    p = np.ascontiguousarray(particles, dtype=PARTICLE)
    return {
        "seed": torch.from_numpy(p["seed"].view(np.int64).copy()).to(device),
        "x": torch.from_numpy(p["x"].copy()).to(device),
        "mu": torch.from_numpy(p["mu"].copy()).to(device),
        "wmc": torch.from_numpy(p["wmc"].copy()).to(device),
        "index": torch.from_numpy(p["index"].copy()).to(device),
    }


def bank_from_device(bank):
    """dict of torch tensors -> numpy PARTICLE array."""
    n = bank["x"].numel()
    out = np.zeros(n, dtype=PARTICLE)
    out["seed"] = bank["seed"].cpu().numpy().view(np.uint64)
    out["x"] = bank["x"].cpu().numpy()
    out["mu"] = bank["mu"].cpu().numpy()
    out["wmc"] = bank["wmc"].cpu().numpy()
    out["index"] = bank["index"].cpu().numpy()
    return out


class Context:
    """A device context (stream + reusable device scratch) of libmcgpu."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().mcgpu_context_create(device, C.byref(self._h))
        if rc != OK:
            raise McgpuError(rc, status_string(rc))
        self.device = device

    def close(self):
        if self._h:
            lib().mcgpu_context_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self):
        return lib().mcgpu_context_error(self._h).decode()

    def check(self, rc, allow=()):
        if rc != OK and rc not in allow:
            raise McgpuError(rc, f"{status_string(rc)}: {self.error()}")
        return rc

    # ------------------------------------------------------------ transport
    def simulate(self, particles, sigs, absorption_rates, weights_absorbed, min_index, max_index,
                 dx, particle_min_weight, options=None, allow=()):
        """Host-buffer drop-in (mcgpu_simulate).  `particles` (PARTICLE array)
        and `weights_absorbed` (float64) are updated in place.  Returns
        (result int32 array indexed by position, stats dict, status)."""
        assert particles.dtype == PARTICLE and particles.flags.c_contiguous
        assert weights_absorbed.dtype == np.float64 and weights_absorbed.flags.c_contiguous
        sigs = np.ascontiguousarray(sigs, np.float64)
        ar = np.ascontiguousarray(absorption_rates, np.float64)
        n = len(particles)
        result = np.zeros(n, np.int32)
        st = Stats()
        opt = options if options is not None else make_options()
        rc = lib().mcgpu_simulate(self._h, n, particles.ctypes.data, sigs.ctypes.data,
                                  ar.ctypes.data, weights_absorbed.ctypes.data, min_index,
                                  max_index, dx, particle_min_weight, result.ctypes.data,
                                  C.byref(opt), C.byref(st))
        self.check(rc, allow)
        return result, st.as_dict(), rc

    def simulate_device(self, bank, sigs, absorption_rates, weights_absorbed, min_index,
                        max_index, dx, particle_min_weight, result=None, events=None,
                        collisions=None, options=None, stream=None, allow=()):
        """Device-resident transport (mcgpu_simulate_device) on a dict-of-tensors
        bank; tensors are updated in place.  Returns (stats dict, status)."""
        import torch
        d, m = self.device, max_index - min_index
        n = check_bank(bank, "bank", device=d)
        for name, t in (("sigs", sigs), ("absorption_rates", absorption_rates),
                        ("weights_absorbed", weights_absorbed)):
            check_tensor(t, name, torch.float64, m, d)
        check_tensor(result, "result", torch.int8, n, d, optional=True)
        check_tensor(events, "events", (torch.int64, torch.uint64), n, d, optional=True)
        check_tensor(collisions, "collisions", (torch.int64, torch.uint64), n, d, optional=True)
        b = _bank_struct(bank)
        st = Stats()
        opt = options if options is not None else make_options()
        rc = lib().mcgpu_simulate_device(self._h, n, C.byref(b), _ptr(sigs), _ptr(absorption_rates),
                                         _ptr(weights_absorbed), min_index, max_index, dx,
                                         particle_min_weight, _ptr(result), _ptr(events),
                                         _ptr(collisions), C.byref(opt), C.byref(st),
                                         _stream(stream))
        self.check(rc, allow)
        return st.as_dict(), rc

    # ------------------------------------------------------------ adjoint
    def forward_cells(self, bank, sig_a, sig_i, weights_absorbed, min_index, max_index, dx,
                      particle_min_weight, g=None, history_score=None, result=None,
                      options=None, stream=None):
        """mcgpu_forward_cells: transport with per-cell (sig_a, sig_i) given
        directly; optionally stores each history's weighted tally."""
        import torch
        d, m = self.device, max_index - min_index
        n = check_bank(bank, "bank", device=d)
        for name, t in (("sig_a", sig_a), ("sig_i", sig_i),
                        ("weights_absorbed", weights_absorbed)):
            check_tensor(t, name, torch.float64, m, d)
        check_tensor(g, "g", torch.float64, m, d, optional=True)
        check_tensor(history_score, "history_score", torch.float64, n, d, optional=True)
        check_tensor(result, "result", torch.int8, n, d, optional=True)
        b = _bank_struct(bank)
        st = Stats()
        opt = options if options is not None else make_options()
        rc = lib().mcgpu_forward_cells(self._h, bank["x"].numel(), C.byref(b), _ptr(sig_a),
                                       _ptr(sig_i), _ptr(weights_absorbed), min_index, max_index,
                                       dx, particle_min_weight, _ptr(g), _ptr(history_score),
                                       _ptr(result), C.byref(opt), C.byref(st), _stream(stream))
        self.check(rc)
        return st.as_dict()

    def adjoint_cells(self, initial_bank, sig_a, sig_i, min_index, max_index, dx,
                      particle_min_weight, g, grad_sig_a, grad_sig_i, history_score=None,
                      options=None, stream=None):
        """mcgpu_adjoint_cells: adds dT/dsig_a, dT/dsig_i of T = sum g*tally."""
        import torch
        d, m = self.device, max_index - min_index
        n = check_bank(initial_bank, "initial_bank", device=d)
        for name, t in (("sig_a", sig_a), ("sig_i", sig_i), ("g", g),
                        ("grad_sig_a", grad_sig_a), ("grad_sig_i", grad_sig_i)):
            check_tensor(t, name, torch.float64, m, d)
        check_tensor(history_score, "history_score", torch.float64, n, d, optional=True)
        b = _bank_struct(initial_bank)
        st = Stats()
        opt = options if options is not None else make_options()
        rc = lib().mcgpu_adjoint_cells(self._h, initial_bank["x"].numel(), C.byref(b), _ptr(sig_a),
                                       _ptr(sig_i), min_index, max_index, dx, particle_min_weight,
                                       _ptr(g), _ptr(history_score), _ptr(grad_sig_a),
                                       _ptr(grad_sig_i), C.byref(opt), C.byref(st),
                                       _stream(stream))
        self.check(rc)
        return st.as_dict()

    def forward_cells_batch(self, slabs, bank, sig_a, sig_i, weights_absorbed, g=None,
                            history_score=None, result=None, options=None, stream=None):
        """mcgpu_forward_cells_batch over slabs (tuples as in Batch)."""
        import torch
        d = self.device
        n = check_bank(bank, "bank", device=d)
        m = _check_slabs(slabs, n)
        check_tensor(sig_a, "sig_a", torch.float64, None, d)
        m = max(m, sig_a.numel())
        for name, t in (("sig_a", sig_a), ("sig_i", sig_i),
                        ("weights_absorbed", weights_absorbed)):
            check_tensor(t, name, torch.float64, m, d)
        check_tensor(g, "g", torch.float64, m, d, optional=True)
        check_tensor(history_score, "history_score", torch.float64, n, d, optional=True)
        check_tensor(result, "result", torch.int8, n, d, optional=True)
        arr = (Slab * len(slabs))(*[Slab(*s) for s in slabs])
        b = _bank_struct(bank)
        st = Stats()
        opt = options if options is not None else make_options()
        rc = lib().mcgpu_forward_cells_batch(self._h, len(slabs), C.cast(arr, C.c_void_p),
                                             sig_a.numel(), C.byref(b), _ptr(sig_a), _ptr(sig_i),
                                             _ptr(weights_absorbed), _ptr(g), _ptr(history_score),
                                             _ptr(result), C.byref(opt), C.byref(st),
                                             _stream(stream))
        self.check(rc)
        return st.as_dict()

    def adjoint_cells_batch(self, slabs, initial_bank, sig_a, sig_i, g, grad_sig_a, grad_sig_i,
                            history_score=None, options=None, stream=None):
        """mcgpu_adjoint_cells_batch over slabs (tuples as in Batch)."""
        import torch
        d = self.device
        n = check_bank(initial_bank, "initial_bank", device=d)
        m = _check_slabs(slabs, n)
        check_tensor(sig_a, "sig_a", torch.float64, None, d)
        m = max(m, sig_a.numel())
        for name, t in (("sig_a", sig_a), ("sig_i", sig_i), ("g", g),
                        ("grad_sig_a", grad_sig_a), ("grad_sig_i", grad_sig_i)):
            check_tensor(t, name, torch.float64, m, d)
        check_tensor(history_score, "history_score", torch.float64, n, d, optional=True)
        arr = (Slab * len(slabs))(*[Slab(*s) for s in slabs])
        b = _bank_struct(initial_bank)
        st = Stats()
        opt = options if options is not None else make_options()
        rc = lib().mcgpu_adjoint_cells_batch(self._h, len(slabs), C.cast(arr, C.c_void_p),
                                             sig_a.numel(), C.byref(b), _ptr(sig_a), _ptr(sig_i),
                                             _ptr(g), _ptr(history_score), _ptr(grad_sig_a),
                                             _ptr(grad_sig_i), C.byref(opt), C.byref(st),
                                             _stream(stream))
        self.check(rc)
        return st.as_dict()

    # ------------------------------------------------------------ cycle loop
    def sync_cycle(self, ranks, sigs, absorption_rates, weights_absorbed, nb_particles_per_cycle,
                   nb_particles, max_cycles=0, options=None, statistics_cycle_time=None,
                   cycle_by_cycle=False):
        """mcgpu_sync_cycle: the synchronous multi-rank loop, every rank on
        this GPU.  `ranks`: one dict per rank with min_index, max_index, dx,
        particle_min_weight, left_border, right_border, cell_offset,
        particles (PARTICLE array, the initial bank), and optionally the lazy
        source nb_particles_create, seed, x_ini, wmc and nb_disabled.
        sigs / absorption_rates / weights_absorbed: host float64 arrays of
        all ranks' cells; weights_absorbed is added to in place.  Returns
        (per-rank output dicts, stats dict); with statistics_cycle_time the
        stats dict also holds "windows", the loop's statistics windows.
        cycle_by_cycle: never run a single rank's cycles concurrently."""
        assert weights_absorbed.dtype == np.float64 and weights_absorbed.flags.c_contiguous
        sigs = np.ascontiguousarray(sigs, np.float64)
        ar = np.ascontiguousarray(absorption_rates, np.float64)
        arr = (CycleRank * len(ranks))()
        keep = []
        for k, d in zip(arr, ranks):
            p = np.ascontiguousarray(d.get("particles", np.zeros(0, PARTICLE)), PARTICLE)
            keep.append(p)
            k.min_index, k.max_index = d["min_index"], d["max_index"]
            k.dx, k.particle_min_weight = d["dx"], d["particle_min_weight"]
            k.left_border, k.right_border = int(d["left_border"]), int(d["right_border"])
            k.cell_offset = d["cell_offset"]
            k.particles, k.n_particles = (p.ctypes.data if len(p) else None), len(p)
            k.nb_particles_create = d.get("nb_particles_create", 0)
            k.seed = d.get("seed", 0)
            k.x_ini, k.wmc = d.get("x_ini", 0.0), d.get("wmc", 0.0)
            k.nb_disabled = d.get("nb_disabled", 0)
        st = CycleStats()
        st.cycle_by_cycle = 1 if cycle_by_cycle else 0
        windows = []
        if statistics_cycle_time is not None:
            def _on_window(_user, w):
                windows.append(w.contents.as_dict())
            cb = CYCLE_WINDOW_FN(_on_window)
            st.statistics_cycle_time = statistics_cycle_time
            st.on_window = cb
        opt = options if options is not None else Options()
        rc = lib().mcgpu_sync_cycle(self._h, len(ranks), arr, sigs.ctypes.data, ar.ctypes.data,
                                    weights_absorbed.ctypes.data, len(weights_absorbed),
                                    nb_particles_per_cycle, nb_particles, max_cycles,
                                    C.byref(opt), C.byref(st))
        self.check(rc)
        out = [{f: getattr(k, f) for f in ("nb_disabled", "n_left", "n_right", "n_killed",
                                           "n_remaining", "nb_particles_create", "seed")}
               for k in arr]
        d = st.as_dict()
        if statistics_cycle_time is not None:
            d["windows"] = windows
        return out, d

    # ------------------------------------------------------------ helpers
    def rng_real_step(self, seeds, reals, stream=None):
        import torch
        check_tensor(seeds, "seeds", (torch.int64, torch.uint64), None, self.device)
        check_tensor(reals, "reals", torch.float64, seeds.numel(), self.device)
        self.check(lib().mcgpu_rng_real_step(self._h, seeds.numel(), _ptr(seeds), _ptr(reals),
                                             _stream(stream)))

    def rng_seed_chain(self, seed, out, stream=None):
        import torch
        check_tensor(out, "out", (torch.int64, torch.uint64), None, self.device)
        self.check(lib().mcgpu_rng_seed_chain(self._h, seed, out.numel(), _ptr(out),
                                              _stream(stream)))

    def create_particles(self, seed, first, bank, x_ini, wmc, dx, stream=None):
        """mcgpu_create_particles: fill the device bank (dict of tensors, n =
        its length) with source particles first .. first+n-1."""
        check_bank(bank, "bank", device=self.device)
        b = _bank_struct(bank)
        self.check(lib().mcgpu_create_particles(self._h, seed, first, bank["x"].numel(), x_ini,
                                                wmc, dx, C.byref(b), _stream(stream)))

    def math_eval(self, which, x, out, stream=None):
        w = {"log": 0, "exp": 1}[which]
        import torch
        check_tensor(x, "x", torch.float64, None, self.device)
        check_tensor(out, "out", torch.float64, x.numel(), self.device)
        self.check(lib().mcgpu_math_eval(self._h, w, x.numel(), _ptr(x), _ptr(out),
                                         _stream(stream)))


class Batch:
    """A batched plan (mcgpu_batch_*): many independent slabs transported by
    one persistent launch replayed from a captured hipGraph.

    slabs: sequence of (min_index, max_index, dx, particle_min_weight,
    cell_offset, particle_offset, n); sigs / absorption_rates / weights: the
    concatenated per-cell device arrays; bank: concatenated device bank."""

    def __init__(self, ctx, slabs, sigs, absorption_rates, weights, bank, result=None,
                 options=None):
        import torch
        self.ctx = ctx
        d = ctx.device
        n = check_bank(bank, "bank", device=d)
        m = _check_slabs(slabs, n)
        check_tensor(weights, "weights", torch.float64, None, d)
        m = max(m, weights.numel())
        for name, t in (("sigs", sigs), ("absorption_rates", absorption_rates),
                        ("weights", weights)):
            check_tensor(t, name, torch.float64, m, d)
        check_tensor(result, "result", torch.int8, n, d, optional=True)
        arr = (Slab * len(slabs))(*[Slab(*s) for s in slabs])
        self._keep = (sigs, absorption_rates, weights, bank, result)
        b = _bank_struct(bank)
        self._h = C.c_void_p()
        opt = options if options is not None else make_options()
        # the plan runs on its own stream: finish the torch work that built the inputs
        try:
            import torch
            torch.cuda.synchronize()
        except ImportError:
            pass
        rc = lib().mcgpu_batch_create(ctx._h, len(slabs), C.cast(arr, C.c_void_p), _ptr(sigs),
                                      _ptr(absorption_rates), _ptr(weights), weights.numel(),
                                      C.byref(b), _ptr(result), C.byref(opt), C.byref(self._h))
        ctx.check(rc)

    def run(self, stream=None, allow=()):
        st = Stats()
        rc = lib().mcgpu_batch_run(self._h, _stream(stream), C.byref(st))
        self.ctx.check(rc, allow)
        return st.as_dict()

    def autotune(self, n_candidates=0):
        times = (C.c_double * 16)()
        chosen = Options()
        self.ctx.check(lib().mcgpu_batch_autotune(self._h, n_candidates, times, C.byref(chosen)))
        n = n_candidates if n_candidates > 0 else 8
        return [times[i] for i in range(n)], options_dict(chosen)

    def options(self):
        o = Options()
        self.ctx.check(lib().mcgpu_batch_options(self._h, C.byref(o)))
        return options_dict(o)

    def close(self):
        if self._h:
            lib().mcgpu_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
