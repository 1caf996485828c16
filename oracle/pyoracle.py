"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the CPU oracle (liboracle.so)
and of the reference's own FP32 code compiled into oracle/_ref/libref.so.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline; the
product (mc-mpi_amd/) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

# Particle layouts (include/types/particle.hpp:7-17): FP32 24 B, FP64 40 B.
P32 = np.dtype([("seed", "<u8"), ("x", "<f4"), ("mu", "<f4"), ("wmc", "<f4"), ("index", "<i4")])
P64 = np.dtype([("seed", "<u8"), ("x", "<f8"), ("mu", "<f8"), ("wmc", "<f8"), ("index", "<i4"),
                ("pad_", "<i4")])
assert P32.itemsize == 24 and P64.itemsize == 40

SIGS, AR, WEIGHTS, PARTICLES, LEFT, RIGHT = range(6)
MATH_CR, MATH_GLIBC = 0, 1

_lib = None
_ref = None


def build(quiet=True):
    """Compile liboracle.so (and _ref/libref.so when the reference is present)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _open(path):
    L = C.CDLL(path)
    vp, i64, i32, dbl, flt = C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_float
    for pfx, real in (("orc32_", flt), ("orc64_", dbl)):
        getattr(L, pfx + "layer_new").restype = vp
        getattr(L, pfx + "layer_new").argtypes = [real, real, i32, i32, real]
        getattr(L, pfx + "decompose_domain").restype = vp
        getattr(L, pfx + "decompose_domain").argtypes = [real, real, real, i32, i32, i32, i32, real]
        getattr(L, pfx + "layer_free").argtypes = [vp]
        getattr(L, pfx + "create_particles").argtypes = [vp, real, real, i32, C.c_uint64]
        getattr(L, pfx + "simulate").argtypes = [vp, i32, i32]
        getattr(L, pfx + "dump_WA").argtypes = [vp, C.c_char_p]
        getattr(L, pfx + "dump_WA").restype = i32
        getattr(L, pfx + "layer_array").argtypes = [vp, i32, C.POINTER(vp)]
        getattr(L, pfx + "layer_array").restype = i64
        getattr(L, pfx + "layer_int").argtypes = [vp, i32]
        getattr(L, pfx + "layer_int").restype = i64
        getattr(L, pfx + "layer_real").argtypes = [vp, i32]
        getattr(L, pfx + "layer_real").restype = dbl
        getattr(L, pfx + "layer_push").argtypes = [vp, vp, i64]
        getattr(L, pfx + "layer_push").restype = i32
        getattr(L, pfx + "simulate_bank").restype = i32
        getattr(L, pfx + "layer_tally_fx").argtypes = [vp, vp, vp]
        getattr(L, pfx + "sync_run").argtypes = [vp, i32, i32, i32, i32, i64]
        getattr(L, pfx + "sync_run").restype = i64
    L.orc64_layer_set_tally_exact.argtypes = [vp, i32]
    L.orc64_simulate_bank_fx.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, i32]
    L.orc64_simulate_bank_fx.restype = i32
    L.orc32_simulate_bank.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, i32]
    L.orc_set_math_mode.argtypes = [i32]
    L.orc_rnd_seed_chain.argtypes = [C.c_uint64, i64, vp]
    L.orc32_rnd_real_step.argtypes = [vp, i64, vp]
    L.orc64_rnd_real_step.argtypes = [vp, i64, vp]
    L.orc64_math_eval.argtypes = [i32, vp, i64, vp]
    L.orc_fx_convert.argtypes = [vp, i64, vp, vp]
    L.orc_fx_finalize.argtypes = [vp, vp, i64, vp]
    L.orc_math_unresolved.restype = C.c_ulonglong
    L.orc64_forward_cells.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, i32]
    L.orc64_forward_cells.restype = i32
    L.orc64_adjoint_cells.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, i32]
    L.orc64_adjoint_cells.restype = i32
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _open(LIB_PATH)
    return _lib


def build_mutant(k, outdir):
    """Compile the oracle with planted mistake number k (ORC_MUTANT in
    orc_impl.inc) into outdir and return its ctypes handle.  Tests only."""
    path = os.path.join(outdir, f"liboracle_mutant{k}.so")
    cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           f"-DORC_MUTANT={k}", "-o", path, os.path.join(HERE, "oracle.c"),
           os.path.join(HERE, "orc_math.c"), "-lquadmath", "-lm"]
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError(out.stderr)
    return _open(path)


class using:
    """Context manager: route this module's calls to another oracle build."""

    def __init__(self, handle):
        self.handle = handle

    def __enter__(self):
        global _lib
        self.saved = lib()
        _lib = self.handle
        return self.handle

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own FP32 Layer (oracle/_ref/libref.so)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        vp, i64, i32, flt = C.c_void_p, C.c_int64, C.c_int, C.c_float
        R.ref_layer_decompose.restype = vp
        R.ref_layer_decompose.argtypes = [flt, flt, flt, i32, i32, i32, i32, flt]
        R.ref_layer_new.restype = vp
        R.ref_layer_new.argtypes = [flt, flt, i32, i32, flt]
        R.ref_layer_free.argtypes = [vp]
        R.ref_layer_simulate.argtypes = [vp, i32, i32]
        R.ref_layer_dump_WA.argtypes = [vp]
        R.ref_layer_array.argtypes = [vp, i32, C.POINTER(vp)]
        R.ref_layer_array.restype = i64
        R.ref_layer_int.argtypes = [vp, i32]
        R.ref_layer_int.restype = i64
        R.ref_layer_dx.argtypes = [vp]
        R.ref_layer_dx.restype = C.c_double
        R.ref_layer_push.argtypes = [vp, vp, i64]
        R.ref_transport_bank.argtypes = [vp, vp, i64, i32, i32, vp, vp, vp]
        R.ref_rnd_real.argtypes = [C.POINTER(C.c_ulonglong)]
        R.ref_rnd_real.restype = flt
        R.ref_rnd_seed.argtypes = [C.POINTER(C.c_ulonglong)]
        R.ref_rnd_seed.restype = C.c_ulonglong
        _ref = R
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data


class _LayerBase:
    """Common accessors for oracle / reference Layer handles."""

    dtype = None
    real = None

    def _array(self, which):
        p = C.c_void_p()
        n = self._arr_fn(self.h, which, C.byref(p))
        if n <= 0:
            return np.zeros(0, dtype=self.real if which <= WEIGHTS else self.dtype)
        dt = self.real if which <= WEIGHTS else self.dtype
        buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(p.value)
        return np.frombuffer(buf, dtype=dt, count=n)

    @property
    def sigs(self):
        return self._array(SIGS)

    @property
    def absorption_rates(self):
        return self._array(AR)

    @property
    def weights_absorbed(self):
        return self._array(WEIGHTS)

    @property
    def particles(self):
        return self._array(PARTICLES)

    @property
    def particles_left(self):
        return self._array(LEFT)

    @property
    def particles_right(self):
        return self._array(RIGHT)

    @property
    def nb_disabled(self):
        return int(self._int_fn(self.h, 0))

    @property
    def nb_active(self):
        return int(self._int_fn(self.h, 1))

    @property
    def left_border(self):
        return bool(self._int_fn(self.h, 2))

    @property
    def right_border(self):
        return bool(self._int_fn(self.h, 3))


class OracleLayer(_LayerBase):
    """Oracle restatement of Layer (precision 32 or 64)."""

    def __init__(self, handle, precision):
        self.h = handle
        self.precision = precision
        self.pfx = "orc32_" if precision == 32 else "orc64_"
        self.dtype = P32 if precision == 32 else P64
        self.real = np.float32 if precision == 32 else np.float64
        L = lib()
        self._arr_fn = getattr(L, self.pfx + "layer_array")
        self._int_fn = getattr(L, self.pfx + "layer_int")

    @classmethod
    def new(cls, precision, x_min, x_max, index_start, m, min_w):
        pfx = "orc32_" if precision == 32 else "orc64_"
        return cls(getattr(lib(), pfx + "layer_new")(x_min, x_max, index_start, m, min_w), precision)

    @classmethod
    def decompose_domain(cls, precision, x_min, x_max, x_ini, world_size, world_rank, nb_cells,
                         nb_particles, min_w):
        pfx = "orc32_" if precision == 32 else "orc64_"
        h = getattr(lib(), pfx + "decompose_domain")(x_min, x_max, x_ini, world_size, world_rank,
                                                     nb_cells, nb_particles, min_w)
        return cls(h, precision)

    def __del__(self):
        try:
            getattr(lib(), self.pfx + "layer_free")(self.h)
        except Exception:
            pass

    def create_particles(self, x_ini, wmc, n, seed):
        getattr(lib(), self.pfx + "create_particles")(self.h, x_ini, wmc, n, seed)

    def simulate(self, n, nthread=1):
        getattr(lib(), self.pfx + "simulate")(self.h, n, nthread)

    def dump_WA(self, path):
        return getattr(lib(), self.pfx + "dump_WA")(self.h, path.encode())

    def push(self, bank):
        bank = np.ascontiguousarray(bank, dtype=self.dtype)
        getattr(lib(), self.pfx + "layer_push")(self.h, bank.ctypes.data, len(bank))

    def set_tally_exact(self, on=True):
        assert self.precision == 64
        lib().orc64_layer_set_tally_exact(self.h, 1 if on else 0)

    def tally_exact_total(self):
        """Exact mode: the tally summed exactly over all simulate calls,
        rounded once to double."""
        m = int(self._int_fn(self.h, 5))
        pairs = np.zeros(2 * m, np.uint64)
        flags = np.zeros(m, np.uint32)
        lib().orc64_layer_tally_fx(self.h, pairs.ctypes.data, flags.ctypes.data)
        return fx_finalize(pairs, flags)

    @property
    def dx(self):
        return getattr(lib(), self.pfx + "layer_real")(self.h, 2)

    @property
    def x_min(self):
        return getattr(lib(), self.pfx + "layer_real")(self.h, 0)

    @property
    def x_max(self):
        return getattr(lib(), self.pfx + "layer_real")(self.h, 1)

    @property
    def events(self):
        return int(self._int_fn(self.h, 7))

    @property
    def collisions(self):
        return int(self._int_fn(self.h, 8))


def sync_run(precision, x_min, x_max, x_ini, world_size, nb_cells, nb_particles, min_w,
             nb_particles_per_cycle, nthread=1, tally_exact=False, max_cycles=10_000_000):
    """The synchronous worker loop (src/worker_sync.cpp:24-135) for all ranks
    of one decomposition in this process.  Returns (layers, cycles)."""
    layers = [OracleLayer.decompose_domain(precision, x_min, x_max, x_ini, world_size, r,
                                           nb_cells, nb_particles, min_w)
              for r in range(world_size)]
    if tally_exact:
        for L in layers:
            L.set_tally_exact(True)
    arr = (C.c_void_p * world_size)(*[L.h for L in layers])
    pfx = "orc32_" if precision == 32 else "orc64_"
    cycles = getattr(lib(), pfx + "sync_run")(C.cast(arr, C.c_void_p), world_size, nb_particles,
                                             nb_particles_per_cycle, nthread, max_cycles)
    if cycles < 0:
        raise RuntimeError(f"oracle sync_run failed ({cycles})")
    return layers, cycles


class RefLayer(_LayerBase):
    """The reference's own FP32 Layer, through oracle/_ref/libref.so."""

    dtype = P32
    real = np.float32

    def __init__(self, handle):
        self.h = handle
        R = ref()
        self._arr_fn = R.ref_layer_array
        self._int_fn = R.ref_layer_int

    @classmethod
    def new(cls, x_min, x_max, index_start, m, min_w):
        return cls(ref().ref_layer_new(x_min, x_max, index_start, m, min_w))

    @classmethod
    def decompose_domain(cls, x_min, x_max, x_ini, world_size, world_rank, nb_cells, nb_particles,
                         min_w):
        return cls(ref().ref_layer_decompose(x_min, x_max, x_ini, world_size, world_rank, nb_cells,
                                             nb_particles, min_w))

    def __del__(self):
        try:
            ref().ref_layer_free(self.h)
        except Exception:
            pass

    def simulate(self, n, nthread=1):
        ref().ref_layer_simulate(self.h, n, nthread)

    def push(self, bank):
        bank = np.ascontiguousarray(bank, dtype=P32)
        ref().ref_layer_push(self.h, bank.ctypes.data, len(bank))

    def dump_WA_in(self, directory):
        """Layer::dump_WA writes ./WA.out; run it inside `directory`."""
        cwd = os.getcwd()
        os.chdir(directory)
        try:
            ref().ref_layer_dump_WA(self.h)
        finally:
            os.chdir(cwd)
        return os.path.join(directory, "WA.out")

    @property
    def dx(self):
        return ref().ref_layer_dx(self.h)

    def transport_bank(self, bank, index_start, m):
        """Per-history reference transport of a bank (processed back to front)."""
        bank = np.array(bank, dtype=P32, copy=True)
        n = len(bank)
        result = np.zeros(n, np.int32)
        events = np.zeros(n, np.uint64)
        tally = np.zeros(len(self.weights_absorbed), np.float32)
        ref().ref_transport_bank(self.h, bank.ctypes.data, n, index_start, m, result.ctypes.data,
                                 events.ctypes.data, tally.ctypes.data)
        return bank, result, events, tally


# ------------------------------------------------------------------ bank level

def params(precision, index_start, m, dx, min_w):
    if precision == 32:
        class Params(C.Structure):
            _fields_ = [("index_start", C.c_int32), ("m", C.c_int32), ("dx", C.c_float),
                        ("min_weight", C.c_float)]
    else:
        class Params(C.Structure):
            _fields_ = [("index_start", C.c_int32), ("m", C.c_int32), ("dx", C.c_double),
                        ("min_weight", C.c_double)]
    return Params(index_start, m, dx, min_w)


def simulate_bank(precision, index_start, m, dx, min_w, sigs, ar, bank, nthread=1,
                  weights=None, exact=True, math=MATH_CR):
    """Transport a bank (in place on a copy) like the OpenMP loop of
    Layer::simulate.  Returns dict with the final bank, result, events,
    collisions, the real_t tally and (FP64) the exact fixed-point tally."""
    L = lib()
    real = np.float32 if precision == 32 else np.float64
    dt = P32 if precision == 32 else P64
    bank = np.array(bank, dtype=dt, copy=True)
    n = len(bank)
    sigs = np.ascontiguousarray(sigs, dtype=real)
    ar = np.ascontiguousarray(ar, dtype=real)
    w = np.zeros(m, real) if weights is None else np.array(weights, dtype=real, copy=True)
    result = np.zeros(n, np.int32)
    events = np.zeros(n, np.uint64)
    coll = np.zeros(n, np.uint64)
    prm = params(precision, index_start, m, dx, min_w)
    out = dict(bank=bank, result=result, events=events, collisions=coll, weights=w)
    if precision == 32:
        rc = L.orc32_simulate_bank(C.addressof(prm), sigs.ctypes.data, ar.ctypes.data,
                                   bank.ctypes.data, n, w.ctypes.data, None, None,
                                   result.ctypes.data, events.ctypes.data, coll.ctypes.data,
                                   nthread)
    else:
        L.orc_set_math_mode(math)
        fx = np.zeros(2 * m, np.uint64) if exact else None
        fl = np.zeros(m, np.uint32) if exact else None
        rc = L.orc64_simulate_bank_fx(C.addressof(prm), sigs.ctypes.data, ar.ctypes.data,
                                      bank.ctypes.data, n, w.ctypes.data, _ptr(fx), _ptr(fl),
                                      result.ctypes.data, events.ctypes.data, coll.ctypes.data,
                                      nthread)
        L.orc_set_math_mode(MATH_CR)
        if exact:
            out["fx"] = fx
            out["fx_flags"] = fl
            out["tally_exact"] = fx_finalize(fx, fl)
    if rc != 0:
        raise MemoryError("oracle simulate_bank failed")
    return out


def fx_finalize(fx, flags):
    m = len(flags)
    out = np.zeros(m, np.float64)
    lib().orc_fx_finalize(np.ascontiguousarray(fx, np.uint64).ctypes.data,
                          np.ascontiguousarray(flags, np.uint32).ctypes.data, m, out.ctypes.data)
    return out


def fx_convert(values):
    v = np.ascontiguousarray(values, np.float64)
    out = np.zeros(2 * len(v), np.uint64)
    fl = np.zeros(len(v), np.uint32)
    lib().orc_fx_convert(v.ctypes.data, len(v), out.ctypes.data, fl.ctypes.data)
    return out, fl


def rnd_seed_chain(seed, n):
    out = np.zeros(n, np.uint64)
    lib().orc_rnd_seed_chain(seed, n, out.ctypes.data)
    return out


def rnd_real_step(seeds, precision=64):
    s = np.array(seeds, dtype=np.uint64, copy=True)
    out = np.zeros(len(s), np.float32 if precision == 32 else np.float64)
    fn = lib().orc32_rnd_real_step if precision == 32 else lib().orc64_rnd_real_step
    fn(s.ctypes.data, len(s), out.ctypes.data)
    return s, out


def math_eval(which, x, math=MATH_CR):
    """which: 'log' or 'exp'; correctly rounded (default) or glibc."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(len(x), np.float64)
    lib().orc_set_math_mode(math)
    lib().orc64_math_eval(0 if which == "log" else 1, x.ctypes.data, len(x), out.ctypes.data)
    lib().orc_set_math_mode(MATH_CR)
    return out


def math_unresolved():
    return int(lib().orc_math_unresolved())


def forward_cells(index_start, m, dx, min_w, sig_a, sig_i, bank, g=None, nthread=8):
    """FP64 forward with per-cell (sig_a, sig_i); returns final bank, result,
    exact tally and (if g) per-history scores."""
    L = lib()
    bank = np.array(bank, dtype=P64, copy=True)
    n = len(bank)
    sa = np.ascontiguousarray(sig_a, np.float64)
    si = np.ascontiguousarray(sig_i, np.float64)
    fx = np.zeros(2 * m, np.uint64)
    fl = np.zeros(m, np.uint32)
    gg = None if g is None else np.ascontiguousarray(g, np.float64)
    score = None if g is None else np.zeros(n, np.float64)
    result = np.zeros(n, np.int32)
    prm = params(64, index_start, m, dx, min_w)
    rc = L.orc64_forward_cells(C.addressof(prm), sa.ctypes.data, si.ctypes.data, bank.ctypes.data,
                               n, fx.ctypes.data, fl.ctypes.data, _ptr(gg), _ptr(score),
                               result.ctypes.data, nthread)
    if rc:
        raise MemoryError("oracle forward_cells failed")
    return dict(bank=bank, result=result, tally_exact=fx_finalize(fx, fl), score=score)


def adjoint_cells(index_start, m, dx, min_w, sig_a, sig_i, initial, g, score=None, nthread=8):
    """FP64 adjoint (same estimator as the GPU); returns (grad_a, grad_i)."""
    L = lib()
    init = np.ascontiguousarray(initial, dtype=P64)
    sa = np.ascontiguousarray(sig_a, np.float64)
    si = np.ascontiguousarray(sig_i, np.float64)
    gg = np.ascontiguousarray(g, np.float64)
    sc = None if score is None else np.ascontiguousarray(score, np.float64)
    ga = np.zeros(2 * m, np.uint64)
    gi = np.zeros(2 * m, np.uint64)
    fl = np.zeros(m, np.uint32)
    prm = params(64, index_start, m, dx, min_w)
    rc = L.orc64_adjoint_cells(C.addressof(prm), sa.ctypes.data, si.ctypes.data, init.ctypes.data,
                               len(init), gg.ctypes.data, _ptr(sc), ga.ctypes.data,
                               gi.ctypes.data, fl.ctypes.data, nthread)
    if rc:
        raise MemoryError("oracle adjoint_cells failed")
    return fx_finalize(ga, fl), fx_finalize(gi, fl)
