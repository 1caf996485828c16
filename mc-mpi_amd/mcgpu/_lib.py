"""ctypes declarations of the C ABI in include/mcgpu.h (libmcgpu.so).

The library is loaded from the package directory (mc-mpi_amd/libmcgpu.so).
If it is missing the import fails with an explicit error: there is no
fallback path.
"""
import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MCGPU_LIB selects another build of the library (e.g. for A/B timing)
LIB_PATH = os.environ.get("MCGPU_LIB") or os.path.join(PKG_ROOT, "libmcgpu.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_NOMEM = -3
ERR_NO_DEVICE = -4
ERR_EVENT_CAP = -5
ERR_RANGE = -6

KERNEL_HISTORY = 0
KERNEL_EVENT = 1

NOALLOC_MAX_CELLS = 65536  # include/mcgpu.h MCGPU_NOALLOC_MAX_CELLS

# include/mcgpu.h mcgpu_particle (include/types/particle.hpp:7-17, FP64): 40 B
PARTICLE = np.dtype([("seed", "<u8"), ("x", "<f8"), ("mu", "<f8"), ("wmc", "<f8"),
                     ("index", "<i4"), ("pad_", "<i4")])
assert PARTICLE.itemsize == 40


class Bank(C.Structure):
    _fields_ = [("seed", C.c_void_p), ("x", C.c_void_p), ("mu", C.c_void_p), ("wmc", C.c_void_p),
                ("index", C.c_void_p)]


class Options(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("block_size", C.c_int32), ("grid_blocks", C.c_int32),
                ("events_per_refill", C.c_int32), ("compaction_interval", C.c_int32),
                ("particles_per_lane", C.c_int32), ("max_events_per_history", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("events", C.c_uint64), ("collisions", C.c_uint64), ("n_left", C.c_int64),
                ("n_right", C.c_int64), ("n_killed", C.c_int64), ("kernel_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Slab(C.Structure):
    _fields_ = [("min_index", C.c_int32), ("max_index", C.c_int32), ("dx", C.c_double),
                ("particle_min_weight", C.c_double), ("cell_offset", C.c_int64),
                ("particle_offset", C.c_int64), ("n", C.c_int64)]


class CycleRank(C.Structure):
    """include/mcgpu.h mcgpu_cycle_rank: one rank's layer in mcgpu_sync_cycle."""
    _fields_ = [("min_index", C.c_int32), ("max_index", C.c_int32), ("dx", C.c_double),
                ("particle_min_weight", C.c_double), ("left_border", C.c_int32),
                ("right_border", C.c_int32), ("cell_offset", C.c_int64),
                ("particles", C.c_void_p), ("n_particles", C.c_int64),
                ("nb_particles_create", C.c_int64), ("seed", C.c_uint64), ("x_ini", C.c_double),
                ("wmc", C.c_double), ("nb_disabled", C.c_int64), ("n_left", C.c_int64),
                ("n_right", C.c_int64), ("n_killed", C.c_int64), ("n_remaining", C.c_int64)]


class CycleWindow(C.Structure):
    """include/mcgpu.h mcgpu_cycle_window: one statistics window of the loop."""
    _fields_ = [("starttime", C.c_double), ("endtime", C.c_double), ("time_comp", C.c_double),
                ("time_send", C.c_double), ("time_recv", C.c_double), ("time_idle", C.c_double),
                ("nb_cycles", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


CYCLE_WINDOW_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(CycleWindow))


class CycleStats(C.Structure):
    _fields_ = [("cycles", C.c_int64), ("events", C.c_uint64), ("collisions", C.c_uint64),
                ("kernel_ms", C.c_double), ("wall_ms", C.c_double),
                ("statistics_cycle_time", C.c_double), ("on_window", CYCLE_WINDOW_FN),
                ("window_user", C.c_void_p), ("cycle_by_cycle", C.c_int32), ("pad_", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k in ("cycles", "events", "collisions", "kernel_ms",
                                              "wall_ms")}


class McgpuError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"mcgpu status {status}: {message}")
        self.status = status


_lib = None

# name -> (restype, argtypes); every symbol include/mcgpu.h declares
vp, i32, i64, u64, dbl, cp = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.c_char_p
SIGNATURES = {
    "mcgpu_context_create": (i32, [C.c_int, C.POINTER(vp)]),
    "mcgpu_context_destroy": (i32, [vp]),
    "mcgpu_status_string": (cp, [C.c_int]),
    "mcgpu_context_error": (cp, [vp]),
    "mcgpu_default_options": (None, [C.POINTER(Options)]),
    "mcgpu_version": (cp, []),
    "mcgpu_simulate": (i32, [vp, i32, vp, vp, vp, vp, i32, i32, dbl, dbl, vp,
                             C.POINTER(Options), C.POINTER(Stats)]),
    "mcgpu_simulate_device": (i32, [vp, i64, C.POINTER(Bank), vp, vp, vp, i32, i32, dbl, dbl,
                                    vp, vp, vp, C.POINTER(Options), C.POINTER(Stats), vp]),
    "mcgpu_simulate_device_noalloc": (i32, [i64, C.POINTER(Bank), vp, vp, vp, i32, i32, dbl,
                                            dbl, vp, vp, vp, C.POINTER(Options),
                                            C.POINTER(Stats), vp]),
    "mcgpu_batch_create": (i32, [vp, i32, vp, vp, vp, vp, i64, C.POINTER(Bank), vp,
                                 C.POINTER(Options), C.POINTER(vp)]),
    "mcgpu_batch_run": (i32, [vp, vp, C.POINTER(Stats)]),
    "mcgpu_batch_destroy": (i32, [vp]),
    "mcgpu_batch_options": (i32, [vp, C.POINTER(Options)]),
    "mcgpu_batch_autotune": (i32, [vp, i32, vp, C.POINTER(Options)]),
    "mcgpu_forward_cells": (i32, [vp, i64, C.POINTER(Bank), vp, vp, vp, i32, i32, dbl, dbl, vp,
                                  vp, vp, C.POINTER(Options), C.POINTER(Stats), vp]),
    "mcgpu_adjoint_cells": (i32, [vp, i64, C.POINTER(Bank), vp, vp, i32, i32, dbl, dbl, vp, vp,
                                  vp, vp, C.POINTER(Options), C.POINTER(Stats), vp]),
    "mcgpu_forward_cells_batch": (i32, [vp, i32, vp, i64, C.POINTER(Bank), vp, vp, vp, vp, vp,
                                        vp, C.POINTER(Options), C.POINTER(Stats), vp]),
    "mcgpu_adjoint_cells_batch": (i32, [vp, i32, vp, i64, C.POINTER(Bank), vp, vp, vp, vp, vp,
                                        vp, C.POINTER(Options), C.POINTER(Stats), vp]),
    "mcgpu_rng_real_step": (i32, [vp, i64, vp, vp, vp]),
    "mcgpu_rng_seed_chain": (i32, [vp, u64, i64, vp, vp]),
    "mcgpu_math_eval": (i32, [vp, C.c_int, i64, vp, vp, vp]),
    "mcgpu_create_particles": (i32, [vp, u64, i64, i64, C.c_double, C.c_double, C.c_double, vp,
                                     vp]),
    "mcgpu_sync_cycle": (i32, [vp, i32, C.POINTER(CycleRank), vp, vp, vp, i64, i32, i64, i64,
                               C.POINTER(Options), C.POINTER(CycleStats)]),
}


def lib():
    """Load libmcgpu.so; raises ImportError when the HIP library is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not found: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()' or make -C mc-mpi_amd)")
        # libmcgpu.so and torch both need libamdhip64.so.7; the process can
        # hold only one library of that name, and torch initialises its GPUs
        # only with the runtime it ships.  Load torch's first so both share it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def status_string(status):
    return lib().mcgpu_status_string(status).decode()
