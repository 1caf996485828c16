"""Synthetic inputs of the five benchmark workloads (cross sections, source
banks, layer geometry), generated deterministically from integer seeds.

Random numbers come from splitmix64 (Steele, Lea, Flood 2014) evaluated with
wrapping uint64 arithmetic in numpy; uniforms are (x >> 11) * 2^-53.  The
per-cell log-normal draws use Box-Muller with Python's `math` functions
element by element (at most 10^4 cells per layer), and the particle banks use
only IEEE-exact operations (sqrt, +, *), so the generated arrays are the same
on every machine.

Cross sections are stored the way the reference stores them
(include/layer/layer.hpp:101-104): sigs = total cross section and
absorption_rates = absorbed fraction, from which the transport step forms
sig_a = sigs*ar and sig_i = sigs*(1.0 - ar) (src/layer.cpp:131-133).

Particle seeds are the reference's seed chain rnd_seed(&s) from s = 5127801
(src/layer.cpp:36, src/random.cpp:24-29), the i-th particle taking the i-th
output.
"""
import math

import numpy as np

from ._lib import PARTICLE

MASK63 = np.uint64(0x7FFFFFFFFFFFFFFF)
RNGS_G = np.uint64(5177284530976225183)
RNGS_C = np.uint64(2096348467109453893)
ROOT_SEED = 5127801


def splitmix64(seed, n, offset=0):
    """n outputs of splitmix64 starting at state seed + (offset+1)*gamma."""
    with np.errstate(over="ignore"):
        gamma = np.uint64(0x9E3779B97F4A7C15)
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * gamma
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, n, offset=0):
    """uniforms in [0, 1) with 53 random bits."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def _u64(v):
    """uint64 constant as the int64 with the same bits (torch has no uint64 arithmetic)."""
    return v - (1 << 64) if v >= 1 << 63 else v


def _srl(z, k):
    """Logical right shift of int64 tensors holding uint64 bits."""
    return (z >> k) & ((1 << (64 - k)) - 1)


def uniform_torch(seed, n, offset=0, device="cuda"):
    """uniform(seed, n, offset) computed with torch on `device`: the same
    splitmix64 outputs in wrapping int64 arithmetic (two's complement, so the
    bits equal the uint64 ones), the same (x >> 11) * 2^-53."""
    import torch
    i = torch.arange(offset + 1, offset + n + 1, dtype=torch.int64, device=device)
    z = _u64(seed) + i * _u64(0x9E3779B97F4A7C15)
    z = (z ^ _srl(z, 30)) * _u64(0xBF58476D1CE4E5B9)
    z = (z ^ _srl(z, 27)) * _u64(0x94D049BB133111EB)
    z = z ^ _srl(z, 31)
    return _srl(z, 11).to(torch.float64) * 2.0 ** -53


def lognormal_cells(seed, n, mu, sigma):
    """n log-normal(mu, sigma) values via Box-Muller with Python math."""
    u = uniform(seed, 2 * n)
    out = np.empty(n, np.float64)
    for i in range(n):
        u1 = 1.0 - u[2 * i]  # (0, 1]
        u2 = u[2 * i + 1]
        z = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
        out[i] = math.exp(mu + sigma * z)
    return out


def seed_chain(n, root=ROOT_SEED, start=0):
    """Outputs start+1 .. start+n of rnd_seed from `root` (affine jump-ahead,
    vectorised; bit-identical to calling rnd_seed n times)."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        A = np.full(n, RNGS_G, np.uint64)
        B = np.full(n, RNGS_C, np.uint64)
        RA = np.ones(n, np.uint64)
        RB = np.zeros(n, np.uint64)
        while np.any(k):
            odd = (k & np.uint64(1)).astype(bool)
            RA = np.where(odd, A * RA, RA)
            RB = np.where(odd, A * RB + B, RB)
            B = A * B + B
            A = A * A
            k = k >> np.uint64(1)
        return (RA * np.uint64(root) + RB) & MASK63


class Workload:
    """One layer + source bank: everything a Layer::simulate call consumes."""

    def __init__(self, name, x_min, x_max, index_start, m, dx, min_weight, sigs, ar, bank,
                 left_border, right_border, note=""):
        self.name = name
        self.x_min, self.x_max = x_min, x_max
        self.index_start, self.m, self.dx = index_start, m, dx
        self.min_weight = min_weight
        self.sigs, self.absorption_rates = sigs, ar
        self.bank = bank
        self.left_border, self.right_border = left_border, right_border
        self.note = note

    @property
    def min_index(self):
        return self.index_start

    @property
    def max_index(self):
        return self.index_start + self.m


def _bank(n, x, index, mu, wmc, seed_start=0):
    b = np.zeros(n, dtype=PARTICLE)
    b["seed"] = seed_chain(n, start=seed_start)
    b["x"] = x
    b["index"] = index
    b["mu"] = mu
    b["wmc"] = wmc
    return b


def slab100(n=100_000):
    """1-D slab [0,1], 100 cells; sigma_a = 0.5 + 0.5 x_c, sigma_s = 1 - 0.5 x_c
    (so sigs = 1.5 and ar = sigma_a / 1.5); n particles at x = 0, mu = 1,
    weight 1; minimum weight 1e-2."""
    m = 100
    dx = (1.0 - 0.0) / m  # Layer::dx (src/layer.cpp:47)
    xc = (np.arange(m) + 0.5) * dx
    sigs = np.full(m, 1.5)
    ar = (0.5 + 0.5 * xc) / 1.5
    bank = _bank(n, 0.0, 0, 1.0, 1.0)
    return Workload("slab100", 0.0, 1.0, 0, m, dx, 1e-2, sigs, ar, bank, True, True)


def slab10k(n=100_000_000, seed=2, with_bank=True):
    """10 000 cells over [0,1]; sigma_a ~ LogNormal(0, 0.5), sigma_s ~
    LogNormal(0.5, 0.5) per cell; n particles at x = 0 with mu = sqrt(U),
    weight 1; minimum weight 1e-3.  with_bank=False skips the bank (the
    benchmark builds it on the device, see slab10k_mu)."""
    m = 10_000
    dx = 1.0 / m
    sa = lognormal_cells(seed * 1000 + 1, m, 0.0, 0.5)
    ss = lognormal_cells(seed * 1000 + 2, m, 0.5, 0.5)
    sigs = sa + ss
    ar = sa / sigs
    bank = _bank(n, 0.0, 0, slab10k_mu(n, seed), 1.0) if with_bank else None
    return Workload("slab10k", 0.0, 1.0, 0, m, dx, 1e-3, sigs, ar, bank, True, True)


def slab10k_mu(n, seed=2, offset=0):
    """Directions of the 10,000-cell slab's source particles offset .. offset+n-1."""
    return np.sqrt(uniform(seed * 1000 + 3, n, offset))


def layer_rank2of5(n=20_000_000, seed=3):
    """One rank's share of a 5-way decomposed 5000-cell slab: rank 2 owns
    cells [2000, 3000) = [0.4, 0.6).  Geometry from decompose_domain
    (src/layer.cpp:17-42) in FP64: dx_g = 1/5000, x_min = 2000 dx_g,
    x_max = 3000 dx_g, layer dx = (x_max - x_min)/1000.  Cross sections: the
    100-cell slab's law at the layer's cell centres x_min + i dx + dx/2.  Inflow: n/2
    particles at x_min (index 2000, mu = +sqrt(U)) and n/2 at x_max (index
    2999, mu = -sqrt(U)), weights U(0.1, 1); minimum weight 1e-2."""
    nb_cells, R, rank = 5000, 5, 2
    cells_per_layer = nb_cells // R
    start = rank * cells_per_layer
    m = cells_per_layer
    dx_g = (1.0 - 0.0) / float(nb_cells)
    x_min = 0.0 + start * dx_g
    x_max = 0.0 + (start + m) * dx_g
    dx = (x_max - x_min) / m
    xc = x_min + np.arange(m) * dx + 0.5 * dx
    sigs = np.full(m, 1.5)
    ar = (0.5 + 0.5 * xc) / 1.5
    u = uniform(seed * 1000 + 1, n)
    w = 0.1 + 0.9 * uniform(seed * 1000 + 2, n)
    half = n // 2
    x = np.where(np.arange(n) < half, x_min, x_max)
    index = np.where(np.arange(n) < half, start, start + m - 1).astype(np.int32)
    mu = np.where(np.arange(n) < half, np.sqrt(u), -np.sqrt(u))
    bank = _bank(n, x, index, mu, w)
    return Workload("layer_rank2of5", x_min, x_max, start, m, dx, 1e-2, sigs, ar, bank, False, False)


def adjoint1000(n=10_000_000, seed=5):
    """1000 cells over [0,1]; sigma_a, sigma_s ~ LogNormal(0, 0.5); n particles
    at x = 0 with mu = sqrt(U), weight 1; minimum weight 1e-3.  Target of the
    gradient: absorbed weight in cells 800-899."""
    m = 1000
    dx = 1.0 / m
    sa = lognormal_cells(seed * 1000 + 1, m, 0.0, 0.5)
    ss = lognormal_cells(seed * 1000 + 2, m, 0.0, 0.5)
    sigs = sa + ss
    ar = sa / sigs
    mu = np.sqrt(uniform(seed * 1000 + 3, n))
    bank = _bank(n, 0.0, 0, mu, 1.0)
    wl = Workload("adjoint1000", 0.0, 1.0, 0, m, dx, 1e-3, sigs, ar, bank, True, True)
    wl.sigma_a, wl.sigma_s = sa, ss
    wl.target = np.zeros(m)
    wl.target[800:900] = 1.0
    return wl


def adjoint1000_batch(n_slabs=64, n_per_slab=10_000_000):
    """The 1,000-cell adjoint slab batched: n_slabs independent 1000-cell slabs on [0,1]; slab s
    has its own LogNormal(0, 0.5) sigma_a, sigma_s (seeds 5000*s + 5001 and
    5000*s + 5002, so slab 0 has the 1,000-cell adjoint slab's cross sections) and n_per_slab
    particles at x = 0, weight 1, minimum weight 1e-3; the target is cells
    800-899 of every slab.  Particle i of the concatenated bank takes output
    i+1 of the reference seed chain and mu = sqrt(uniform(5004, .)[i]).
    Returns (slabs, sigma_a, sigma_s, g, total_particles) in the layout of
    mcgpu_forward_cells_batch (slab tuples as in slabs4096_layout)."""
    m = 1000
    slabs, sa, ss, g = [], [], [], []
    for s in range(n_slabs):
        sa.append(lognormal_cells(5000 * s + 5001, m, 0.0, 0.5))
        ss.append(lognormal_cells(5000 * s + 5002, m, 0.0, 0.5))
        t = np.zeros(m)
        t[800:900] = 1.0
        g.append(t)
        slabs.append((0, m, 1.0 / m, 1e-3, s * m, s * n_per_slab, n_per_slab))
    return slabs, np.concatenate(sa), np.concatenate(ss), np.concatenate(g), n_slabs * n_per_slab


def adjoint1000_batch_mu(n, offset=0):
    """Directions of the batched 1,000-cell slabs' particles offset .. offset+n-1."""
    return np.sqrt(uniform(5004, n, offset))


def slabs4096_layout(n_slabs=4096, seed=4, particle_scale=1.0):
    """4096 independent slabs on [0,1]: cells log-uniform in [32, 8192],
    particles log-uniform in [1e3, 1e6] (times particle_scale, at least 1),
    per cell sigma_t ~ LogNormal(0, 1) and scattering ratio c ~ U(0.1, 0.99)
    stored as sigs = sigma_t, ar = 1 - c; source at x = 0, mu = sqrt(U),
    weight 1; minimum weight 1e-3.  Slab s owns cells [cell_offset, +m) of the
    concatenated sigs/ar and particles [particle_offset, +n) of the
    concatenated bank, whose seeds are outputs particle_offset+1.. of the
    reference seed chain.  Returns (slabs, sigs, ar, total_particles) with
    slabs as tuples (min_index, max_index, dx, min_weight, cell_offset,
    particle_offset, n)."""
    u = uniform(seed * 1000 + 1, 2 * n_slabs)
    ms = np.floor(np.exp(np.log(32.0) + u[0::2] * (np.log(8192.0) - np.log(32.0)))).astype(np.int64)
    ns = np.floor(np.exp(np.log(1e3) + u[1::2] * (np.log(1e6) - np.log(1e3))) * particle_scale)
    ns = np.maximum(ns, 1).astype(np.int64)
    total_cells = int(ms.sum())
    st = lognormal_cells(seed * 1000 + 2, total_cells, 0.0, 1.0)
    c = 0.1 + 0.89 * uniform(seed * 1000 + 3, total_cells)
    slabs = []
    coff = poff = 0
    for s in range(n_slabs):
        m, n = int(ms[s]), int(ns[s])
        slabs.append((0, m, 1.0 / m, 1e-3, coff, poff, n))
        coff += m
        poff += n
    return slabs, st, 1.0 - c, poff


def slabs4096_mu(n, offset=0, seed=4):
    """Directions of the 4096-slab batch's source particles offset .. offset+n-1."""
    return np.sqrt(uniform(seed * 1000 + 4, n, offset))


def slabs4096_bank(total_particles, offset=0, count=None, seed=4):
    """Host bank of the 4096-slab batch's particles offset .. offset+count-1."""
    count = total_particles - offset if count is None else count
    return _bank(count, 0.0, 0, slabs4096_mu(count, offset, seed), 1.0, seed_start=offset)
