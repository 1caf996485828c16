"""Pins of the oracle's adjoint (orc64_forward_cells / orc64_adjoint_cells),
independent of its own formulas:

  * a pure absorber at normal incidence is deterministic, and its target
    tally has a closed-form gradient;
  * the per-history scores of the forward add up to the weighted tally;
  * replaying a history to recompute its score gives the same gradient as
    passing the stored score;
  * central finite differences with common random numbers (same bank and
    seeds) agree with the adjoint within 3 standard errors (sig_a with
    Richardson extrapolation, whose truncation error is O(h^4));
  * oracles with a planted adjoint mistake fail these checks.
"""
import numpy as np
import pytest

import pyoracle as po
from mcgpu import configs


def _bank(n, mu, root=12345, w=1.0):
    b = np.zeros(n, po.P64)
    b["seed"] = configs.seed_chain(n, root=root)
    b["mu"] = mu
    b["wmc"] = w
    return b


def _absorber_case():
    m, dx = 50, 1.0 / 50
    sa = np.random.default_rng(3).uniform(0.5, 3.0, m)
    si = np.zeros(m)                    # sig_i <= EPS: flights are never sampled
    g = np.zeros(m)
    g[30:40] = 1.0
    return m, dx, sa, si, g, _bank(64, 1.0)


def _absorber_gradient(m, dx, sa, n):
    """T = W_29 - W_39 per particle with W_c = exp(-dx sum_{k<=c} sa_k):
    dT/dsa_j = dx (W_39 [j <= 39] - W_29 [j <= 29])."""
    W = np.exp(-dx * np.cumsum(sa))
    j = np.arange(m)
    return n * dx * (W[39] * (j <= 39) - W[29] * (j <= 29))


def test_pure_absorber_gradient_closed_form(oracle):
    m, dx, sa, si, g, bank = _absorber_case()
    f = oracle.forward_cells(0, m, dx, 0.0, sa, si, bank, g=g)
    W = np.exp(-dx * np.cumsum(sa))
    assert np.allclose(f["score"], W[29] - W[39], rtol=1e-13, atol=0)
    ga, gi = oracle.adjoint_cells(0, m, dx, 0.0, sa, si, bank, g, score=f["score"])
    want = _absorber_gradient(m, dx, sa, len(bank))
    assert np.allclose(ga, want, rtol=1e-12, atol=1e-15)
    assert np.all(ga[40:] == 0) and np.all(gi == 0)


def _small_case(n):
    m, dx = 20, 1.0 / 20
    rng = np.random.default_rng(11)
    sa = rng.lognormal(0.0, 0.5, m)
    si = rng.lognormal(0.5, 0.5, m)
    g = np.zeros(m)
    g[12:16] = 1.0
    bank = _bank(n, np.sqrt(rng.uniform(0.0, 1.0, n)))
    return m, dx, sa, si, g, bank


def test_history_scores_add_up_to_the_weighted_tally(oracle):
    m, dx, sa, si, g, bank = _small_case(20_000)
    f = oracle.forward_cells(0, m, dx, 1e-3, sa, si, bank, g=g)
    total = float(np.dot(g, f["tally_exact"]))
    assert abs(f["score"].sum() - total) <= 1e-12 * abs(total)


def test_replay_without_scores_equals_stored_scores(oracle):
    m, dx, sa, si, g, bank = _small_case(5_000)
    f = oracle.forward_cells(0, m, dx, 1e-3, sa, si, bank, g=g)
    a1, i1 = oracle.adjoint_cells(0, m, dx, 1e-3, sa, si, bank, g, score=f["score"])
    a2, i2 = oracle.adjoint_cells(0, m, dx, 1e-3, sa, si, bank, g, score=None)
    assert np.array_equal(a1, a2) and np.array_equal(i1, i2)


def _fd_z(lib, which, cells, n=200_000, K=20):
    """Largest |mean(adjoint - FD)| / SE over the cells, histories split into
    K groups; sig_a: Richardson-extrapolated central differences at relative
    steps 1e-4 and 2e-4; sig_i: central differences at relative step 1e-2
    (a sig_i change also switches some events between collision and crossing,
    which a much smaller step would not sample)."""
    with po.using(lib):
        m, dx, sa, si, g, bank = _small_case(n)
        f = po.forward_cells(0, m, dx, 1e-3, sa, si, bank, g=g)
        e = np.linspace(0, n, K + 1).astype(int)
        adj = np.array([po.adjoint_cells(0, m, dx, 1e-3, sa, si, bank[e[k]:e[k + 1]], g,
                                         score=f["score"][e[k]:e[k + 1]])[which == "i"]
                        for k in range(K)])
        base = sa if which == "a" else si

        def fd(c, rel):
            up, dn = base.copy(), base.copy()
            up[c] += rel * base[c]
            dn[c] -= rel * base[c]
            args = (up, si) if which == "a" else (sa, up)
            s_up = po.forward_cells(0, m, dx, 1e-3, *args, bank, g=g)["score"]
            args = (dn, si) if which == "a" else (sa, dn)
            s_dn = po.forward_cells(0, m, dx, 1e-3, *args, bank, g=g)["score"]
            h = (s_up - s_dn) / (up[c] - dn[c])
            return np.array([h[e[k]:e[k + 1]].sum() for k in range(K)])

        worst = 0.0
        for c in cells:
            if which == "a":
                est = (4 * fd(c, 1e-4) - fd(c, 2e-4)) / 3
            else:
                est = fd(c, 1e-2)
            d = adj[:, c] - est
            se = d.std(ddof=1) / np.sqrt(K)
            slack = 1e-11 * abs(est.sum()) / K
            worst = max(worst, abs(d.mean()) / (se + slack))
        return worst


CELLS = (2, 10, 13, 15, 18)


@pytest.mark.parametrize("which", ["a", "i"])
def test_adjoint_within_3_sigma_of_finite_differences(oracle, which):
    z = _fd_z(oracle.lib(), which, CELLS)
    assert z < 3.0, z


@pytest.mark.parametrize("mutant,pin", [(6, "closed_form"), (7, "fd_i"), (8, "fd_i")])
def test_adjoint_pins_catch_planted_mistakes(oracle, mutant, pin, tmp_path):
    lib = oracle.build_mutant(mutant, str(tmp_path))
    if pin == "closed_form":
        with po.using(lib):
            m, dx, sa, si, g, bank = _absorber_case()
            f = po.forward_cells(0, m, dx, 0.0, sa, si, bank, g=g)
            ga, _ = po.adjoint_cells(0, m, dx, 0.0, sa, si, bank, g, score=f["score"])
        assert not np.allclose(ga, _absorber_gradient(m, dx, sa, len(bank)), rtol=1e-6)
    else:
        assert _fd_z(lib, "i", CELLS) > 3.0
