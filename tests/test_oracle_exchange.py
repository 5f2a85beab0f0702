"""Pins the oracle's replica exchange (ParallelSampleSMP, Samplers.py:589-669,
1880-1893): schedule shape, always-swap at equal temperatures, the swap
probability min(1, exp((1/T_a - 1/T_b)(chi_a - chi_b))) in closed form, and
invariance of the tempered product measure under exchange moves (a sign slip
in the improvement sum breaks it)."""
import numpy as np

from oracle import hmc_oracle as O


def _gauss1d(T):
    return O.Tempered(O.Normal(np.zeros((1, 1)), 1.0), T)


def test_schedule_rows_hold_distinct_chains():
    s = O.exchange_schedule(np.random.default_rng(1), 7, 20, 2)
    assert s.shape == (10, 6)
    for row in s:
        assert len(set(row.tolist())) == 6 and row.min() >= 0 and row.max() < 7
    assert np.array_equal(s, O.exchange_schedule(np.random.default_rng(1), 7, 20, 2))


def test_equal_temperatures_always_swap():
    dists = [_gauss1d(2.0), _gauss1d(2.0)]
    xs = [np.array([[0.3]]), np.array([[-1.7]])]
    chis = [d.misfit(x) for d, x in zip(dists, xs)]
    sw = O.exchange_event(dists, xs, chis, [0, 1], [0.999999])
    assert sw == [True] and xs[0][0, 0] == -1.7 and xs[1][0, 0] == 0.3
    assert chis[0] == dists[0].misfit(xs[0])


def test_swap_probability_closed_form():
    Ta, Tb = 1.0, 3.0
    xa, xb = np.array([[0.4]]), np.array([[2.5]])
    chi = lambda x: 0.5 * float(x[0, 0] ** 2)
    p = min(1.0, np.exp((1 / Ta - 1 / Tb) * (chi(xa) - chi(xb))))
    u = np.random.default_rng(2).uniform(size=20000)
    hits = 0
    for uk in u:
        dists = [_gauss1d(Ta), _gauss1d(Tb)]
        xs = [xa.copy(), xb.copy()]
        chis = [dists[0].misfit(xs[0]), dists[1].misfit(xs[1])]
        hits += O.exchange_event(dists, xs, chis, [0, 1], [uk])[0]
    assert abs(hits / len(u) - p) < 4 * np.sqrt(p * (1 - p) / len(u))


def test_exchange_preserves_tempered_product_measure():
    Ta, Tb = 1.0, 4.0
    rng = np.random.default_rng(3)
    n = 40000
    xa = rng.normal(0, np.sqrt(Ta), n)
    xb = rng.normal(0, np.sqrt(Tb), n)
    u = rng.uniform(size=n)
    da, db = _gauss1d(Ta), _gauss1d(Tb)
    out_a, out_b = np.empty(n), np.empty(n)
    for k in range(n):
        xs = [np.array([[xa[k]]]), np.array([[xb[k]]])]
        chis = [da.misfit(xs[0]), db.misfit(xs[1])]
        O.exchange_event([da, db], xs, chis, [0, 1], [u[k]])
        out_a[k], out_b[k] = xs[0][0, 0], xs[1][0, 0]
    se = lambda v: v * np.sqrt(2.0 / n)
    assert abs(out_a.var() - Ta) < 4 * se(Ta)
    assert abs(out_b.var() - Tb) < 4 * se(Tb)
    # the exchanges did happen
    assert np.mean(out_a != xa) > 0.2
