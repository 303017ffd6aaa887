"""Ensemble statistics of the engine against closed forms and against the
reference's own Fortran.

* Langmuir adsorption/desorption (the quickstart model, k_ads = k_des =
  100 s^-1, 20 x 20 sites, independent two-state sites): the stationary
  coverage is k_a / (k_a + k_d) = 0.5 with binomial variance, and the
  adsorption turnover per site is k_a k_d / (k_a + k_d) = 50 s^-1; the
  ensemble means of 512 replicas agree within 3 sigma.
* ZGB regimes (Ziff, Gulari and Barshad, Phys. Rev. Lett. 56, 2553 (1986)):
  below y1 ~ 0.39 the surface poisons with O, above y2 ~ 0.525 with CO, in
  between it stays reactive and makes CO2.
* RuO2 (the reference's default.xml model): 64 replicas on the engine and
  32 replicas of the reference's generated Fortran (built by
  oracle/build_ref.py, other seeds) run the same number of steps from the
  empty lattice; the ensemble means of the CO and O coverages and of the
  CO2 turnover agree within 3 sigma (Welch).
* Pairwise lateral interactions (the lat_int workload's model on a 4 x 4
  periodic lattice, eps = -0.05, 0, 0.05 eV, kT = 0.05 eV): detailed
  balance makes the stationary law Boltzmann in the number of occupied
  neighbour pairs, so every process's time-averaged number of enabled
  sites is known exactly by enumerating the 2^16 states
  (tests/pairwise_exact.py); 256 replicas per eps agree within 3 sigma
  family-wise.  At eps = 0 this is the Langmuir law (coverage 1/2,
  independent sites).
"""
import numpy as np
import pytest

from kmos_amd import workloads as W
from kmos_amd.compiler import compile_model

pytestmark = pytest.mark.gpu


def _batch(engine, cm, n, size, rates, seed0):
    m = engine.Model(cm)
    b = engine.Batch(m, [list(size) + [1]] * n)
    b.init(np.arange(seed0, seed0 + n, dtype=np.uint32), np.zeros(n, np.uint32))
    b.set_rates(np.tile(np.asarray(rates, np.float64), (n, 1)))
    return m, b


def test_langmuir_coverage_and_turnover(engine):
    cm = compile_model(W.quickstart_project(), "local_smart")
    co = cm.species_names.index("CO")
    ads = cm.process_names.index("CO_adsorption")
    n, V = 512, 400
    m, b = _batch(engine, cm, n, (20, 20), [100.0, 100.0], 1)
    b.run(20000)                                   # ~50 events per site: stationary
    t0 = np.array([b.scalars(r)[0] for r in range(n)])
    p0 = np.array([b.procstat(r)[ads] for r in range(n)])
    b.run(40000)
    t1 = np.array([b.scalars(r)[0] for r in range(n)])
    p1 = np.array([b.procstat(r)[ads] for r in range(n)])
    theta = np.array([(b.lattice(r) == co).mean() for r in range(n)])
    sigma = np.sqrt(0.5 * 0.5 / V / n)
    assert abs(theta.mean() - 0.5) < 3 * sigma, (theta.mean(), sigma)
    tof = (p1 - p0) / ((t1 - t0) * V)
    se = tof.std(ddof=1) / np.sqrt(n)
    assert abs(tof.mean() - 50.0) < 3 * se, (tof.mean(), se)
    b.close()


def test_zgb_regimes(engine):
    cm = compile_model(W.zgb_project(), "local_smart")
    names = cm.process_names
    co, o = cm.species_names.index("CO"), cm.species_names.index("O")
    ox = [i for i, nm in enumerate(names) if nm.startswith("CO_oxidation")]
    out = {}
    for y in (0.30, 0.45, 0.60):
        rates = [y if nm == "CO_adsorption" else 1e10 if nm.startswith("CO_oxidation") else (1 - y) / 2
                 for nm in names]
        m, b = _batch(engine, cm, 32, (48, 48), rates, 7)
        b.run(200000)
        st = np.array([b.scalars(r)[4] for r in range(32)])
        lat = [b.lattice(r) for r in range(32)]
        out[y] = (np.mean([(x == co).mean() for x in lat]), np.mean([(x == o).mean() for x in lat]),
                  np.array([b.procstat(r)[ox].sum() for r in range(32)]), st)
        b.close()
    print({y: (v[0], v[1], int(np.sum(v[3] != 0))) for y, v in out.items()})
    assert out[0.30][1] > 0.8                      # O-poisoned
    assert out[0.60][0] > 0.8                      # CO-poisoned
    theta_co, theta_o, co2, st = out[0.45]
    assert np.all(st == 0) and theta_co < 0.1 and theta_o < 0.95 and np.all(co2 > 50000)


def test_ruo2_ensemble_matches_reference_fortran(engine):
    from refbin import driver_path, run_reference
    if driver_path("ruo2_local_smart") is None:
        pytest.skip("reference build missing (oracle/build_ref.py)")
    cm = compile_model(W.ruo2_project(), "local_smart")
    names, P = cm.process_names, cm.nr_of_proc
    rates = np.array([1.0 if "_adsorption" in nm else 0.5 if nm.startswith("co_desorption")
                      else 0.1 if "_desorption" in nm else 2.0 if "_diffusion" in nm else 5.0
                      for nm in names])
    co, o = cm.species_names.index("co"), cm.species_names.index("oxygen")
    react = [i for i, nm in enumerate(names) if nm.startswith("reaction")]
    size, steps, V = (20, 20), 20000, 800

    def obs(lattice, procstat, t):
        return ((lattice == co).mean(), (lattice == o).mean(), procstat[react].sum() / (t * V))

    n = 64
    m, b = _batch(engine, cm, n, size, rates, 1)
    b.run(steps)
    gpu = np.array([obs(b.lattice(r), b.procstat(r), b.scalars(r)[0]) for r in range(n)])
    b.close()
    ref = []
    for k in range(32):
        res = run_reference("ruo2_local_smart", size, 1001 + k, 0, steps, rates, P, mode=1)
        assert res["kmc_step"] == steps
        ref.append(obs(res["lattice"], res["procstat"], res["kmc_time"]))
    ref = np.array(ref)
    for j, what in enumerate(("CO coverage", "O coverage", "CO2 turnover")):
        a, c = gpu[:, j], ref[:, j]
        se = np.sqrt(a.var(ddof=1) / a.size + c.var(ddof=1) / c.size)
        assert abs(a.mean() - c.mean()) < 3 * se, (what, a.mean(), c.mean(), se)


def test_pairwise_stationary_law_exact(engine):
    from pairwise_exact import check_time_averages, enabled_means
    pt = W.pairwise_project()
    cm = compile_model(pt, "lat_int")
    epss = (-0.05, 0.0, 0.05)
    n = 256
    rates = np.concatenate([np.tile(W.pairwise_rates(pt, e, 0.05), (n, 1)) for e in epss])
    m = engine.Model(cm)
    N = n * len(epss)
    b = engine.Batch(m, [[4, 4, 1]] * N)
    b.init(np.full(N, 5, np.uint32), np.arange(N, dtype=np.uint32))
    b.set_rates(rates)
    b.run(20000)
    i0 = np.array([b.integ_rates(r) for r in range(N)])
    t0 = np.array([b.scalars(r)[0] for r in range(N)])
    b.run(200000)
    i1 = np.array([b.integ_rates(r) for r in range(N)])
    t1 = np.array([b.scalars(r)[0] for r in range(N)])
    assert all(b.scalars(r)[4] == 0 for r in range(N))
    assert np.all(b.check_lists() == 0)
    b.close()
    for k, e in enumerate(epss):
        sl = slice(k * n, (k + 1) * n)
        check_time_averages(cm.process_names, rates[sl], i0[sl], i1[sl], t0[sl], t1[sl],
                            enabled_means(4, e, 0.05), where="GPU eps=%g" % e)
