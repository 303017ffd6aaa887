"""Exact stationary averages of the pairwise lattice gas by enumeration.

The pairwise model (adsorption at 1/s, desorption at 1/s * exp(n eps/kT),
hops at 100/s * exp(n eps/kT), n = occupied nearest neighbours of the
leaving particle other than the hop target) obeys detailed balance with the
Boltzmann law pi(s) ~ exp(-eps B(s) / kT), B = the number of occupied
nearest-neighbour pairs, at zero chemical potential:
  adsorption/desorption at a site with n occupied neighbours:
    k_ads / k_des = exp(-n eps/kT) = pi(occupied) / pi(empty);
  a hop from a site with n other occupied neighbours to one with m:
    k_fwd / k_rev = exp((n - m) eps/kT), the bond change being m - n.
On an L x L periodic lattice all 2^(L*L) states are enumerated, so the mean
number of sites at which each process variant is enabled is exact.
"""
import itertools

import numpy as np

NN = [(1, 0), (0, 1), (-1, 0), (0, -1)]
DIRS = {"right": (1, 0), "up": (0, 1), "left": (-1, 0), "down": (0, -1)}


def enabled_means(L, eps, kT):
    """{process name: exact stationary mean of the number of sites at
    which that process is enabled} on an L x L periodic lattice."""
    V = L * L
    states = np.arange(1 << V, dtype=np.int64)
    occ = ((states[:, None] >> np.arange(V)) & 1).astype(bool)      # [S, V]

    def idx(x, y):
        return (y % L) * L + (x % L)
    xs, ys = np.meshgrid(np.arange(L), np.arange(L), indexing="xy")
    xs, ys = xs.ravel(), ys.ravel()
    site = idx(xs, ys)
    nb = {d: idx(xs + d[0], ys + d[1]) for d in NN}
    bonds = occ[:, site] & occ[:, nb[(1, 0)]]
    bonds = bonds.sum(1) + (occ[:, site] & occ[:, nb[(0, 1)]]).sum(1)
    w = np.exp(-eps * bonds / kT)
    w /= w.sum()
    out = {"A_adsorption": float(w @ (~occ).sum(1))}
    o = occ[:, site]
    for conf in itertools.product([0, 1], repeat=4):
        m = o.copy()
        for bit, d in zip(conf, NN):
            m &= occ[:, nb[d]] == bool(bit)
        out["A_desorption_" + "".join(map(str, conf))] = float(w @ m.sum(1))
    for dname, d in DIRS.items():
        others = [e for e in NN if e != d]
        for conf in itertools.product([0, 1], repeat=3):
            m = o & ~occ[:, nb[d]]
            for bit, e in zip(conf, others):
                m &= occ[:, nb[e]] == bool(bit)
            out["A_diffusion_%s_" % dname + "".join(map(str, conf))] = float(w @ m.sum(1))
    return out


def check_time_averages(names, rates, integ0, integ1, t0, t1, exact, where=""):
    """Per process, the replicas' time-averaged enabled-site counts
    (integ_rates difference / (rate x time)) against the exact means.

    The replicas are independent, so their spread gives the standard error
    of each mean.  All processes are compared at once, so each must lie
    within the Student-t bound that keeps the family-wise level at 3 sigma
    (two-sided 0.27 %, Bonferroni over the processes).  Returns the t
    values."""
    from scipy import stats
    avg = (integ1 - integ0) / (rates * (t1 - t0)[:, None])           # [n, P]
    n, k = avg.shape
    bound = stats.t.ppf(1.0 - 0.0027 / (2 * k), n - 1)
    ts = []
    for p, nm in enumerate(names):
        mean = avg[:, p].mean()
        se = avg[:, p].std(ddof=1) / np.sqrt(n)
        t = (mean - exact[nm]) / se if se > 0 else (0.0 if mean == exact[nm] else np.inf)
        ts.append(t)
        assert abs(t) < bound, "%s %s: mean %.5g exact %.5g se %.3g t %.2f (bound %.2f)" % (
            where, nm, mean, exact[nm], se, t, bound)
    return ts
