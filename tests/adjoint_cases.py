"""Shared adjoint test cases: a small betatron loss and a finite-difference
estimator of its gradient (used by the CPU and GPU adjoint tests)."""
import numpy as np

from oracle import oracle
from synchrad_amd import workloads

KEV = workloads.KEV_PER_INV_M


def adjoint_case(comp="total", n_particles=2, n_steps=300, seed=5):
    wl = workloads.config5(n_particles=n_particles)
    tr = [[a[:n_steps] for a in t[:6]] + [t[6] * (1.0 + 0.3 * i), 0] for i, t in enumerate(wl.track_list())]
    axes = oracle.make_axes([(1.0 * KEV, 8.0 * KEV), (0.0, 0.01), (0.0, 2 * np.pi), (12, 5, 4)], "far")
    dt = wl.batch.dt
    keys = oracle.COMP_KEYS[comp]
    S0 = oracle.spectrum("far", comp, tr, axes, dt)
    rng = np.random.default_rng(seed)
    T = {k: S0[k] * (1 + 0.1 * rng.standard_normal(S0[k].shape)) for k in keys}
    W = {k: rng.uniform(0.5, 1.5, S0[k].shape) for k in keys}
    G = {k: 2 * W[k] * (S0[k] - T[k]) for k in keys}
    return tr, axes, dt, S0, T, W, G


class FD:
    """Finite differences of L with respect to one sample of one particle."""

    def __init__(self, comp, tr, axes, dt, S0, T, W, library=None):
        self.comp, self.tr, self.axes, self.dt = comp, tr, axes, dt
        self.S0, self.T, self.W, self.lib = S0, T, W, library
        self.keys = oracle.COMP_KEYS[comp]
        self.Sp0 = [oracle.spectrum("far", comp, [t], axes, dt, library=library) for t in tr]
        self.wmax = 2 * np.pi * axes["omega"].max()
        self.nmax = np.array([np.sin(axes["theta"].max())] * 2 + [1.0])

    def diff(self, p, comp_idx, s, h):
        def Sp(delta):
            t = list(self.tr[p])
            if comp_idx == 6:
                t[6] = t[6] + delta
            else:
                t[comp_idx] = t[comp_idx].copy()
                t[comp_idx][s] += delta
            return oracle.spectrum("far", self.comp, [t], self.axes, self.dt, library=self.lib)
        Spp, Spm = Sp(h), Sp(-h)
        tot = 0.0
        for k in self.keys:
            rest = self.S0[k] - self.Sp0[p][k] - self.T[k]
            tot += np.sum(self.W[k] * (Spp[k] - Spm[k]) * (Spp[k] + Spm[k] + 2 * rest))
        return tot / (2 * h)

    def best(self, p, comp_idx, s):
        if comp_idx == 6:
            h0 = 1e-3 * abs(self.tr[p][6])
        elif comp_idx < 3:
            h0 = 1e-2 / (self.wmax * self.nmax[comp_idx])
        else:
            h0 = 1e-3 * max(abs(self.tr[p][comp_idx][s]), 1.0)
        D = [self.diff(p, comp_idx, s, h0 * 0.25 ** k) for k in range(6)]
        R = [(16 * D[k + 1] - D[k]) / 15 for k in range(5)]   # h ratio 4: O(h^2) term cancels
        k = int(np.argmin([abs(R[i] - R[i + 1]) for i in range(4)]))
        return R[k]


PROBES = [(0, 0, 10), (0, 2, 50), (1, 1, 120), (0, 3, 7), (1, 5, 200), (0, 4, 299), (1, 2, 0),
          (0, 1, 33), (1, 4, 250), (1, 0, 298), (0, 5, 150)]
