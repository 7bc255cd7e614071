"""Seeded synthetic track sets: the five benchmark configurations.

All lengths are metres, ``timeStep`` is c*dt in metres and ``omega`` is 1/lambda
in 1/m (the kernels multiply it by 2*pi, synchrad/calc.py:494-495).  The RNG is
``numpy.random.default_rng(1000 + config number)``.  Far-field tracks are
time-staggered as the reference's examples are: momenta at t_i, positions at
t_i + dt/2 (tests/test_undulator_analytic.py:46-48).

Each builder returns a :class:`Workload` holding the CSR host batch and the
calculator input; ``scale`` < 1 shrinks particles and grid for CPU-oracle
parity runs while keeping the physics of the configuration.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import engine

KEV_PER_INV_M = 8.0655439e8        # 1 keV photon <-> 1/lambda in 1/m
C_LIGHT = 299792458.0
E_CHARGE = 1.602176634e-19
EPS0 = 8.8541878128e-12
M_E = 9.1093837015e-31


@dataclass
class Workload:
    name: str
    batch: engine.HostBatch
    calc_input: Dict
    comp: str
    mode: str = "far"
    L_screen: Optional[float] = None
    notes: str = ""
    extra: Dict = field(default_factory=dict)

    @property
    def axes(self) -> Dict[str, np.ndarray]:
        """Axes exactly as SynchRad._init_args builds them (double)."""
        g = self.calc_input["grid"]
        No, N1, Nph = g[-1]
        omin, omax = g[0]
        if "logGrid" in self.calc_input.get("Features", []):
            d = np.log(omax / omin) / (No - 1.0)
            omega = omin * np.exp(d * np.arange(No))
        else:
            omega = np.r_[omin:omax:No * 1j]
        ax1 = np.r_[g[1][0]:g[1][1]:N1 * 1j]
        if self.screen == "cartesian":
            # a Cartesian near-field screen: x and y both inclusive
            y = np.r_[g[2][0]:g[2][1]:Nph * 1j]
            return {"omega": omega.astype(np.double), "x": ax1.astype(np.double), "y": y.astype(np.double)}
        phi = g[2][0] + (g[2][1] - g[2][0]) / Nph * np.arange(Nph)
        key = "theta" if self.mode == "far" else "radius"
        return {"omega": omega.astype(np.double), key: ax1.astype(np.double),
                "phi": phi.astype(np.double)}

    @property
    def screen(self) -> str:
        return self.extra.get("screen", "polar")

    def device_grid(self, device=None):
        """The workload's grid as an engine.DeviceGrid."""
        ax = self.axes
        if self.screen == "cartesian":
            return engine.DeviceGrid("near", ax["omega"], ax["x"], ax["y"], L_screen=self.L_screen, device=device,
                                     screen="cartesian")
        ax1 = ax["theta"] if self.mode == "far" else ax["radius"]
        return engine.DeviceGrid(self.mode, ax["omega"], ax1, ax["phi"], L_screen=self.L_screen, device=device)

    @property
    def n_grid(self) -> int:
        g = self.calc_input["grid"][-1]
        return int(g[0] * g[1] * g[2])

    @property
    def evals(self) -> int:
        """spectral points x step intervals: N_grid * sum_p (n_p - 1)."""
        return self.n_grid * self.batch.n_intervals()

    def track_list(self) -> List[list]:
        """Reference-style ``[x, y, z, ux, uy, uz, w, it_start]`` lists (views)."""
        hb = self.batch
        out = []
        for p in range(hb.n_particles):
            a, b = int(hb.offsets[p]), int(hb.offsets[p + 1])
            out.append([hb.x[a:b], hb.y[a:b], hb.z[a:b], hb.ux[a:b], hb.uy[a:b], hb.uz[a:b],
                        float(hb.weight[p]), int(hb.it_start[p])])
        return out


def _csr(lengths: np.ndarray) -> np.ndarray:
    off = np.zeros(lengths.size + 1, dtype=np.int64)
    np.cumsum(lengths, out=off[1:])
    return off


def undulator_omega1(gamma: float, K: float, lambda_u: float) -> float:
    """Fundamental 1/lambda on axis: 2 gamma^2 / (lambda_u (1 + K^2/2))."""
    return 2.0 * gamma ** 2 / (lambda_u * (1.0 + K ** 2 / 2.0))


def planar_undulator(gamma: float, K: float, lambda_u: float, n_steps: int, cdt: float,
                     psi: float = 0.0, stagger: bool = True, x0=0.0, y0=0.0, thx=0.0, thy=0.0):
    """Analytic planar-undulator orbit: u_x = K cos(k_u s + psi) + gamma thx,
    x = (K/gamma k_u) sin(k_u s' + psi) + thx s', z = beta_bar s' - K^2/(8 gamma^2 k_u)
    sin(2(k_u s' + psi)), with s' = s + cdt/2 when staggered."""
    ku = 2 * np.pi / lambda_u
    s = np.arange(n_steps, dtype=np.float64) * cdt
    sp = s + 0.5 * cdt if stagger else s
    ux = K * np.cos(ku * s + psi) + gamma * thx
    uy = np.full_like(s, gamma * thy)
    uz = np.sqrt(gamma ** 2 - 1.0 - ux ** 2 - uy ** 2)
    bbar = 1.0 - (1.0 + K ** 2 / 2.0) / (2.0 * gamma ** 2)
    x = x0 + (K / (gamma * ku)) * (np.sin(ku * sp + psi) - np.sin(psi)) + thx * sp
    y = y0 + thy * sp
    z = bbar * sp - (K ** 2 / (8.0 * gamma ** 2 * ku)) * (np.sin(2 * (ku * sp + psi)) - np.sin(2 * psi))
    return x, y, z, ux, uy, uz


def helical_undulator(gamma: float, K: float, lambda_u: float, n_steps: int, cdt: float,
                      psi: float = 0.0, stagger: bool = True):
    """Helical orbit: u_perp = K (cos, sin)(k_u s + psi), |u_perp| = K, z = beta_bar s'."""
    ku = 2 * np.pi / lambda_u
    s = np.arange(n_steps, dtype=np.float64) * cdt
    sp = s + 0.5 * cdt if stagger else s
    ux = K * np.cos(ku * s + psi)
    uy = K * np.sin(ku * s + psi)
    uz = np.full_like(s, np.sqrt(gamma ** 2 - 1.0 - K ** 2))
    bbar = 1.0 - (1.0 + K ** 2) / (2.0 * gamma ** 2)
    x = (K / (gamma * ku)) * (np.sin(ku * sp + psi) - np.sin(psi))
    y = -(K / (gamma * ku)) * (np.cos(ku * sp + psi) - np.cos(psi))
    z = bbar * sp
    return x, y, z, ux, uy, uz


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
def config1(scale: float = 1.0) -> Workload:
    """Single electron, planar undulator: gamma=100, K=1, lambda_u=1 cm, 50 periods
    x 32 steps; far 'total', FP64; 256 omega on [0.5, 1.5] w1 x 32 theta on
    [0, 1/gamma] x 32 phi on [0, 2 pi)."""
    gamma, K, lu = 100.0, 1.0, 0.01
    n = 1600
    cdt = lu / 32
    x, y, z, ux, uy, uz = planar_undulator(gamma, K, lu, n, cdt)
    w1 = undulator_omega1(gamma, K, lu)
    no, nt, nph = (max(8, int(256 * scale)), max(4, int(32 * scale)), max(4, int(32 * scale)))
    hb = engine.pack_csr(np.array([0, n]), x, y, z, ux, uy, uz, np.array([1.0]), dt=cdt)
    ci = {"grid": [(0.5 * w1, 1.5 * w1), (0.0, 1.0 / gamma), (0.0, 2 * np.pi), (no, nt, nph)],
          "dtype": "double"}
    return Workload("undulator-1e", hb, ci, "total", extra={"gamma": gamma, "K": K, "lambda_u": lu,
                                                            "omega1": w1, "periods": 50})


def betatron_push(gamma0: np.ndarray, x0, y0, ux0, uy0, kp: float, cdt: float, n_steps: int,
                  force_scale: float = 1.0):
    """Leapfrog in s = ct for F = -(m w_p^2/2) r_perp:
    u^{i+1/2} = u^{i-1/2} - (k_p^2/2) r_perp^i cdt,  r^{i+1} = r^i + cdt u^{i+1/2}/gamma^{i+1/2};
    stored u[i] = u^{i-1/2}, x[i] = r^i (arrays (Np, n_steps))."""
    Np = gamma0.size
    uz0 = np.sqrt(gamma0 ** 2 - 1.0 - ux0 ** 2 - uy0 ** 2)
    X = np.empty((Np, n_steps)); Y = np.empty((Np, n_steps)); Z = np.empty((Np, n_steps))
    UX = np.empty((Np, n_steps)); UY = np.empty((Np, n_steps)); UZ = np.empty((Np, n_steps))
    x, y, z = x0.copy(), y0.copy(), np.zeros(Np)
    ux, uy, uz = ux0.copy(), uy0.copy(), uz0.copy()
    k2 = 0.5 * kp ** 2 * force_scale
    for i in range(n_steps):
        X[:, i], Y[:, i], Z[:, i] = x, y, z
        UX[:, i], UY[:, i], UZ[:, i] = ux, uy, uz
        ux = ux - k2 * x * cdt
        uy = uy - k2 * y * cdt
        g = np.sqrt(1.0 + ux * ux + uy * uy + uz * uz)
        x = x + cdt * ux / g
        y = y + cdt * uy / g
        z = z + cdt * uz / g
    return X, Y, Z, UX, UY, UZ


def _betatron_beam(rng, Np: int):
    gamma0 = rng.normal(400.0, 16.0, Np)
    x0 = rng.normal(0.0, 1e-6, Np)
    y0 = rng.normal(0.0, 1e-6, Np)
    ux0 = rng.normal(0.0, 2.0, Np)
    uy0 = rng.normal(0.0, 2.0, Np)
    return gamma0, x0, y0, ux0, uy0


def plasma_kp(n_e: float = 5e24) -> float:
    wp = np.sqrt(n_e * E_CHARGE ** 2 / (EPS0 * M_E))
    return wp / C_LIGHT


def config2(scale: float = 1.0, n_particles: Optional[int] = None, n_steps: int = 8000,
            grid: Optional[Tuple[int, int, int]] = None) -> Workload:
    """Betatron beam: 2000 e-, gamma ~ N(400, 4%), x0,y0 ~ N(0, 1 um), ux,uy ~ N(0, 2);
    n_e = 5e18 cm^-3, leapfrog 8000 steps at dt = 2 pi/(200 w_p); far 'total';
    512 omega on [1, 75] keV x 64 theta on [0, 40 mrad] x 64 phi."""
    rng = np.random.default_rng(1002)
    Np_full = 2000
    gamma0, x0, y0, ux0, uy0 = _betatron_beam(rng, Np_full)
    Np = n_particles if n_particles is not None else max(1, int(round(Np_full * scale)))
    kp = plasma_kp()
    cdt = 2 * np.pi / (200 * kp)
    X, Y, Z, UX, UY, UZ = betatron_push(gamma0[:Np], x0[:Np], y0[:Np], ux0[:Np], uy0[:Np], kp, cdt,
                                        n_steps)
    lengths = np.full(Np, n_steps, dtype=np.int64)
    hb = engine.pack_csr(_csr(lengths), X.ravel(), Y.ravel(), Z.ravel(), UX.ravel(), UY.ravel(),
                         UZ.ravel(), np.ones(Np), dt=cdt)
    if grid is None:
        grid = (max(8, int(512 * scale)), max(4, int(64 * scale)), max(4, int(64 * scale)))
    ci = {"grid": [(1.0 * KEV_PER_INV_M, 75.0 * KEV_PER_INV_M), (0.0, 0.04), (0.0, 2 * np.pi), grid],
          "dtype": "float"}
    return Workload("betatron", hb, ci, "total", extra={"kp": kp, "n_steps": n_steps})


def config3(scale: float = 1.0, n_particles: Optional[int] = None) -> Workload:
    """Uneven-length batch: 20000 e-, n_p log-uniform on [200, 20000]; planar or
    helical (50/50) undulator orbits with K ~ U(0.5, 3), gamma ~ U(50, 500), entry
    phase ~ U(0, 2 pi), lambda_u = 1 cm, 24 steps/period, weights ~ U(0.5, 1.5);
    far 'cartesian', split; logGrid 256 omega x 32 theta on [0, 0.06] x 32 phi."""
    rng = np.random.default_rng(1003)
    Np_full = 20000
    lengths_full = np.round(np.exp(rng.uniform(np.log(200), np.log(20000), Np_full))).astype(np.int64)
    helical = rng.uniform(size=Np_full) < 0.5
    Ks = rng.uniform(0.5, 3.0, Np_full)
    gammas = rng.uniform(50.0, 500.0, Np_full)
    psis = rng.uniform(0.0, 2 * np.pi, Np_full)
    ws = rng.uniform(0.5, 1.5, Np_full)
    Np = n_particles if n_particles is not None else max(1, int(round(Np_full * scale)))
    lu = 0.01
    cdt = lu / 24
    lengths = lengths_full[:Np]
    off = _csr(lengths)
    cols = [np.empty(int(off[-1])) for _ in range(6)]
    for p in range(Np):
        fn = helical_undulator if helical[p] else planar_undulator
        arrs = fn(gammas[p], Ks[p], lu, int(lengths[p]), cdt, psi=psis[p])
        for c in range(6):
            cols[c][off[p]:off[p + 1]] = arrs[c]
    hb = engine.pack_csr(off, *cols, ws[:Np], dt=cdt)
    wmin = 0.5 * undulator_omega1(50.0, 3.0, lu)
    wmax = 2.0 * undulator_omega1(500.0, 0.5, lu)
    no, nt, nph = (max(8, int(256 * scale)), max(4, int(32 * scale)), max(4, int(32 * scale)))
    ci = {"grid": [(wmin, wmax), (0.0, 0.06), (0.0, 2 * np.pi), (no, nt, nph)],
          "dtype": "float", "Features": ["logGrid"]}
    return Workload("uneven-csr", hb, ci, "cartesian")


def config4(scale: float = 1.0, n_particles: Optional[int] = None, screen: str = "polar") -> Workload:
    """Near field: 500 e-, planar undulator gamma = 200(1 + N(0, 0.5%)), K=2,
    lambda_u = 2 cm, 100 periods x 40 steps; offsets N(0, 20 um), angles N(0, 10 urad);
    screen at L = 7 m (5 m past the exit); 128 omega on [0.8, 1.2] w1 x 64 R on
    [0, 7.5 cm] x 64 phi (the reference's polar screen), or with
    screen='cartesian' 64 x 64 points over +-7.5 cm (+-3/gamma at 5 m); FP64
    distance, FP32 amplitude."""
    rng = np.random.default_rng(1004)
    Np_full = 500
    g0, K, lu = 200.0, 2.0, 0.02
    n = 4000
    cdt = lu / 40
    gammas = g0 * (1.0 + rng.normal(0.0, 0.005, Np_full))
    xs0 = rng.normal(0.0, 20e-6, Np_full)
    ys0 = rng.normal(0.0, 20e-6, Np_full)
    thx = rng.normal(0.0, 10e-6, Np_full)
    thy = rng.normal(0.0, 10e-6, Np_full)
    Np = n_particles if n_particles is not None else max(1, int(round(Np_full * scale)))
    cols = [np.empty(Np * n) for _ in range(6)]
    for p in range(Np):
        arrs = planar_undulator(gammas[p], K, lu, n, cdt, stagger=False, x0=xs0[p], y0=ys0[p],
                                thx=thx[p], thy=thy[p])
        for c in range(6):
            cols[c][p * n:(p + 1) * n] = arrs[c]
    hb = engine.pack_csr(_csr(np.full(Np, n)), *cols, np.ones(Np), dt=cdt)
    w1 = undulator_omega1(g0, K, lu)
    no, nr, nph = (max(8, int(128 * scale)), max(4, int(64 * scale)), max(4, int(64 * scale)))
    if screen == "cartesian":
        ci = {"grid": [(0.8 * w1, 1.2 * w1), (-0.075, 0.075), (-0.075, 0.075), (no, nr, nph)],
              "dtype": "float", "mode": "near"}
        return Workload("near-undulator-xy", hb, ci, "total", mode="near", L_screen=7.0,
                        extra={"omega1": w1, "screen": "cartesian"})
    ci = {"grid": [(0.8 * w1, 1.2 * w1), (0.0, 0.075), (0.0, 2 * np.pi), (no, nr, nph)],
          "dtype": "float", "mode": "near"}
    return Workload("near-undulator", hb, ci, "total", mode="near", L_screen=7.0,
                    extra={"omega1": w1})


def config5(scale: float = 1.0, n_particles: Optional[int] = None, target: bool = False) -> Workload:
    """Adjoint workload: 256 betatron e- (config 2 set-up), 4000 steps, tracks
    rounded to fp32 then used as exact; 128 omega on [1, 75] keV x 32 theta on
    [0, 40 mrad] x 32 phi.  ``target=True`` pushes with 1% stronger focusing."""
    rng = np.random.default_rng(1005)
    Np_full = 256
    gamma0, x0, y0, ux0, uy0 = _betatron_beam(rng, Np_full)
    Np = n_particles if n_particles is not None else max(1, int(round(Np_full * scale)))
    kp = plasma_kp()
    cdt = 2 * np.pi / (200 * kp)
    n_steps = 4000
    X, Y, Z, UX, UY, UZ = betatron_push(gamma0[:Np], x0[:Np], y0[:Np], ux0[:Np], uy0[:Np], kp, cdt,
                                        n_steps, force_scale=1.01 if target else 1.0)
    cols = [a.ravel().astype(np.float32).astype(np.float64) for a in (X, Y, Z, UX, UY, UZ)]
    hb = engine.pack_csr(_csr(np.full(Np, n_steps)), *cols, np.ones(Np), dt=cdt)
    no, nt, nph = (max(8, int(128 * scale)), max(4, int(32 * scale)), max(4, int(32 * scale)))
    ci = {"grid": [(1.0 * KEV_PER_INV_M, 75.0 * KEV_PER_INV_M), (0.0, 0.04), (0.0, 2 * np.pi),
                   (no, nt, nph)], "dtype": "float"}
    return Workload("adjoint-betatron", hb, ci, "total")


CONFIGS = {1: config1, 2: config2, 3: config3, 4: config4, 5: config5}
