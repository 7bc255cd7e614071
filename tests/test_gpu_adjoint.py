"""Adjoint kernel (through the C ABI and the autograd Function) against the
FD-pinned FP64 oracle adjoint and against finite differences of the GPU
forward itself."""
import numpy as np
import pytest
import torch

from adjoint_cases import FD, PROBES, adjoint_case
from oracle import oracle
from synchrad_amd import autograd, engine, workloads

pytestmark = pytest.mark.gpu

KEV = workloads.KEV_PER_INV_M


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gpu_grads(tr, axes, dt, G, comp, precision, tuning=None):
    hb = engine.pack_tracks(tr, dt=dt)
    batch = engine.DeviceBatch(hb)
    grid = engine.DeviceGrid("far", axes["omega"], axes["theta"], axes["phi"])
    dS = [torch.from_numpy(np.ascontiguousarray(G[k].swapaxes(-1, -3))).cuda()
          for k in oracle.COMP_KEYS[comp]]
    g = engine.far_adjoint(batch, grid, dS, comp=comp, precision=precision, tuning=tuning)
    torch.cuda.synchronize()
    g = [t.cpu().numpy() for t in g]
    off = hb.offsets
    per = [[g[c][off[p]:off[p + 1]] for c in range(6)] for p in range(hb.n_particles)]
    return per, g[6]


def compare(per, gw, ref, ref_w, tol, label):
    for c in range(6):
        got = np.concatenate([p[c] for p in per])
        want = np.concatenate([p[c] for p in ref])
        r = rel_l2(got, want)
        print(f"{label} grad[{c}] rel-L2 {r:.3e}", flush=True)
        assert r <= tol, (c, r)
    r = rel_l2(gw, ref_w)
    print(f"{label} grad w rel-L2 {r:.3e}", flush=True)
    assert r <= tol


@pytest.mark.parametrize("comp", ["total", "cartesian"])
@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("float", 1e-4)])
def test_adjoint_matches_oracle_adjoint(comp, precision, tol):
    tr, axes, dt, S0, T, W, G = adjoint_case(comp)
    ref, ref_w = oracle.far_adjoint(comp, tr, axes, dt, G)
    per, gw = gpu_grads(tr, axes, dt, G, comp, precision)
    compare(per, gw, ref, ref_w, tol, f"small {comp} {precision}")


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("float", 1e-4)])
def test_adjoint_with_active_filter(precision, tol):
    """Full 1-75 keV x 40 mrad range: many steps are filtered; the mask is
    held fixed in both implementations."""
    wl = workloads.config5(n_particles=3)
    tr = [[a[:600] for a in t[:6]] + [t[6], 0] for t in wl.track_list()]
    axes = oracle.make_axes([(1.0 * KEV, 75.0 * KEV), (0.0, 0.04), (0.0, 2 * np.pi), (40, 8, 6)], "far")
    dt = wl.batch.dt
    S0 = oracle.spectrum("far", "total", tr, axes, dt)["total"]
    rng = np.random.default_rng(3)
    G = {"total": rng.standard_normal(S0.shape) * S0.mean()}
    ref, ref_w = oracle.far_adjoint("total", tr, axes, dt, G)
    per, gw = gpu_grads(tr, axes, dt, G, "total", precision)
    compare(per, gw, ref, ref_w, tol, f"filtered {precision}")


def test_adjoint_f64_phases_beyond_table_range():
    """FP64 adjoint with phases of 1.26e13..2.51e13 rad (tracks 1.25e6 m behind
    the origin along z, so tau = t - n.x is ~1.25e6 m), on both sides of
    2^44 = 1.76e13 rad, where the FP64 sin/cos switches from its direct table
    reduction to a 2 pi pre-reduction; against the oracle adjoint."""
    dt = 0.01 / 16
    tr = []
    for i, (g, K) in enumerate([(100.0, 1.0), (120.0, 0.8)]):
        arrs = list(workloads.planar_undulator(g, K, 0.01, 400, dt, psi=0.3 * i))
        arrs[2] = arrs[2] - 1.25e6
        tr.append(arrs + [1.0 + 0.5 * i, 0])
    axes = oracle.make_axes([(1.6e6, 3.2e6), (0.0, 0.01), (0.0, 2 * np.pi), (48, 4, 4)], "far")
    w = 2 * np.pi * axes["omega"]
    assert w.min() * 1.25e6 < 2.0 ** 44 < w.max() * 1.25e6
    S0 = oracle.spectrum("far", "total", tr, axes, dt)["total"]
    assert (S0 > 0).all()
    rng = np.random.default_rng(5)
    G = {"total": rng.standard_normal(S0.shape) * S0.mean()}
    ref, ref_w = oracle.far_adjoint("total", tr, axes, dt, G)
    per, gw = gpu_grads(tr, axes, dt, G, "total", "double")
    compare(per, gw, ref, ref_w, 1e-10, "phases beyond 2^44")


@pytest.mark.parametrize("groups,waves", [(1, 1), (3, 2), (7, 4)])
def test_adjoint_decomposition_does_not_change_the_result(groups, waves):
    tr, axes, dt, S0, T, W, G = adjoint_case("total")
    a, aw = gpu_grads(tr, axes, dt, G, "total", "double", engine.Tuning(slices=groups, waves_per_block=waves))
    b, bw = gpu_grads(tr, axes, dt, G, "total", "double")
    # other direction groups regroup the FP64 sums, and the gradient sums
    # cancel strongly (terms ~1e2 x the result): agreement to ~1e-13 is all
    # that FP64 rounding allows, so the bar is 1e-11
    for p in range(len(a)):
        for c in range(6):
            assert rel_l2(a[p][c], b[p][c]) < 1e-11
    assert rel_l2(aw, bw) < 1e-11


def test_autograd_function_and_finite_differences():
    """loss.backward() through far_spectrum gives the oracle gradient, and
    agrees with finite differences of the GPU FP64 forward to 1e-6."""
    tr, axes, dt, S0, T, W, G = adjoint_case("total")
    hb = engine.pack_tracks(tr, dt=dt)
    grid = engine.DeviceGrid("far", axes["omega"], axes["theta"], axes["phi"])
    dev = grid.device
    leaves = [torch.tensor(getattr(hb, c), device=dev, requires_grad=True)
              for c in ("x", "y", "z", "ux", "uy", "uz", "weight")]
    Td = torch.from_numpy(np.ascontiguousarray(T["total"].swapaxes(-1, -3))).to(dev)
    Wd = torch.from_numpy(np.ascontiguousarray(W["total"].swapaxes(-1, -3))).to(dev)
    S = autograd.far_spectrum(*leaves, hb, grid, comp="total", precision="double")[0]
    loss = torch.sum(Wd * (S - Td) ** 2)
    loss.backward()
    ref, ref_w = oracle.far_adjoint("total", tr, axes, dt, G)
    off = hb.offsets
    for c in range(6):
        got = leaves[c].grad.cpu().numpy()
        want = np.concatenate([r[c] for r in ref])
        assert rel_l2(got, want) < 1e-10
    assert rel_l2(leaves[6].grad.cpu().numpy(), ref_w) < 1e-10

    # finite differences of the GPU forward (one particle's spectrum varied)
    batch_rest = None
    fd = FD("total", tr, axes, dt, S0, T, W)

    def gpu_single(p, comp_idx, s, delta):
        t = list(tr[p])
        if comp_idx == 6:
            t[6] = t[6] + delta
        else:
            t[comp_idx] = t[comp_idx].copy()
            t[comp_idx][s] += delta
        b = engine.DeviceBatch(engine.pack_tracks([t], dt=dt))
        o = engine.spectrum(b, grid, comp="total", precision="double")[0]
        return {"total": o.cpu().numpy().swapaxes(-1, -3)}

    fd.diff_spectrum = gpu_single
    worst = 0.0
    for (p, c, s) in PROBES[:6]:
        def diff(p_, c_, s_, h):
            Spp, Spm = gpu_single(p_, c_, s_, h), gpu_single(p_, c_, s_, -h)
            rest = S0["total"] - fd.Sp0[p_]["total"] - T["total"]
            return np.sum(W["total"] * (Spp["total"] - Spm["total"]) * (Spp["total"] + Spm["total"] + 2 * rest)) / (2 * h)
        fd.diff = diff
        est = fd.best(p, c, s)
        got = leaves[c].grad[off[p] + s].item()
        err = abs(est - got) / abs(got)
        worst = max(worst, err)
        assert err < 1e-6, (p, c, s, est, got)
    print(f"GPU adjoint vs GPU-forward finite differences: worst {worst:.2e}")


def test_adjoint_betatron_full_grid_f64_vs_split():
    """Config-5 grid (128 x 32 x 32), 4 particles x 1000 steps: the split
    adjoint against the FP64 adjoint, and the FP64 adjoint against the oracle."""
    wl = workloads.config5(n_particles=4)
    tr = [[a[:1000] for a in t[:6]] + [t[6], 0] for t in wl.track_list()]
    ax = wl.axes
    dt = wl.batch.dt
    hb = engine.pack_tracks(tr, dt=dt)
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    batch = engine.DeviceBatch(hb)
    S = engine.spectrum(batch, grid, comp="total", precision="double")[0]
    Wd = torch.sin(torch.as_tensor(ax["theta"], device=grid.device))[None, None, :, None] * torch.ones_like(S)
    dS = [2 * Wd * S * 0.01]
    g64 = engine.far_adjoint(batch, grid, dS, precision="double")
    g32 = engine.far_adjoint(batch, grid, dS, precision="float")
    torch.cuda.synchronize()
    for a, b in zip(g32, g64):
        r = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        print(f"split vs f64 adjoint rel-L2 {r:.3e}", flush=True)
        assert r < 1e-4
    G = {"total": dS[0].cpu().numpy().swapaxes(-1, -3)}
    ref, ref_w = oracle.far_adjoint("total", tr, ax, dt, G)
    for c in range(6):
        want = np.concatenate([r[c] for r in ref])
        assert rel_l2(g64[c].cpu().numpy(), want) < 1e-10
    assert rel_l2(g64[6].cpu().numpy(), ref_w) < 1e-10


def test_adjoint_config5_fd_50_coordinates():
    """The adjoint workload at full size (256 e-, 4000 fp32-rounded steps,
    128 x 32 x 32 over 1-75 keV x 40 mrad); loss = sum W (S - T)^2 with T from
    tracks pushed with 1% stronger focusing and W = sin(theta).

    * loss.backward() through far_spectrum (split precision) agrees with the
      FP64 adjoint to 1e-4 rel-L2, and its forward with the FP64 forward;
    * the FP64 adjoint agrees with FP64 central finite differences of the loss
      at 50 seeded coordinates (x, y, z, ux, uy, uz of a step, or a weight)
      to 1e-6 relative, Richardson-extrapolated over a ladder of steps.

    A weight's differences come from the GPU FP64 forward of the particle
    with the weight moved.  A sample's position or momentum enters at most
    two terms of the particle's sum: x_s only tau_s = t_s - n.x_s (linear in
    x) and the filter tests of steps s and s+1; u_s the amplitudes of steps
    s-1 and s.  Moving them through the whole FP64 forward cannot resolve
    their derivatives at this scale: tau carries ~5e-20 m of rounding
    (|t| ~ 3e-4 m) while any position step above ~1e-15 m flips a filter
    decision somewhere on the grid and makes L jump (the adjoint
    differentiates the piece on which the mask is fixed), and the 4000-term
    sums bury a momentum's effect in their rounding (1e-7 .. 5e-6 measured).
    So those differences are formed with the mask held: the particle's FP64
    complex field (GPU 'cartesian_complex' forward) with the affected terms
    replaced by their values at the moved coordinate (the phase moved
    exactly by -+ omega n_c h for a position; momentum-moved amplitudes in
    extended precision), the change of |F|^2 written out so that nothing
    large cancels."""
    wl = workloads.config5()
    tg = workloads.config5(target=True)
    hb, ax, dt = wl.batch, wl.axes, wl.batch.dt
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    dev = grid.device
    S64 = engine.spectrum(engine.DeviceBatch(hb, dev), grid, comp="total", precision="double")[0]
    T = engine.spectrum(engine.DeviceBatch(tg.batch, dev), grid, comp="total", precision="double")[0]
    Wt = torch.sin(torch.as_tensor(ax["theta"], device=dev))[None, None, :, None].expand_as(S64).contiguous()

    cols = ("x", "y", "z", "ux", "uy", "uz", "weight")
    leaves = [torch.tensor(getattr(hb, c), device=dev, requires_grad=True) for c in cols]
    S = autograd.far_spectrum(*leaves, hb, grid, comp="total", precision="float")[0]
    r = float(torch.linalg.norm(S.detach() - S64) / torch.linalg.norm(S64))
    print(f"config5 split forward vs FP64 rel-L2 {r:.3e}", flush=True)
    assert r < 1e-4
    torch.sum(Wt * (S - T) ** 2).backward()
    G = 2 * Wt * (S64 - T)
    g64 = engine.far_adjoint(engine.DeviceBatch(hb, dev), grid, [G], precision="double")
    torch.cuda.synchronize()
    for c in range(7):
        r = float(torch.linalg.norm(leaves[c].grad - g64[c]) / torch.linalg.norm(g64[c]))
        print(f"config5 split adjoint vs FP64 adjoint, {cols[c]}: rel-L2 {r:.3e}", flush=True)
        assert r < 1e-4
    g64 = [g.cpu().numpy() for g in g64]

    # grid on the host exactly as the device holds it, in the device order (phi, theta, omega)
    om = grid.omega.cpu().numpy()[None, None, :]
    st, ct = grid.sin_theta.cpu().numpy()[None, :, None], grid.cos_theta.cpu().numpy()[None, :, None]
    sp, cp = grid.sin_phi.cpu().numpy()[:, None, None], grid.cos_phi.cpu().numpy()[:, None, None]
    nvec = (st * cp, st * sp, ct * np.ones_like(cp))
    rng = np.random.default_rng(50)
    off = hb.offsets
    tl = wl.track_list()
    single = {}

    def own(p):
        if p not in single:
            b = engine.DeviceBatch(engine.pack_tracks([tl[p]], dt=dt), dev)
            Sp = engine.spectrum(b, grid, comp="total", precision="double")[0]
            cc = engine.spectrum(b, grid, comp="cartesian_complex", precision="double")
            scale = np.sqrt(float(hb.weight[p])) * dt
            F = [torch.complex(cc[2 * c], cc[2 * c + 1]).cpu().numpy()[0] / scale for c in range(3)]
            single[p] = (b, Sp, F)
        return single[p]

    def phase(p, k):
        """FP64 phase of sample k of particle p, formed as the kernels form it."""
        a = int(off[p]) + k
        t = float(np.uint32(int(hb.it_start[p]) + k)) * dt
        return om * (t - (hb.x[a] * nvec[0] + hb.y[a] * nvec[1] + hb.z[a] * nvec[2]))

    def mom(p, k):
        a = int(off[p]) + k
        return np.array([hb.ux[a], hb.uy[a], hb.uz[a]])

    LD = np.longdouble
    nvl = [a.astype(LD) for a in nvec]

    def amp(u0, u1):
        """A(n) of a step from its two momenta (synchrad/kernel_farfield.cl:76-94),
        in extended precision: 1 - n.bbar ~ 3e-6 loses 5 digits to cancellation,
        which would otherwise swamp the momentum differences."""
        u0, u1, one = u0.astype(LD), u1.astype(LD), LD(1)
        b0 = u0 * (one / np.sqrt(one + (u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2])))
        b1 = u1 * (one / np.sqrt(one + (u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2])))
        acc = (b1 - b0) * (one / LD(dt))
        mb = LD(0.5) * (b1 + b0)
        c2 = one / (one - (mb[0] * nvl[0] + mb[1] * nvl[1] + mb[2] * nvl[2]))
        c1 = (acc[0] * nvl[0] + acc[1] * nvl[1] + acc[2] * nvl[2]) * c2 * c2
        return [c1 * (nvl[c] - mb[c]) - c2 * acc[c] for c in range(3)]

    def expo(p, s):
        """mask_s exp(i phase_s): the filter decision of the unperturbed track."""
        ph = phase(p, s)
        prev = phase(p, s - 1) if s > 0 else 0.0
        return np.where(np.abs(ph - prev) < np.pi, np.exp(1j * ph), 0.0)

    worst = 0.0
    for i in range(50):
        p = int(rng.integers(hb.n_particles))
        c = int(rng.integers(7))
        n_p = int(off[p + 1] - off[p])
        s = 0 if c == 6 else int(rng.integers(n_p - 1 if c < 3 else n_p))
        b, Sp0, F = own(p)
        rest2 = 2 * (S64 - Sp0 - T)
        v = float(getattr(hb, cols[c])[int(off[p]) + s]) if c < 6 else float(hb.weight[p])
        wdt2 = float(hb.weight[p]) * dt * dt
        if c < 6:
            Wn = Wt[0].cpu().numpy()
            restn = rest2[0].cpu().numpy()
            F2 = sum(np.abs(F[k]) ** 2 for k in range(3))
        if c < 3:
            t0 = [A.astype(np.float64) * expo(p, s) for A in amp(mom(p, s), mom(p, s + 1))]
            FbT = sum(np.conj(F[k]) * t0[k] for k in range(3))
            T2 = sum(np.abs(t0[k]) ** 2 for k in range(3))
            h0 = 1e-3 / (om.max() * np.abs(nvec[c]).max())

            def D(h):
                # S_p(+-h) = w dt^2 |F + t0 (exp(-+i d) - 1)|^2, d = omega n_c h
                d = om * nvec[c] * h
                dS = 4.0 * wdt2 * np.sin(d) * FbT.imag
                q = np.sin(0.5 * d) ** 2
                Ssum = wdt2 * (2.0 * F2 - 8.0 * q * FbT.real + 8.0 * q * T2)
                return float(np.sum(Wn * dS * (Ssum + restn))) / (2 * h)
        elif c < 6:
            # u_s enters the amplitudes of steps s-1 and s (those that exist)
            steps = [q for q in (s - 1, s) if 0 <= q <= n_p - 2]
            ex = {q: expo(p, q) for q in steps}
            h0 = 1e-3 * max(abs(v), 1.0)

            def D(h):
                he = (v + h) - v
                diff = [0.0, 0.0, 0.0]      # d(+h) - d(-h)
                summ = [0.0, 0.0, 0.0]      # d(+h) + d(-h)
                for q in steps:
                    us = [mom(p, q), mom(p, q + 1)]
                    j = 0 if q == s else 1
                    up = [u.copy() for u in us]
                    um = [u.copy() for u in us]
                    up[j][c - 3] = v + he
                    um[j][c - 3] = v - he
                    A0, Ap, Am = amp(*us), amp(*up), amp(*um)
                    for k in range(3):
                        diff[k] = diff[k] + (Ap[k] - Am[k]).astype(np.float64) * ex[q]
                        summ[k] = summ[k] + (Ap[k] + Am[k] - 2 * A0[k]).astype(np.float64) * ex[q]
                dS = wdt2 * sum(2.0 * (np.conj(F[k]) * diff[k]).real + (diff[k] * np.conj(summ[k])).real
                                for k in range(3))
                dp2 = sum(np.abs(0.5 * (summ[k] + diff[k])) ** 2 for k in range(3))
                dm2 = sum(np.abs(0.5 * (summ[k] - diff[k])) ** 2 for k in range(3))
                Ssum = wdt2 * (2.0 * F2 + sum(2.0 * (np.conj(F[k]) * summ[k]).real for k in range(3)) + dp2 + dm2)
                return float(np.sum(Wn * dS * (Ssum + restn))) / (2 * he)
        else:
            arr = b.weight
            h0 = 1e-3 * abs(v)

            def D(h):
                he = (v + h) - v
                arr[0] = v + he
                Spp = engine.spectrum(b, grid, comp="total", precision="double")[0]
                arr[0] = v - he
                Spm = engine.spectrum(b, grid, comp="total", precision="double")[0]
                arr[0] = v
                return float(torch.sum(Wt * (Spp - Spm) * (Spp + Spm + rest2))) / (2 * he)

        Ds = [D(h0 * 0.25 ** k) for k in range(6)]
        R = [(16 * Ds[k + 1] - Ds[k]) / 15 for k in range(5)]
        k = int(np.argmin([abs(R[j] - R[j + 1]) for j in range(4)]))
        est = R[k]
        got = float(g64[c][off[p] + s]) if c < 6 else float(g64[6][p])
        err = abs(est - got) / abs(got) if got != 0.0 else abs(est)
        worst = max(worst, err)
        print(f"config5 FD #{i:2d} particle {p:3d} {cols[c]:6s} step {s:4d}: adjoint {got:+.10e} "
              f"FD {est:+.10e} rel {err:.2e} (h0 {h0:.3e})", flush=True)
        assert err < 1e-6, (i, p, c, s, est, got)
    print(f"config5 FP64 adjoint vs FP64 central FD, 50 coordinates: worst {worst:.2e}", flush=True)


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("float", 1e-5)])
def test_adjoint_chunks_of_particles(precision, tol):
    """On a 512 x 64 x 64 grid one particle's fields take ~100 MB, so 30
    particles are processed in two chunks of the ~2 GiB field workspace.
    Each sample's gradient depends only on its own particle's field, so the
    batch must agree with the particles run one at a time (to the FP64
    rounding of the regrouped sums; the split path's own rounding)."""
    wl = workloads.config2(n_particles=30, n_steps=200)
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    hb = wl.batch
    rng = np.random.default_rng(7)
    G = torch.from_numpy(rng.standard_normal((1, ax["phi"].size, ax["theta"].size, ax["omega"].size))).cuda()
    full = engine.far_adjoint(engine.DeviceBatch(hb), grid, [G], precision=precision)
    torch.cuda.synchronize()
    full = [t.cpu().numpy() for t in full]
    one = [[] for _ in range(7)]
    for t in wl.track_list():
        g = engine.far_adjoint(engine.DeviceBatch(engine.pack_tracks([t], dt=hb.dt)), grid, [G], precision=precision)
        torch.cuda.synchronize()
        for c in range(7):
            one[c].append(g[c].cpu().numpy())
    for c in range(7):
        r = rel_l2(full[c], np.concatenate(one[c]))
        print(f"chunked adjoint {precision} grad[{c}] vs one particle at a time: rel-L2 {r:.3e}", flush=True)
        assert r < tol


@pytest.mark.parametrize("comp", ["total", "cartesian"])
@pytest.mark.parametrize("precision", ["double", "float"])
def test_fields_forward_equals_spectrum(comp, precision):
    """The spectra formed from the kept per-particle fields equal the spectrum
    kernel's (w dt^2 |F|^2 with the record's expressions, in particle order),
    and each particle's stored field gives its own spectrum."""
    wl = workloads.config2(n_particles=5, n_steps=1500, grid=(96, 8, 6))
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    batch = engine.DeviceBatch(wl.batch)
    ref = engine.spectrum(batch, grid, comp=comp, precision=precision, tuning=engine.Tuning(slices=1))
    F, outs = engine.far_fields(batch, grid, comp=comp, precision=precision)
    torch.cuda.synchronize()
    for a, b in zip(outs, ref):
        r = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        print(f"fields-derived {comp} {precision}: rel-L2 {r:.3e}", flush=True)
        assert r < 1e-14
    # particle 2 alone
    t2 = wl.track_list()[2]
    one = engine.spectrum(engine.DeviceBatch(engine.pack_tracks([t2], dt=wl.batch.dt)), grid, comp=comp,
                          precision=precision)
    f = F[:, 2, :].reshape(3, 2, -1)
    sq = (f ** 2).sum(dim=1) * (wl.batch.weight[2] * wl.batch.dt ** 2)
    got = [sq.sum(dim=0)] if comp == "total" else [sq[c] for c in range(3)]
    for a, b in zip(got, one):
        assert float(torch.linalg.norm(a - b.reshape(-1)) / torch.linalg.norm(b)) < 1e-13


@pytest.mark.parametrize("precision,tol", [("double", 1e-11), ("float", 1e-6)])
def test_adjoint_from_kept_fields_equals_recomputed(precision, tol):
    """The backward given the forward's fields equals the one that recomputes
    them, here over two chunks of the recomputing path's field workspace
    (30 particles x 512 x 64 x 64).  The two paths split the directions into
    different numbers of groups (one chunk of 30 particles vs chunks of ~20),
    so the FP64 sums are regrouped; the momentum gradients, whose terms cancel,
    show that as ~1e-13."""
    wl = workloads.config2(n_particles=30, n_steps=200)
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    batch = engine.DeviceBatch(wl.batch)
    rng = np.random.default_rng(11)
    G = torch.from_numpy(rng.standard_normal((1, ax["phi"].size, ax["theta"].size, ax["omega"].size))).cuda()
    F, _ = engine.far_fields(batch, grid, comp="total", precision=precision)
    a = engine.far_adjoint(batch, grid, [G], precision=precision, fields=F)
    b = engine.far_adjoint(batch, grid, [G], precision=precision)
    torch.cuda.synchronize()
    for c in range(7):
        r = float(torch.linalg.norm(a[c] - b[c]) / torch.linalg.norm(b[c]))
        print(f"kept-fields adjoint {precision} grad[{c}]: rel-L2 {r:.3e}", flush=True)
        assert r < tol


def test_autograd_with_and_without_kept_fields(monkeypatch):
    """loss.backward() gives the same gradients whether the forward keeps the
    fields or the backward recomputes them."""
    wl = workloads.config2(n_particles=4, n_steps=1200, grid=(64, 6, 4))
    ax = wl.axes
    grid = engine.DeviceGrid("far", ax["omega"], ax["theta"], ax["phi"])
    hb = wl.batch
    rng = np.random.default_rng(3)
    W = torch.from_numpy(rng.random((1, ax["phi"].size, ax["theta"].size, ax["omega"].size))).cuda()

    def grads():
        leaves = [torch.tensor(getattr(hb, c), device="cuda", requires_grad=True)
                  for c in ("x", "y", "z", "ux", "uy", "uz", "weight")]
        S = autograd.far_spectrum(*leaves, hb, grid, comp="total", precision="double")[0]
        (W * S ** 2).sum().backward()
        return S.detach(), [t.grad for t in leaves]

    S_keep, g_keep = grads()
    monkeypatch.setattr(autograd, "KEEP_FIELDS_BYTES", 0)
    S_re, g_re = grads()
    torch.cuda.synchronize()
    assert float(torch.linalg.norm(S_keep - S_re) / torch.linalg.norm(S_re)) < 1e-14
    for a, b in zip(g_keep, g_re):
        assert float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) < 1e-12


@pytest.mark.parametrize("comp", ["total", "cartesian"])
@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("float", 1e-4)])
def test_adjoint_beyond_one_omega_tile(comp, precision, tol):
    """A 1100-omega grid: the coefficient tables are held 512 omegas at a time
    (three tiles, the last partial), and the gradients still equal the
    oracle adjoint's."""
    wl = workloads.config5(n_particles=2)
    tr = [[a[:240] for a in t[:6]] + [t[6] * (1.0 + 0.4 * i), 0] for i, t in enumerate(wl.track_list())]
    axes = oracle.make_axes([(1.0 * KEV, 30.0 * KEV), (0.0, 0.02), (0.0, 2 * np.pi), (1100, 3, 2)], "far")
    dt = wl.batch.dt
    keys = oracle.COMP_KEYS[comp]
    S0 = oracle.spectrum("far", comp, tr, axes, dt)
    rng = np.random.default_rng(17)
    G = {k: rng.standard_normal(S0[k].shape) * S0[k].mean() for k in keys}
    ref, ref_w = oracle.far_adjoint(comp, tr, axes, dt, G)
    per, gw = gpu_grads(tr, axes, dt, G, comp, precision)
    compare(per, gw, ref, ref_w, tol, f"1100 omega {comp} {precision}")
