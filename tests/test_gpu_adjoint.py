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


@pytest.mark.parametrize("groups,waves", [(1, 1), (3, 2), (7, 4)])
def test_adjoint_decomposition_does_not_change_the_result(groups, waves):
    tr, axes, dt, S0, T, W, G = adjoint_case("total")
    a, aw = gpu_grads(tr, axes, dt, G, "total", "double", engine.Tuning(slices=groups, waves_per_block=waves))
    b, bw = gpu_grads(tr, axes, dt, G, "total", "double")
    for p in range(len(a)):
        for c in range(6):
            assert rel_l2(a[p][c], b[p][c]) < 1e-13
    assert rel_l2(aw, bw) < 1e-13


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
