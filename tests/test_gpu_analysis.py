"""The analysis helpers on the device copy of the spectra agree with the same
helpers on the host arrays (synchrad/utils.py:23-152 defines them)."""
import numpy as np
import pytest
import torch

from synchrad_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,comp", [("far", "total"), ("far", "cartesian"), ("near", "total"),
                                       ("far", "cartesian_complex")])
def test_reductions_on_device_match_host(mode, comp):
    from synchrad.calc import SynchRad
    if mode == "far":
        wl = workloads.config2(n_particles=3, n_steps=1500, grid=(48, 12, 10))
    else:
        wl = workloads.config4(scale=0.125, n_particles=2)
    calc = SynchRad(dict(wl.calc_input))
    calc.calculate_spectrum(particleTracks=wl.track_list(), timeStep=wl.batch.dt, comp=comp, L_screen=wl.L_screen,
                            nSnaps=2, sigma_particle=1e-11, verbose=False)
    key = next(iter(calc.Data["radiation"]))
    assert torch.is_tensor(calc._slice(key, -1)), "the device copy should be used right after the call"
    kw = dict(lambda0_um=0.8)
    dev = {"full": calc.get_full_spectrum(**kw), "espec": calc.get_energy_spectrum(**kw),
           "energy": calc.get_energy(phot_num=True), "spot": calc.get_spot(**kw),
           "spot_k0": calc.get_spot(k0=float(calc.Args["omega"][7]), iteration=0),
           "cart": calc.get_spot_cartesian(th_part=0.8, bins=(64, 48))[0]}
    calc.analysis_on_device = False
    assert not torch.is_tensor(calc._slice(key, -1))
    host = {"full": calc.get_full_spectrum(**kw), "espec": calc.get_energy_spectrum(**kw),
            "energy": calc.get_energy(phot_num=True), "spot": calc.get_spot(**kw),
            "spot_k0": calc.get_spot(k0=float(calc.Args["omega"][7]), iteration=0),
            "cart": calc.get_spot_cartesian(th_part=0.8, bins=(64, 48))[0]}
    for name in dev:
        a, b = np.asarray(dev[name]), np.asarray(host[name])
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / scale if scale > 0 else np.abs(a).max()
        print(f"{mode} {comp} {name}: device vs host {err:.2e}", flush=True)
        assert a.shape == b.shape and err < 1e-12
    if comp == "cartesian_complex":
        z = calc.get_full_spectrum(comp="x")
        calc.analysis_on_device = True
        assert np.iscomplexobj(z) and np.allclose(calc.get_full_spectrum(comp="x"), z, rtol=1e-13)
    # a replaced host array is what the helpers then read
    calc.analysis_on_device = True
    calc.Data["radiation"][key] = calc.Data["radiation"][key] * 2.0
    assert not torch.is_tensor(calc._slice(key, -1))


@pytest.mark.parametrize("dtype", ["double", "float"])
def test_repeated_calls_replay_the_captured_launch(dtype):
    """Repeated calculate_spectrum calls with the same sizes refill the device
    batch and replay a captured graph; each result equals a fresh calculator's
    direct call on the same tracks, bit for bit, and the analysis helpers read
    the replayed spectra."""
    from synchrad.calc import SynchRad
    calc = SynchRad(dict(workloads.config2(n_particles=2, n_steps=900, grid=(64, 8, 8)).calc_input, dtype=dtype))
    for seed in range(4):
        wl = workloads.config2(n_particles=3, n_steps=900, grid=(64, 8, 8))
        tracks = wl.track_list()
        rng = np.random.default_rng(seed)
        for t in tracks:
            t[6] = float(rng.uniform(0.5, 2.0))
            t[0] = t[0] + 1e-7 * rng.standard_normal(t[0].size)
        calc.calculate_spectrum([list(t) for t in tracks], timeStep=wl.batch.dt, comp="cartesian", nSnaps=2,
                                verbose=False)
        fresh = SynchRad(dict(wl.calc_input, dtype=dtype))
        fresh.calculate_spectrum([list(t) for t in tracks], timeStep=wl.batch.dt, comp="cartesian", nSnaps=2,
                                 verbose=False)
        for k in ("x", "y", "z"):
            assert np.array_equal(calc.Data["radiation"][k], fresh.Data["radiation"][k]), (seed, k)
        assert calc.get_energy() == pytest.approx(fresh.get_energy(), rel=1e-13)
        assert (calc._replay["graph"] is not None) == (seed >= 1)
