"""Analysis of computed spectra: the reference's ``Utilities`` methods.

What the methods return is fixed by synchrad/utils.py:23-166; the code here is
built from the definitions below, not from that file's statements.

* The raw outputs are R = sum_p w_p dt^2 |F_p|^2 (or, for the coherent
  ``*_complex`` runs, the field sums themselves, whose squares add up to R).
  The radiated energy per unit frequency and solid angle (far field) or per
  unit frequency and screen area (near field) is P * R with the prefactor
  P = alpha/(4 pi^2) far and alpha pi/4/(2 pi)^2 near (synchrad/utils.py:54-58);
  photon numbers divide by omega, and ``lambda0_um`` converts to J through
  J_in_um/lambda0 (utils.py:66-71).
* Every integral is a fixed linear functional of the spectrum: quadrature
  weights that depend on the axes only, applied as one weighted sum.
    - over omega: the trapezoid rule on the omega nodes (utils.py:101, 112);
    - over the far-field polar angle: the trapezoid rule in theta on the
      midpoints of the theta nodes, of the midpoint average of the two
      neighbouring values times sin(theta_mid) (utils.py:81-86); spreading each
      midpoint's weight half to each neighbour gives one weight per node;
    - over the near-field radius: the trapezoid rule of r R(r) (utils.py:89-91);
    - over phi: the sum of the nodes times the phi step.
* ``get_spot_cartesian`` resamples a (theta|r, phi) map onto a Cartesian mesh
  by linear interpolation over the Delaunay triangles of the polar nodes, zero
  outside their hull (scipy's griddata 'linear', utils.py:140-151).  The
  triangles and each mesh point's barycentric weights depend on the axes
  only; they are formed once on the host and reused.

Where the reductions run: right after ``calculate_spectrum`` the spectra are
still resident on the HIP device, and every reduction above is applied there
(torch tensors) before the small result is copied back; after a spectrum file
is reloaded, or when ``analysis_on_device`` is False, the same weights are
applied to the host arrays of ``Data['radiation']``.  Results are returned as
numpy arrays either way.
"""
from __future__ import annotations

from typing import Dict, Tuple

import numpy as np
from scipy.constants import c, e, epsilon_0, hbar, m_e
from scipy.constants import alpha as alpha_fs
from scipy.spatial import Delaunay

# physical constants, synchrad/utils.py:16-19
J_in_um = 2e6 * np.pi * hbar * c
r_e = e ** 2 / (4 * np.pi * epsilon_0 * m_e * c ** 2)
omega_1m = 2 * np.pi * c
energy_1m_eV = omega_1m * hbar / e

# |F|^2 -> energy per unit frequency and solid angle (far) or area (near)
PREFACTOR = {"far": alpha_fs / (4 * np.pi ** 2), "near": alpha_fs * np.pi / 4 / (2 * np.pi) ** 2}


# ---------------------------------------------------------------------------
# quadrature weights (host, float64; they depend on the axes only)
# ---------------------------------------------------------------------------
def trapezoid_weights(x: np.ndarray) -> np.ndarray:
    """w with sum_i w_i f(x_i) = the trapezoid rule of f over the nodes x (in
    their order, so a descending axis integrates to the negative); one node
    integrates to 0."""
    x = np.asarray(x, dtype=np.double)
    w = np.zeros(x.size)
    if x.size > 1:
        h = np.diff(x)
        w[:-1] += 0.5 * h
        w[1:] += 0.5 * h
    return w


def polar_angle_weights(theta: np.ndarray) -> np.ndarray:
    """Per-node weights of the far field's polar integral: the trapezoid rule
    over the theta midpoints of the midpoint-averaged integrand times
    sin(theta_mid)."""
    theta = np.asarray(theta, dtype=np.double)
    w = np.zeros(theta.size)
    if theta.size > 1:
        mid = 0.5 * (theta[1:] + theta[:-1])
        w_mid = trapezoid_weights(mid) * np.sin(mid)
        w[:-1] += 0.5 * w_mid
        w[1:] += 0.5 * w_mid
    return w


def radial_weights(radius: np.ndarray) -> np.ndarray:
    """Per-node weights of the near field's radial integral of r R(r)."""
    radius = np.asarray(radius, dtype=np.double)
    return trapezoid_weights(radius) * radius


def nearest_node_rule(omega: np.ndarray, k0: float) -> int:
    """The omega node ``get_spot`` reads for a frequency k0: the first node
    that is not below k0, or the node after it when that one lies closer to k0
    (synchrad/utils.py:119-124 compares only those two).  Past the last node
    the reference raises IndexError; here the last node is used."""
    above = int(np.count_nonzero(np.asarray(omega) < k0))
    last = np.asarray(omega).size - 1
    if above >= last:
        return last
    return above + 1 if abs(omega[above + 1] - k0) < abs(omega[above] - k0) else above


class PolarResampler:
    """Linear interpolation of a polar (r, phi) node set onto a Cartesian
    mesh over [-m, m]^2 (m = r_part * max r): each mesh point's Delaunay
    triangle and barycentric weights, found once."""

    def __init__(self, r: np.ndarray, phi: np.ndarray, r_part: float, bins: Tuple[int, int]):
        r = np.asarray(r, dtype=np.double)
        phi = np.asarray(phi, dtype=np.double)
        rr = np.repeat(r, phi.size)                     # node (i, j) -> flat i * n_phi + j
        pp = np.tile(phi, r.size)
        nodes = np.column_stack([rr * np.cos(pp), rr * np.sin(pp)])
        self.half_width = r_part * r.max()
        m = self.half_width
        mesh = np.mgrid[-m:m:bins[0] * 1j, -m:m:bins[1] * 1j]
        self.shape = mesh[0].shape
        targets = np.column_stack([mesh[0].ravel(), mesh[1].ravel()])
        tri = Delaunay(nodes)
        simplex = tri.find_simplex(targets)
        affine = tri.transform[simplex]
        lam = np.einsum("pij,pj->pi", affine[:, :2], targets - affine[:, 2])
        weights = np.column_stack([lam, 1.0 - lam.sum(axis=1)])
        corners = tri.simplices[simplex]
        outside = simplex < 0
        weights[outside] = 0.0
        corners[outside] = 0
        self.corners, self.weights = corners, weights
        self._dev: Dict[str, tuple] = {}

    def extent(self) -> np.ndarray:
        m = self.half_width
        return np.array([-m, m, -m, m])

    def __call__(self, values):
        """values: (n_r, n_phi), numpy or a torch tensor on any device."""
        if _is_tensor(values):
            import torch
            key = str(values.device)
            if key not in self._dev:
                self._dev[key] = (torch.as_tensor(self.corners, device=values.device),
                                  torch.as_tensor(self.weights, device=values.device))
            corners, weights = self._dev[key]
            flat = values.reshape(-1)
            return (flat[corners] * weights.to(flat.dtype)).sum(-1).reshape(self.shape)
        flat = np.asarray(values).reshape(-1)
        return (flat[self.corners] * self.weights).sum(-1).reshape(self.shape)


def _is_tensor(a) -> bool:
    return type(a).__module__.startswith("torch")


def _host(a):
    return a.detach().cpu().numpy() if _is_tensor(a) else a


def _like(weights: np.ndarray, ref):
    """Axis weights as an array of ref's kind (numpy or a device tensor)."""
    if _is_tensor(ref):
        import torch
        return torch.as_tensor(weights, dtype=torch.float64, device=ref.device)
    return weights


class Utilities:
    """Post-processing of ``self.Data['radiation']`` (the reference's
    ``synchrad.utils.Utilities``; its methods and keyword arguments)."""

    #: run the reductions on the device copy of the spectra while it is current
    analysis_on_device = True

    # ------------------------------------------------------------ sources
    def _slice(self, key: str, iteration):
        """One snapshot of one output as (n_omega, n_theta|n_r, n_phi): the
        device tensor while it still holds Data['radiation'][key] (the
        device layout is (nSnaps, nPhi, n1, nOmega)), else the host array."""
        host = self.Data["radiation"][key]
        dev = getattr(self, "_rad_device", None)
        mirrors = getattr(self, "_rad_device_mirrors", {})
        if self.analysis_on_device and dev is not None and key in dev and mirrors.get(key) is host:
            return dev[key][iteration].permute(2, 1, 0)
        return np.asarray(host[iteration], dtype=np.double)

    def _scaled(self, spect_filter=None, phot_num=False, lambda0_um=None, normalize_to_weights=False,
                comp="total", iteration=-1):
        coherent = self.Args["comp"].rsplit("_", 1)[-1] == "complex"
        keys = list(self.Data["radiation"].keys())
        if comp != "total":
            if coherent:
                re, im = self._slice(comp + "re", iteration), self._slice(comp + "im", iteration)
                if _is_tensor(re):
                    import torch
                    val = torch.complex(re.contiguous(), im.contiguous())
                else:
                    val = re.astype(np.complex128) + 1j * im
            else:
                val = self._slice(comp, iteration)
        else:
            parts = [self._slice(k, iteration) for k in keys]
            val = parts[0] * parts[0] if coherent else parts[0]
            for p in parts[1:]:
                val = val + (p * p if coherent else p)
        val = val * PREFACTOR.get(self.Args["mode"], 1.0)
        if spect_filter is not None:
            val = val * _like(np.asarray(spect_filter, dtype=np.double), val) \
                if _is_tensor(val) else val * spect_filter
        if normalize_to_weights:
            val = val / self.total_weight
        if phot_num:
            val = val / _like(np.asarray(self.Args["omega"], dtype=np.double), val)[:, None, None]
        elif lambda0_um is not None:
            val = val * (J_in_um / lambda0_um)
        return val

    def _angular(self, val):
        """Integral over the directions (far) or the screen (near) of a
        (n_omega, n1, n_phi) spectrum: one value per omega."""
        if self.Args["mode"] == "far":
            w1 = polar_angle_weights(self.Args["theta"])
        else:
            w1 = radial_weights(self.Args["radius"])
        return (val * _like(w1, val)[None, :, None]).sum((1, 2)) * self.Args["dph"]

    # ------------------------------------------------------------ API
    def get_full_spectrum(self, spect_filter=None, phot_num=False, lambda0_um=None,
                          normalize_to_weights=False, comp="total", iteration=-1):
        """The scaled spectrum of one snapshot, (n_omega, n_theta|n_r, n_phi):
        the sum of the outputs (of their squares for coherent runs) or one
        component (complex for coherent runs)."""
        return _host(self._scaled(spect_filter, phot_num, lambda0_um, normalize_to_weights, comp, iteration))

    def get_energy_spectrum(self, spect_filter=None, phot_num=False, lambda0_um=None, **kw_args):
        """Energy (or photon number) per unit frequency: the angular integral."""
        val = self._scaled(spect_filter=spect_filter, phot_num=phot_num, lambda0_um=lambda0_um, **kw_args)
        return _host(self._angular(val))

    def get_energy(self, spect_filter=None, phot_num=False, lambda0_um=None, **kw_args):
        """Total energy (or photon number): the frequency integral of
        get_energy_spectrum."""
        val = self._angular(self._scaled(spect_filter=spect_filter, phot_num=phot_num, lambda0_um=lambda0_um,
                                         **kw_args))
        total = (val * _like(trapezoid_weights(self.Args["omega"]), val)).sum()
        return float(_host(total))

    def _spot(self, k0, spect_filter, phot_num, lambda0_um, kw_args):
        val = self._scaled(spect_filter=spect_filter, phot_num=phot_num, lambda0_um=lambda0_um, **kw_args)
        if k0 is not None:
            return val[nearest_node_rule(np.asarray(self.Args["omega"]), k0)]
        if val.shape[0] > 1:
            w = trapezoid_weights(self.Args["omega"])
            return (val * _like(w, val)[:, None, None]).sum(0)
        # one frequency: its value times the unit element dw = [1]
        return val[0] * _like(np.asarray(self.Args["dw"], dtype=np.double), val)

    def get_spot(self, k0=None, spect_filter=None, phot_num=False, lambda0_um=None, **kw_args):
        """The angular map (n_theta|n_r, n_phi) integrated over frequency, or
        at the node of frequency k0."""
        return _host(self._spot(k0, spect_filter, phot_num, lambda0_um, kw_args))

    def get_spot_cartesian(self, k0=None, th_part=1.0, bins=(200, 200), spect_filter=None,
                           phot_num=False, lambda0_um=None, **kw_args):
        """get_spot resampled onto a bins[0] x bins[1] Cartesian mesh over
        +-th_part * max(theta|r); returns the map and its extent."""
        mode = self.Args["mode"]
        if mode not in ("far", "near"):
            raise ValueError("get_spot_cartesian covers the 'far' and 'near' modes")
        r = self.Args["theta"] if mode == "far" else self.Args["radius"]
        key = (mode, np.asarray(r).tobytes(), np.asarray(self.Args["phi"]).tobytes(), float(th_part),
               tuple(int(b) for b in bins))
        cache = self.__dict__.setdefault("_resamplers", {})
        if key not in cache:
            cache[key] = PolarResampler(r, self.Args["phi"], th_part, bins)
        res = cache[key]
        spot = self._spot(k0, spect_filter, phot_num, lambda0_um, kw_args)
        return _host(res(spot)), res.extent()

    def get_spectral_axis(self):
        """The frequency nodes, Args['omega'] (synchrad/utils.py:154-166)."""
        return self.Args["omega"]

    def exportToVTK(self, *args, **kw_args):
        """VTK export needs tvtk, which this image does not have; like the
        reference without it (synchrad/utils.py:168-215), it only says so."""
        print("TVTK API is not found")

