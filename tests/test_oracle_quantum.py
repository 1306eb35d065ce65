"""The quantum thermostat's filter in the oracle against the mathematics it
follows (quantum_noise_CPU.cpp:83-113): H is the cosine transform, over the
2 Nf points m = -Nf .. Nf-1, of the spectrum

    H~(w) = sqrt(hbar |w| (1/2 + 1/(exp(hbar |w| / T) - 1))) * (w h/2) / sin(w h/2),

sampled at w_k = k pi / (h Nf), h = M dt / 2, with H~(0) = sqrt(T). The
transform is its own inverse up to 2 Nf on this grid, so transforming H back
must give the spectrum, evaluated here independently with numpy; H is even
about m = 0. A wrong sign, scale, time step or index in the oracle fails
one of these."""
import numpy as np
import pytest

from oracle import orcpy
from spilady_amd import abi, configs

HBAR = 6.58211899e-16


def _filter(T_kelvin, nf, m, step=1e-15):
    at = configs.make_bcc_fe(5, T_kelvin, 3)
    v = configs.fe_variables(T_kelvin, 3, thermostats=True, step=step)
    v.update({"quantumnoise": 1, "Msteps_quantum": m, "Nfrequency_quantum": nf})
    p = abi.params_from_variables(v, at.box, at.natom)
    o = orcpy.Oracle(p, at, compute_initial_forces=True)
    H = o.quantum_H()
    o.close()
    return H, p.temperature, m * step / 2.0


@pytest.mark.parametrize("T_kelvin,nf,m", [(10.0, 8, 3), (300.0, 150, 50), (3000.0, 40, 7)])
def test_filter_inverts_to_the_spectrum(T_kelvin, nf, m):
    H, T, h = _filter(T_kelvin, nf, m)
    assert H.size == 2 * nf
    k = np.arange(-nf, nf)
    w = k * np.pi / (h * nf)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        x = HBAR * np.abs(w)
        spec = np.sqrt(x * (0.5 + 1.0 / np.expm1(x / T))) * (w * h / 2) / np.sin(w * h / 2)
    spec[k == 0] = np.sqrt(T)
    back = np.array([np.sum(H * np.cos(np.pi * kk * k / nf)) for kk in k])
    np.testing.assert_allclose(back, spec, rtol=0, atol=1e-12 * np.abs(spec).max())
    # even about m = 0: H[Nf + j] = H[Nf - j]
    np.testing.assert_allclose(H[nf + 1:], H[nf - 1:0:-1], rtol=0, atol=1e-15 * np.abs(H).max())
    # the m = 0 tap is the spectrum's mean, sqrt(T) among the rest
    assert H[nf] == pytest.approx(spec.sum() / (2 * nf), rel=1e-12)


def test_filter_classical_limit():
    """At T >> hbar w_max the spectrum tends to sqrt(T) (w h/2)/sin(w h/2):
    hbar w (1/2 + 1/expm1(hbar w/T)) = T (1 + (hbar w/T)^2/12 + ...)."""
    nf, m = 20, 4
    H, T, h = _filter(1e7, nf, m)
    k = np.arange(-nf, nf)
    w = k * np.pi / (h * nf)
    with np.errstate(divide="ignore", invalid="ignore"):
        lim = np.sqrt(T) * (w * h / 2) / np.sin(w * h / 2)
    lim[k == 0] = np.sqrt(T)
    back = np.array([np.sum(H * np.cos(np.pi * kk * k / nf)) for kk in k])
    xmax = HBAR * np.pi / h / T
    np.testing.assert_allclose(back, lim, rtol=xmax ** 2 / 20 + 1e-12)
