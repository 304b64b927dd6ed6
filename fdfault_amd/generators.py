"""Seeded synthetic inputs of the benchmark workloads, written in the
reference's binary side-file layouts (native-endian float64, C order):

* :func:`gaussian_random_field` -- stationary field with Gaussian covariance
  C(r) = std^2 exp(-r^2 / l^2), by spectral filtering of white noise;
* :func:`write_load_file` -- friction load file ``[3][n0][n1]`` (sn, s2, s3),
  read by friction::read_load (friction.cpp:507-567);
* :func:`self_similar_profile` -- band-limited self-similar rough-fault
  profile (Hurst exponent H, amplitude/wavelength ratio) for curved faults.
"""
from __future__ import annotations

import numpy as np


def gaussian_random_field(shape, dx, corr_len, std, seed):
    """Zero-mean field on a regular grid (spacing dx, any dimension) with
    covariance std^2 exp(-r^2/corr_len^2); exactly `std` sample deviation."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    spec = np.fft.fftn(noise)
    k2 = np.zeros(shape)
    for ax, n in enumerate(shape):
        k = 2 * np.pi * np.fft.fftfreq(n, d=dx)
        sh = [1] * len(shape)
        sh[ax] = n
        k2 = k2 + (k ** 2).reshape(sh)
    # power spectrum of a Gaussian covariance exp(-r^2/l^2) ~ exp(-k^2 l^2 / 4);
    # the amplitude filter is its square root
    field = np.real(np.fft.ifftn(spec * np.exp(-k2 * corr_len ** 2 / 8.0)))
    field -= field.mean()
    field *= std / field.std()
    return field


def write_load_file(path, sn, s2, s3):
    arr = np.stack([np.asarray(sn, float), np.asarray(s2, float), np.asarray(s3, float)])
    np.ascontiguousarray(arr).tofile(path)
    return path


def self_similar_profile(n, dx, hurst, amp_ratio, lmin, lmax, seed):
    """Random profile h(x) with power spectrum ~ k^-(2H+1) between
    wavelengths lmin and lmax, scaled so rms(h) = amp_ratio * lmax / (2 sqrt 2)
    (the usual amplitude-to-wavelength definition of rough-fault studies)."""
    rng = np.random.default_rng(seed)
    k = 2 * np.pi * np.fft.rfftfreq(n, d=dx)
    amp = np.zeros_like(k)
    band = (k >= 2 * np.pi / lmax) & (k <= 2 * np.pi / lmin)
    amp[band] = k[band] ** (-(2 * hurst + 1) / 2.0)
    phase = rng.uniform(0, 2 * np.pi, size=k.shape)
    h = np.fft.irfft(amp * np.exp(1j * phase), n=n)
    h -= h.mean()
    target = amp_ratio * lmax / (2 * np.sqrt(2))
    if h.std() > 0:
        h *= target / h.std()
    return h
