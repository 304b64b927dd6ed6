"""Seeded synthetic inputs of the benchmark workloads, written in the
reference's binary side-file layouts (native-endian float64, C order):

* :func:`gaussian_random_field` -- stationary field with Gaussian covariance
  C(r) = std^2 exp(-r^2 / l^2), by spectral filtering of white noise;
* :func:`write_load_file` -- friction load file ``[3][n0][n1]`` (sn, s2, s3),
  read by friction::read_load (friction.cpp:507-567);
* :func:`self_similar_profile` -- band-limited self-similar rough-fault
  profile (Hurst exponent H, amplitude/wavelength ratio) for curved faults;
* :func:`fault_surfaces` -- the block boundary curves of a two-block 2-D
  problem whose shared face follows a fault profile (surface files
  ``[2][n][1]``, surface.cpp:11-50).
"""
from __future__ import annotations

import os

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


def fault_surfaces(d, nx, ny, dx, h):
    """Boundary curves of two stacked 2-D blocks (nx x ny points each, spacing
    dx km) whose shared face is the fault y = Ly + h(x): block 1 gets it as
    its top face (3), block 2 as its bottom face (2); side faces are straight
    lines between the corners.  Returns [[4 paths of block 1], [4 of block 2]]."""
    Lx, Ly = (nx - 1) * dx, (ny - 1) * dx
    xs = np.linspace(0., Lx, nx)
    yf = Ly + np.asarray(h, float)
    paths = []
    for blk in (0, 1):
        faces = []
        for face in range(4):
            if face in (0, 1):
                x = 0. if face == 0 else Lx
                yc = yf[0 if face == 0 else -1]
                yb, yt = (0., yc) if blk == 0 else (yc, 2 * Ly)
                arr = np.zeros((2, ny, 1))
                arr[0] = x
                arr[1, :, 0] = np.linspace(yb, yt, ny)
            else:
                if (blk == 0 and face == 3) or (blk == 1 and face == 2):
                    yy = yf
                else:
                    yy = np.full(nx, 0. if face == 2 else 2 * Ly)
                arr = np.zeros((2, nx, 1))
                arr[0, :, 0] = xs
                arr[1, :, 0] = yy
            pth = os.path.join(d, "surf_b%d_f%d.bin" % (blk, face))
            np.ascontiguousarray(arr).tofile(pth)
            faces.append(pth)
        paths.append(faces)
    return paths
