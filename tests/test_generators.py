"""Property tests of the seeded input generators behind the benchmark
workloads (fdfault_amd/generators.py, fdfault_amd/ensemble.py); the
reference has no generators, so each is checked against what its definition
fixes: the covariance of the Gaussian random field, the spectrum and rms of
the self-similar profile, and the binary layouts the reference's readers
expect (friction.cpp:507-567 load files, surface.cpp:11-50 curves)."""
import os

import numpy as np

from fdfault_amd.ensemble import ensemble_workload
from fdfault_amd.generators import fault_surfaces, gaussian_random_field, self_similar_profile, write_load_file


def _autocorr(field, lag, axis):
    a = np.moveaxis(field, axis, 0)
    return float(np.mean(a[lag:] * a[:-lag]) / np.mean(a * a))


def test_grf_moments_and_seed():
    f = gaussian_random_field((256, 192), 0.1, 1.0, 5.0, 11)
    assert abs(f.mean()) < 1e-12
    assert abs(f.std() - 5.0) < 1e-12
    assert np.array_equal(f, gaussian_random_field((256, 192), 0.1, 1.0, 5.0, 11))
    assert not np.array_equal(f, gaussian_random_field((256, 192), 0.1, 1.0, 5.0, 12))


def test_grf_gaussian_covariance():
    """Empirical correlation at lags l/2, l, 2l against exp(-r^2/l^2)
    (0.779, 0.368, 0.018), averaged over seeds and both axes; a filter with
    the wrong width (say exp(-k^2 l^2 / 4), the power spectrum itself) would
    give exp(-r^2 / (2 l^2)): 0.882, 0.607, 0.135."""
    l, dx = 1.0, 0.0625  # 16 points per correlation length
    lags = {8: np.exp(-0.25), 16: np.exp(-1.0), 32: np.exp(-4.0)}
    acc = {k: [] for k in lags}
    for seed in range(6):
        f = gaussian_random_field((512, 512), dx, l, 1.0, seed)
        for k in lags:
            acc[k] += [_autocorr(f, k, 0), _autocorr(f, k, 1)]
    for k, want in lags.items():
        got = float(np.mean(acc[k]))
        print("lag %d: %.3f vs %.3f" % (k, got, want))
        assert abs(got - want) < 0.04, (k, got, want)


def test_self_similar_profile_spectrum_and_rms():
    n, dx, H, ratio, lmin, lmax = 4096, 0.025, 1.0, 1e-2, 0.2, 20.0
    h = self_similar_profile(n, dx, H, ratio, lmin, lmax, 7)
    assert abs(h.mean()) < 1e-12
    assert abs(h.std() - ratio * lmax / (2 * np.sqrt(2))) < 1e-12
    k = 2 * np.pi * np.fft.rfftfreq(n, d=dx)
    amp = np.abs(np.fft.rfft(h))
    band = (k >= 2 * np.pi / lmax) & (k <= 2 * np.pi / lmin)
    out = ~band & (k > 0)
    assert amp[out].max() < 1e-9 * amp[band].max()  # band-limited
    # power ~ k^-(2H+1): amplitude slope -(2H+1)/2 in log-log, exactly (random phases only)
    slope = np.polyfit(np.log(k[band]), np.log(amp[band]), 1)[0]
    assert abs(slope + (2 * H + 1) / 2) < 1e-9, slope
    assert np.array_equal(h, self_similar_profile(n, dx, H, ratio, lmin, lmax, 7))


def test_load_file_layout(tmp_path):
    """[3][n0][n1] float64, C order: sn, then s2, then s3 (friction.cpp:555-567)."""
    n0, n1 = 5, 7
    sn = np.arange(n0 * n1, dtype=float).reshape(n0, n1)
    s2, s3 = sn + 100., sn + 200.
    p = write_load_file(str(tmp_path / "l.bin"), sn, s2, s3)
    a = np.fromfile(p).reshape(3, n0, n1)
    assert np.array_equal(a[0], sn) and np.array_equal(a[1], s2) and np.array_equal(a[2], s3)
    assert a[0, 1, 0] == 7.0  # C order: the second index is the fastest of the two


def test_fault_surfaces_share_the_profile(tmp_path):
    nx, ny, dx = 41, 21, 0.1
    h = self_similar_profile(nx, dx, 1.0, 1e-2, 0.4, 4.0, 3)
    paths = fault_surfaces(str(tmp_path), nx, ny, dx, h)
    Ly = (ny - 1) * dx
    top1 = np.fromfile(paths[0][3]).reshape(2, nx, 1)
    bot2 = np.fromfile(paths[1][2]).reshape(2, nx, 1)
    assert np.array_equal(top1, bot2)
    assert np.allclose(top1[1, :, 0], Ly + h)
    for blk in (0, 1):  # side faces meet the fault at its ends
        left = np.fromfile(paths[blk][0]).reshape(2, ny, 1)
        right = np.fromfile(paths[blk][1]).reshape(2, ny, 1)
        end = -1 if blk == 0 else 0
        assert np.isclose(left[1, end, 0], Ly + h[0]) and np.isclose(right[1, end, 0], Ly + h[-1])
        assert np.all(np.diff(left[1, :, 0]) > 0) and np.all(left[0] == 0.)


def test_ensemble_workload_sizes_and_loads(tmp_path):
    paths, shapes = ensemble_workload(str(tmp_path / "a"), n=6, nt=5, nmin=31, nmax=61)
    paths2, shapes2 = ensemble_workload(str(tmp_path / "b"), n=6, nt=5, nmin=31, nmax=61)
    assert shapes == shapes2 and len(paths) == 6
    for (nx, ny), p in zip(shapes, paths):
        assert 31 <= nx <= 61 and 30 <= ny <= 62
        assert os.path.exists(p)
        lf = np.fromfile(os.path.join(os.path.dirname(p), "load.bin")).reshape(3, nx)
        assert np.all(lf[0] == -120.) and np.all(lf[2] == 0.)
        assert abs(lf[1].mean() - 70.) < 1e-9 and abs(lf[1].std() - 3.) < 1e-9
