"""Seeded ensemble of uneven 2-D slip-weakening ruptures (the batched
workload): per instance a grid of nx x ny points with nx, ny uniform on
[nmin, nmax] (two blocks of nx x ny//2 split by a planar fault at mid-y),
dx = 50 m, shear prestress 70 MPa plus a Gaussian random field (std 3 MPa,
correlation length 1 km, seeded per instance) over -120 MPa normal stress,
and a 3 km nucleation patch (+11.6 MPa) at a random position along the
fault.  Friction: slip weakening mu_s 0.677, mu_d 0.525, dc 0.4 m; SBP
order 4 interior (sbporder 3); RK4; absorbing outer boundaries.
"""
from __future__ import annotations

import os

import numpy as np

from .generators import gaussian_random_field, write_load_file
from .problems import Pert, two_block_fault


def ensemble_workload(d, n=64, seed=2024, nmin=201, nmax=1201, dx=0.05, nt=2000, sbporder=3, sizes=None):
    """Write the .in and load files of n instances under directory d; returns
    (paths, shapes) with shapes[i] = (nx, ny) of instance i's grid."""
    rng = np.random.default_rng(seed)
    paths, shapes = [], []
    os.makedirs(d, exist_ok=True)
    for i in range(n):
        if sizes is not None:
            nx, ny = sizes[i]
        else:
            nx = int(rng.integers(nmin, nmax + 1))
            ny = int(rng.integers(nmin, nmax + 1))
        nyb = max(ny // 2, 13)
        Lx = (nx - 1) * dx
        tau = 70.0 + gaussian_random_field((nx,), dx, 1.0, 3.0, seed * 1000 + i)
        sub = os.path.join(d, "ens%02d" % i)
        os.makedirs(os.path.join(sub, "data"), exist_ok=True)
        lf = write_load_file(os.path.join(sub, "load.bin"), np.full(nx, -120.0), tau, np.zeros(nx))
        xc = float(rng.uniform(0.25, 0.75)) * Lx
        loads = [Pert("boxcar", 0., xc, 1.5, 0., 0., (0., 11.6, 0.))]
        p = two_block_fault(ndim=2, nx=nx, ny=nyb, dx=dx, nt=nt, sbporder=sbporder, loads=loads,
                            datadir=os.path.join(sub, "data") + "/", name="ens%02d" % i, loadfile=lf,
                            front=("V", 0.001))
        paths.append(p.write(os.path.join(sub, "ens%02d.in" % i)))
        shapes.append((nx, 2 * nyb))
    return paths, shapes
