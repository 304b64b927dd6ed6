"""Exact plane S wave for order-of-accuracy tests of the bulk scheme.

A pulse vy = g(x - cs t), sxy = -rho cs g(x - cs t) (vx = sxx = syy = 0)
solves the 2-D mode II (and 3-D) elastic equations exactly.  It is set as
the initial state of two blocks glued by a locked interface (both sides
carry the same fields, so the interface penalty vanishes) and compared after
T seconds in a band around the interface, far from the outer y (and z)
faces, whose absorbing penalty does not match a wave running along them.
The pulse cos^8 on a compact support is smooth enough for the orders tested.

With lam_var > 0 the blocks read a material file (the per-point material
branch of block.cpp:1048-1278): rho and G stay uniform, so the wave speed and
impedance do not change, and lambda varies smoothly in x and y.  The plane S
wave stays exact, since lambda multiplies only div v = dvx/dx + dvy/dy = 0,
so a mix-up of lambda and G in the per-point branch breaks the order.  (A
lambda random per point keeps the interior order but lowers the boundary
order of sbporder 3 from 3.3 to 1.9: the rough coefficient scatters the
boundary's own error.)
"""
from __future__ import annotations

import numpy as np

from cases import write_case

RHO, LAM, G = 2.67, 32.04, 32.04
CS = np.sqrt(G / RHO)
ZS = RHO * CS


def pulse(xi, x0, a):
    u = (xi - x0) / a
    return np.where(np.abs(u) < 1, np.cos(0.5 * np.pi * u) ** 8, 0.)


def error(open_solver, tmpdir, dx, sbporder, T, Lx, Ly, x0=2.7, a=2.0, band=1.5, ndim=2, Lz=None,
          lam_var=0.0):
    """Max |vy - exact| over the band |y - Ly| < band (and |z - Lz/2| < band
    in 3-D) after round(T/dt) steps.  open_solver(path) returns an object
    with dt, array(), set_array() and step() (the oracle or the library).
    lam_var > 0: lambda = LAM (1 + lam_var u) with the smooth
    u = (1 + sin(6 x / Lx) cos(2.5 y / Ly)) / 2, from a material file."""
    nx, ny = int(round(Lx / dx)) + 1, int(round(Ly / dx)) + 1
    kw = dict(ndim=ndim)
    nz = 1
    if ndim == 3:
        nz = kw["nz"] = int(round(Lz / dx)) + 1
    if lam_var > 0:
        xi = (np.arange(nx) * dx)[:, None, None]
        j = np.arange(2 * ny)
        yj = np.where(j < ny, j * dx, Ly + (j - ny) * dx)[None, :, None]  # two blocks share y = Ly
        u = 0.5 * (1 + np.sin(6 * xi / Lx) * np.cos(2.5 * yj / Ly))
        mat = np.empty((3, nx, 2 * ny, nz))
        mat[0] = RHO
        mat[1] = LAM * (1 + lam_var * u)
        mat[2] = G
        mp = tmpdir + "_mat.bin"
        mat.tofile(mp)
        kw["matfile"] = mp
    p = write_case(tmpdir, "pw", nx=nx, ny=ny, dx=dx, law="locked", loads=[], front=None, nt=1,
                   sbporder=sbporder, bc="absorbing", **kw)
    o = open_solver(p)
    try:
        nt = int(round(T / o.dt))
        x = o.array("x")
        f = np.zeros(o.array("f").shape)
        g = pulse(x[0], x0, a)
        f[1] = g             # vy
        f[3 if ndim == 2 else 4] = -ZS * g  # sxy
        o.set_array("f", f)
        o.set_array("df", np.zeros(o.array("df").shape))
        o.step(nt)
        vy = o.array("f")[1]
        exact = pulse(x[0] - CS * nt * o.dt, x0, a)
        sel = np.abs(x[1] - Ly) < band
        if ndim == 3:
            sel &= np.abs(x[2] - 0.5 * Lz) < band
        return float(np.abs(vy - exact)[sel].max())
    finally:
        o.close()


def order(e_coarse, e_fine):
    return float(np.log2(e_coarse / e_fine))
