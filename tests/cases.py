"""Problem files shared by the parity tests: small variants of the
reference's example problems (problems/test2d.in, problems/tpv4.in) that
exercise every branch of the hot path -- modes, SBP orders, RK orders,
boundary types, interface laws, plasticity, dissipation, heterogeneous
stress/material and curved (file) block surfaces."""
from __future__ import annotations

import os

import numpy as np

from fdfault_amd.problems import Output, Pert, two_block_fault

STZ_PARAMS = dict(chi0=0.0, statefile="none",
                  perts=[Pert("constant", 0., 0., 0., 0., 0., (1.e-6, 0.3, 0.01, 0.1, 1.e4, 1.e-3, 0.1, 0.4, 1.e4))])


def write_case(tmpdir: str, name: str, **kw) -> str:
    d = os.path.join(tmpdir, name)
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    kw.setdefault("datadir", os.path.join(d, "data") + "/")
    kw.setdefault("name", name)
    p = two_block_fault(**kw)
    return p.write(os.path.join(d, name + ".in"))


def curved_surfaces(tmpdir, nx, ny, dx, amp=0.05, ndim=2, nz=1):
    """Block surfaces for a sinusoidal fault y = Ly + amp*sin(2 pi x / Lx):
    block 1 (below) gets the curve as its top face, block 2 as its bottom face
    (surface files [ndim][n][1], surface.cpp:11-50)."""
    Lx = (nx - 1) * dx
    Ly = (ny - 1) * dx
    xs = np.linspace(0., Lx, nx)
    yf = Ly + amp * np.sin(2 * np.pi * xs / Lx)
    paths = []
    for blk, y0 in ((0, 0.), (1, Ly)):
        faces = []
        for face in range(4):
            if face in (0, 1):  # x faces: curve in y
                x = 0. if face == 0 else Lx
                yb = 0. if blk == 0 else yf[0 if face == 0 else -1]
                yt = yf[0 if face == 0 else -1] if blk == 0 else 2 * Ly
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
            pth = os.path.join(tmpdir, "surf_b%d_f%d.bin" % (blk, face))
            np.ascontiguousarray(arr).tofile(pth)
            faces.append(pth)
        paths.append(faces)
    return paths


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def reference_example(tmpdir: str, name: str, nt: int, datadir: str, text: str = None) -> str:
    """The reference's example problem `name` (tests/golden/<name>.in, a
    verbatim copy of problems/<name>.in, or the given text) with its data
    directory set to `datadir`, its step count cut to `nt` and each output
    unit's last step clipped to nt; every other token is unchanged."""
    if text is None:
        text = open(os.path.join(GOLDEN, name + ".in")).read()
    lines = text.split("\n")
    out, sec, k = [], None, 0
    for ln in lines:
        t = ln.strip()
        if t.startswith("["):
            sec, k = t, 0
            out.append(ln)
            continue
        if sec == "[fdfault.problem]" and t:
            k += 1
            if k == 2:
                ln = datadir
            elif k == 3:
                ln = str(nt)
        elif sec == "[fdfault.outputlist]" and t:
            k += 1
            if k % 6 == 3:  # "tmin tmax tstride" of a unit
                tm, tM, ts = t.split()
                ln = "%s %d %s" % (tm, min(int(tM), nt), ts)
        out.append(ln)
    os.makedirs(os.path.join(tmpdir, name), exist_ok=True)
    path = os.path.join(tmpdir, name, name + ".in")
    open(path, "w").write("\n".join(out))
    return path




GOLDEN_PY = os.path.join(GOLDEN, "python")


def python_example(tmpdir: str, name: str, nt: int, datadir: str) -> str:
    """The reference's Python-written example `name` (tests/golden/python/
    <name>.in, written by the unchanged fdfault module from
    problems/<name>.py) shortened like reference_example, with its side-file
    paths (written relative to the reference checkout as problems/...)
    pointed at the fixture directory."""
    text = open(os.path.join(GOLDEN_PY, name + ".in")).read().replace("problems/", GOLDEN_PY + "/")
    return reference_example(tmpdir, name, nt, datadir, text=text)
