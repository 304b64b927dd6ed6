"""Writers of fdfault ``.in`` problem files (text format of the reference's
parsers: problem.cpp:27-44, domain.cpp:47-70, fields.cpp:36-41,
block.cpp:65-85, interface.cpp:42-48, friction.cpp:97-123,
slipweak.cpp:63-90, stz.cpp:72-107, outputlist.cpp:40-66,
frontlist.cpp:40-44).

These helpers build the synthetic rupture problems used by the tests and the
benchmark: 2-D mode II / mode III and 3-D planar faults between two blocks,
with slip-weakening, STZ, frictionless or locked interfaces, optional
Drucker-Prager plasticity and optional curved (file) surfaces.  They write the
same tokens the reference's unchanged Python module writes, so a file made here
is read identically by the reference binary and by this package.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np


def _f(x: float) -> str:
    return repr(float(x))


@dataclass
class Pert:
    """One spatial perturbation (pert.cpp:7-43): type t0 x0 dx y0 dy + amplitudes."""
    kind: str
    t0: float
    x0: float
    dx: float
    y0: float
    dy: float
    amps: Sequence[float]

    def text(self) -> str:
        vals = [self.t0, self.x0, self.dx, self.y0, self.dy] + list(self.amps)
        return self.kind + " " + " ".join(_f(v) for v in vals)


@dataclass
class Output:
    name: str
    field: str
    t: Sequence[int]          # tm tp ts
    x: Sequence[int]          # xm xp xs
    y: Sequence[int]
    z: Sequence[int] = (0, 0, 1)


@dataclass
class Problem:
    name: str = "prob"
    datadir: str = "data/"
    nt: int = 10
    dt: float = 0.0
    ttot: float = 0.0
    cfl: float = 0.3
    ninfo: int = 1000
    rkorder: int = 4
    ndim: int = 2
    mode: int = 2
    nx_block: Sequence[Sequence[int]] = ((101,), (51, 51), (1,))
    sbporder: int = 4
    material: str = "elastic"
    iftypes: Sequence[str] = ("slipweak",)
    s0: Sequence[float] = (0., 0., 0., 0., 0., 0.)
    loadfile: str = "none"
    matfile: str = "none"
    plastic_tensor: int = 0
    cdiss: Optional[float] = 0.0
    # per block: dict(coords=(i,j,k), rho, lam, g, [mu c beta eta], x0, l, bc[2ndim], surf[2ndim])
    blocks: List[dict] = field(default_factory=list)
    # per interface: dict(dir, b1, b2, loads=[Pert], loadfile, sw=[Pert] | stz=dict(...), paramfile)
    interfaces: List[dict] = field(default_factory=list)
    outputs: List[Output] = field(default_factory=list)
    front: Optional[Sequence] = None   # ("V", 0.001)

    def nx(self):
        return [sum(b) for b in self.nx_block]

    def text(self) -> str:
        L = ["[fdfault.problem]", self.name, self.datadir, str(self.nt), _f(self.dt), _f(self.ttot),
             _f(self.cfl), str(self.ninfo), str(self.rkorder), ""]
        nx = self.nx()
        L += ["[fdfault.domain]", str(self.ndim), str(self.mode), " ".join(map(str, nx)),
              " ".join(str(len(b)) for b in self.nx_block)]
        for b in self.nx_block:
            L.append(" ".join(map(str, b)))
        L += [str(len(self.iftypes)), " ".join(self.iftypes), str(self.sbporder), self.material, ""]
        L += ["[fdfault.fields]", " ".join(_f(v) for v in self.s0), self.loadfile, self.matfile,
              str(self.plastic_tensor), ""]
        for b in self.blocks:
            c = b["coords"]
            L.append("[fdfault.block%d%d%d]" % tuple(c))
            mat = [b.get("rho", 2.67), b.get("lam", 32.04), b.get("g", 32.04)]
            if self.material == "plastic":
                mat += [b.get("mu", 0.5735), b.get("c", 0.), b.get("beta", 0.2867), b.get("eta", 0.2775)]
            L.append(" ".join(_f(v) for v in mat))
            L.append(" ".join(_f(v) for v in b["x0"][: self.ndim]))
            L.append(" ".join(_f(v) for v in b["l"][: self.ndim]))
            L += list(b["bc"])
            L += list(b.get("surf", ["none"] * (2 * self.ndim)))
            L.append("")
        if self.cdiss is not None:
            L += ["[fdfault.operator]", _f(self.cdiss), ""]
        for i, itf in enumerate(self.interfaces):
            L += ["[fdfault.interface%d]" % i, itf["dir"], "%d %d %d" % tuple(itf["b1"]),
                  "%d %d %d" % tuple(itf["b2"]), ""]
            t = self.iftypes[i]
            if t in ("frictionless", "slipweak", "stz"):
                loads = itf.get("loads", [])
                L += ["[fdfault.friction]", str(len(loads))] + [p.text() for p in loads]
                L += [itf.get("loadfile", "none"), ""]
            if t == "slipweak":
                sw = itf.get("sw", [])
                L += ["[fdfault.slipweak]", str(len(sw))] + [p.text() for p in sw]
                L += [itf.get("paramfile", "none"), ""]
            if t == "stz":
                st = itf["stz"]
                L += ["[fdfault.stz]", _f(st["chi0"]), st.get("statefile", "none"), str(len(st["perts"]))]
                L += [p.text() for p in st["perts"]]
                L += [st.get("paramfile", "none"), ""]
        L.append("[fdfault.outputlist]")
        for o in self.outputs:
            L += [o.name, o.field, "%d %d %d" % tuple(o.t), "%d %d %d" % tuple(o.x),
                  "%d %d %d" % tuple(o.y), "%d %d %d" % tuple(o.z)]
        L += ["", ""]
        L.append("[fdfault.frontlist]")
        if self.front:
            L += ["1", self.front[0], _f(self.front[1])]
        else:
            L.append("0")
        return "\n".join(L) + "\n"

    def write(self, path: str) -> str:
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        with open(path, "w") as fh:
            fh.write(self.text())
        return path


def two_block_fault(ndim=2, mode=2, nx=101, ny=51, nz=51, dx=0.1, nt=10, sbporder=4, rkorder=4,
                    law="slipweak", material="elastic", bc="absorbing", cdiss=0.0, datadir="data/",
                    name="prob", s0=None, loads=None, sw=None, stz=None, outputs=None, front=("V", 0.001),
                    surf=None, plastic=None, cfl=0.3, loadfile="none", paramfile="none", matfile="none",
                    hetstressfile="none", plastic_tensor=0) -> Problem:
    """Two blocks split by a planar fault normal to y (the SCEC-style set-up of
    the reference's problems/test2d.in and problems/tpv4.in).  ``ny`` points per
    block in y; grid spacing ``dx`` km."""
    Lx = (nx - 1) * dx
    Ly = (ny - 1) * dx
    Lz = (nz - 1) * dx
    nxb = ((nx,), (ny, ny), (nz,) if ndim == 3 else (1,))
    x0a = [0., 0., 0.]
    x0b = [0., Ly, 0.]
    lens = [Lx, Ly, Lz]
    nb = 2 * ndim
    bcs_a = [bc] * nb
    bcs_b = [bc] * nb
    bcs_a[3] = "none"
    bcs_b[2] = "none"
    blk = []
    for c, x0, bcs in (((0, 0, 0), x0a, bcs_a), ((0, 1, 0), x0b, bcs_b)):
        b = dict(coords=c, x0=x0, l=lens, bc=bcs)
        if plastic:
            b.update(plastic)
        blk.append(b)
    if surf is not None:
        blk[0]["surf"] = surf[0]
        blk[1]["surf"] = surf[1]
    if s0 is None:
        s0 = (0., 0., 0., 0., 0., 0.)
    xc = 0.5 * Lx
    zc = 0.5 * Lz
    if loads is None:
        if law in ("slipweak", "frictionless", "stz"):
            loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
                     Pert("boxcar", 0., xc, 1.5 * max(dx * 5, 0.3), zc, (1.5 * max(dx * 5, 0.3) if ndim == 3 else 0.),
                          (0., 11.6, 0.))]
        else:
            loads = []
    if sw is None:
        sw = [Pert("constant", 0., 0., 0., 0., 0., (0.4, 0.677, 0.525, 0., 0., 0.))]
    itf = dict(dir="y", b1=(0, 0, 0), b2=(0, 1, 0), loads=loads, sw=sw, loadfile=loadfile, paramfile=paramfile)
    if law == "stz":
        itf["stz"] = stz
    p = Problem(name=name, datadir=datadir, nt=nt, cfl=cfl, rkorder=rkorder, ndim=ndim, mode=mode,
                nx_block=nxb, sbporder=sbporder, material=material, iftypes=(law,), s0=s0,
                cdiss=cdiss, blocks=blk, interfaces=[itf], outputs=list(outputs or []), front=front,
                loadfile=hetstressfile, matfile=matfile, plastic_tensor=plastic_tensor)
    return p
