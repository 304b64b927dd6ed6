"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests and no golden outputs, and its C++ needs MPI
(absent here), so the oracle is pinned by independent checks chosen so that
a plausible slip (a dropped term, wrong sign or index, swapped constant)
fails one of them:

* SBP tables (fd.cpp:33-155): derivative rows exact on polynomials, the
  summation-by-parts identity Q + Q^T = diag(-1,0..0,1) with H from h0;
* RK tables (rk.cpp:8-98): order conditions via the stability polynomial and
  stage times via exact integration of polynomials in t;
* interface/friction (interface.cpp:1017-1028, friction.cpp:293-435):
  characteristic invariants of the locked solve, |traction| = strength,
  frictionless S = 0;
* STZ root (utilities.cpp:89-180): residual and brute-force bisection;
* plasticity (block.cpp:591-697): the viscoplastic return-map identity;
* rotations (boundary.cpp:505-568): against R S R^T;
* bulk RHS (block.cpp:1048-2050): exact rates for linear velocity and stress
  fields, free-stream preservation on curved grids, dissipation null space;
* absorbing SAT (boundary.cpp:265-503): vanishes on an outgoing wave;
* load shapes (pert.cpp:59-148): closed forms at t = 0;
* fronts (front.cpp:103-125): brute force from the slip-rate history;
* outputs (outputunit.cpp): absolute stress at t = 0 equals s0, file shapes;
* dt (problem.cpp:94-134): cfl * h / cs on a Cartesian grid.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from cases import curved_surfaces, write_case
from fdfault_amd.problems import Output, Pert

L = O.lib()
dp = C.POINTER(C.c_double)


def _d(a):
    return a.ctypes.data_as(dp)


def fd_tables(s):
    fd = np.zeros(63)
    di = np.zeros(63)
    h0 = C.c_double()
    assert L.orc_fd_coeffs(s, _d(fd), _d(di), C.byref(h0)) == 0
    return fd.reshape(7, 9), di.reshape(7, 9), h0.value


def d_matrix(s, n):
    c, _, _ = fd_tables(s)
    D = np.zeros((n, n))
    nb = 2 * (s - 1)
    for i in range(n):
        if i < nb:
            D[i, :3 * (s - 1)] = c[i + 1, :3 * (s - 1)]
        elif i < n - nb:
            D[i, i - s + 1:i + s] = c[0, :2 * s - 1]
        else:
            r = n - i
            for t in range(3 * (s - 1)):
                D[i, n - 1 - t] -= c[r, t]
    return D


@pytest.mark.parametrize("s", [2, 3, 4])
def test_sbp_rows_exact_on_polynomials(s):
    n = 40
    D = d_matrix(s, n)
    x = np.arange(n, dtype=float)
    nb = 2 * (s - 1)
    for p in range(0, 2 * (s - 1) + 1):
        err = D @ x ** p - (p * x ** (p - 1) if p else 0 * x)
        scale = max(1.0, np.abs(x ** p).max())
        assert np.abs(err[nb:n - nb]).max() <= 1e-9 * scale, (s, p)
        if p <= s - 1:
            assert np.abs(err[:nb]).max() <= 1e-9 * scale, (s, p)
            assert np.abs(err[n - nb:]).max() <= 1e-9 * scale, (s, p)


@pytest.mark.parametrize("s", [2, 3, 4])
def test_sbp_summation_by_parts(s):
    n = 40
    D = d_matrix(s, n)
    _, _, h0 = fd_tables(s)
    h = np.ones(n)
    h[0] = h[-1] = h0
    nb = 2 * (s - 1)
    # recover the remaining boundary weights from the off-diagonal symmetry
    for i in range(1, nb):
        h[i] = h[n - 1 - i] = -h[0] * D[0, i] / D[i, 0] if D[i, 0] != 0 else h[i]
    for i in range(1, nb):
        # weights from any nonzero pair with an interior-row point are equivalent
        j = next(j for j in range(n) if j != i and D[i, j] != 0 and D[j, i] != 0 and (j >= nb or h[j] != 1.0 or j == 0))
        h[i] = h[n - 1 - i] = -h[j] * D[j, i] / D[i, j]
    Q = np.diag(h) @ D
    B = np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    tol = 1e-13 if s < 4 else 2e-12
    assert np.abs(Q + Q.T - B).max() <= tol
    assert (h > 0).all()


@pytest.mark.parametrize("s", [2, 3, 4])
def test_dissipation_annihilates_low_degree(s):
    _, di, _ = fd_tables(s)
    n = 40
    x = np.arange(n, dtype=float)
    nb = 2 * (s - 1)
    Dd = np.zeros((n, n))
    for i in range(n):
        if i < nb:
            Dd[i, :3 * (s - 1)] = di[i + 1, :3 * (s - 1)]
        elif i < n - nb:
            Dd[i, i - s + 1:i + s] = di[0, :2 * s - 1]
        else:
            for t in range(3 * (s - 1)):
                Dd[i, n - 1 - t] += di[n - i, t]
    for p in range(s - 1 if s > 2 else 1):
        assert np.abs(Dd @ x ** p).max() <= 1e-9 * max(1.0, (x ** p).max()), p


def rk(order):
    ns = C.c_int()
    A, B, Cc = np.zeros(5), np.zeros(5), np.zeros(6)
    assert L.orc_rk_coeffs(order, C.byref(ns), _d(A), _d(B), _d(Cc)) == 0
    return ns.value, A, B, Cc


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_rk_order_conditions(order):
    """One 2N-storage step on y' = z y gives R(z); its Taylor coefficients
    must be 1/k! up to the order (stability polynomial of an order-p method)."""
    ns, A, B, Cc = rk(order)
    u = np.zeros(order + 2)
    u[0] = 1.0
    du = np.zeros_like(u)
    for st in range(ns):
        du = A[st] * du + np.concatenate([[0.0], u[:-1]])  # + z * u
        u = u + B[st] * du
    fact = [1.0 / np.math.factorial(k) for k in range(order + 1)] if hasattr(np, "math") else None
    import math
    for k in range(order + 1):
        assert abs(u[k] - 1.0 / math.factorial(k)) <= 1e-12, (order, k)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_rk_stage_times(order):
    """Stage times C: y' = t^m integrated exactly for m < order."""
    ns, A, B, Cc = rk(order)
    for m in range(order):
        y, dy, dt = 0.0, 0.0, 1.0
        for st in range(ns):
            dy = A[st] * dy + dt * (Cc[st] * dt) ** m
            y = y + B[st] * dy
        assert abs(y - 1.0 / (m + 1)) <= 1e-12, (order, m)


def test_locked_characteristics():
    rng = np.random.default_rng(3)
    for _ in range(50):
        s1, s2, v1, v2 = rng.normal(size=4) * [50, 50, 1, 1]
        z1, z2 = rng.uniform(1, 20, 2)
        io = np.array([s1, s2, v1, v2])
        L.orc_solve_locked(_d(io), z1, z2)
        S1, S2, V1, V2 = io
        assert S1 == S2 and V1 == V2
        assert abs((S1 - z1 * V1) - (s1 - z1 * v1)) <= 1e-12 * max(1, abs(s1))
        assert abs((S2 + z2 * V2) - (s2 + z2 * v2)) <= 1e-12 * max(1, abs(s2))


def test_rotation_matches_tensor_algebra():
    rng = np.random.default_rng(4)
    for _ in range(20):
        Rm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        nn, t1, t2 = Rm[0].copy(), Rm[1].copy(), Rm[2].copy()
        b = rng.normal(size=9)
        v = b[:3]
        S = np.array([[b[3], b[4], b[5]], [b[4], b[6], b[7]], [b[5], b[7], b[8]]])
        out = np.zeros(9)
        L.orc_rotate(_d(b), _d(nn), _d(t1), _d(t2), 1, _d(out))
        vr = Rm @ v
        Sr = Rm @ S @ Rm.T
        ref = np.array([vr[0], vr[1], vr[2], Sr[0, 0], Sr[0, 1], Sr[0, 2], Sr[1, 1], Sr[1, 2], Sr[2, 2]])
        np.testing.assert_allclose(out, ref, atol=1e-12)
        back = np.zeros(9)
        L.orc_rotate(_d(out), _d(nn), _d(t1), _d(t2), 0, _d(back))
        np.testing.assert_allclose(back, b, atol=1e-12)


def _bisect(p, lo, hi):
    f = lambda m: L.orc_stz_func(m, _d(p))
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) < 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_stz_newton_root():
    rng = np.random.default_rng(5)
    n = 0
    for _ in range(200):
        snc = -rng.uniform(50, 150)
        phi = rng.uniform(10, 120)
        eta = rng.uniform(1, 5)
        chi = rng.uniform(0.05, 0.1)
        p = np.array([phi, eta, snc, chi, 1e-6, 0.3, 0.01, rng.uniform(0.1, 0.6)])
        st = C.c_int()
        mu = L.orc_stz_mu(_d(p), C.byref(st))
        assert st.value == 0
        hi = -phi / snc
        if p[7] >= hi:
            assert mu == hi
            continue
        n += 1
        assert p[7] <= mu <= hi
        ref = _bisect(p, p[7], hi)
        assert abs(mu - ref) <= 1e-12 * abs(ref)
    assert n > 50


def test_plastic_return_map_identity():
    rng = np.random.default_rng(6)
    mu, c, beta, eta, dt = 0.6, 0.5, 0.3, 1.6, 0.01
    lam, G = 32.04, 32.04
    k = lam + 2.0 / 3.0 * G
    yielded = 0
    for _ in range(200):
        s = rng.normal(size=6) * 40 - np.array([80, 0, 0, 80, 0, 80]) * rng.uniform(0, 1.5)
        sin = np.zeros(14)
        sin[:6] = s
        sin[6] = 0.1  # gammap
        out = np.zeros(14)
        L.orc_plastic_flow(dt, _d(sin), k, G, mu, c, beta, eta, _d(out))
        sig = s[[0, 3, 5]].sum() / 3
        tau = np.sqrt(((s[0] - s[3]) ** 2 + (s[3] - s[5]) ** 2 + (s[5] - s[0]) ** 2) / 6 + s[1] ** 2 + s[2] ** 2 + s[4] ** 2)
        if tau - max(0.0, c - mu * sig) <= 0:
            assert out[7] == 0.0 and np.array_equal(out[:6], s)
            continue
        yielded += 1
        so = out[:6]
        sig2 = so[[0, 3, 5]].sum() / 3
        tau2 = np.sqrt(((so[0] - so[3]) ** 2 + (so[3] - so[5]) ** 2 + (so[5] - so[0]) ** 2) / 6 + so[1] ** 2 + so[2] ** 2 + so[4] ** 2)
        Lm = out[7]
        assert Lm > 0
        if tau * mu * beta * k > (mu * sig - c) * G:
            assert abs(tau2 + mu * sig2 - c - eta * Lm) <= 1e-10 * max(1.0, tau)
        else:
            assert abs(tau2 - eta * Lm) <= 1e-10 * max(1.0, tau)
        assert abs(out[6] - (0.1 + dt * Lm)) <= 1e-14
        dev_in = s - sig * np.array([1, 0, 0, 1, 0, 1])
        dev_out = so - sig2 * np.array([1, 0, 0, 1, 0, 1])
        np.testing.assert_allclose(dev_out * tau, dev_in * tau2, atol=1e-9 * tau * tau)
    assert yielded > 30


# --------------------------------------------------------------------------
# bulk operator pins through one RK stage on a quiet problem


def quiet_case(tmp_path, name, **kw):
    """Locked interface, no loads, 'none' external boundaries: df after one
    stage from df = 0 is exactly dt * RHS of the bulk operator."""
    kw.setdefault("law", "locked")
    kw.setdefault("bc", "none")
    kw.setdefault("front", None)
    kw.setdefault("nt", 1)
    kw.setdefault("rkorder", 1)
    kw.setdefault("loads", [])
    return write_case(str(tmp_path), name, **kw)


def _coords(o):
    x = o.array("x")
    return x


@pytest.mark.parametrize("sbp", [2, 3, 4])
def test_linear_velocity_gives_exact_stress_rate_2d(tmp_path, sbp):
    p = quiet_case(tmp_path, "lin2", nx=31, ny=16, dx=0.1, sbporder=sbp)
    o = O.Oracle(p)
    x = _coords(o)
    X, Y = x[0], x[1]
    a, b, c, d = 0.3, -0.7, 1.1, 0.4
    f = np.zeros(o.array("f").shape)
    f[0] = a * X + b * Y  # vx
    f[1] = c * X + d * Y  # vy
    o.set_array("f", f)
    o.set_array("df", np.zeros(o.array("df").shape))
    o.stage(0, 0)
    df = o.array("df")
    lam = G = 32.04
    dt = o.dt
    np.testing.assert_allclose(df[2], dt * ((lam + 2 * G) * a + lam * d), rtol=1e-10)
    np.testing.assert_allclose(df[3], dt * G * (b + c), rtol=1e-10)
    np.testing.assert_allclose(df[4], dt * ((lam + 2 * G) * d + lam * a), rtol=1e-10)
    assert np.abs(df[:2]).max() <= 1e-10


def test_linear_stress_gives_exact_acceleration_3d(tmp_path):
    p = quiet_case(tmp_path, "lin3", ndim=3, nx=21, ny=13, nz=15, dx=0.1)
    o = O.Oracle(p)
    x = _coords(o)
    X, Y, Z = x
    f = np.zeros(o.array("f").shape)
    # sxx sxy sxz syy syz szz with gradients
    f[3] = 2.0 * X
    f[4] = 1.5 * Y + 0.5 * X
    f[5] = -1.0 * Z
    f[6] = 3.0 * Y
    f[7] = 0.25 * Z + 0.5 * Y
    f[8] = -2.0 * Z + X
    o.set_array("f", f)
    o.set_array("df", np.zeros(o.array("df").shape))
    o.stage(0, 0)
    df = o.array("df")
    dt, rho = o.dt, 2.67
    np.testing.assert_allclose(df[0], dt / rho * (2.0 + 1.5 - 1.0), rtol=1e-10)
    np.testing.assert_allclose(df[1], dt / rho * (0.5 + 3.0 + 0.25), rtol=1e-10)
    np.testing.assert_allclose(df[2], dt / rho * (0.0 + 0.5 - 2.0), rtol=1e-10, atol=1e-12)
    assert np.abs(df[3:]).max() <= 1e-10


def test_free_stream_on_curved_grid(tmp_path):
    surf = curved_surfaces(str(tmp_path), 41, 21, 0.1, amp=0.15)
    p = quiet_case(tmp_path, "fs", nx=41, ny=21, dx=0.1, surf=surf)
    o = O.Oracle(p)
    f = np.zeros(o.array("f").shape)
    f[:] = np.array([0.3, -0.2, 5.0, -3.0, 7.0])[:, None, None, None]
    o.set_array("f", f)
    o.set_array("df", np.zeros(o.array("df").shape))
    o.stage(0, 0)
    assert np.abs(o.array("df")).max() <= 1e-11


def test_dissipation_null_on_linear_fields(tmp_path):
    p0 = quiet_case(tmp_path, "d0", nx=31, ny=16, dx=0.1, cdiss=0.0)
    p1 = quiet_case(tmp_path, "d1", nx=31, ny=16, dx=0.1, cdiss=0.3)
    res = []
    for p in (p0, p1):
        o = O.Oracle(p)
        x = _coords(o)
        f = np.zeros(o.array("f").shape)
        f[0] = 0.4 * x[0] - 0.1 * x[1]
        f[2] = 2.0 * x[1]
        o.set_array("f", f)
        o.set_array("df", np.zeros(o.array("df").shape))
        o.stage(0, 0)
        res.append(o.array("df"))
    assert np.abs(res[0] - res[1]).max() <= 1e-11


def test_absorbing_sat_vanishes_on_outgoing_wave(tmp_path):
    """Uniform state with sxx = -Zp vx on the x-minus face (outward normal
    -x): the normal characteristic entering there is zero, so the absorbing
    SAT adds nothing on that face (and the bulk RHS of a uniform state is 0)."""
    p = write_case(str(tmp_path), "abs", nx=31, ny=16, dx=0.1, law="locked", loads=[], front=None, nt=1, rkorder=1,
                   bc="absorbing")
    o = O.Oracle(p)
    rho, lam, G = 2.67, 32.04, 32.04
    zp = rho * np.sqrt((lam + 2 * G) / rho)
    vx = 0.2
    f = np.zeros(o.array("f").shape)
    f[0] = vx
    f[2] = zp * vx  # sxx: outgoing through x-minus (n = -x): s_nn = -zp v_n with v_n = -vx
    o.set_array("f", f)
    o.set_array("df", np.zeros(o.array("df").shape))
    o.stage(0, 0)
    df = o.array("df")
    ny = o.nx[1]
    # x-minus face rows away from the y/z edges
    assert np.abs(df[:, 0, 2:ny - 2, :]).max() <= 1e-12
    # the x-plus face sees an incoming wave and must be penalised
    assert np.abs(df[:, -1, 2:ny - 2, :]).max() > 1e-6


def test_load_shapes_at_t0(tmp_path):
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 0., 0.)),
             Pert("gaussian", 0., 2.0, 0.5, 0., 0., (10., 0., 0.)),
             Pert("boxcar", 0., 1.0, 0.3, 0., 0., (5., 0., 0.)),
             Pert("ellipse", 0., 2.5, 0.4, 0., 0., (-3., 0., 0.)),
             Pert("linear", 0., 0.5, 2.0, 0., 0., (1., 0., 0.)),
             Pert("constant", 1.0, 0., 0., 0., 0., (1000., 0., 0.))]  # ramp t0 > 0: zero at t = 0
    p = write_case(str(tmp_path), "shape", nx=41, ny=13, dx=0.1, law="frictionless", loads=loads, front=None)
    o = O.Oracle(p)
    x = np.linspace(0.0, 4.0, 41)
    ref = -120.0 + 10.0 * np.exp(-((x - 2.0) / 0.5) ** 2)
    box = np.where(np.abs(np.abs(x - 1.0) - 0.3) < 1e-14, 0.5, np.where(np.abs(x - 1.0) < 0.3, 1.0, 0.0))
    ref += 5.0 * box
    ref += -3.0 * (((x - 2.5) / 0.4) ** 2 <= 1.0)
    ref += 1.0 * (x / 2.0 + 0.5)
    np.testing.assert_allclose(o.iface(0, "Sn"), ref, rtol=1e-13, atol=1e-12)
    assert np.all(o.iface(0, "S") == 0.0)  # frictionless: no shear strength


def test_friction_traction_equals_strength(tmp_path):
    p = write_case(str(tmp_path), "sw", nx=81, ny=41, dx=0.1, nt=30)
    o = O.Oracle(p)
    o.step(30)
    V, S, Sn, Vx, Sx = (o.iface(0, w) for w in ("V", "S", "Sn", "Vx", "Sx"))
    assert V.max() > 0.1
    np.testing.assert_allclose(np.abs(Sx), S, rtol=1e-12)
    np.testing.assert_allclose(np.abs(Vx), V, rtol=1e-12, atol=1e-15)
    slipping = V > 0
    assert np.all(S[slipping] >= 0.525 * np.abs(Sn[slipping]) * (1 - 1e-12))
    assert np.all(S <= 0.677 * np.abs(Sn) * (1 + 1e-12))


def test_front_brute_force(tmp_path):
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 4.0, 0.75, 0., 0., (0., 20., 0.))]
    p = write_case(str(tmp_path), "fr", nx=81, ny=41, dx=0.1, nt=80, front=("V", 0.05), loads=loads)
    o = O.Oracle(p)
    t_first = np.full(81, -1.0)
    V0 = o.iface(0, "V")
    t_first[V0 >= 0.05] = 0.0
    for i in range(80):
        o.step(1)
        V = o.iface(0, "V")
        new = (t_first < 0) & (V >= 0.05)
        t_first[new] = (i + 1) * o.dt
    np.testing.assert_array_equal(o.front(0), t_first)
    assert (t_first > 0).sum() > 3


def test_outputs_absolute_stress_and_shapes(tmp_path):
    outs = [Output("sxx", "sxx", (0, 4, 2), (0, 30, 3), (0, 31, 1), (0, 0, 1)),
            Output("V", "V", (0, 4, 1), (0, 30, 1), (16, 16, 1), (0, 0, 1))]
    s0 = (-100., 52., 0., -120., 0., 0.)
    p = write_case(str(tmp_path), "out", nx=31, ny=16, dx=0.1, nt=4, s0=s0, outputs=outs, loads=[])
    assert O.run(p) == 0
    d = os.path.join(str(tmp_path), "out", "data")
    meta = {}
    exec(open(os.path.join(d, "out_sxx.py")).read(), meta)
    a = np.fromfile(os.path.join(d, "out_sxx_sxx.dat")).reshape(meta["nt"], meta["nx"], meta["ny"], meta["nz"])
    assert (meta["nt"], meta["nx"], meta["ny"], meta["nz"]) == (3, 11, 32, 1)
    assert np.all(a[0] == -100.0)
    t = np.fromfile(os.path.join(d, "out_sxx_t.dat"))
    assert t.shape == (3,) and t[0] == 0.0
    meta = {}
    exec(open(os.path.join(d, "out_V.py")).read(), meta)
    v = np.fromfile(os.path.join(d, "out_V_V.dat"))
    assert v.size == meta["nt"] * meta["nx"] * meta["ny"] * meta["nz"] == 5 * 31


def test_3d_outputs_absolute_stress(tmp_path):
    outs = [Output(n, n, (0, 0, 1), (0, 20, 5), (0, 25, 1), (0, 14, 7)) for n in ("sxx", "sxy", "sxz", "syy", "syz", "szz")]
    s0 = (-100., 52., 3., -120., 4., -55.)
    p = write_case(str(tmp_path), "o3", ndim=3, nx=21, ny=13, nz=15, dx=0.1, nt=1, s0=s0, outputs=outs, loads=[],
                   law="locked", front=None)
    assert O.run(p) == 0
    d = os.path.join(str(tmp_path), "o3", "data")
    for val, n in zip(s0, ("sxx", "sxy", "sxz", "syy", "syz", "szz")):
        a = np.fromfile(os.path.join(d, "o3_%s_%s.dat" % (n, n)))
        assert a.size > 0 and np.all(a == val), n


def test_dt_closed_form(tmp_path):
    p = write_case(str(tmp_path), "dt", nx=41, ny=21, dx=0.04, cfl=0.3)
    o = O.Oracle(p)
    cs = np.sqrt(32.04 / 2.67)
    assert abs(o.dt - 0.3 * 0.04 / cs) <= 1e-15
