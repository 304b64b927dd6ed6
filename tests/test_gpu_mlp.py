"""The learned STZ root proposal on the GPU.

* The training kernels' gradient and the inference kernel against the CPU
  oracle's network (FP32 GPU vs FP64 CPU).
* GPU training fits a smooth target.
* End to end: roots logged from a plain STZ run train the network on the GPU;
  on a different seed, the proposal plus Newton finish needs far fewer Newton
  iterations than the plain bracketed solve, falls back rarely, and gives the
  plain solver's results (and the oracle's) within 1e-10.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from cases import STZ_PARAMS, write_case
from fdfault_amd.problems import Pert

pytestmark = pytest.mark.gpu

GROUPS = {"W1": (0, 256), "b1": (256, 320), "W2": (320, 4416), "b2": (4416, 4480), "W3": (4480, 4544),
          "b3": (4544, 4545)}


def _rand_params(seed, scale=0.2):
    rng = np.random.default_rng(seed)
    P = np.zeros(4555, np.float32)
    P[:4545] = rng.normal(0., scale, 4545)
    return P


def test_gpu_gradient_matches_oracle():
    from fdfault_amd import lib
    rng = np.random.default_rng(1)
    P = _rand_params(2)
    X = rng.normal(size=(1000, 4)).astype(np.float32)
    T = rng.normal(size=1000).astype(np.float32)
    G, loss = lib.mlp_grad(P, X, T)
    Go, lo = O.mlp_grad(P.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - lo) <= 1e-5 * max(1.0, abs(lo))
    for name, (a, b) in GROUPS.items():
        r = np.abs(G[a:b] - Go[a:b]).max() / max(np.abs(Go[a:b]).max(), 1e-30)
        print(name, "rel", r, flush=True)
        assert r <= 2e-4, (name, r)


def test_gpu_predict_matches_oracle():
    from fdfault_amd import lib
    rng = np.random.default_rng(3)
    P = _rand_params(4, 0.3)
    X = rng.normal(size=(4000, 4)).astype(np.float32)
    Y = lib.mlp_predict(P, X)
    Yo = O.mlp_forward(P.astype(np.float64), X.astype(np.float64))
    assert np.abs(Y - Yo).max() <= 2e-5 * max(1.0, np.abs(Yo).max())


def test_gpu_predict_mfma_matches_oracle():
    """The fault kernel's evaluation (hidden layer on the matrix cores, FP16
    operands, FP32 accumulation) against the FP64 oracle: within the FP16
    rounding of weights and activations.  The sizes leave a partial last wave
    (lanes past n evaluate a dummy input)."""
    from fdfault_amd import lib
    rng = np.random.default_rng(6)
    P = _rand_params(8, 0.3)
    for n in (4000, 77, 1):
        X = rng.normal(size=(n, 4)).astype(np.float32)
        Y = lib.mlp_predict(P, X, mfma=True)
        Yo = O.mlp_forward(P.astype(np.float64), X.astype(np.float64))
        err = np.abs(Y - Yo).max()
        print("n", n, "max |mfma - oracle|", err, "max |y|", np.abs(Yo).max(), flush=True)
        assert err <= 5e-3 * max(1.0, np.abs(Yo).max())
        assert err > 0.0 or n == 1  # FP16 operands: not the FP32 path's numbers


def test_gpu_training_fits_smooth_target():
    from fdfault_amd import lib
    rng = np.random.default_rng(5)
    n = 20000
    x = rng.uniform(-1, 1, size=(n, 4))
    rows = np.empty((n, 5))
    rows[:, :4] = x * np.array([10., 20., 0.01, 1.]) + np.array([70., -120., 0.07, 4.6])
    rows[:, 4] = np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - x[:, 3] ** 2
    P, losses = lib.mlp_train(rows, epochs=30, batch=256, lr=3e-3, seed=7)
    print("losses", losses[0], losses[-1], flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.02 * losses[0]
    assert P.size == lib.mlp_param_count()


def test_training_independent_of_row_order():
    """The fault kernel logs solves in atomic order; training depends on the
    set of rows only (rows sorted before the statistics and the shuffle,
    fixed-order gradient reductions): bit-identical parameters."""
    from fdfault_amd import lib
    rng = np.random.default_rng(8)
    n = 5000
    x = rng.uniform(-1, 1, size=(n, 4))
    rows = np.empty((n, 5))
    rows[:, :4] = x * np.array([10., 20., 0.01, 1.]) + np.array([70., -120., 0.07, 4.6])
    rows[:, 4] = np.sin(2 * x[:, 0]) + x[:, 3]
    P1, l1 = lib.mlp_train(rows, epochs=3, batch=256, lr=3e-3, seed=7)
    P2, l2 = lib.mlp_train(rows[rng.permutation(n)], epochs=3, batch=256, lr=3e-3, seed=7)
    assert np.array_equal(P1, P2)


def _stz_case(tmp, name, seed, nx=161, ny=41, nt=40):
    d = os.path.join(str(tmp), name)
    os.makedirs(d, exist_ok=True)
    chi = np.random.default_rng(seed).uniform(0.06, 0.08, size=(nx, 1))
    sp = os.path.join(d, "state.bin")
    chi.tofile(sp)
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.))]
    return write_case(str(tmp), name, nx=nx, ny=ny, dx=0.1, nt=nt, law="stz",
                      stz=dict(STZ_PARAMS, chi0=0., statefile=sp), loads=loads)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_stz_proposal_gain_and_parity(tmp_path):
    from fdfault_amd import lib
    nt = 40
    train_case = _stz_case(tmp_path, "stz_a", seed=1, nt=nt)
    a = lib.Solver(train_case)
    a.stz_log(400000)
    a.step(nt)
    rows = a.stz_log_read()
    a.close()
    print("logged solves", rows.shape, flush=True)
    assert rows.shape[0] > 10000 and np.all(np.isfinite(rows))
    P, losses = lib.mlp_train(rows, epochs=40, batch=512, lr=3e-3, seed=0)
    print("train loss", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.05 * losses[0]

    case = _stz_case(tmp_path, "stz_b", seed=2, nt=nt)
    plain = lib.Solver(case)
    plain.step(nt)
    sp = plain.stz_stats()
    prop = lib.Solver(case)
    prop.stz_set_model(P)
    prop.step(nt)
    sm = prop.stz_stats()
    mean_plain, mean_prop = sp[1] / sp[0], sm[1] / sm[0]
    print("plain: solves %d mean iterations %.2f; proposal: mean %.2f, used %d, fallbacks %d"
          % (sp[0], mean_plain, mean_prop, sm[2], sm[3]), flush=True)
    assert sp[0] == sm[0] and sp[3] == 0 and sp[2] == 0
    assert sm[2] >= 0.95 * sm[0]
    assert mean_prop < 0.6 * mean_plain
    o = O.Oracle(case)
    o.step(nt)
    for w in ("U", "V", "S", "state"):
        assert _rel(prop.iface(0, w), plain.iface(0, w)) <= 1e-10, w
        assert _rel(prop.iface(0, w), o.iface(0, w)) <= 1e-10, w
    assert _rel(prop.array("f"), o.array("f")) <= 1e-10
    # the network off again: every solve is the plain one
    prop.stz_set_model(None)
    prop.step(1)
    s1 = prop.stz_stats()
    assert s1[0] > 0 and s1[2] == 0 and s1[3] == 0
    o.close()
    plain.close()
    prop.close()


def _stz3d_case(tmp, name, seed, nx=41, ny=16, nz=25, nt=20):
    d = os.path.join(str(tmp), name)
    os.makedirs(d, exist_ok=True)
    chi = np.random.default_rng(seed).uniform(0.06, 0.08, size=(nx, nz))
    sp = os.path.join(d, "state.bin")
    chi.tofile(sp)
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.))]
    return write_case(str(tmp), name, ndim=3, nx=nx, ny=ny, nz=nz, dx=0.1, nt=nt, law="stz",
                      stz=dict(STZ_PARAMS, chi0=0., statefile=sp), loads=loads)


@pytest.mark.parametrize("ndim,precision", [(3, 0), (2, 1)])
def test_stz_proposal_other_dimension_and_precision(tmp_path, ndim, precision):
    """The proposal in the two fault-kernel forms the other tests leave out:
    3-D with FP64 bulk and 2-D with FP32 bulk.  Trained on plain solves of
    one seed, used on another: fewer Newton iterations, (almost) no
    fallbacks, and the plain solver's results (FP64 bulk: within 1e-10;
    FP32 bulk: within 1e-5, the FP32 fields' rounding), every solve to the
    plain solver's tolerance."""
    from fdfault_amd import lib
    nt = 20
    mk = (lambda n, s: _stz3d_case(tmp_path, n, s, nt=nt)) if ndim == 3 else \
         (lambda n, s: _stz_case(tmp_path, n, seed=s, nx=121, ny=31, nt=nt))
    a = lib.Solver(mk("train", 1), precision=precision)
    a.stz_log(400000)
    a.step(nt)
    rows = a.stz_log_read()
    a.close()
    print("ndim", ndim, "precision", precision, "logged", rows.shape, flush=True)
    P, losses = lib.mlp_train(rows, epochs=40, batch=512, lr=3e-3, seed=0)
    case = mk("use", 2)
    plain = lib.Solver(case, precision=precision)
    plain.step(nt)
    sp = plain.stz_stats()
    prop = lib.Solver(case, precision=precision)
    prop.stz_set_model(P)
    prop.step(nt)
    sm = prop.stz_stats()
    mp, mm = sp[1] / sp[0], sm[1] / sm[0]
    print("plain %.2f iterations per solve; proposal %.2f, used %d of %d, fallbacks %d"
          % (mp, mm, sm[2], sm[0], sm[3]), flush=True)
    assert sp[0] == sm[0] and sp[2] == 0 and sp[3] == 0
    assert sm[2] >= 0.95 * sm[0] and sm[3] <= 0.05 * sm[0]
    assert mm < 0.7 * mp
    tol = 1e-10 if precision == 0 else 1e-5
    for w in ("U", "V", "S", "state"):
        r = _rel(prop.iface(0, w), plain.iface(0, w))
        print(w, "proposal vs plain", r, flush=True)
        assert r <= tol, w
    plain.close()
    prop.close()
