"""The differentiable, batched engine (include/esp_diff.h, calibration.py).

Forward: a dense bonded specimen run by the differentiable engine follows the
oracle's run of the same model (bond breaks equal, sampled stresses and
positions within 1e-9).  Gradient: d loss / d (k_n, k_s, mu) against central
finite differences of the oracle's loss (relative step 1e-6, only parameters
whose +- perturbations change no break event), d loss / d cohesion zero.
Checkpoint intervals change no gradient bit; a batch of specimens gives each
specimen's single-specimen loss and gradient; the torch.autograd.Function
returns the adjoint's gradient."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS, EVERY = 300, 20
THETA0 = np.array([1e5, 3e4, 50.0, 0.6])


@pytest.fixture(scope="module")
def cal(pkg):
    return importlib.import_module("esys_particle_amd.calibration")


@pytest.fixture(scope="module")
def spm(pkg):
    return importlib.import_module("esys_particle_amd.specimens")


@pytest.fixture(scope="module")
def spec2k(spm):
    return spm.dense_bonded_compression(n=2000, seed=3)


def model_of(cal, specs, K=100):
    return cal.DiffModel.from_specimens(specs, STEPS, EVERY, checkpoint_every=K)


def test_forward_follows_oracle(pkg, orc, cal, spec2k):
    m = model_of(cal, [spec2k])
    print("case: diff forward 2000 particles, %d steps" % STEPS, flush=True)
    st, loss = m.forward(THETA0, np.zeros((1, m.nsamples)))
    L, so, br = cal.plain_loss(orc.oracle_lsm_class(pkg), m, THETA0, np.zeros(m.nsamples))
    assert len(br) > 50  # the case breaks bonds
    np.testing.assert_array_equal(m.bond_breaks(), br)
    np.testing.assert_allclose(st[0], so, rtol=1e-9, atol=1e-9 * np.abs(so).max())
    assert abs(loss[0] - L) <= 1e-9 * L
    o = orc.oracle_lsm_class(pkg)()
    cal.plain_setup(o, m, THETA0)
    o.runTimeSteps(STEPS)
    assert np.max(np.abs(m.particle_data("pos") - o.getParticleData("pos"))) <= 1e-9


def test_gradient_matches_oracle_finite_differences(pkg, orc, cal, spec2k):
    m = model_of(cal, [spec2k])
    true = THETA0 * np.array([1.08, 0.93, 1.0, 0.85])
    target, _ = m.forward(true, np.zeros((1, m.nsamples)))
    _, loss = m.forward(THETA0, target)
    g = m.backward()[0]
    print("case: gradient 2000 particles, loss %.6g, grad %s" % (loss[0], g), flush=True)
    assert g[2] == 0.0
    cls = orc.oracle_lsm_class(pkg)
    L0, _, b0 = cal.plain_loss(cls, m, THETA0, target[0])
    checked = 0
    for k in (0, 1, 2, 3):
        h = 1e-6 * THETA0[k]
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[k] += h
        tm[k] -= h
        Lp, _, bp = cal.plain_loss(cls, m, tp, target[0])
        Lm, _, bm = cal.plain_loss(cls, m, tm, target[0])
        if not (np.array_equal(bp, b0) and np.array_equal(bm, b0)):
            print("parameter %d: a perturbation flips break events, excluded" % k)
            continue
        fd = (Lp - Lm) / (tp[k] - tm[k])
        print("parameter %d: adjoint %.9e  finite difference %.9e" % (k, g[k], fd))
        if k == 2:
            assert fd == 0.0  # cohesion moves break events only
        else:
            assert abs(g[k] - fd) <= 1e-6 * abs(fd) + 1e-12 * abs(L0) / THETA0[k]
        checked += 1
    assert checked >= 3


def test_checkpoint_interval_changes_nothing(cal, spec2k):
    out = []
    for K in (1, 37, 400):
        m = model_of(cal, [spec2k], K=K)
        target = np.full((1, m.nsamples), 5.0)
        _, loss = m.forward(THETA0, target)
        out.append((loss.tobytes(), m.backward().tobytes()))
    assert out[0] == out[1] == out[2]


def test_batch_equals_single_specimens(cal, spm):
    specs = [spm.dense_bonded_compression(n=n, seed=s, height=12.0)
             for n, s in ((500, 5), (900, 6), (1300, 7))]
    thetas = np.array([THETA0, THETA0 * [1.1, 0.9, 1.0, 1.2], THETA0 * [0.8, 1.1, 1.1, 0.7]])
    mb = model_of(cal, specs)
    target = np.full((3, mb.nsamples), 3.0)
    _, lb = mb.forward(thetas, target)
    gb = mb.backward()
    for s, sp in enumerate(specs):
        ms = model_of(cal, [sp])
        _, ls = ms.forward(thetas[s:s + 1], target[s:s + 1])
        gs = ms.backward()
        np.testing.assert_allclose(lb[s], ls[0], rtol=1e-12)
        np.testing.assert_allclose(gb[s], gs[0], rtol=1e-10, atol=1e-14)


def test_autograd_function(cal, spec2k):
    import torch
    m = model_of(cal, [spec2k])
    target = torch.full((1, m.nsamples), 4.0, dtype=torch.float64)
    theta = torch.tensor([THETA0], dtype=torch.float64, requires_grad=True)
    loss = cal.CalibrationLoss.apply(theta, m, target)
    loss.sum().backward()
    g = m.backward()
    np.testing.assert_array_equal(theta.grad.numpy(), g)
