"""The differentiable, batched engine (include/esp_diff.h, calibration.py).

Forward: a dense bonded specimen run by the differentiable engine follows the
oracle's run of the same model (bond breaks equal, sampled stresses and
positions within 1e-9).

Gradient, two checks:
- against the tangent-linear run (forward mode, esp_diff_forward_tangent): the
  adjoint's d loss / d theta along each parameter and a random direction
  within 1e-10 (relative), on a bonded specimen and on a slipping granular one;
- against central finite differences of the oracle's loss at relative steps
  1e-5 and 1e-6, only where neither perturbation flips a bond break, a contact,
  a slip or a wall contact (branch digests of the same perturbed runs).  Such
  differences resolve the derivative only to their own noise: a 1e-6 change of
  k_n moves a particle less than one ulp of its coordinates in the first
  steps, and the spread between the two steps measures that noise, so the
  adjoint must lie within twice it of the 1e-5 difference (plus 1e-6 relative).
d loss / d cohesion is zero (cohesion moves break events only).  Checkpoint
intervals change no gradient bit; a batch of specimens gives each specimen's
loss and gradient bitwise as a model holding that specimen alone at the same
position; the torch.autograd.Function returns the adjoint's gradient."""
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


def model_of(cal, specs, K=100, x_start=0.0):
    return cal.DiffModel.from_specimens(specs, STEPS, EVERY, checkpoint_every=K, x_start=x_start)


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


def _perturbed(m, cls, theta, k, rel, target, ref_digest, ref_breaks, cal):
    """Oracle loss at theta +- rel theta_k, and whether either perturbation
    flips a branch (breaks, contacts, slip, wall contacts) in the engine's run
    of the same parameters."""
    h = rel * theta[k]
    tp, tm = theta.copy(), theta.copy()
    tp[k] += h
    tm[k] -= h
    flip = False
    for t in (tp, tm):
        m.forward(t, target[None])
        flip |= not (np.array_equal(m.branch_digest(), ref_digest)
                     and np.array_equal(m.bond_breaks(), ref_breaks))
    Lp, _, _ = cal.plain_loss(cls, m, tp, target)
    Lm, _, _ = cal.plain_loss(cls, m, tm, target)
    return (Lp - Lm) / (tp[k] - tm[k]), flip


def test_gradient_matches_oracle_finite_differences(pkg, orc, cal, spec2k):
    m = model_of(cal, [spec2k])
    m.track_branches()
    true = THETA0 * np.array([1.08, 0.93, 1.0, 0.85])
    target, _ = m.forward(true, np.zeros((1, m.nsamples)))
    _, loss = m.forward(THETA0, target)
    dig, brk = m.branch_digest(), m.bond_breaks()
    g = m.backward()[0]
    print("case: gradient 2000 particles, loss %.6g, grad %s" % (loss[0], g), flush=True)
    assert g[2] == 0.0 and g[0] != 0.0
    cls = orc.oracle_lsm_class(pkg)
    checked = 0
    for k in (0, 1, 2, 3):
        fd5, f5 = _perturbed(m, cls, THETA0, k, 1e-5, target[0], dig, brk, cal)
        fd6, f6 = _perturbed(m, cls, THETA0, k, 1e-6, target[0], dig, brk, cal)
        if f5 or f6:
            print("parameter %d: a perturbation flips a branch, excluded" % k)
            continue
        print("parameter %d: adjoint %.9e  finite differences %.9e (1e-5) %.9e (1e-6)"
              % (k, g[k], fd5, fd6))
        if k == 2:
            assert fd5 == 0.0 and fd6 == 0.0  # cohesion moves break events only
        else:
            assert abs(g[k] - fd5) <= 2.0 * abs(fd6 - fd5) + 1e-6 * abs(fd5)
        checked += 1
    assert checked >= 2


def _tangent_check(m, theta, target, tol=1e-10):
    g = m.backward()
    rng = np.random.default_rng(1)
    dirs = [np.eye(4)[k] for k in range(4)] + [rng.normal(size=4) * theta[0]]
    for dth in dirs:
        d = np.tile(dth, (m.nspec, 1))
        _, _, dl = m.forward_tangent(theta, d, target)
        adj = (g * d).sum(1)
        print("direction %s: adjoint %s tangent %s" % (dth, adj, dl), flush=True)
        np.testing.assert_allclose(adj, dl, rtol=tol, atol=1e-300)
    return g


def test_adjoint_matches_tangent_linear_bonded(cal, spec2k):
    m = model_of(cal, [spec2k], K=37)
    true = THETA0 * np.array([1.08, 0.93, 1.0, 0.85])
    target, _ = m.forward(true, np.zeros((1, m.nsamples)))
    m.forward(THETA0, target)
    g = _tangent_check(m, THETA0[None], target)
    assert g[0, 0] != 0.0 and g[0, 1] != 0.0


def test_adjoint_matches_tangent_linear_slipping(cal, spm):
    """Weak bonds and fast walls: most bonds break, contacts slip, so the
    friction coefficient's gradient is exercised."""
    spec = spm.dense_bonded_compression(n=1500, seed=9, coh_mean=2.0, coh_sd=0.2, wall_speed=20.0)
    theta = np.array([1e5, 3e4, 2.0, 0.3])
    m = model_of(cal, [spec], K=50)
    target, _ = m.forward(theta * [1.05, 0.9, 1.0, 1.2], np.zeros((1, m.nsamples)))
    m.forward(theta, target)
    assert len(m.bond_breaks()) > 500
    g = _tangent_check(m, theta[None], target)
    assert g[0, 3] != 0.0


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
        # alone at the same x position (coordinates round alike), bitwise equal
        ms = model_of(cal, [sp], x_start=mb.starts[s])
        _, ls = ms.forward(thetas[s:s + 1], target[s:s + 1])
        gs = ms.backward()
        assert lb[s] == ls[0]
        np.testing.assert_array_equal(gb[s], gs[0])


def test_autograd_function(cal, spec2k):
    import torch
    m = model_of(cal, [spec2k])
    target = torch.full((1, m.nsamples), 4.0, dtype=torch.float64)
    theta = torch.tensor(np.array([THETA0]), dtype=torch.float64, requires_grad=True)
    loss = cal.CalibrationLoss.apply(theta, m, target)
    loss.sum().backward()
    g = m.backward()
    np.testing.assert_array_equal(theta.grad.numpy(), g)


def test_queries_after_adjoint_and_digest_opt_in(cal, spm):
    """The adjoint sweep recomputes earlier states in the model's arrays; the
    queries after it still see the forward run's final state and break steps.
    Branch digests exist only when tracked, and tracking changes no result."""
    spec = spm.dense_bonded_compression(n=1500, seed=9, coh_mean=2.0, coh_sd=0.2, wall_speed=20.0)
    theta = np.array([[1e5, 3e4, 2.0, 0.3]])
    m = model_of(cal, [spec], K=50)
    target = np.zeros((1, m.nsamples))
    st, loss = m.forward(theta, target)
    brk, pos, q = m.bond_breaks(), m.particle_data("pos"), m.particle_data("quat")
    print("case: queries after adjoint, %d breaks" % len(brk), flush=True)
    assert len(brk) > 500
    with pytest.raises(RuntimeError):
        m.branch_digest()
    g = m.backward()
    np.testing.assert_array_equal(m.bond_breaks(), brk)
    np.testing.assert_array_equal(m.particle_data("pos"), pos)
    np.testing.assert_array_equal(m.particle_data("quat"), q)
    m.track_branches()
    st2, loss2 = m.forward(theta, target)
    dig = m.branch_digest()
    assert dig.shape == (STEPS,) and np.any(dig != 0)
    np.testing.assert_array_equal(st2, st)
    np.testing.assert_array_equal(loss2, loss)
    np.testing.assert_array_equal(m.backward(), g)
