"""GPU checks of the learned Barnes-Hut correction, through the C-ABI.

* forward and gradients of the BF16-MFMA network against the FP64 oracle,
  both with the oracle's products rounded to bfloat16 like the GPU's (tight)
  and in plain FP64 (the BF16 tolerance);
* walk features against the oracle's walk;
* training on one particle set, evaluation on another: the gain over the
  coarse walk without the network (ablation);
* the spot check: it reports the true error, and a failing check returns
  exactly the plain theta = 0.5 walk;
* an untrained network fails loudly.
"""
import numpy as np
import pytest

from grav_sim_amd import ic
from grav_sim_amd import lib as L
from oracle import mlp as NN
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(autouse=True)
def _default_policy():
    L.learned_set_policy(1e-3, 1024, 0.5)
    yield
    L.learned_set_policy(1e-3, 1024, 0.5)


@pytest.mark.parametrize("rows", [1, 17, 1000])
def test_forward_matches_oracle(rows):
    w = NN.random_params(11)
    L.learned_set_params(w.astype(np.float32))
    X = np.random.default_rng(rows).normal(size=(rows, 16)).astype(np.float32)
    Y = L.learned_forward(X)
    Ye, _ = NN.forward(w, X, emulate=True)
    Y64, _ = NN.forward(w, X)
    print(f"rows {rows}: vs bf16-emulating oracle {_rel(Y, Ye):.2e}, vs FP64 {_rel(Y, Y64):.2e}")
    assert _rel(Y, Ye) < 1e-3
    assert _rel(Y, Y64) < 3e-2


@pytest.mark.parametrize("rows", [37, 512, 4096, 5000])
def test_gradients_match_oracle(rows):
    w = NN.random_params(12)
    L.learned_set_params(w.astype(np.float32))
    rng = np.random.default_rng(100 + rows)
    X = rng.normal(size=(rows, 16)).astype(np.float32)
    T = rng.normal(size=(rows, 3)).astype(np.float32)
    g, loss = L.learned_gradients(X, T)
    ge, le = NN.gradients(w, X, T, emulate=True)
    g64, l64 = NN.gradients(w, X, T)
    assert abs(loss - le) <= 1e-3 * le and abs(loss - l64) <= 3e-2 * l64
    k = 0
    for name, shp in NN.SHAPES:
        sz = int(np.prod(shp))
        e1, e2 = _rel(g[k:k + sz], ge[k:k + sz]), _rel(g[k:k + sz], g64[k:k + sz])
        print(f"rows {rows} {name}: vs emulating oracle {e1:.2e}, vs FP64 {e2:.2e}")
        assert e1 < 5e-3, name
        assert e2 < 5e-2, name
        k += sz


def test_walk_features_match_oracle():
    x, v, m = ic.two_spheres(8192, 41)
    eps, theta, G = 1e-3, 0.8, 1.0
    a, F = L.learned_features(x, m, G=G, eps=eps, theta=theta, leaf=1)
    t = O.build_tree(x, m, 1)
    q = O.octree_quad(x, m, t)
    o = O.octree_oct(x, m, t, q)
    a_o, cnt, wf = O.bh_walk(x, m, t, theta, G=G, eps=eps, want_counts=True, quad=q, features=True, oct=o)
    Fo = NN.features16(a_o, wf, G)
    assert _rel(a, a_o) < 1e-6
    assert np.allclose(F[:, 8], Fo[:, 8], rtol=1e-6) and np.allclose(F[:, 9], Fo[:, 9], rtol=1e-6)
    for k in range(16):
        scale = np.sqrt(np.mean(Fo[:, k] ** 2))
        if scale == 0.0:
            assert np.all(F[:, k] == 0.0), k
            continue
        err = np.max(np.abs(F[:, k] - Fo[:, k])) / scale
        assert err < 1e-3, (k, err)


def test_learned_acceleration_matches_oracle_pipeline():
    """Method 18 end to end against the oracle: oracle walk and features,
    the network in FP64 with bfloat16-rounded products, the correction in
    the local frame."""
    x, v, m = ic.two_spheres(8192, 45)
    eps, theta, G, ts = 1e-3, 0.8, 1.0, 0.5
    t = O.build_tree(x, m, 1)
    q = O.octree_quad(x, m, t)
    o = O.octree_oct(x, m, t, q)
    a_o, wf = O.bh_walk(x, m, t, theta, G=G, eps=eps, quad=q, features=True, oct=o)
    Fo = NN.features16(a_o, wf, G)
    mu = Fo.mean(axis=0).astype(np.float32)
    sd = Fo.std(axis=0).astype(np.float32)
    sd[sd == 0] = 1.0
    w = (NN.random_params(14) * 0.5).astype(np.float32)
    L.learned_set_params(w, mu, sd, ts)
    L.learned_set_policy(1.0, 0, 0.5)
    a_g = L.acceleration(x, m, 18, theta=theta, eps=eps, leaf=1)
    Xn = (Fo.astype(np.float32) - mu) / sd
    Y, _ = NN.forward(w.astype(np.float64), Xn, emulate=True)
    a_ref = NN.apply_correction(a_o, wf, Y, G, ts)
    corr = np.linalg.norm(a_ref - a_o - G * wf[:, 0:3], axis=1) / np.linalg.norm(a_o, axis=1)
    e = np.linalg.norm(a_g - a_ref, axis=1) / np.linalg.norm(a_ref, axis=1)
    print(f"network correction median {np.median(corr):.2e}; GPU vs oracle median {np.median(e):.2e} "
          f"p99 {np.quantile(e, 0.99):.2e} max {e.max():.2e}")
    assert np.median(corr) > 1e-3  # the network's part is not negligible
    assert np.median(e) < 2e-5 and np.quantile(e, 0.99) < 2e-4 and e.max() < 2e-3


def test_untrained_network_fails_loudly():
    L.learned_reset()
    x, v, m = ic.two_spheres(4096, 42)
    with pytest.raises(RuntimeError, match="no trained network"):
        L.acceleration(x, m, 18, theta=0.8, leaf=1)


def test_forced_fallback_is_the_group_quadrupole_walk_checked_again():
    """A spot check that fails replaces the result by the group walk with
    quadrupoles at the fallback angle and checks that again: the status
    then reports the fallback's own p99, and 2 when it misses the tolerance
    too (here a tolerance nothing meets).  The fallback walks the tree the
    learned method built (FP32 second moments rounded from its FP64 ones);
    method 21 forms its FP32 moments directly, so the two agree to FP32
    rounding of the quadrupole terms."""
    x, v, m = ic.two_spheres(1 << 15, 43)
    L.learned_set_params(NN.random_params(13).astype(np.float32) * 1e-3)
    L.learned_set_policy(1e-12, 256, 0.5)
    a = L.acceleration(x, m, 18, theta=0.8, leaf=1)
    p99, fell_back = L.learned_status()
    assert fell_back == 2 and p99 > 1e-12
    a_fb = L.acceleration(x, m, 21, theta=0.5, leaf=1)
    assert L.last_walk_interactions() > 0
    e = np.linalg.norm(a - a_fb, axis=1) / np.linalg.norm(a_fb, axis=1)
    assert np.median(e) <= 1e-7 and e.max() <= 1e-5, (np.median(e), e.max())
    # the reported p99 is the fallback's: the spot check of the group walk itself
    L.learned_set_policy(1.0, 256, 0.5)
    sim = L.Sim(x, v, m, 1.0)
    med, p99_fb, mx = sim.accel_error(L.param(21, theta=0.5, leaf=1), samples=256)
    sim.close()
    assert abs(p99 - p99_fb) <= 1e-3 * p99_fb + 1e-6, (p99, p99_fb)
    # a tolerance the fallback meets: status 1
    L.learned_set_policy(1.5 * p99, 256, 0.5)
    L.learned_set_params(NN.random_params(13).astype(np.float32) * 1e3)  # a network far off: fails the check
    L.acceleration(x, m, 18, theta=0.8, leaf=1)
    p99b, fell_back = L.learned_status()
    assert fell_back == 1 and p99b <= 1.5 * p99


@pytest.fixture(scope="module")
def trained():
    x, v, m = ic.two_spheres(1 << 17, 51)
    loss = L.learned_train(x, m, eps=1e-3, theta=0.8, leaf=1, samples=1 << 16, epochs=20, batch=1024,
                           lr=2e-3, seed=5)
    rep = L.learned_training_report()
    print(f"training loss {loss:.4f} (3 = predicting zero); held out: {rep}")
    return loss, rep


def test_training_reduces_loss(trained):
    loss, rep = trained
    assert loss < 3.0
    assert 0.0 <= rep["alpha"] <= 1.0 and rep["p99_net"] <= rep["p99_without"]


def test_learned_gain_over_coarse_walk(trained):
    """Ablation on a particle set neither training nor calibration has seen.
    The output scale is chosen on a second, independent realisation (seed
    54); on a third (seed 52) the learned correction -- the exact quadrupole
    and octupole sums of the accepted nodes plus the network's prediction of
    what remains -- must beat its quadrupole sum alone (method 19) by at
    least 15 % in p99 and 20 % in the median.  Against the octupole sums
    without the network (the oracle's octupole walk) it measures 3 % better
    in the median and 5 % worse in p99 (the network's share is chosen on
    another realisation): bounded here at 2 % and 10 %."""
    xc, vc, mc = ic.two_spheres(1 << 17, 54)
    L.learned_calibrate(xc, mc, eps=1e-3, theta=0.8, leaf=1, samples=1 << 16, seed=9)
    rep = L.learned_training_report()
    x, v, m = ic.two_spheres(1 << 17, 52)
    L.learned_set_policy(1.0, 0, 0.5)  # no spot check: the network's own result
    sim = L.Sim(x, v, m, 1.0)
    coarse = sim.accel_error(L.param(3, theta=0.8, eps=1e-3, leaf=1), samples=4096)
    learned = sim.accel_error(L.param(18, theta=0.8, eps=1e-3, leaf=1), samples=4096)
    quad = sim.accel_error(L.param(19, theta=0.8, eps=1e-3, leaf=1), samples=4096)
    # the octupole walk of the oracle on the same strided sample, against the
    # GPU FP64 direct sum the sampled errors above use
    t = O.build_tree(x, m, 1)
    q = O.octree_quad(x, m, t)
    o = O.octree_oct(x, m, t, q)
    idx = (np.arange(4096) * len(m)) // 4096
    a_oct = O.bh_walk(x, m, t, 0.8, eps=1e-3, quad=q, oct=o)[idx]
    a_ref = L.acceleration(x, m, 1, eps=1e-3)[idx]
    e_oct = np.linalg.norm(a_oct - a_ref, axis=1) / np.linalg.norm(a_ref, axis=1)
    octupole = (np.median(e_oct), np.percentile(e_oct, 99))
    sim.close()
    print(f"calibration {rep}; theta=0.8 median/p99: coarse {coarse[0]:.2e}/{coarse[1]:.2e}, "
          f"learned {learned[0]:.2e}/{learned[1]:.2e}, quadrupole sum only {quad[0]:.2e}/{quad[1]:.2e}, "
          f"octupole sum only (oracle) {octupole[0]:.2e}/{octupole[1]:.2e}")
    assert learned[0] < 0.6 * coarse[0] and learned[1] < 0.6 * coarse[1]
    assert learned[0] <= 0.8 * quad[0] and learned[1] <= 0.85 * quad[1]
    assert learned[0] <= 1.02 * octupole[0] and learned[1] <= 1.10 * octupole[1]
    assert rep["p99_net"] <= rep["p99_without"]


def test_spot_check_reports_true_error(trained):
    x, v, m = ic.two_spheres(1 << 17, 53)
    L.learned_set_policy(1.0, 1024, 0.5)  # checked, never falls back
    sim = L.Sim(x, v, m, 1.0)
    med, p99, mx = sim.accel_error(L.param(18, theta=0.8, eps=1e-3, leaf=1), samples=1024)
    p99_spot, fell_back = L.learned_status()
    sim.close()
    assert not fell_back
    # same strided sample; FP32 split-precision labels vs FP64 labels
    assert abs(p99_spot - p99) <= 1e-3 * p99 + 1e-6, (p99_spot, p99)


def test_zero_network_is_the_octupole_walk():
    """With all parameters zero the network adds nothing: the result is the
    coarse walk plus its quadrupole and octupole sums, the oracle's walk with
    both terms (up to the FP32 accumulation of those sums)."""
    x, v, m = ic.two_spheres(1 << 15, 44)
    L.learned_set_params(np.zeros(L.learned_num_params(), np.float32))
    L.learned_set_policy(1.0, 0, 0.5)
    a = L.acceleration(x, m, 18, theta=0.8, eps=1e-3, leaf=1)
    t = O.build_tree(x, m, 1)
    q = O.octree_quad(x, m, t)
    o = O.octree_oct(x, m, t, q)
    a_o = O.bh_walk(x, m, t, 0.8, eps=1e-3, quad=q, oct=o)
    e = np.linalg.norm(a - a_o, axis=1) / np.linalg.norm(a_o, axis=1)
    assert np.median(e) < 1e-6 and e.max() < 1e-4, (np.median(e), e.max())
