"""Pins of the oracle's group walk (orc_bh_walk_group), CPU only.

The group walk has no counterpart in the reference; it is pinned by the
cases that reduce to what is pinned elsewhere:
* groups of one particle are the per-particle walk (src/acceleration_barnes_hut.c
  acceptance rule), force and interaction count alike;
* theta = 0 opens every node: the direct sum, n - 1 interactions each;
* a group accepts a node only where every particle of the group would accept
  it, so no particle has fewer interactions than in the per-particle walk --
  also for the particles of an outlier cloud, whose wide runs are split into
  groups of one particle.

Its local expansions (orc_bh_walk_group_le, orc_far_add / orc_far_eval) are
pinned by the Taylor property itself -- the second-order expansion of a sum
of softened point-mass fields differs from the exact field by O(|d|^3), so a
wrong term of order 0, 1 or 2 would leave an error of order |d|^0..2 --, by
the limit phi -> 0 (no node expanded: the plain group walk bit for bit) and
by the accuracy against the FP64 direct sum, which an expansion error would
raise.
"""
import numpy as np
import pytest

from grav_sim_amd import ic
from oracle import oracle as O


def _cloud(n, seed):
    x, _, m = ic.two_spheres(n, seed)
    rng = np.random.default_rng(seed)
    x[:40] = rng.uniform(-60.0, 60.0, size=(40, 3))  # a sparse outlier halo
    return x, m


@pytest.mark.parametrize("theta", [0.5, 0.9])
def test_groups_of_one_are_the_per_particle_walk(theta):
    x, m = _cloud(6000, 1)
    t = O.build_tree(x, m, 1)
    a1, c1 = O.bh_walk(x, m, t, theta, eps=1e-3, want_counts=True)
    ag, cg = O.bh_walk_group(x, m, t, theta, 1, eps=1e-3)
    assert np.array_equal(cg, c1)
    assert np.abs(ag - a1).max() <= 1e-12 * np.abs(a1).max()


@pytest.mark.parametrize("group", [8, 64])
def test_theta_zero_is_the_direct_sum(group):
    x, m = _cloud(3000, 2)
    t = O.build_tree(x, m, 1)
    ag, cg = O.bh_walk_group(x, m, t, 0.0, group, eps=1e-3)
    ad = O.pairwise(x, m, eps=1e-3)
    assert np.all(cg == len(m) - 1)
    assert np.abs(ag - ad).max() <= 1e-11 * np.abs(ad).max()


@pytest.mark.parametrize("theta", [0.4, 0.7, 1.0])
def test_group_interactions_never_below_the_per_particle_walk(theta):
    x, m = _cloud(20000, 3)
    t = O.build_tree(x, m, 1)
    _, c1 = O.bh_walk(x, m, t, theta, eps=1e-3, want_counts=True)
    ag, cg = O.bh_walk_group(x, m, t, theta, 64, eps=1e-3)
    assert np.all(cg >= c1)
    ad = O.pairwise(x, m, eps=1e-3)
    err = np.linalg.norm(ag - ad, axis=1) / np.linalg.norm(ad, axis=1)
    a1 = O.bh_walk(x, m, t, theta, eps=1e-3)
    err1 = np.linalg.norm(a1 - ad, axis=1) / np.linalg.norm(ad, axis=1)
    assert np.percentile(err, 99) <= np.percentile(err1, 99)


@pytest.mark.parametrize("eps", [0.0, 0.3])
def test_far_expansion_error_is_third_order(eps):
    """Sum of 7 softened sources expanded about the origin: halving the
    offset divides the error by 8 (7.5-8.5 at these offsets), in every
    direction; the value at d = 0 is the exact field."""
    rng = np.random.default_rng(3)
    y = rng.normal(size=(7, 3))
    y = 4.0 * y / np.linalg.norm(y, axis=1)[:, None]  # sources 4 away: c - com = y
    M = rng.uniform(0.5, 2.0, size=7)
    F = O.far_moments(y, M, eps)

    def exact(d):
        r = y + d  # evaluation point minus source position
        s2 = np.einsum("ij,ij->i", r, r) + eps * eps
        return -(M / s2 ** 1.5) @ r

    assert np.allclose(O.far_eval(F, np.zeros(3)), exact(np.zeros(3)), rtol=1e-14, atol=0)
    for _ in range(6):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        e = [np.linalg.norm(O.far_eval(F, h * u) - exact(h * u)) for h in (0.2, 0.1, 0.05)]
        assert 7.5 < e[0] / e[1] < 8.5 and 7.5 < e[1] / e[2] < 8.5, e


def test_expansion_ratio_to_zero_is_the_group_walk():
    """phi -> 0 expands nothing in groups of nonzero width (the plain group
    walk bit for bit); a group of one particle has a box of zero width, so
    every node it accepts is expanded, exactly (d = 0: the field itself)."""
    x, m = _cloud(6000, 4)
    t = O.build_tree(x, m, 1)
    a0, c0 = O.bh_walk_group(x, m, t, 0.6, 64, eps=1e-3)
    a1, c1, far = O.bh_walk_group(x, m, t, 0.6, 64, eps=1e-3, phi=1e-12)
    assert np.array_equal(c1, c0)
    # only the particles of singleton groups (split outlier runs) differ
    differ = np.any(a1 != a0, axis=1)
    assert 0 < differ.sum() <= 60 and far > 0
    assert np.abs(a1 - a0).max() <= 1e-13 * np.abs(a0).max()


def test_expansions_keep_the_group_walk_accuracy():
    """Two spheres, N = 2^14, theta = 0.5: with phi = 0.1 a good share of
    the accepted nodes goes through the expansions, the interaction counts
    are the group walk's, and the errors against the FP64 direct sum stay
    within 3 % of it."""
    x, _, m = ic.two_spheres(1 << 14, 5)
    t = O.build_tree(x, m, 1)
    a_d = O.pairwise(x, m, eps=1e-3)
    a0, c0 = O.bh_walk_group(x, m, t, 0.5, 64, eps=1e-3)
    a1, c1, far = O.bh_walk_group(x, m, t, 0.5, 64, eps=1e-3, phi=0.1)
    assert np.array_equal(c1, c0)
    # (group, node) pairs against interactions / 64 (~ pairs of full groups)
    assert far > 0.1 * c0.sum() / 64.0
    e0 = np.linalg.norm(a0 - a_d, axis=1) / np.linalg.norm(a_d, axis=1)
    e1 = np.linalg.norm(a1 - a_d, axis=1) / np.linalg.norm(a_d, axis=1)
    assert np.median(e1) <= 1.03 * np.median(e0) and np.quantile(e1, 0.99) <= 1.03 * np.quantile(e0, 0.99)


def test_far_opening_angle_with_quadrupoles_keeps_the_accuracy():
    """Far nodes accepted at 1.25 theta, with their quadrupole values at the
    group centre: fewer interactions than the group walk at theta, and its
    median and p99 errors against the FP64 direct sum are kept (within 1 %)."""
    x, _, m = ic.two_spheres(1 << 14, 5)
    t = O.build_tree(x, m, 1)
    q = O.octree_quad(x, m, t)
    a_d = O.pairwise(x, m, eps=1e-3)
    for theta in (0.4, 0.5):
        a0, c0 = O.bh_walk_group(x, m, t, theta, 64, eps=1e-3)
        a1, c1, far = O.bh_walk_group(x, m, t, theta, 64, eps=1e-3, phi=0.1, theta_far=1.25 * theta, quad_far=q)
        assert c1.sum() < c0.sum() and np.all(c1 <= c0) and far > 0
        e0 = np.linalg.norm(a0 - a_d, axis=1) / np.linalg.norm(a_d, axis=1)
        e1 = np.linalg.norm(a1 - a_d, axis=1) / np.linalg.norm(a_d, axis=1)
        assert np.median(e1) <= 1.01 * np.median(e0) and np.quantile(e1, 0.99) <= 1.01 * np.quantile(e0, 0.99)
