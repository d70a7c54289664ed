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
