"""Pins of the FP64 oracle of the learned correction (CPU only).

The network and the walk features have no counterpart in the reference, so
the oracle is pinned by what the mathematics fixes: finite-difference
gradients, the GELU closed form, torch's bfloat16 rounding, identities that
tie the walk features to the quadrupole walk and to the interaction counts,
and a configuration whose walk accepts one known node.
"""
import numpy as np
import pytest
import torch

from grav_sim_amd import ic
from oracle import mlp as NN
from oracle import oracle as O


def _batch(seed, rows=64):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(rows, NN.F))
    T = rng.normal(size=(rows, NN.O))
    return X, T


def test_param_count_matches_layout():
    assert NN.NPARAM == 16 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 3 + 3 == 35587
    w = NN.random_params(0)
    assert np.array_equal(NN.pack(NN.unpack(w)), w)


def test_gelu_closed_form_and_derivative():
    # tanh-form GELU at known points: 0 -> 0, odd part, large |z| limits
    assert NN.gelu(np.array(0.0)) == 0.0
    z1 = 0.5 * (1.0 + np.tanh(np.sqrt(2.0 / np.pi) * (1.0 + 0.044715)))
    assert abs(NN.gelu(np.array(1.0)) - z1) < 1e-15
    assert abs(NN.gelu(np.array(1.0)) - 0.8411919906082768) < 1e-12
    z = np.linspace(-6, 6, 241)
    assert np.allclose(NN.gelu(z) - NN.gelu(-z), z, atol=1e-12)  # gelu(z) - gelu(-z) = z
    assert abs(NN.gelu(np.array(40.0)) - 40.0) < 1e-12 and abs(NN.gelu(np.array(-40.0))) < 1e-12
    h = 1e-6
    fd = (NN.gelu(z + h) - NN.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - NN.gelu_grad(z))) < 1e-8


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.normal(size=20000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, -3.0e38, 1e-40], np.float32)])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(NN.bf16(a), ref)


@pytest.mark.parametrize("emulate", [False])
def test_gradients_vs_finite_differences(emulate):
    w = NN.random_params(1)
    X, T = _batch(2, rows=32)
    g, loss = NN.gradients(w, X, T, emulate)
    assert abs(loss - NN.loss(w, X, T)) < 1e-12 * max(1.0, loss)
    rng = np.random.default_rng(4)
    k0 = 0
    for name, shp in NN.SHAPES:
        sz = int(np.prod(shp))
        for k in k0 + rng.choice(sz, size=min(sz, 12), replace=False):
            h = 1e-5
            wp, wm = w.copy(), w.copy()
            wp[k] += h
            wm[k] -= h
            fd = (NN.loss(wp, X, T) - NN.loss(wm, X, T)) / (2 * h)
            assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(fd), (name, k, fd, g[k])
        k0 += sz


def test_emulated_products_round_operands_only():
    # with operands already representable in bf16, emulation changes nothing
    w = NN.bf16(NN.random_params(5))
    X, _ = _batch(6, rows=16)
    X = NN.bf16(X)
    p = NN.unpack(w)
    Z1 = X @ p["W1"] + p["b1"]
    Y0, c0 = NN.forward(w, X, emulate=False)
    Y1, c1 = NN.forward(w, X, emulate=True)
    assert np.array_equal(c0["Z1"], Z1) and np.array_equal(c1["Z1"], Z1)
    assert np.linalg.norm(Y1 - Y0) / np.linalg.norm(Y0) < 2e-2


def test_walk_feature_identities():
    x, v, m = ic.two_spheres(3000, 31)
    t = O.build_tree(x, m, 1)
    q = O.octree_quad(x, m, t)
    o = O.octree_oct(x, m, t, q)
    for theta, eps in ((0.8, 0.0), (0.5, 1e-2)):
        G = 2.5
        a_m, cnt = O.bh_walk(x, m, t, theta, G=G, eps=eps, want_counts=True)
        a_f, cnt_f, wf = O.bh_walk(x, m, t, theta, G=G, eps=eps, want_counts=True, quad=q, features=True, oct=o)
        a_q = O.bh_walk(x, m, t, theta, G=G, eps=eps, quad=q)
        a_o = O.bh_walk(x, m, t, theta, G=G, eps=eps, quad=q, oct=o)
        # features mode keeps the monopole acceleration unchanged
        assert np.array_equal(a_f, a_m) and np.array_equal(cnt_f, cnt)
        # quadrupole sum = quadrupole walk - monopole walk (divided by G)
        assert np.allclose(G * wf[:, 0:3], a_q - a_m, rtol=1e-9, atol=1e-12 * np.abs(a_m).max())
        # octupole sum = octupole walk - quadrupole walk (divided by G)
        assert np.allclose(G * wf[:, 10:13], a_o - a_q, rtol=1e-7, atol=1e-12 * np.abs(a_m).max())
        # every interaction is either an accepted node or a particle pair
        assert np.array_equal(wf[:, 8] + wf[:, 9], cnt.astype(float))
        # |v| <= s3 (per node |term| = M l^3 |R| / r^6 <= M l^3 / r^5)
        assert np.all(np.linalg.norm(wf[:, 3:6], axis=1) <= wf[:, 7] * (1 + 1e-12))
        assert np.all(wf[:, 13:] == 0.0)
        # without third moments the octupole sums stay zero
        _, wq = O.bh_walk(x, m, t, theta, G=G, eps=eps, quad=q, features=True)
        assert np.all(wq[:, 10:] == 0.0) and np.array_equal(wq[:, :10], wf[:, :10])


def test_walk_features_theta_zero_all_pairs():
    x, v, m = ic.plummer(500, 32)
    t = O.build_tree(x, m, 1)
    q = O.octree_quad(x, m, t)
    a, wf = O.bh_walk(x, m, t, 0.0, quad=q, features=True)
    assert np.all(wf[:, 8] == 0.0) and np.all(wf[:, 9] == 499.0)
    assert np.all(wf[:, 0:8] == 0.0)


def test_walk_features_single_far_node_closed_form():
    """A cluster far from one probe particle: the probe's walk accepts the
    cluster as one node at depth 1, so its features follow from the node's
    mass, centre of mass, side and quadrupole in closed form."""
    rng = np.random.default_rng(7)
    cl = rng.uniform(-0.5, 0.5, size=(100, 3)) + np.array([-400.0, -400.0, -400.0])
    x = np.vstack([cl, [[600.0, 600.0, 600.0]]])
    m = np.concatenate([rng.uniform(0.5, 1.5, 100), [1.0]])
    t = O.build_tree(x, m, 1)
    q = O.octree_quad(x, m, t)
    o = O.octree_oct(x, m, t, q)
    eps = 0.3
    a, cnt, wf = O.bh_walk(x, m, t, 0.5, eps=eps, want_counts=True, quad=q, features=True, oct=o)
    f = wf[100]
    assert f[8] == 1.0 and f[9] == 0.0 and cnt[100] == 1
    M = m[:100].sum()
    com = (m[:100, None] * cl).sum(axis=0) / M
    R = x[100] - com
    r2 = R @ R + eps * eps
    r = np.sqrt(r2)
    # the cluster fills one depth-1 octant: side width / 2, l / d ~ 0.29 < 0.5
    l = t["width"] / 2.0
    d1 = np.flatnonzero(t["depth"] == 1)
    assert sorted(t["count"][d1].tolist()) == [1, 100]
    assert abs(f[6] - M * l ** 2 / r2 ** 2) < 1e-12 * f[6]
    assert abs(f[7] - M * l ** 3 / r ** 5) < 1e-12 * f[7]
    assert np.allclose(f[3:6], M * l ** 3 / r ** 6 * R, rtol=1e-12, atol=0)
    d = cl - com
    Q = (m[:100, None, None] * d[:, :, None] * d[:, None, :]).sum(axis=0)
    qa = 3 * Q @ R / r ** 5 - 7.5 * (R @ Q @ R) * R / r ** 7 + 1.5 * np.trace(Q) * R / r ** 5
    assert np.allclose(f[0:3], qa, rtol=1e-9, atol=1e-12 * np.abs(qa).max())
    # the octupole term: -(1/6) T_abc d_a d_b d_c d_d (1/r) of the cluster's
    # third moments, written with numpy's einsum
    T = np.einsum("i,ia,ib,ic->abc", m[:100], d, d, d)
    w = np.einsum("abd,a,b->d", T, R, R)
    tv = np.einsum("aad->d", T)
    oa = -17.5 * (w @ R) * R / r ** 9 + 7.5 * ((tv @ R) * R + w) / r ** 7 - 1.5 * tv / r ** 5
    assert np.allclose(f[10:13], oa, rtol=1e-9, atol=1e-12 * np.abs(oa).max())


def _rotation(seed):
    qm, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return qm * np.sign(np.linalg.det(qm))  # proper rotation


def test_local_frame_orthonormal_right_handed():
    rng = np.random.default_rng(8)
    a, q, v = rng.normal(size=(3, 200, 3))
    q[:20] = 0.0                 # falls back to v
    q[20:30] = 3.0 * a[20:30]    # parallel to a: falls back to v
    q[30:40] = 0.0
    v[30:40] = 0.0               # falls back to a coordinate axis
    e1, e2, e3 = NN.local_frame(a, q, v)
    E = np.stack([e1, e2, e3], axis=1)
    assert np.allclose(E @ E.transpose(0, 2, 1), np.eye(3), atol=1e-14)
    assert np.allclose(np.linalg.det(E), 1.0)
    assert np.allclose(e1, a / np.linalg.norm(a, axis=1)[:, None])
    # e2 lies in span(a, q) with a positive q component where q is usable
    sel = slice(40, None)
    assert np.allclose(np.sum(np.cross(a[sel], q[sel]) * e2[sel], axis=1), 0.0, atol=1e-12)
    assert np.all(np.sum(q[sel] * e2[sel], axis=1) > 0)
    assert np.allclose(np.sum(np.cross(a[:30], v[:30]) * e2[:30], axis=1), 0.0, atol=1e-12)


def test_features16_rotation_invariant_and_definitions():
    rng = np.random.default_rng(9)
    a = rng.normal(size=(100, 3))
    wf = np.zeros((100, 16))
    wf[:, 0:6] = rng.normal(size=(100, 6)) * 1e-2
    wf[:, 6:10] = np.abs(rng.normal(size=(100, 4)))
    wf[:, 7] = np.linalg.norm(wf[:, 3:6], axis=1) * 1.5  # |v| <= s3
    wf[:, 10:13] = rng.normal(size=(100, 3)) * 1e-3
    f = NN.features16(a, wf, 2.0)
    R = _rotation(10)
    wr = wf.copy()
    wr[:, 0:3] = wf[:, 0:3] @ R.T
    wr[:, 3:6] = wf[:, 3:6] @ R.T
    wr[:, 10:13] = wf[:, 10:13] @ R.T
    fr = NN.features16(a @ R.T, wr, 2.0)
    assert np.allclose(fr, f, rtol=1e-12, atol=1e-13)
    an = np.linalg.norm(a, axis=1)
    # the frame components reassemble |q| and |v|; q has no e3 component
    assert np.allclose(np.hypot(f[:, 0], f[:, 1]) * an, 2.0 * np.linalg.norm(wf[:, 0:3], axis=1))
    assert np.allclose(np.linalg.norm(f[:, 2:5], axis=1) * an, 2.0 * np.linalg.norm(wf[:, 3:6], axis=1))
    assert np.allclose(f[:, 0] * an ** 2, 2.0 * np.sum(wf[:, 0:3] * a, axis=1))
    assert np.allclose(f[:, 7], np.log10(an))
    # the octupole components reassemble |o| / s3
    assert np.allclose(np.linalg.norm(f[:, 13:16], axis=1), np.linalg.norm(wf[:, 10:13], axis=1) / wf[:, 7])
    assert np.allclose(10 ** f[:, 10], 2.0 * wf[:, 7] / an)


def test_apply_correction_rotates_back():
    rng = np.random.default_rng(11)
    a = rng.normal(size=(50, 3))
    wf = np.abs(rng.normal(size=(50, 16)))
    Y = rng.normal(size=(50, 3))
    out = NN.apply_correction(a, wf, Y, 1.5, 0.25)
    e1, e2, e3 = NN.local_frame(a, wf[:, 0:3], wf[:, 3:6])
    d = (out - a - 1.5 * (wf[:, 0:3] + wf[:, 10:13])) / (1.5 * 0.25 * wf[:, 7])[:, None]
    assert np.allclose(np.stack([np.sum(d * e, axis=1) for e in (e1, e2, e3)], axis=1), Y)
    # equivariance: rotating the inputs rotates the correction
    R = _rotation(12)
    wr = wf.copy()
    wr[:, 0:3] = wf[:, 0:3] @ R.T
    wr[:, 3:6] = wf[:, 3:6] @ R.T
    wr[:, 10:13] = wf[:, 10:13] @ R.T
    assert np.allclose(NN.apply_correction(a @ R.T, wr, Y, 1.5, 0.25), out @ R.T)
