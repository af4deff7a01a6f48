"""Topology helpers: the reference's own known-answer tests
(tests/test_topology_updates.py:7-55, re-expressed) against both the oracle and
the product, plus closed forms for Adam, the projection and the blur adjoint."""
import numpy as np
import pytest

from beamz_amd.optimization import topology as T
from beamz_amd.optimization.optimizers import Optimizer
from oracle import topo_ref as O

IMPLS = [("oracle", O.apply_density_update), ("product", lambda *a, **k: T.apply_density_update(*a, **k))]


@pytest.mark.parametrize("name,fn", IMPLS)
def test_kat_masked_region_changes(name, fn):
    density = np.zeros((3, 3)); mask = np.zeros_like(density, dtype=bool); mask[1, 1] = True
    new, applied, delta, cn, mu = fn(density, np.ones_like(density), mask, learning_rate=0.2)
    assert new[1, 1] == pytest.approx(0.2) and cn > 0 and mu == pytest.approx(cn)
    assert np.all(new[~mask] == 0.0) and np.all(delta[~mask] == 0.0) and applied[1, 1] == pytest.approx(1.0)


@pytest.mark.parametrize("name,fn", IMPLS)
def test_kat_clip_and_mask(name, fn):
    density = np.full((3, 3), 0.95); mask = np.zeros_like(density, dtype=bool); mask[1, 1] = True
    new, _, delta, cn, mu = fn(density, np.ones_like(density), mask, learning_rate=0.5)
    assert new[1, 1] == pytest.approx(1.0) and np.all(new[~mask] == density[~mask])
    assert np.all(delta[~mask] == 0.0) and cn == pytest.approx(mu)


@pytest.mark.parametrize("name,fn", IMPLS)
def test_kat_zero_gradient(name, fn):
    density = np.random.default_rng(0).random((4, 4)); mask = np.ones_like(density, dtype=bool)
    new, applied, delta, cn, mu = fn(density, np.zeros_like(density), mask, learning_rate=0.3)
    assert np.allclose(new, density) and np.allclose(applied, 0) and np.allclose(delta, 0) and cn == 0 and mu == 0


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([3.0, -0.5, 1e-3])
    st = {}
    u = O.adam_update(g, 0.1, st)
    # m_hat = g, v_hat = g^2 at t = 1  ->  -lr * g / (|g| + eps)
    np.testing.assert_allclose(u, -0.1 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    opt = Optimizer(method="adam", learning_rate=0.1)
    np.testing.assert_allclose(opt.step(g.copy()), u, rtol=1e-15)
    u2o = O.adam_update(g * 2, 0.1, st)
    np.testing.assert_allclose(opt.step(g * 2), u2o, rtol=1e-15)


def test_projection_identities():
    for beta in (2.0, 8.0, 32.0):
        p, d = O.smoothed_heaviside(np.array([0.0, 0.5, 1.0]), beta, 0.5)
        np.testing.assert_allclose(p, [0.0, 0.5, 1.0], atol=1e-14)
        x = np.linspace(0.05, 0.95, 7)
        fd = (O.smoothed_heaviside(x + 1e-6, beta, 0.5)[0] - O.smoothed_heaviside(x - 1e-6, beta, 0.5)[0]) / 2e-6
        tol = 1e-6 * np.abs(fd).max()
        np.testing.assert_allclose(O.smoothed_heaviside(x, beta, 0.5)[1], fd, rtol=1e-6, atol=tol)
        np.testing.assert_allclose(T.smoothed_heaviside(x, beta, 0.5)[1], fd, rtol=1e-6, atol=tol)


def test_box_blur_backprop_is_adjoint():
    rng = np.random.default_rng(1)
    mask = np.zeros((12, 14), bool); mask[2:10, 3:12] = True
    x = np.where(mask, rng.random(mask.shape), 0.0)
    y = rng.random(mask.shape)
    for r in (1, 2):
        b, w = O.masked_box_blur(x, mask, r)
        # the blur uses edge padding of the masked input; inside a zero-bordered mask that equals zero
        # padding, so the operator is linear in x and its transpose is the backprop
        lhs = np.sum(b * y)
        rhs = np.sum(x * O.masked_box_blur_backprop(y, mask, w, r))
        assert lhs == pytest.approx(rhs, rel=1e-12)
        pb, pw = T.masked_box_blur(x, mask, r)
        np.testing.assert_allclose(pb, b, rtol=1e-14)
        np.testing.assert_allclose(T.masked_box_blur_backprop(y, mask, pw, r),
                                   O.masked_box_blur_backprop(y, mask, w, r), rtol=1e-14)


def test_cone_filter_adjoint_and_normalisation():
    rng = np.random.default_rng(2)
    x = rng.random((15, 17)); y = rng.random((15, 17))
    np.testing.assert_allclose(O.cone_filter(np.ones((9, 9)), 3), 1.0, rtol=1e-14)
    assert np.sum(O.cone_filter(x, 3) * y) == pytest.approx(np.sum(x * O.cone_filter_T(y, 3)), rel=1e-12)
    np.testing.assert_allclose(T.cone_filter(x, 3), O.cone_filter(x, 3), rtol=1e-14)
    np.testing.assert_allclose(T.cone_filter_backprop(y, 3), O.cone_filter_T(y, 3), rtol=1e-14)


def test_overlap_gradient_time_reversal():
    f = [np.full((2, 2), float(k)) for k in range(4)]
    a = [np.full((2, 2), 10.0 ** k) for k in range(4)]
    g = O.overlap_gradient(a, f)
    # sum_k adj[k] * fwd[N-1-k] = 1*3 + 10*2 + 100*1 + 1000*0
    np.testing.assert_allclose(g, 123.0)
