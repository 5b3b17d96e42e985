"""The NumPy bath-distance oracle pinned without the product: closed forms on a one-pole
bath and the finite-difference rule on a function with a known gradient."""
import numpy as np

from oracle import bath


def test_delta_single_pole_closed_form():
    wn = np.pi / 50 * (2 * np.arange(1, 65) - 1)
    d = bath.delta([0.3], [[0.7]], wn)
    assert np.allclose(d[:, 0, 0], 0.49 / (1j * wn - 0.3), rtol=0, atol=1e-15)


def test_distance_single_pole_by_hand():
    rng = np.random.default_rng(0)
    wn = np.pi / 20 * (2 * np.arange(1, 33) - 1)
    tgt = (rng.normal(size=(32, 2, 1, 1)) + 1j * rng.normal(size=(32, 2, 1, 1))) * 0.1
    for p, q, s in ((2.0, 0.0, 0.0), (1.0, 1.0, 0.0), (3.0, 2.0, 0.5)):
        d = 0.25 / (1j * wn + 0.8)
        shift = abs(wn[0]) if s == 0.0 else s
        ref = 0.0
        for n in range(32):
            for sg in range(2):
                ref += abs(d[n] - tgt[n, sg, 0, 0]) ** p / (abs(wn[n] - wn[0]) + shift) ** q
        ref /= 32 * 4   # N_w (2 nc)^2 with nc = 1
        got = bath.distance([-0.8], [[0.5]], wn, tgt, p, q, s)
        assert abs(got - ref) < 1e-14 * max(1.0, ref)


def test_fd_gradient_rule_on_a_quadratic():
    A = np.array([[2.0, 0.3], [0.3, 1.0]])
    f = lambda x: 0.5 * x @ A @ x + x[0] ** 3
    x = np.array([0.7, -1.3])
    exact = A @ x + np.array([3 * x[0] ** 2, 0.0])
    central = bath.fd_gradient_civelli(f, x, 1e-5, idiff=2)
    forward = bath.fd_gradient_civelli(f, x, 1e-5, idiff=1)
    assert np.max(np.abs(central - exact)) < 1e-8
    assert np.max(np.abs(forward - exact)) < 1e-4
    assert np.max(np.abs(forward - exact)) > np.max(np.abs(central - exact))   # O(z) vs O(z^2)
