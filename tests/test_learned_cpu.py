"""CPU checks of the learned-guess reference network (the reference the GPU kernels are compared against)."""
import numpy as np

from oracle.learned_ref import NetRef


def test_reference_gradient_matches_finite_differences():
    net = NetRef(41, C=4, NL=4, H=16, NB=6)
    rng = np.random.default_rng(0)
    p = rng.standard_normal(net.nparams) * 0.3
    X = rng.standard_normal((4, 41))
    Y = rng.standard_normal((4, 6))
    _, g = net.loss_grad(p, X, Y)
    floor = 1e-8 * np.linalg.norm(g)     # central-difference rounding floor
    for i in rng.choice(net.nparams, 40, replace=False):
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (net.loss_grad(p + e, X, Y)[0] - net.loss_grad(p - e, X, Y)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * (abs(g[i]) + abs(fd)) + floor


def test_reference_layout_of_the_default_network():
    net = NetRef(641)
    assert net.L == [641, 321, 161, 81, 41]
    assert net.flat == 32 * 41
    # conv: 1->32, 3 x 32->32, kernel 5; MLP 1312 -> 256 -> 64
    assert net.nparams == (32 * 5 + 32) + 3 * (32 * 32 * 5 + 32) + (256 * 1312 + 256) + (64 * 256 + 64)
