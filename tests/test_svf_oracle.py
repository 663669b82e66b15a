"""The scaling-and-squaring oracle (oracle/svf.py) against closed forms:
exp(0) is the identity, a constant velocity is a translation (dft), zero
steps give identity + v, a velocity and its negation compose to the
identity to second order, and halving the step count with a doubled
velocity equals one squaring more."""
import numpy as np

from oracle import svf as E


def _smooth_vel(shape, amp, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((1,) + shape + (3,))
    F = np.fft.fftn(v, axes=(1, 2, 3))
    k = np.meshgrid(*[np.fft.fftfreq(s) for s in shape], indexing="ij")
    F *= np.exp(-2 * (np.pi * 3) ** 2 * sum(kk ** 2 for kk in k))[None, ..., None]
    v = np.real(np.fft.ifftn(F, axes=(1, 2, 3)))
    return v * amp / np.abs(v).max()


def test_zero_and_constant():
    shape = (6, 5, 4)
    ident = E.identity(shape)[None]
    assert np.allclose(E.exp_forward(np.zeros((1,) + shape + (3,))), ident, atol=1e-14)
    c = np.array([0.3, -1.2, 2.5])
    v = np.broadcast_to(c, (1,) + shape + (3,))
    for order in ("linear", "cubic"):
        assert np.allclose(E.exp_forward(v, interpolation=order, displacement=True), c, atol=1e-12)


def test_zero_steps():
    v = _smooth_vel((6, 5, 4), 0.7, 0)
    assert np.allclose(E.exp_forward(v, steps=0), E.identity((6, 5, 4))[None] + v, atol=1e-14)


def test_inverse_composes_to_identity():
    """Linear (interpolating) splines; cubic B-spline pulls smooth the field,
    as the reference's, so the composition is only approximate there."""
    v = _smooth_vel((16, 16, 16), 0.5, 1)
    fwd = E.exp_forward(v, displacement=True)
    inv = E.exp_forward(v, inverse=True, displacement=True)
    comp = inv + E.pull_vel(fwd, E.identity((16, 16, 16))[None] + inv, 1)
    assert np.abs(comp).max() < 0.05 * np.abs(fwd).max()


def test_step_doubling():
    v = _smooth_vel((8, 7, 6), 0.8, 2)
    a = E.exp_forward(v, steps=5, interpolation="cubic")
    b = E.exp_forward(2 * v, steps=6, interpolation="cubic")   # 2v / 2^6 = v / 2^5, one more squaring
    ident = E.identity((8, 7, 6))[None]
    d = a - ident
    sq = d + E.pull_vel(d, ident + d, 3)
    assert np.allclose(b - ident, sq, atol=1e-12)
