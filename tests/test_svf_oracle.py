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


# --- Jacobians and the analytic backward (oracle/svf.py) -------------------
_PADS = {"dct2": "symmetric", "replicate": "edge", "dct1": "reflect", "dft": "wrap", "zero": "constant"}


def test_central_differences_match_padded_definition():
    """Each bound's end rule equals the central difference of the line
    extended by numpy.pad's matching mode; dst2 by the antisymmetric
    extension x[-1] = -x[0]."""
    rng = np.random.default_rng(30)
    x = rng.standard_normal((3, 7, 2))
    for bound, mode in _PADS.items():
        d = E.diff_central(x, [1], bound)[..., 0]
        p = np.pad(x, ((0, 0), (1, 1), (0, 0)), mode=mode)
        assert np.allclose(d, (p[:, 2:] - p[:, :-2]) / 2, atol=1e-14), bound
    p = np.concatenate([-x[:, :1], x, -x[:, -1:]], 1)
    assert np.allclose(E.diff_central(x, [1], "dst2")[..., 0], (p[:, 2:] - p[:, :-2]) / 2, atol=1e-14)


def test_grid_jacobian_of_a_periodic_field():
    """dft central difference of sin(2 pi k x / n) is cos(2 pi k x / n) sin(2 pi k / n);
    jac[..., i, j] = d psi_i / d x_j."""
    shape = (8, 6, 5)
    idx = E.identity(shape)
    k = 2 * np.pi * np.array([1, 2, 1]) / np.array(shape)
    disp = np.zeros((1,) + shape + (3,))
    disp[0, ..., 1] = np.sin(k[0] * idx[..., 0])          # psi_1 depends on x_0
    jac = E.grid_jacobian(disp, bound="dft", type="disp")
    assert np.allclose(jac[0, ..., 1, 0], np.cos(k[0] * idx[..., 0]) * np.sin(k[0]), atol=1e-12)
    assert np.allclose(jac[0, ..., 0, 0], 1) and np.allclose(jac[0, ..., 1, 1], 1)
    assert np.allclose(jac[0, ..., 0, 1], 0) and np.allclose(jac[0, ..., 2, 0], 0)
    # a transformation field gives the same Jacobian as its displacement
    assert np.allclose(E.grid_jacobian(disp + E.identity(shape)[None], bound="dft"), jac)


def test_forward_jacobian_constant_and_smooth():
    c = np.array([0.7, -1.3, 2.1])
    v = np.broadcast_to(c, (1, 6, 7, 5, 3)).copy()
    disp, jac = E.exp_forward_jac(v, steps=3)
    assert np.allclose(jac, np.eye(3), atol=1e-12) and np.allclose(disp, c, atol=1e-12)
    # smooth field: the composed Jacobian tracks the finite-difference one
    v = _smooth_vel((24, 24, 24), 0.6, 3)
    disp, jac = E.exp_forward_jac(v, steps=4)
    fd = E.grid_jacobian(disp, bound="dft", type="disp")
    assert np.abs(jac - fd).max() <= 0.1 * np.abs(fd - np.eye(3)).max()


def test_backward_constant_velocity_closed_form():
    """Constant velocity c with c / 2^s an integer step u (dft, linear):
    grad_out(x) = 2^-s sum_{m < 2^s} grad(x - m u); inverse: the negated sum
    over x + m u."""
    rng = np.random.default_rng(31)
    s, u = 2, np.array([1, 0, 2])
    v = np.broadcast_to(u * 2.0 ** s, (1, 6, 5, 8, 3)).copy()
    g = rng.standard_normal(v.shape)
    ref = sum(np.roll(g, shift=tuple(m * u), axis=(1, 2, 3)) for m in range(2 ** s)) / 2 ** s
    assert np.allclose(E.exp_backward(v, g, steps=s), ref, atol=1e-12)
    ref_i = -sum(np.roll(g, shift=tuple(-m * u), axis=(1, 2, 3)) for m in range(2 ** s)) / 2 ** s
    assert np.allclose(E.exp_backward(v, g, steps=s, inverse=True), ref_i, atol=1e-12)
    # with a Hessian: J = I, det = 1, so it is summed the same way
    h = rng.standard_normal(v.shape[:-1] + (6,))
    _, hh = E.exp_backward(v, g, h, steps=s)
    assert np.allclose(hh, sum(np.roll(h, shift=tuple(m * u), axis=(1, 2, 3)) for m in range(2 ** s)) / 2 ** s,
                       atol=1e-12)


def test_backward_tracks_finite_differences():
    """For a small smooth velocity the analytic (DARTEL) gradient of
    L(v) = <exp(v), g> agrees with a central finite difference to first order."""
    rng = np.random.default_rng(32)
    v = _smooth_vel((16, 16, 16), 0.3, 4)
    g = rng.standard_normal(v.shape)
    d = _smooth_vel((16, 16, 16), 1.0, 5)
    eps = 1e-4
    lp = (E.exp_forward(v + eps * d, steps=5) * g).sum()
    lm = (E.exp_forward(v - eps * d, steps=5) * g).sum()
    fd = (lp - lm) / (2 * eps)
    an = (E.exp_backward(v, g, steps=5, rotate_grad=True) * d).sum()
    assert abs(an - fd) <= 0.1 * abs(fd)


def test_jhj_is_congruence():
    rng = np.random.default_rng(33)
    J = rng.standard_normal((4, 3, 3))
    h = rng.standard_normal((4, 6))
    H = E._sym(h)
    out = E._sym(E.jhj(J, h))
    for i in range(4):
        assert np.allclose(out[i], J[i] @ H[i] @ J[i].T)
