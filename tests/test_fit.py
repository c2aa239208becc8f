"""The bounded BFGS that replaces ming2 (tools.c:6361-6545), run on the host
through bsa_debug_minimize on objectives whose box-constrained minima are
known in closed form: Rosenbrock's valley, a separable quadratic whose
minimiser is the projection of the unconstrained one, a minimum on a corner
of the box, and an objective that is undefined on part of the box (the line
search must back off, as ming2's does when lfundG fails)."""
import numpy as np
import pytest

from cucodeml_amd import engine


def rosen(x):
    f = np.sum(100 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
    g = np.zeros_like(x)
    g[:-1] += -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
    g[1:] += 200 * (x[1:] - x[:-1] ** 2)
    return f, g


def test_rosenbrock_interior_minimum():
    x, f, info = engine.minimize(rosen, np.array([-1.2, 1.0, -1.2, 1.0, 0.5]), [-5] * 5, [5] * 5)
    assert info["converged"]
    assert np.max(np.abs(x - 1)) < 1e-5
    assert f < 1e-10


def test_rosenbrock_minimum_on_a_bound():
    # with x0 <= 0.5 the minimum is x0 = 0.5, x1 = 0.25 (KKT: df/dx0 < 0 there)
    x, f, info = engine.minimize(rosen, np.array([-1.2, 1.0]), [-5, -5], [0.5, 5])
    assert info["converged"]
    assert np.allclose(x, [0.5, 0.25], atol=1e-7)
    assert abs(f - 0.25) < 1e-12


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_quadratic_minimiser_is_projection(seed):
    rng = np.random.default_rng(seed)
    n = 12
    d = 10 ** rng.uniform(-2, 2, n)            # badly scaled diagonal Hessian
    c = rng.normal(size=n) * 2
    lo, hi = -np.ones(n), np.ones(n)

    def q(x):
        return 0.5 * np.sum(d * (x - c) ** 2), d * (x - c)

    # f-based stopping: |dx|^2 d / 2 < e, so ask for e = 1e-16 to pin x to 1e-6
    x, f, info = engine.minimize(q, rng.uniform(-1, 1, n), lo, hi, e=1e-16)
    assert info["converged"]
    assert np.allclose(x, np.clip(c, lo, hi), atol=1e-6)


def test_corner_minimum_and_start_outside_box():
    # f = sum(x) on [1, 2]^3: minimum at the lower corner; start outside
    x, f, info = engine.minimize(lambda x: (np.sum(x), np.ones_like(x)), np.array([5.0, -3.0, 9.0]),
                                 [1, 1, 1], [2, 2, 2])
    assert np.allclose(x, 1) and abs(f - 3) < 1e-12


def test_undefined_region_is_avoided():
    # log barrier undefined for x <= 0 inside the box; minimum of x - log x at x = 1
    def f(x):
        if np.any(x <= 0):
            return np.inf, np.zeros_like(x)
        return np.sum(x - np.log(x)), 1 - 1 / x

    x, fv, info = engine.minimize(f, np.array([0.05, 4.0, 9.0]), [-10] * 3, [10] * 3)
    assert info["converged"]
    assert np.allclose(x, 1, atol=1e-6)


def test_bad_bounds_rejected():
    with pytest.raises(engine.BsaError):
        engine.minimize(rosen, np.zeros(2), [1, 0], [0, 1])


def _ridge(scale):
    """A likelihood-like objective with one flat direction: Hessian
    eigenvalues from 1e4 down to `scale`, minimum at c."""
    rng = np.random.default_rng(5)
    n = 10
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    lam = np.geomspace(1e4, scale, n)
    A = (Q * lam) @ Q.T
    c = rng.uniform(-0.5, 0.5, n)
    return (lambda x: (0.5 * (x - c) @ A @ (x - c) + 8000.0, A @ (x - c))), c, n


def test_newton_finish_lands_on_a_flat_minimum():
    """BFGS's H_end-type stop leaves a flat direction (curvature 1e-2) short
    of the minimum; the Newton finish (Hessian from differences of the
    gradient, one evaluation per variable) lands on it, and says so."""
    fun, c, n = _ridge(1e-2)
    x0 = np.full(n, 0.9)
    xb, _, ib = engine.minimize(fun, x0, [-1] * n, [1] * n)
    xn, _, inn = engine.minimize(fun, x0, [-1] * n, [1] * n, newton=3)
    assert inn["stop"] == "newton" and inn["converged"]
    assert np.max(np.abs(xn - c)) < 1e-8
    assert np.max(np.abs(xn - c)) <= np.max(np.abs(xb - c))


def test_newton_only_from_a_near_point_and_on_a_bound():
    # max_iter < 0: Newton iterations alone; one coordinate's minimum lies
    # outside the box, so it stays on its bound
    fun, c, n = _ridge(1.0)
    lo, hi = np.full(n, -1.0), np.full(n, 1.0)
    hi[3] = c[3] - 0.2
    x, _, info = engine.minimize(fun, np.clip(c + 0.01, lo, hi), lo, hi, max_iter=-1, newton=4)
    assert info["converged"] and info["stop"] == "newton"
    assert abs(x[3] - hi[3]) < 1e-12
    # the other coordinates minimise the quadratic with x[3] held there
    _, g = fun(x)
    free = np.arange(n) != 3
    assert np.max(np.abs(g[free])) < 1e-6


def _ridge_large(n=300, nblock=5, seed=3):
    """A tree-likelihood-like Hessian on n variables: the bulk of the
    spectrum log-normal about 3e6 (1e4.3 .. 1e8.5), and two flat directions
    (curvature 0.2 and 7.8) inside the last nblock coordinates, as the
    p2 w2 ridge lies in codeml's model parameters after the branch lengths.
    The start is 0.6 and 0.3 along the flat directions from the minimum c
    and 1e-4 off it in the steep ones, as where an FP32 fit stops."""
    rng = np.random.default_rng(seed)
    U = np.zeros((n, n))
    F, _ = np.linalg.qr(rng.normal(size=(nblock, 2)))
    U[n - nblock:, :2] = F
    R = rng.normal(size=(n, n - 2))
    R -= U[:, :2] @ (U[:, :2].T @ R)
    U[:, 2:], _ = np.linalg.qr(R)
    lam = np.concatenate([[0.2, 7.8], 10 ** np.clip(rng.normal(6.5, 0.45, n - 2), 4.3, 8.5)])
    A = (U * lam) @ U.T
    A = 0.5 * (A + A.T)
    c = rng.uniform(-0.5, 0.5, n)
    x0 = c + U[:, :2] @ np.array([0.6, 0.3]) + U[:, 2:] @ (rng.normal(size=n - 2) * 1e-4)
    return (lambda x: (0.5 * (x - c) @ A @ (x - c) + 5e6, A @ (x - c))), A, c, x0


def test_newton_on_many_variables_preconditions_with_the_trailing_block():
    """More than 160 free variables: Newton steps from preconditioned
    conjugate gradients (the Hessian block of the last `block` variables
    formed exactly, a scalar on the rest), at most 5 + 32 gradients each
    plus the line search.  With the block they land on the minimum along
    the flat directions; plain conjugate gradients (block 0) on the same
    budget barely move along them."""
    fun, A, c, x0 = _ridge_large()
    n = len(c)
    lo, hi = -np.ones(n), np.ones(n)
    x, _, info = engine.minimize(fun, x0, lo, hi, max_iter=-1, newton=6, block=5)
    assert info["converged"] and info["stop"] == "newton"
    assert np.max(np.abs(x - c)) < 1e-8
    assert info["evaluations"] <= 1 + 6 * (5 + 32 + 30)
    xp, _, ip = engine.minimize(fun, x0, lo, hi, max_iter=-1, newton=6, block=0)
    assert not ip["converged"] and np.max(np.abs(xp - c)) > 0.1


def test_newton_on_many_variables_with_a_bound():
    # the minimum of the last variable lies outside the box: it stays on
    # its bound and the other variables minimise with it held there
    fun, A, c, x0 = _ridge_large()
    n = len(c)
    lo, hi = -np.ones(n), np.ones(n)
    hi[-1] = c[-1] - 0.2
    x, _, info = engine.minimize(fun, np.clip(x0, lo, hi), lo, hi, max_iter=-1, newton=6,
                                 block=5)
    assert info["converged"] and x[-1] == hi[-1]
    _, g = fun(x)
    assert np.max(np.abs(g[:-1])) < 1e-3 and g[-1] < 0


def test_block_argument_is_checked():
    # nblock names the last variables of x: 0..n
    fun, A, c, x0 = _ridge_large(n=170)
    lo, hi = -np.ones(len(c)), np.ones(len(c))
    for bad in (-1, len(c) + 1):
        with pytest.raises(engine.BsaError):
            engine.minimize(fun, x0, lo, hi, max_iter=-1, newton=1, block=bad)
    x, _, info = engine.minimize(fun, x0, lo, hi, max_iter=-1, newton=1, block=len(c))
    assert np.all(np.isfinite(x)) and info["iterations"] == 1
