"""Pins of the oracle's Newton objective and line search (oracle/__init__.py
scene_energy / line_search): central differences of the energy equal the
assembled gradient the oracle writes with the Hessian (itself pinned in the
element tests), with and without the inertia term; the line search accepts
a Newton direction with a lower energy and halves a rising direction until
it gives up."""
import numpy as np
import pytest


def _fd_grad(orc, s, x, idx, h, xt=None):
    g = []
    for i in idx:
        xp, xm = x.copy().reshape(-1), x.copy().reshape(-1)
        xp[i] += h
        xm[i] -= h
        g.append((orc.scene_energy(s, xp, xt) - orc.scene_energy(s, xm, xt)) / (2 * h))
    return np.array(g)


@pytest.mark.parametrize("kind", ["tets", "cloth_contact"])
def test_energy_gradient_is_the_assembled_gradient(pkg, orc, kind):
    if kind == "tets":
        s = pkg.scenes.tet_grid_scene("e_grid4", 4, n_pairs=0)
        h = 1e-8
    else:
        s = pkg.scenes.cloth_contact_scene(6, kappa=100.0)
        h = 1e-10
    _, _, _, g = orc.triplets(s, with_grad=True)
    rng = np.random.default_rng(0)
    idx = rng.choice(3 * s.n, 24, replace=False)
    xt = s.x + rng.uniform(-1e-4, 1e-4, s.x.shape)
    for x_tilde in (None, xt):
        want = g[idx].copy()
        if x_tilde is not None:
            want += np.repeat(s.mass, 3)[idx] * (s.x - x_tilde).reshape(-1)[idx]
        fd = _fd_grad(orc, s, s.x, idx, h, x_tilde)
        np.testing.assert_allclose(fd, want, rtol=2e-5, atol=1e-6 * np.abs(want).max())


def test_line_search_accepts_newton_and_rejects_ascent(pkg, orc):
    s = pkg.scenes.cloth_contact_scene(8, kappa=100.0)
    _, _, _, g = orc.triplets(s, with_grad=True)
    xt = s.x.copy()
    # Newton direction of the objective at x (x~ = x: the gradient is the elastic + contact one)
    s2 = s.__class__(**{**s.__dict__, "rhs": g})
    d = orc.solve_scene(s2, 32, tol=1e-8).x.reshape(-1, 3)
    a, x, e0, e, stopped = orc.line_search(s, d, xt)
    assert not stopped and e < e0 and 0 < a <= 1.0
    a, x, e0, e, stopped = orc.line_search(s, -d, xt)
    assert stopped and np.array_equal(x, s.x) and a <= 0.5 ** 9
