"""Pins of the oracle's feasible-step sizes (oracle/contact_oracle.c):

* additive CCD of one point-triangle and one edge-edge pair closing at a
  known speed: conservative advancement with eta = 1 - slackness stops after
  its second step at exactly slackness * t_contact (the closed form of the
  method for a uniform approach);
* the self step of a two-sheet scene never reaches a contact: after the
  step every pair is still apart (a brute-force detection with a tiny dHat
  finds no pair at zero distance), and the step is not vacuous;
* the ground step and the tet injectivity step against their closed forms:
  a vertex at height h closing at speed v; a tet whose apex moves toward the
  opposite face (volume linear in t).
"""
import numpy as np
import pytest


def _one_pair_scene(kind, h, v):
    if kind == "pt":
        x = np.array([[0.3, 0.3, h], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
        faces = np.array([[1, 2, 3]])
        edges = np.zeros((0, 2))
        surf = np.array([0])
        d = np.zeros_like(x)
        d[0, 2] = v  # x - t d: the point moves down at speed v
    else:
        x = np.array([[-1.0, 0.0, h], [1.0, 0.0, h], [0.0, -1.0, 0.0], [0.0, 1.0, 0.0]])
        faces = np.zeros((0, 3))
        edges = np.array([[0, 1], [2, 3]])
        surf = np.zeros(0)
        d = np.zeros_like(x)
        d[0:2, 2] = v
    return x, d, faces, edges, surf


@pytest.mark.parametrize("kind", ["pt", "ee"])
@pytest.mark.parametrize("slack", [0.8, 0.5])
def test_accd_closed_form(orc, kind, slack):
    h, v = 3e-4, 1e-3
    x, d, faces, edges, surf = _one_pair_scene(kind, h, v)
    step, nc = orc.ccd_self_step(x, d, 1.0, faces, edges, surf, 1e-6, slack)
    assert nc == 1
    assert step == pytest.approx(slack * h / v, rel=1e-12)


def test_self_step_keeps_every_pair_apart(pkg, orc):
    geo = pkg.scenes.contact_sheets(12)
    n1 = geo.n // 2
    d = np.zeros_like(geo.x)
    d[n1:, 2] = 2 * geo.dhat
    d += np.random.default_rng(3).uniform(-1e-4, 1e-4, d.shape)
    step, nc = orc.ccd_self_step(geo.x, d, 1.0, geo.faces, geo.edges, geo.surf, geo.dhat ** 2, 0.8)
    assert nc > 100 and 0.01 < step < 1.0
    xs = geo.x - step * d
    # nothing within 1e-8 of contact after the step
    pr, _, c = orc.contact_detect(xs, geo.rest, geo.faces, geo.edges, geo.surf, 1e-16)
    assert c[0] == 0
    # the step stops short of the sheets' nearest approach: sheet B falls by 2 dhat per unit t, its
    # lowest vertex starts at most gap + jitter above sheet A's highest, so contact comes before
    # t = (zB_min - zA_max + 2e-4) / (2 dhat) and the step is at most slackness times that
    za, zb = geo.x[:n1, 2].max(), geo.x[n1:, 2].min()
    assert step <= 0.8 * (zb - za + 2e-4) / (2 * geo.dhat - 1e-4)


def test_ground_step_closed_form(orc):
    nrm = np.array([0.0, 1.0, 0.0])
    x = np.array([[0.0, 0.5, 0.0], [1.0, 0.2, 0.0], [2.0, 0.9, 0.0]])
    d = np.array([[0.0, 1.0, 0.0], [0.0, 0.1, 0.0], [0.0, -1.0, 0.0]])  # x - t d moves the first two down
    # distances to y = 0.1 are 0.4, 0.1, 0.8; contact times 0.4 and 1.0
    assert orc.ground_step(x, np.arange(3), nrm, 0.1, d, 0.8) == pytest.approx(0.8 * 0.4, rel=1e-14)


@pytest.mark.parametrize("slack", [0.8, 0.3])
def test_injective_step_closed_form(orc, slack):
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.2, 0.3, 0.5]])
    tets = np.array([[0, 1, 2, 3]])
    d = np.zeros_like(x)
    d[3, 2] = 0.25  # the apex falls toward the base; volume zero at t = 2
    # det(t) = det0 (1 - t / 2) = slack det0  ->  t = 2 (1 - slack)
    for er in (1e-20, 1.0):  # cubic branch and the reference's low-order branch
        assert orc.injective_step(x, tets, d, slack, er) == pytest.approx(min(1.0, 2 * (1 - slack)), rel=1e-9)
