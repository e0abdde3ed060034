"""Oracle checks fixed by the mathematics rather than by the fixtures.

The golden fixtures pin the oracle to the reference bit for bit on three
boxes; these tests pin the same functions against closed forms and
invariants, chosen so that a plausible mistake (a dropped term, a wrong
sign or index, a transposed operand) fails one of them:

  deposit (source/core/plasma_currentKernel.c:20-279)
      discrete continuity (rho1 - rho0)/tau + div J = 0 at round-off for a
      single marker crossing 0, 1, 2 and 3 cell faces in one step;
  Boris push (source/core/plasma.c:133-171)
      uniform E: p(n tau) = p0 + n tau (q/m) E exactly;
      uniform H: |p| conserved and the rotation per step equal to
      2 atan(|t|), t = tau (q/m) H / (2 gamma);
  Yee update (source/core/em.c:55-184)
      a standing wave E_y = sin(k x) evolves with the exact discrete
      solution of the Yee leapfrog (grid curl-curl eigenvalue
      (2 sin(k h / 2) / h)^2), started from H = 0 as the main loop starts.
"""
import numpy as np
import pytest

from oracle_lib import MARKER_DTYPE, Oracle

PERIODIC = [0, 0, 0, 0, 0, 0]


def box(n, h=0.05):
    n = np.array(n)
    hh = np.full(3, h)
    tau = 0.45 * h / np.sqrt(3.0)
    return n, hh, n * hh, tau


def zero_fields(n):
    nodes = int(np.prod(n + 1))
    return np.zeros(3 * nodes), np.zeros(3 * nodes)


def one_marker(x, u, rho=-1e-9, qm=-1.0):
    M = np.zeros(1, MARKER_DTYPE)
    M["x"], M["y"], M["z"] = x
    M["vx"], M["vy"], M["vz"] = u
    M["rho"], M["qDivM"], M["id"] = rho, qm, 1
    return M


def node_view(F, n):
    return F.reshape(n[0] + 1, n[1] + 1, n[2] + 1, 3)


@pytest.mark.parametrize("x_cells,u", [
    ((3.40, 3.50, 3.60), (0.01, -0.02, 0.015)),   # stays in its cell
    ((3.97, 3.50, 3.60), (0.9, 0.0, 0.0)),        # crosses x
    ((3.97, 3.02, 3.60), (0.9, -0.9, 0.0)),       # crosses x and y
    ((3.97, 3.02, 3.98), (0.9, -0.9, 0.9)),       # crosses all three
    ((3.03, 3.98, 3.01), (-0.7, 0.8, -0.6)),      # all three, other signs
])
def test_single_marker_continuity(x_cells, u):
    n, h, L, tau = box((8, 8, 8))
    o = Oracle(n, h, L, tau, PERIODIC)
    E, H = zero_fields(n)
    o.set_fields(E, H)
    x = np.array(x_cells) * h
    rho = -1.0
    o.set_markers(one_marker(x, u, rho=rho))
    o.init()
    _, g = o.step(gauss=True)
    M = o.markers()
    moved = np.floor(np.array([M["x"][0], M["y"][0], M["z"][0]]) / h) - np.floor(x / h)
    crossings = int(np.count_nonzero(moved))
    expect = sum(1 for a in range(3) if np.floor(x_cells[a] + 1e-12) !=
                 np.floor((x[a] + u[a] / np.sqrt(1 + np.dot(u, u)) * tau) / h[a]))
    assert crossings == expect
    scale = abs(rho) / tau
    assert g[0] < 1e-12 * scale, (g, crossings)


def test_uniform_E_momentum_is_linear():
    n, h, L, tau = box((8, 8, 8))
    o = Oracle(n, h, L, tau, PERIODIC)
    E, H = zero_fields(n)
    Ev = np.array([0.3, -0.2, 0.1])
    node_view(E, n)[...] = Ev
    o.set_fields(E, H)
    u0 = np.array([0.05, 0.02, -0.01])
    qm = -1.0
    o.set_markers(one_marker(np.array([4.3, 4.4, 4.5]) * h, u0, rho=-1e-12, qm=qm))
    o.init()
    steps = 40
    for _ in range(steps):
        o.step()
    M = o.markers()
    u = np.array([M["vx"][0], M["vy"][0], M["vz"][0]])
    expect = u0 + steps * tau * qm * Ev
    assert np.allclose(u, expect, rtol=1e-12, atol=1e-14), (u, expect)


def test_uniform_H_rotation():
    n, h, L, tau = box((8, 8, 8))
    o = Oracle(n, h, L, tau, PERIODIC)
    E, H = zero_fields(n)
    Hz = 40.0
    node_view(H, n)[..., 2] = Hz
    o.set_fields(E, H)
    u0 = np.array([0.6, 0.0, 0.25])
    qm = -1.0
    o.set_markers(one_marker(np.array([4.5, 4.5, 4.5]) * h, u0, rho=-1e-12, qm=qm))
    o.init()
    steps = 30
    angles = []
    for _ in range(steps):
        o.step()
        M = o.markers()
        u = np.array([M["vx"][0], M["vy"][0], M["vz"][0]])
        assert abs(np.linalg.norm(u) - np.linalg.norm(u0)) < 1e-13
        assert abs(u[2] - u0[2]) < 1e-15
        angles.append(np.arctan2(u[1], u[0]))
    gamma = np.sqrt(1 + np.dot(u0, u0))
    t = tau * qm * Hz / (2 * gamma)
    # the rotation of p about H per step: right-handed about -t (charge sign)
    dtheta = np.unwrap(np.diff([0.0] + angles))
    assert np.allclose(dtheta, -2 * np.arctan(t), rtol=0, atol=1e-12), (dtheta[:3], -2 * np.arctan(t))


def test_yee_standing_wave_dispersion():
    n, h, L, tau = box((32, 8, 8))
    o = Oracle(n, h, L, tau, PERIODIC)
    E, H = zero_fields(n)
    k = 2 * np.pi / L[0] * 3
    xi = np.arange(n[0] + 1) * h[0]
    profile = np.sin(k * xi)
    node_view(E, n)[..., 1] = profile[:, None, None]
    o.set_fields(E, H)
    o.set_markers(np.zeros(0, MARKER_DTYPE))
    o.init()
    # Main-loop order (source/core/main.c:365-425): two H half steps with
    # E^n, then E^{n+1} = E^n + tau curl H^{n+1}.  For the mode, with
    # W = 2 sin(k h / 2) / h the grid curl-curl eigenvalue, the amplitude
    # obeys q[n+1] - 2 q[n] + q[n-1] = -(W tau)^2 q[n], q[0] = 1 and (H = 0
    # at the start) q[1] = 1 - (W tau)^2:
    #   q[n] = cos(n theta) + B sin(n theta),  cos(theta) = 1 - (W tau)^2 / 2,
    #   B = -((W tau)^2 / 2) / sin(theta).
    W = 2 * np.sin(k * h[0] / 2) / h[0]
    c = 1 - (W * tau) ** 2 / 2
    theta = np.arccos(c)
    B = -((W * tau) ** 2 / 2) / np.sin(theta)
    for step in range(1, 61):
        o.step()
        if step % 15 == 0:
            Ey = node_view(o.fields()[0], n)[..., 1]
            q = np.cos(step * theta) + B * np.sin(step * theta)
            assert np.max(np.abs(Ey - (q * profile)[:, None, None])) < 1e-12, step
            # the continuum dispersion (w = k) is off by far more than round-off
            qc = np.cos(step * k * tau) - (k * tau) ** 2 / 2 / np.sin(k * tau) * np.sin(step * k * tau)
            assert np.max(np.abs(Ey - (qc * profile)[:, None, None])) > 1e-6
