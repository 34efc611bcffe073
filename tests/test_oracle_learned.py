"""Pins the binary64 oracle of the learned collision step, its training
targets and the 4x restriction (oracle/learned_oracle.c) with closed forms,
each chosen so that a plausible slip fails it:

  one-moment states        a swapped or mis-normalised moment row, a
                           correction that leaks into other moments
  linear pass-through net  a wrong input normalisation (weight index, s_in,
                           equilibrium without the Guo velocity)
  clamp over a tau sweep   a wrong bound, an unstable relaxation factor
  mask                     corrections next to solids or open edges
  constant-modulation      a wrong residual, regulariser or clamp in the
  transition               training targets
  uniform shear            a wrong non-equilibrium rescale 4 tau_c / tau_f
  random restriction       wrong block sums, mass weighting, Guo shift
"""
import numpy as np
import pytest

from oracle import oracle as O
from simuverse_amd import workloads as W

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, -1, 0, 1, -1, -1, 1, 1])
WT = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
# orthogonal moment rows of the correction (Lallemand & Luo 2000)
M_E = np.array([-4, -1, -1, -1, -1, 2, 2, 2, 2], float)
M_XX = (EX * EX - EY * EY).astype(float)
M_XY = (EX * EY).astype(float)
ROWS = [M_E, M_XX, M_XY]
NORM2 = [36.0, 4.0, 4.0]
# the remaining rows of the basis: mass, momentum and the three ghost moments
M_QX = np.array([0, -2, 0, 2, 0, 1, -1, -1, 1], float)
M_QY = np.array([0, 0, 2, 0, -2, -1, -1, 1, 1], float)
M_EPS = np.array([4, -2, -2, -2, -2, 1, 1, 1, 1], float)
OTHER = [np.ones(9), EX.astype(float), EY.astype(float), M_QX, M_QY, M_EPS]

NP = 5385
B1, W2, B2, W3, B3 = 576, 640, 4736, 4800, 5376


def feq(rho, ux, uy):
    eu = EX[:, None] * np.ravel(ux)[None] + EY[:, None] * np.ravel(uy)[None]
    usq = np.ravel(ux) ** 2 + np.ravel(uy) ** 2
    return WT[:, None] * np.ravel(rho)[None] * (1 + 3 * eu + 4.5 * eu**2 - 1.5 * usq)


def const_net(y):
    """A network whose output is the constant y (all weights zero)."""
    p = np.zeros(NP)
    p[B3:B3 + len(y)] = y
    return p


def linear_net(j, k, out=0):
    """A network that passes input j, times k, to output `out`: one hidden
    unit per layer biased to +20, where GELU(z) = z to binary64 precision,
    the others biased to -40, where GELU is 0; the output bias removes the
    offset."""
    p = np.zeros(NP)
    p[0 * 9 + j] = k
    p[B1] = 20.0
    p[B1 + 1:W2] = -40.0
    p[W2 + 0] = 1.0  # W2[0][0]
    p[B2 + 1:W3] = -40.0
    p[W3 + out * 64 + 0] = 1.0
    p[B3 + out] = -20.0
    return p


def uniform_state(nx, ny, f):
    return np.repeat(np.asarray(f, float)[:, None], nx * ny, axis=1).reshape(9, ny, nx)


@pytest.mark.parametrize("tau", [0.505, 0.6, 1.0, 2.5])
@pytest.mark.parametrize("a", [0, 1, 2], ids=["e", "xx", "xy"])
def test_one_moment_state_is_modulated_in_that_moment_only(tau, a):
    """f = feq(rho, u) + eps m_a / |m_a|^2 in every cell of a periodic box:
    the learned step leaves (1 - omega + alpha_a) eps in moment a, zero in
    the other two, and mass, momentum and the ghost moments as plain BGK."""
    nx, ny = 6, 5
    rho0, u0 = 1.03, (0.04, -0.02)
    eps = 2e-3
    omega = 1 / tau
    b = O.modulation_bound(omega)
    for alpha in (0.6 * b, -0.3 * b):
        s_out = 0.5
        y = np.zeros(9)
        y[a] = alpha / s_out
        f = feq(rho0, *u0)[:, 0] + eps * ROWS[a] / NORM2[a]
        post = uniform_state(nx, ny, f)
        p = O.phys_params(nx, ny, tau)
        out, learned = O.learned_step(p, np.zeros((ny, nx), bool), post, const_net(y), 0.7, s_out)
        assert learned.all()
        bgk = O.phys_collide(p, f)
        q = out[:, 2, 3] - feq(rho0, *u0)[:, 0]
        for bb in range(3):
            expect = (1 - omega + alpha) * eps if bb == a else 0.0
            assert abs(ROWS[bb] @ q - expect) < 1e-15, (bb, ROWS[bb] @ q, expect)
        for m in OTHER:
            assert abs(m @ out[:, 2, 3] - m @ bgk) < 1e-15
        assert np.all(out == out[:, :1, :1])


def test_linear_network_pins_the_input_normalisation():
    """With y_0 = k x_j (x_j = (f_j - feq_j) / (w_j s_in), feq at the Guo
    velocity) the correction's e-moment is clamp(s_out k x_j) n_e and its
    other moments vanish, for every input j: on a periodic lattice with a
    body force and a random state."""
    nx, ny = 9, 7
    tau, g = 0.56, (3e-4, -2e-4)
    s_in, s_out = 3e-3, 0.8
    omega = 1 / tau
    b = O.modulation_bound(omega)
    rng = np.random.default_rng(4)
    post = feq(np.ones(nx * ny), 0.03 * np.ones(nx * ny), 0.01 * np.ones(nx * ny)) \
        + rng.uniform(-2e-3, 2e-3, (9, nx * ny))
    post = post.reshape(9, ny, nx)
    p = O.phys_params(nx, ny, tau, force=g)
    solid = np.zeros((ny, nx), bool)
    plain = O.PhysState(p, solid, post)
    plain.run(1)
    fg = np.stack([O.PhysState(p, solid, post).gather(x, y) for y in range(ny) for x in range(nx)], axis=1)
    rho = fg.sum(axis=0)
    ux = (EX @ fg + 0.5 * g[0]) / rho
    uy = (EY @ fg + 0.5 * g[1]) / rho
    q = fg - feq(rho, ux, uy)
    n_e = M_E @ q
    for j in range(9):
        k = 0.06 * (1 + 0.1 * j)
        out, learned = O.learned_step(p, solid, post, linear_net(j, k), s_in, s_out)
        assert learned.all()
        corr = (out - plain.f).reshape(9, -1)
        x_j = q[j] / (WT[j] * s_in)
        alpha = np.clip(s_out * k * x_j, -b, b)
        assert 0 < np.count_nonzero(np.abs(s_out * k * x_j) > b) < nx * ny  # both sides of the clamp
        np.testing.assert_allclose(M_E @ corr, alpha * n_e, rtol=1e-9, atol=1e-18)
        for m in [M_XX, M_XY] + OTHER:
            assert np.max(np.abs(m @ corr)) < 1e-16


@pytest.mark.parametrize("tau", [0.5005, 0.505, 0.52, 0.6, 0.8, 1.0, 1.49, 1.5, 2.0, 3.0, 10.0])
def test_clamp_keeps_every_relaxation_factor_inside_unit_interval(tau):
    """A saturating network (outputs +-1e3) drives alpha to the bound b:
    the non-equilibrium of each moment comes out times 1 - omega +- b,
    which lies in (-1, 1) for every tau, and b = (2 - omega)/2 for tau <= 1
    (the bound the step was designed with) and omega/2 above."""
    omega = 1 / tau
    b = O.modulation_bound(omega)
    assert b == pytest.approx((2 - omega) / 2 if tau <= 1 else omega / 2, rel=1e-15)
    eps = 1e-3
    p = O.phys_params(3, 3, tau)
    solid = np.zeros((3, 3), bool)
    for a in range(3):
        for sign in (1.0, -1.0):
            y = np.zeros(9)
            y[a] = sign * 1e3
            f = feq(1.0, 0.02, 0.0)[:, 0] + eps * ROWS[a] / NORM2[a]
            out, _ = O.learned_step(p, solid, uniform_state(3, 3, f), const_net(y), 1.0, 1.0)
            factor = ROWS[a] @ (out[:, 1, 1] - feq(1.0, 0.02, 0.0)[:, 0]) / eps
            assert factor == pytest.approx(1 - omega + sign * b, abs=1e-11)
            assert -1 < factor < 1


def test_mask_channel_with_obstacles():
    """Cells with a solid in their 3x3 neighbourhood, and the inlet and
    outlet columns of a channel, take plain BGK bit for bit; every other
    fluid cell is corrected."""
    nx, ny = 40, 24
    solid = W.cylinder_channel_solid(nx, ny, cx=14, cy=12, r=4)
    f0 = W.rest_state(nx, ny, solid, amplitude=2e-3, seed=9)
    p = O.phys_params(nx, ny, 0.6, x_edge=1, inlet_ux=0.04, force=(1e-5, 0.0))
    rng = np.random.default_rng(1)
    net = rng.standard_normal(NP) * 0.3
    out, learned = O.learned_step(p, solid, f0, net, 1e-3, 0.2)
    plain = O.PhysState(p, solid, f0)
    plain.run(1)
    expect = np.zeros((ny, nx), bool)
    for y in range(ny):
        for x in range(1, nx - 1):
            expect[y, x] = not any(solid[(y + dy) % ny, x + dx] for dy in (-1, 0, 1) for dx in (-1, 0, 1))
    assert np.array_equal(learned, expect)
    same = np.all(out == plain.f, axis=0)
    fluid = ~solid
    assert np.all(same[fluid & ~expect])
    assert not np.any(same[expect])
    assert np.all(out[:, solid] == 0)


@pytest.mark.parametrize("force", [(0.0, 0.0), (2e-4, 1e-4)], ids=["noforce", "force"])
def test_transition_targets_recover_a_constant_modulation(force):
    """s1 made from s0 by a learned step with constant modulations alpha:
    the residual of plain BGK is exactly the correction, so the targets are
    clamp(alpha_a n_a^2 / (n_a^2 + eps2)) / s_out with eps2 = 1e-2 (n_xx^2 +
    n_xy^2 + n_e^2 / 9), and the inputs are (f - feq) / (w s_in); candidates
    next to a solid are skipped."""
    nx, ny = 14, 11
    tau = 0.53
    b = O.modulation_bound(1 / tau)
    alpha = np.array([0.5 * b, -0.8 * b, 0.9 * b])
    s_in, s_out = 2e-3, 0.25
    rng = np.random.default_rng(8)
    solid = np.zeros((ny, nx), bool)
    solid[5, 6] = True
    n = nx * ny
    s0 = feq(1 + rng.uniform(-0.01, 0.01, n), rng.uniform(-0.05, 0.05, n), rng.uniform(-0.05, 0.05, n)) \
        + rng.uniform(-1e-3, 1e-3, (9, n))
    s0 = s0.reshape(9, ny, nx)
    # post-collision state whose gather is s0 (inverse streaming), one
    # learned step, then streaming: s1_i(x) = post_i(x - e_i)
    prev = np.stack([np.roll(np.roll(s0[i], -EY[i], 0), -EX[i], 1) for i in range(9)])
    p = O.phys_params(nx, ny, tau, force=force)
    post, _ = O.learned_step(p, solid, prev, const_net(alpha / s_out), s_in, s_out)
    s1 = np.stack([np.roll(np.roll(post[i], EY[i], 0), EX[i], 1) for i in range(9)])
    X, T = O.transition_samples(s0, s1, solid, tau, force, s_in, s_out, stride=2, offset=1)
    cells = [c for c in range(1, n, 2)
             if not any(solid[(c // nx + dy) % ny, (c % nx + dx) % nx] for dy in (-1, 0, 1) for dx in (-1, 0, 1))]
    assert X.shape == (len(cells), 9) and len(cells) < (n - 1 + 1) // 2
    f = s0.reshape(9, -1)[:, cells]
    rho = f.sum(axis=0)
    ux = (EX @ f + 0.5 * force[0]) / rho
    uy = (EY @ f + 0.5 * force[1]) / rho
    q = f - feq(rho, ux, uy)
    # q is a difference of O(0.1) values: ~1e-17 absolute, 1e-11 of x
    np.testing.assert_allclose(X, (q / (WT[:, None] * s_in)).T, rtol=1e-9, atol=1e-10)
    nm = [r @ q for r in ROWS]
    eps2 = 1e-2 * (nm[1] ** 2 + nm[2] ** 2 + nm[0] ** 2 / 9)
    for a in range(3):
        expect = np.clip(alpha[a] * nm[a] ** 2 / (nm[a] ** 2 + eps2), -b, b) / s_out
        np.testing.assert_allclose(T[:, a], expect, rtol=1e-9, atol=1e-13)
    assert np.all(T[:, 3:] == 0)


def test_restriction_of_uniform_shear_rescales_by_four_tau_ratio():
    nx, ny = 16, 12
    tau_f, tau_c = 0.52, 0.505
    rho0, u0 = 1.02, (0.03, -0.01)
    delta = 1e-3 * M_XY / 4 + 4e-4 * M_XX / 4 + 2e-4 * M_E / 36
    f = feq(rho0, *u0)[:, 0] + delta
    p = O.phys_params(nx, ny, tau_f)
    fc, sc, uxc, uyc = O.restrict4_state(p, np.zeros((ny, nx), bool), uniform_state(nx, ny, f), tau_c)
    assert not sc.any()
    np.testing.assert_allclose(uxc, u0[0], rtol=1e-13)
    np.testing.assert_allclose(uyc, u0[1], rtol=1e-13)
    got = fc.reshape(9, -1) - feq(rho0, *u0)
    np.testing.assert_allclose(got, np.repeat((4 * tau_c / tau_f * delta)[:, None], got.shape[1], 1),
                               rtol=1e-11, atol=1e-17)


def test_restriction_of_random_state_with_solids_and_force():
    """Coarse solid iff fewer than 8 of 16 fine cells are fluid; rho_c is
    the block mean of rho, u_c the mass-weighted block mean of the Guo
    velocity; the coarse Guo velocity (sum e fc + g_c/2)/rho_c equals u_c;
    the coarse non-equilibrium moments e, xx, xy are 4 tau_c/tau_f times the
    block means of the fine ones."""
    nx, ny = 32, 24
    tau_f, tau_c = 0.52, 0.505
    g, gc = (2e-5, -1e-5), (8e-5, -4e-5)
    scene = W.periodic_disks(nx, ny, 0.12, 2, 4, seed=6)
    solid = scene.solid.copy()
    solid[0:4, 8:12] = True  # one fully solid block
    solid[8, 20:24] = True   # blocks with a few solid cells
    post = W.rest_state(nx, ny, solid, amplitude=2e-3, seed=5)
    p = O.phys_params(nx, ny, tau_f, force=g)
    fc, sc, uxc, uyc = O.restrict4_state(p, solid, post, tau_c, gc)
    st = O.PhysState(p, solid, post)
    rho, ux, uy = st.macro()
    fg = np.zeros((9, ny, nx))
    for y in range(ny):
        for x in range(nx):
            if not solid[y, x]:
                fg[:, y, x] = st.gather(x, y)
    fl = ~solid
    blk = lambda a: a.reshape(a.shape[:-2] + (ny // 4, 4, nx // 4, 4)).sum(axis=(-3, -1))
    nfl = blk(fl.astype(float))
    assert np.array_equal(sc, nfl < 8) and sc.any() and (~sc).any()
    cf = ~sc
    m = blk(np.where(fl, rho, 0.0))
    np.testing.assert_allclose(fc.sum(axis=0)[cf], (m / np.maximum(nfl, 1))[cf], rtol=1e-13)
    np.testing.assert_allclose(uxc[cf], (blk(np.where(fl, rho * ux, 0.0)) / np.where(m > 0, m, 1))[cf], rtol=1e-12)
    np.testing.assert_allclose(uyc[cf], (blk(np.where(fl, rho * uy, 0.0)) / np.where(m > 0, m, 1))[cf], rtol=1e-12)
    rc = fc.sum(axis=0)
    np.testing.assert_allclose(((np.tensordot(EX, fc, 1) + 0.5 * gc[0]) / np.where(cf, rc, 1))[cf], uxc[cf],
                               rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(((np.tensordot(EY, fc, 1) + 0.5 * gc[1]) / np.where(cf, rc, 1))[cf], uyc[cf],
                               rtol=1e-11, atol=1e-15)
    qf = (fg.reshape(9, -1) - feq(np.where(fl, rho, 1).ravel(), ux.ravel(), uy.ravel())).reshape(9, ny, nx)
    qf[:, solid] = 0
    qc = fc.reshape(9, -1) - feq(np.where(cf, rc, 1).ravel(), uxc.ravel(), uyc.ravel())
    for r in ROWS:
        fine_mean = blk(np.tensordot(r, qf, 1)) / np.maximum(nfl, 1)
        np.testing.assert_allclose((r @ qc).reshape(sc.shape)[cf], (4 * tau_c / tau_f * fine_mean)[cf],
                                   rtol=1e-9, atol=1e-16)
    assert np.all(fc[:, sc] == 0) and np.all(uxc[sc] == 0)
