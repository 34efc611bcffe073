"""Pins the binary64 D3Q19 oracle (oracle/d3q19.py) by what the mathematics
fixes: the lattice's moment identities, the rest state, exact mass and
momentum bookkeeping (with bounce-back and with the Guo force), the decay
rate of a shear wave, and the exact Poiseuille profile between half-way
bounce-back walls at tau = 1/2 + sqrt(3/16)."""
import numpy as np

from oracle import d3q19 as D


def test_lattice_tables():
    E, W, INV = D.E, D.W, D.INV
    assert len({tuple(e) for e in E}) == 19
    assert sorted(int((e * e).sum()) for e in E) == [0] + [1] * 6 + [2] * 12
    assert abs(W.sum() - 1.0) < 1e-15
    assert np.allclose((W[:, None] * E).sum(0), 0.0, atol=1e-16)
    second = np.einsum("i,ia,ib->ab", W, E, E)
    assert np.allclose(second, np.eye(3) / 3.0, atol=1e-15)
    fourth = np.einsum("i,ia,ib,ic,id->abcd", W, E, E, E, E)
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9.0
    assert np.allclose(fourth, iso, atol=1e-15)
    assert np.array_equal(E[INV], -E)
    assert np.array_equal(INV[INV], np.arange(19))


def spheres(nx, ny, nz, count, rmax, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    solid = np.zeros((nz, ny, nx), bool)
    for _ in range(count):
        r = int(rng.integers(1, rmax + 1))
        c = rng.integers(0, (nx, ny, nz))
        solid |= (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r
    return solid


def test_rest_state_is_a_fixed_point():
    nx, ny, nz = 10, 8, 6
    solid = spheres(nx, ny, nz, 3, 2, seed=1)
    st = D.D3Q19State(0.7, solid, D.rest_state(nx, ny, nz))
    st.run(10)
    fl = ~solid
    assert np.max(np.abs(st.f[:, fl] - D.W[:, None])) < 1e-15


def test_mass_and_momentum_bookkeeping():
    nx, ny, nz = 9, 7, 8
    rng = np.random.default_rng(3)
    f0 = D.rest_state(nx, ny, nz) * (1.0 + 0.01 * rng.standard_normal((19, nz, ny, nx)))
    # periodic, no solids, no force: mass and momentum exact
    st = D.D3Q19State(0.62, np.zeros((nz, ny, nx), bool), f0)
    r0 = st.reduce()
    st.run(20)
    r1 = st.reduce()
    assert abs(r1[0] - r0[0]) < 1e-12 * r0[0]
    assert np.max(np.abs(r1[1:4] - r0[1:4])) < 1e-12
    # with a force: the first moment grows by g per cell per step
    g = np.array([1e-4, -2e-4, 3e-4])
    st = D.D3Q19State(0.62, np.zeros((nz, ny, nx), bool), f0, force=g)
    r0 = st.reduce()
    st.run(5)
    r1 = st.reduce()
    assert np.max(np.abs((r1[1:4] - r0[1:4]) - 5 * g * nx * ny * nz)) < 1e-12
    # bounce-back conserves mass; momentum changes by the force on the solids
    solid = spheres(nx, ny, nz, 4, 2, seed=5)
    st = D.D3Q19State(0.62, solid, f0, force=g)
    r0 = st.reduce()
    st.step()
    r1 = st.reduce()
    nfl = int((~solid).sum())
    assert abs(r1[0] - r0[0]) < 1e-12 * r0[0]
    # p(t+1) - p(t) = g N_fluid - (momentum handed to the solids by step t+1's bounce)
    assert np.max(np.abs((r1[1:4] - r0[1:4]) - (g * nfl - r1[4:7]))) < 1e-12


def test_shear_wave_decay_rate():
    """u_x = U sin(2 pi y / ny) decays as exp(-nu k^2 t), nu = (tau - 1/2)/3."""
    nx, ny, nz, tau, U = 2, 64, 2, 0.8, 1e-4
    k = 2 * np.pi / ny
    y = np.arange(ny)
    u = np.zeros((3, nz, ny, nx))
    u[0] = U * np.sin(k * y)[None, :, None]
    f0 = D.feq(np.ones((nz, ny, nx)), u)
    st = D.D3Q19State(tau, np.zeros((nz, ny, nx), bool), f0)
    amp = []
    for t in range(301):
        _, uu = st.macro()
        amp.append(float((uu[0, 0, :, 0] * np.sin(k * y)).sum() * 2 / ny))
        st.step()
    nu = (tau - 0.5) / 3.0
    rate = -np.log(amp[300] / amp[100]) / 200.0
    assert abs(rate / (nu * k * k) - 1.0) < 2e-3
    # a wrong viscosity (tau instead of tau - 1/2) would be 2.7x off
    assert abs(rate / (tau / 3 * k * k) - 1.0) > 0.5


def test_poiseuille_exact_between_half_way_walls():
    """Body force g_x between solid planes z = 0 and z = nz - 1: the walls sit
    half-way (z = 1/2 and z = nz - 3/2) and at tau = 1/2 + sqrt(3/16) the
    steady BGK solution is the parabola g/(2 nu) (z - 1/2)(nz - 3/2 - z)."""
    nx, ny, nz = 2, 2, 12
    tau = 0.5 + np.sqrt(3.0 / 16.0)
    nu = (tau - 0.5) / 3.0
    gx = 1e-5
    solid = np.zeros((nz, ny, nx), bool)
    solid[0] = solid[nz - 1] = True
    st = D.D3Q19State(tau, solid, D.rest_state(nx, ny, nz), force=(gx, 0.0, 0.0))
    st.run(3000)
    _, u = st.macro()
    z = np.arange(1, nz - 1)
    exact = gx / (2 * nu) * (z - 0.5) * (nz - 1.5 - z)
    got = u[0, 1:nz - 1, 0, 0]
    assert np.max(np.abs(got - exact)) < 1e-9 * exact.max()
    assert np.max(np.abs(u[1:, 1:nz - 1])) < 1e-9 * exact.max()
