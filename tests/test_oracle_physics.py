"""Pins of the oracle's spectral, diagnostic and obstacle functions: Philox known-answer
vectors, isotropic-turbulence IC invariants, forcing (Parseval, mode count, injected
energy), dissipation integrals, the Towers chi (planar closed form, sphere volume),
penalization per-cell closed form and surface forces. CPU only."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cell_centres(shape, h):
    Z, Y, X = shape
    z, y, x = np.meshgrid((np.arange(Z) + 0.5) * h, (np.arange(Y) + 0.5) * h, (np.arange(X) + 0.5) * h,
                          indexing="ij")
    return x, y, z


def test_philox_known_answers(oracle):
    kat = np.load(os.path.join(GOLD, "philox_kat.npz"))["kat"]
    for row in kat:
        row = [int(v) for v in row]
        assert oracle.philox4x32_10(row[0:4], row[4:6]) == row[6:10]


# ---------------------------------------------------------- isotropic turbulence IC

def test_hit_ic_invariants(oracle):
    s = oracle.OracleSim((4, 4, 4))
    s.ic_hit(k0=4.0, u_rms=1.0, seed=20260101)
    u, v, w = s.velocity()
    N = u.shape[0]
    # u_rms = 1 per component on average, zero mean
    assert np.sqrt((u ** 2 + v ** 2 + w ** 2).mean() / 3) == pytest.approx(1.0, rel=1e-12)
    for f in (u, v, w):
        assert abs(f.mean()) < 1e-12
    # spectrally solenoidal: k . u_hat = 0 for every mode (numpy FFT, index wavenumbers)
    U, V, W = (np.fft.fftn(f) for f in (u, v, w))
    k = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    div = kx * U + ky * V + kz * W
    assert np.abs(div).max() < 1e-9 * np.abs(U).max()
    # no Nyquist content
    assert np.abs(U[N // 2, :, :]).max() < 1e-9 and np.abs(U[:, :, N // 2]).max() < 1e-9
    # shell energy follows E(k) ~ k^4 exp(-2 (k/k0)^2) in shape: peak near k0 = 4
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    E = 0.5 * (np.abs(U) ** 2 + np.abs(V) ** 2 + np.abs(W) ** 2)
    shells = np.array([E[(kk >= n - 0.5) & (kk < n + 0.5)].sum() for n in range(1, 12)])
    assert 2 <= int(np.argmax(shells)) + 1 <= 5
    # reproducible from the seed, different for another seed
    s2 = oracle.OracleSim((4, 4, 4))
    s2.ic_hit(k0=4.0, u_rms=1.0, seed=20260101)
    assert np.array_equal(s2.get("u"), u)
    s2.ic_hit(k0=4.0, u_rms=1.0, seed=7)
    assert not np.allclose(s2.get("u"), u)


# ------------------------------------------------------------------- forcing

def test_forcing_stats_parseval_and_injection(oracle):
    s = oracle.OracleSim((4, 4, 4))
    s.ic_hit(k0=4.0, u_rms=1.0, seed=3)
    u, v, w = s.velocity()
    h, nu, dt = s.h, 0.005, 1e-3
    tke_phys = 0.5 * (u ** 2 + v ** 2 + w ** 2).mean()
    # low-mode energy from numpy: 0 < |k|^2 <= 4 (integer k in a 2pi box), 32 modes
    N = u.shape[0]
    k = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    low = (k2 > 0) & (k2 <= 4)
    assert low.sum() == 32
    F = [np.fft.fftn(f) for f in (u, v, w)]
    tke_low = 0.5 * sum((np.abs(X[low]) ** 2).sum() for X in F) / N ** 6
    # enstrophy-based dissipation with central-difference symbols 2 sin(2 pi k/N) / (2h)
    def dc(f, ax):
        return (np.roll(f, -1, ax) - np.roll(f, 1, ax)) / (2 * h)
    ox = dc(w, 1) - dc(v, 0); oy = dc(u, 0) - dc(w, 2); oz = dc(v, 2) - dc(u, 1)
    eps = nu * (ox ** 2 + oy ** 2 + oz ** 2).mean()
    target = 1.1 * tke_phys
    st = s.spectral_forcing(dt, nu, tke_target=target)
    assert st["tke"] == pytest.approx(tke_phys, rel=1e-12)
    assert st["tke_filtered"] == pytest.approx(tke_low, rel=1e-12)
    assert st["eps"] == pytest.approx(eps, rel=1e-12)
    fac = dt * (target - tke_phys) / dt / (2 * tke_low)
    assert st["fac"] == pytest.approx(fac, rel=1e-12)
    # u += fac * P_low u: the low-mode energy grows by ((1+fac)^2 - 1) tke_low exactly,
    # the rest is untouched
    u2, v2, w2 = s.velocity()
    F2 = [np.fft.fftn(f) for f in (u2, v2, w2)]
    tke_low2 = 0.5 * sum((np.abs(X[low]) ** 2).sum() for X in F2) / N ** 6
    assert tke_low2 == pytest.approx((1 + fac) ** 2 * tke_low, rel=1e-11)
    for X, Y in zip(F, F2):
        assert np.abs((Y - X)[~low]).max() < 1e-8 * np.abs(X).max()


def test_forcing_constant_injection_rate(oracle):
    """Without a target the injection is the constant rate r (SpectralForcing.cpp:41-46):
    fac = dt r / (2 tke_low) (:71), and the low modes gain ((1 + fac)^2 - 1) tke_low."""
    s = oracle.OracleSim((4, 4, 4))
    s.ic_hit(k0=4.0, u_rms=1.0, seed=9)
    u, v, w = s.velocity()
    N = u.shape[0]
    k = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    low = (k2 > 0) & (k2 <= 4)
    tke_low = 0.5 * sum((np.abs(np.fft.fftn(f)[low]) ** 2).sum() for f in (u, v, w)) / N ** 6
    dt, r = 2e-3, 0.3
    st = s.spectral_forcing(dt, 0.005, injection_rate=r)
    assert st["fac"] == pytest.approx(dt * r / (2 * tke_low), rel=1e-12)
    u2, v2, w2 = s.velocity()
    tke_low2 = 0.5 * sum((np.abs(np.fft.fftn(f)[low]) ** 2).sum() for f in (u2, v2, w2)) / N ** 6
    assert tke_low2 == pytest.approx((1 + st["fac"]) ** 2 * tke_low, rel=1e-11)


def test_forcing_negative_injection_is_a_no_op(oracle):
    s = oracle.OracleSim((2, 2, 2))
    s.ic_hit(k0=2.0, u_rms=1.0, seed=5)
    u = s.get("u")
    st = s.spectral_forcing(1e-3, 0.01, tke_target=1e-3)
    assert st["fac"] < 0
    assert np.array_equal(s.get("u"), u)


# ----------------------------------------------------------------- diagnostics

def test_dissipation_integrals(oracle):
    """KE of the reference TGV (uMax=1, 2pi box) is (2pi)^3/8 up to midpoint-rule exactness;
    linear momentum of a constant field is U * volume; circulation of any periodic field
    is zero."""
    s = oracle.OracleSim((4, 4, 4))
    s.ic_taylor_green(1.0)
    R = s.dissipation(0.01)
    vol = (2 * np.pi) ** 3
    assert R[18] == pytest.approx(vol / 8, rel=1e-12)
    assert abs(R[0]) < 1e-12 and abs(R[1]) < 1e-12 and abs(R[2]) < 1e-12
    # viscous power of TGV: nu (u.lap u + 2 D:D) integrates to -nu |grad u|^2 + 2 nu |D|^2
    c = oracle.OracleSim((2, 2, 2))
    z = np.zeros(c.shape)
    c.set_velocity(z + 0.5, z - 1.0, z + 2.0)
    R = c.dissipation(0.01)
    assert R[6:9] == pytest.approx(np.array([0.5, -1.0, 2.0]) * vol, rel=1e-13)
    assert R[18] == pytest.approx(0.5 * (0.25 + 1 + 4) * vol, rel=1e-13)
    assert abs(R[16]) < 1e-13 and abs(R[19]) < 1e-12


# ----------------------------------------------------------------------- sphere

def test_towers_chi_planar_closed_form(oracle):
    """A huge sphere is locally a plane: with the SDF d linear in x across the +-2h band,
    Towers' central-difference chi (max(d+h,0) - max(d-h,0))/(2h) is the Heaviside smoothed
    over two cells: clip(1/2 + d/(2h), 0, 1)."""
    s = oracle.OracleSim((4, 2, 2), maxextent=2.0)
    h = s.h
    R = 100.0
    c = np.array([1.0 - R + 0.3 * h, 0.125, 0.125])   # surface at x = 1 + 0.3h
    s.sphere_create(R, c)
    chi = s.get("chi")
    x, y, z = cell_centres(s.shape, h)
    d = R - np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    # interior of the band, away from the box faces: compare along the centre row
    row = chi[4, 4, :]
    dr = d[4, 4, :]
    exp = np.clip(0.5 + dr / (2 * h), 0.0, 1.0)
    band = np.abs(dr) < 1.9 * h
    assert band.sum() >= 3
    np.testing.assert_allclose(row[band], exp[band], atol=1e-4)


def test_sphere_volume(oracle):
    """sum chi h^3 vs 4/3 pi R^3 (the reference's validation mass column: -5.6e-6 relative
    at R/h = 25.6)."""
    s = oracle.OracleSim((8, 8, 8), maxextent=1.0)
    R = 0.2
    s.sphere_create(R, (0.5, 0.5, 0.5))
    info = s.sphere_info()
    assert info["mass"] == pytest.approx(4 / 3 * np.pi * R ** 3, rel=2e-4)
    np.testing.assert_allclose(info["com"], [0.5, 0.5, 0.5], atol=1e-12)
    assert info["npoints"] > 0


def test_hemisphere_volume_centroid_and_flat_face(oracle):
    """The hemisphere (HemiSphereObstacle, Sphere.cpp:53-84): sum chi h^3 converges to
    2/3 pi R^3 and the centroid to 3R/8 behind the flat face at second order (the rim, where
    the SDF min(-dx, R - r) has a kink, costs O(h^2): -3.0e-3 and -6.7e-4 of the volume at
    R/h = 12.8 and 25.6); along the axis through the centre the flat face's chi is Towers'
    planar closed form clip(1/2 - (x - c_x)/(2h))."""
    R = 0.2
    errs = []
    for n in (8, 16):
        s = oracle.OracleSim((n, n, n), maxextent=1.0)
        c = np.array([0.5 + 0.3 / (8 * n), 0.5, 0.5])
        s.sphere_create(R, c, hemi=True)
        info = s.sphere_info()
        errs.append((info["mass"] / (2 / 3 * np.pi * R ** 3) - 1, info["com"][0] - (c[0] - 3 * R / 8)))
        np.testing.assert_allclose(info["com"][1:], [0.5, 0.5], atol=1e-12)
    assert abs(errs[1][0]) < 1e-3 and abs(errs[0][0]) > 3 * abs(errs[1][0])
    assert abs(errs[1][1]) < 2e-5 and abs(errs[0][1]) > 3 * abs(errs[1][1])
    s = oracle.OracleSim((8, 8, 8), maxextent=1.0)
    c = np.array([0.5 + 0.3 / 64, 0.5, 0.5])
    s.sphere_create(R, c, hemi=True)
    h = s.h
    chi = s.get("chi")
    x, _, _ = cell_centres(s.shape, h)
    row, xr = chi[32, 32, :], x[32, 32, :]   # y, z at 32.5 h: near the axis
    dx = xr - c[0]
    band = np.abs(dx) < 1.9 * h
    assert band.sum() >= 3
    np.testing.assert_allclose(row[band], np.clip(0.5 - dx[band] / (2 * h), 0.0, 1.0), atol=1e-12)
    assert (row[dx > 2 * h] == 0).all() and (row[(dx < -2 * h) & (dx > -R + 3 * h)] == 1).all()


def test_penalization_closed_form(oracle):
    s = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    s.sphere_create(0.2, (0.5, 0.5, 0.5))
    rng = np.random.default_rng(4)
    u0, v0, w0 = (rng.standard_normal(s.shape) for _ in range(3))
    s.set_velocity(u0, v0, w0)
    chi = s.get("chi")
    dt, lam, V = 1e-3, 1e6, np.array([-1.0, 0.0, 0.0])
    ft = s.penalize(dt, lam, True, V)
    pf = chi * lam / (1 + lam * dt * chi)
    exp_u = u0 + dt * pf * (V[0] - u0)
    u, v, w = s.velocity()
    mask = chi > 0
    np.testing.assert_allclose(u[mask], exp_u[mask], rtol=1e-12, atol=1e-14)
    assert np.array_equal(u[~mask], u0[~mask])
    h3 = s.h ** 3
    assert ft[0] == pytest.approx((h3 * pf * (V[0] - u0))[mask].sum(), rel=1e-11)
    assert ft[1] == pytest.approx((h3 * pf * (0 - v0))[mask].sum(), rel=1e-11)


def test_surface_force_of_uniform_pressure_vanishes(oracle):
    """-p n summed over a closed surface is zero for constant p (discrete: the surface
    weights are a discrete gradient of chi, which sums to ~0); with zero velocity the
    viscous part vanishes."""
    s = oracle.OracleSim((4, 4, 4), maxextent=1.0)
    s.sphere_create(0.2, (0.5, 0.5, 0.5))
    z = np.zeros(s.shape)
    s.set_velocity(z, z, z)
    s.set("p", z + 3.0)
    q = s.surface_forces(0.01, (-1.0, 0, 0))
    area = 4 * np.pi * 0.2 ** 2
    assert abs(q[0]) < 1e-10 * 3 * area and abs(q[1]) < 1e-10 * 3 * area
    assert np.all(q[6:9] == 0)


def test_spectrum_interp_pins(oracle):
    """EnergySpectrum::interpE (SpectralManip.cpp:96-116): exact on a linear table inside and
    beyond it, the reference's form below the first node (E0 - slope (K1 - k)), 0 for
    negative values."""
    K = np.array([1.0, 2.0, 4.0, 7.0])
    E = 3.0 + 2.0 * K
    for k in (1.0, 1.3, 2.0, 3.7, 6.99, 7.0, 9.5):
        assert oracle.spectrum_interp(K, E, k) == pytest.approx(3.0 + 2.0 * k, rel=1e-14)
    assert oracle.spectrum_interp(K, E, 0.5) == pytest.approx(E[0] - 2.0 * (K[1] - 0.5), rel=1e-14)
    assert oracle.spectrum_interp(K, 5.0 - K, 6.0) == 0.0  # 5 - 6 < 0


def hit_table_energy(shape, h, K, E):
    """Closed form of the table IC's energy: every mode with k_xy != 0 (kx = ky = 0 modes carry
    none) and no Nyquist index holds |u_k|^2 = amp^2 = E(|k|) / (2 pi (|k|/wf)^2), and the
    unnormalized inverse transform gives tke = 1/2 sum_k amp^2 (Parseval)."""
    Z, Y, X = shape
    wf = 2 * np.pi / (X * h)
    kx = np.fft.fftfreq(X, 1.0 / X)
    ky = np.fft.fftfreq(Y, 1.0 / Y)
    kz = np.fft.fftfreq(Z, 1.0 / Z)
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    ok = (2 * np.abs(KX) != X) & (2 * np.abs(KY) != Y) & (2 * np.abs(KZ) != Z) & ((KX != 0) | (KY != 0))
    wfx, wfy, wfz = 2 * np.pi / (X * h), 2 * np.pi / (Y * h), 2 * np.pi / (Z * h)
    kn = np.sqrt((KX * wfx) ** 2 + (KY * wfy) ** 2 + (KZ * wfz) ** 2)[ok]
    Ek = np.interp(kn, K, E)  # the tables here span every |k| of the grid
    return 0.5 * np.sum(Ek / (2 * np.pi * (kn / wf) ** 2))


def test_table_ic_energy_and_divergence(oracle):
    """The IC from a target table keeps the table's energy (closed form above) and is
    spectrally solenoidal (central divergence small against |grad u|)."""
    s = oracle.OracleSim((4, 4, 4))
    K = np.linspace(0.5, 30.0, 40)
    E = K ** 4 * np.exp(-2 * (K / 4.0) ** 2) / 100.0
    s.ic_hit_spectrum(K, E, u_rms=0.0, seed=7)
    u, v, w = s.velocity()
    tke = 0.5 * np.mean(u ** 2 + v ** 2 + w ** 2)
    assert tke == pytest.approx(hit_table_energy(s.shape, s.h, K, E), rel=1e-12)
    s.ic_hit_spectrum(K, E, u_rms=0.7, seed=7)
    u2, v2, w2 = s.velocity()
    assert np.sqrt(np.mean(u2 ** 2 + v2 ** 2 + w2 ** 2) / 3) == pytest.approx(0.7, rel=1e-12)
    np.testing.assert_allclose(u2 / u, 0.7 / np.sqrt(2 * tke / 3), rtol=1e-10)
