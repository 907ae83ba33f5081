"""The oracle's HIT analysis (Analysis.cpp:159-196, SpectralManipFFTW.cpp:110-207) pinned
by quantities computed independently in physical space or by hand: Parseval for tke,
nu <|omega_c|^2> for eps (the central-difference vorticity, whose spectral symbols are the
2 sin(2 pi k/N) of the reference), single-mode shells, the band-limited spectrum summing to
tke, the integral length of a single mode, and the gradient statistics."""
import numpy as np
import pytest


def field(oracle, bpd, fn):
    s = oracle.OracleSim(bpd)
    x = [(np.arange(n) + 0.5) * s.h for n in (bpd[0] * 8, bpd[1] * 8, bpd[2] * 8)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")  # [z][y][x]
    u, v, w = fn(X, Y, Z)
    s.set_velocity(u, v, w)
    return s, (u, v, w)


def test_parseval_and_vorticity_dissipation(oracle):
    rng = np.random.default_rng(3)
    s = oracle.OracleSim((2, 2, 2))
    u, v, w = (rng.standard_normal(s.shape) for _ in range(3))
    s.set_velocity(u, v, w)
    nu = 0.01
    r = s.hit_analysis(nu)
    assert r["tke"] == pytest.approx(0.5 * np.mean(u ** 2 + v ** 2 + w ** 2), rel=1e-12)
    h = s.h
    d = lambda f, ax: (np.roll(f, -1, ax) - np.roll(f, 1, ax)) / (2 * h)  # axes: 0 z, 1 y, 2 x
    ox, oy, oz = d(w, 1) - d(v, 0), d(u, 0) - d(w, 2), d(v, 2) - d(u, 1)
    assert r["eps"] == pytest.approx(nu * np.mean(ox ** 2 + oy ** 2 + oz ** 2), rel=1e-12)
    up = np.sqrt(2 * r["tke"] / 3)
    assert r["uprime"] == pytest.approx(up, rel=1e-14)
    assert r["lambda"] == pytest.approx(np.sqrt(15 * nu / r["eps"]) * up, rel=1e-14)
    assert r["Re_lambda"] == pytest.approx(up * r["lambda"] / nu, rel=1e-14)


@pytest.mark.parametrize("kvec,bin_", [((1, 0, 0), 0), ((2, 0, 0), 1), ((1, 1, 1), 1), ((3, 4, 0), 4),
                                       ((7, 2, 0), 6), ((8, 0, 0), None)])
def test_single_mode_shell(oracle, kvec, bin_):
    """u = cos(k.x) with integer k on 16^3 (nyquist 8, 7 shells): the shell is
    floor(|k| * 7/8); |k| >= 8 lies outside the histogram (there u = sin(8x), the Nyquist
    mode on the cell centres, of energy 1/2). tke = 1/4 and l_integral = 1/|k|."""
    kx, ky, kz = kvec
    f = np.sin if bin_ is None else np.cos
    s, _ = field(oracle, (2, 2, 2), lambda X, Y, Z: (f(kx * X + ky * Y + kz * Z), 0 * X, 0 * X))
    r = s.hit_analysis(0.01)
    E = r["E"]
    assert len(E) == 7
    assert r["tke"] == pytest.approx(0.5 if bin_ is None else 0.25, rel=1e-12)
    assert r["l_integral"] == pytest.approx(1 / np.sqrt(kx * kx + ky * ky + kz * kz), rel=1e-12)
    if bin_ is None:
        assert np.abs(E).max() < 1e-25
    else:
        assert E[bin_] == pytest.approx(0.25, rel=1e-12)
        assert np.abs(np.delete(E, bin_)).max() < 1e-25


def test_band_limited_spectrum_sums_to_tke(oracle):
    """A random field of modes with |k| < 8 only: the shells hold all of tke."""
    rng = np.random.default_rng(4)
    s = oracle.OracleSim((2, 2, 2))
    k = np.fft.fftfreq(16, 1 / 16)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    mask = KX ** 2 + KY ** 2 + KZ ** 2 < 64
    comp = []
    for _ in range(3):
        sp = np.fft.fftn(rng.standard_normal(s.shape)) * mask
        comp.append(np.real(np.fft.ifftn(sp)))
    s.set_velocity(*comp)
    r = s.hit_analysis(0.01)
    assert r["E"].sum() == pytest.approx(r["tke"], rel=1e-12)


def test_gradient_statistics(oracle):
    rng = np.random.default_rng(5)
    s = oracle.OracleSim((2, 2, 2))
    u, v, w = (rng.standard_normal(s.shape) for _ in range(3))
    s.set_velocity(u, v, w)
    r = s.hit_analysis(0.01)
    g2 = 0
    for f in (u, v, w):
        for ax in range(3):
            g2 = g2 + (np.roll(f, -1, ax) - np.roll(f, 1, ax)) ** 2
    g2 = g2 / (4 * s.h ** 2)
    m = np.mean(np.sqrt(g2))
    assert r["grad_mean"] == pytest.approx(m, rel=1e-12)
    assert r["grad_std"] == pytest.approx(np.sqrt(np.mean(g2) - m * m), rel=1e-10)
