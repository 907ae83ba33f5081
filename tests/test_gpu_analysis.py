"""The HIT analysis (cup_hit_analysis: gradient statistics and the spectral analysis of
SpectralManipFFTW::_compute_analysis) against the oracle's, FP64 to 1e-11, on random fields
(grid shapes that exercise the packed kx = 0 / Nx/2 slot and non-cubic boxes), on the seeded
turbulence IC at 128^3, and the FP32 form against FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("tke", "eps", "uprime", "lambda", "Re_lambda", "l_integral", "tau_integral", "grad_mean", "grad_std")


@pytest.mark.parametrize("bpd", [(2, 2, 2), (4, 2, 2), (2, 4, 1), (8, 8, 8)])
def test_analysis_matches_oracle(cup, oracle, bpd):
    print(f"hit analysis bpd={bpd}", flush=True)
    rng = np.random.default_rng(11)
    G = cup.Grid(bpd, precision=8)
    O = oracle.OracleSim(bpd)
    u = tuple(rng.standard_normal(G.shape) for _ in range(3))
    G.set_velocity(*u)
    O.set_velocity(*u)
    rg, ro = G.hit_analysis(0.01), O.hit_analysis(0.01)
    for k in KEYS:
        assert rg[k] == pytest.approx(ro[k], rel=1e-11), k
    np.testing.assert_allclose(rg["E"], ro["E"], rtol=1e-11, atol=1e-14 * ro["E"].max())


def test_analysis_hit_ic_128(cup, oracle):
    """The seeded art-spectrum IC at 128^3: energy in the shells around k0 = 4."""
    bpd = (16, 16, 16)
    G = cup.Grid(bpd, precision=8)
    O = oracle.OracleSim(bpd)
    G.ic_hit(4.0, 1.0, 20260101)
    O.set_velocity(*G.velocity())
    rg, ro = G.hit_analysis(0.005), O.hit_analysis(0.005)
    for k in KEYS:
        assert rg[k] == pytest.approx(ro[k], rel=1e-11), k
    np.testing.assert_allclose(rg["E"], ro["E"], rtol=1e-10, atol=1e-13 * ro["E"].max())
    assert rg["tke"] == pytest.approx(1.5, rel=1e-12)  # u_rms = 1 per component
    assert 1 <= int(np.argmax(rg["E"])) <= 5
    print(f"128^3 IC: Re_lambda {rg['Re_lambda']:.2f}, lambda {rg['lambda']:.4f}, l_integral "
          f"{rg['l_integral']:.4f}, spectrum peak at shell {int(np.argmax(rg['E']))}", flush=True)


def test_analysis_fp32_vs_fp64(cup):
    bpd = (8, 8, 8)
    G8 = cup.Grid(bpd, precision=8)
    G8.ic_hit(4.0, 1.0, 20260101)
    G4 = cup.Grid(bpd, precision=4)
    G4.set_velocity(*G8.velocity())
    r8, r4 = G8.hit_analysis(0.005), G4.hit_analysis(0.005)
    for k in KEYS:
        assert r4[k] == pytest.approx(r8[k], rel=1e-5), k
    big = r8["E"] > 1e-6 * r8["E"].max()
    np.testing.assert_allclose(r4["E"][big], r8["E"][big], rtol=1e-4)
