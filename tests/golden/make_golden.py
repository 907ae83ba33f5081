"""Writes the golden vectors of tests/golden/*.npz with numpy (a library FFT routine and
plain numpy algebra), independent of the oracle and of the device code:

  fft_golden.npz      random complex 8x16x4 and real 16^3 arrays and numpy.fft.fftn of them
  poisson_golden.npz  a random periodic RHS on 16^3 (h = 2pi/16) and the solution of the
                      grid-consistent periodic Poisson problem of
                      PoissonSolverPeriodic.cpp:15-61, computed with numpy.fft
  fft_mixed_golden.npz  lengths that are not powers of two: random complex lines of 12, 24,
                      48 and 96 points and a random real 24 x 40 x 12 array with their
                      numpy.fft transforms, and the Poisson solution of a random RHS on
                      24 x 16 x 40 cells
  philox_kat.npz      Philox-4x32-10 known-answer vectors (Random123 kat_vectors)

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def poisson_numpy(rhs, h):
    """p_hat = r_hat * 4/(N0 h N1 N2) / (2D - 6), D = sum cos(4 pi k/N), 8 modes zeroed,
    with unnormalized forward and backward transforms (FFTW's r2c/c2r convention)."""
    N0, N1, N2 = rhs.shape
    R = np.fft.fftn(rhs)
    i = np.arange(N0)[:, None, None]
    l = np.arange(N1)[None, :, None]
    k = np.arange(N2)[None, None, :]
    D = np.cos(4 * np.pi * i / N0) + np.cos(4 * np.pi * l / N1) + np.cos(4 * np.pi * k / N2)
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = 4.0 / (N0 * h * N1 * N2) / (2 * D - 6)
    P = R * fac
    for a in (0, N0 // 2):
        for b in (0, N1 // 2):
            for c in (0, N2 // 2):
                P[a, b, c] = 0
    return np.real(np.fft.ifftn(P) * (N0 * N1 * N2))


def main():
    rng = np.random.default_rng(12345)
    a = rng.standard_normal((8, 16, 4)) + 1j * rng.standard_normal((8, 16, 4))
    r = rng.standard_normal((16, 16, 16))
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), a=a, A=np.fft.fftn(a), r=r, R=np.fft.fftn(r))
    rhs = rng.standard_normal((16, 16, 16))
    h = 2 * np.pi / 16
    np.savez_compressed(os.path.join(HERE, "poisson_golden.npz"), rhs=rhs, h=h, p=poisson_numpy(rhs, h))
    mixed = {}
    for n in (12, 24, 48, 96):
        line = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        mixed[f"line{n}"] = line
        mixed[f"LINE{n}"] = np.fft.fft(line)
    m = rng.standard_normal((24, 40, 12))
    mixed["m"] = m
    mixed["M"] = np.fft.fftn(m)
    rhs2 = rng.standard_normal((24, 16, 40))
    h2 = 2 * np.pi / 40
    mixed["rhs"] = rhs2
    mixed["h"] = h2
    mixed["p"] = poisson_numpy(rhs2, h2)
    np.savez_compressed(os.path.join(HERE, "fft_mixed_golden.npz"), **mixed)
    kat = np.array([
        [0, 0, 0, 0, 0, 0, 0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8],
        [0xffffffff] * 6 + [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd],
        [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0,
         0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1],
    ], dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "philox_kat.npz"), kat=kat)


if __name__ == "__main__":
    main()
