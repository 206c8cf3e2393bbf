"""Pins of the offshore wave-boundary generator (xbeach_gpu_amd/csrc/
xbg_bnd.cpp: makjonswap / GenWGnLBW / GenCstWave of makjonswap.cpp).

The reference's generator needs FFTW and netCDF headers and does not build
here, and its tests hold no vectors for it, so the checks are the closed
forms the construction guarantees (spectral moments, normalisations, the
JONSWAP shape of Hasselmann et al. 1973), the textbook transforms (numpy's
inverse DFT, scipy's Hilbert transform), the C++ standard library's
minstd_rand0 + generate_canonical recipe that the reference's random draws
use, and statistics of the generated records."""
import math

import numpy as np
import pytest
import scipy.signal

from xbeach_gpu_amd import bnd_jonswap as bj
from xbeach_gpu_amd import default_params

RHO, G = 1025.0, 9.81


def moments(sp):
    df = sp.freq[1] - sp.freq[0]
    dd = sp.dir[1] - sp.dir[0]
    Sf = sp.S.sum(axis=0) * dd
    Sd = sp.S.sum(axis=1) * df
    return Sf, Sd, df, dd


@pytest.mark.parametrize("Hs,Tp,Dp,s,gam", [(1.5, 10.0, 270.0, 10.0, 3.3), (3.0, 12.0, 260.0, 20.0, 3.3),
                                            (0.8, 6.0, 300.0, 35.0, 1.0), (2.0, 1.0, 270.0, 2.0, 7.0)])
def test_jonswap_moments_and_shape(Hs, Tp, Dp, s, gam):
    sp = bj.jonswap(Hs, Tp, Dp, s, gam)
    Tpe = max(Tp, 1.5)                         # makjonswap.cpp:79 clamp
    fp = 1 / Tpe
    assert sp.freq[0] == pytest.approx(fp / 20, rel=1e-6)
    assert len(sp.freq) == math.ceil((3 * fp - fp / 20) / (fp / 20)) and len(sp.dir) == 90
    Sf, Sd, df, dd = moments(sp)
    # 4 sqrt(m0) = Hs and the directional distribution integrates to one
    assert 4 * math.sqrt(Sf.sum() * df) == pytest.approx(Hs, rel=1e-12)
    assert (Sd * dd).sum() / (Sf.sum() * df) == pytest.approx(1.0, rel=1e-12)
    # JONSWAP shape relative to the peak bin (f = fp exactly at index 19)
    x = sp.freq / fp
    sig = np.where(x < 1, np.float32(0.07), np.float32(0.09)).astype(np.float64)
    shape = x ** -5 * np.exp(-1.25 * x ** -4) * gam ** np.exp(-(x - 1) ** 2 / (2 * sig ** 2))
    assert np.allclose(Sf / Sf[19], shape / shape[19], rtol=1e-6)
    # cos^2s spreading around the grid direction of Dp, s clamped to [10, 1000]
    sc = max(round(min(s, 1000.0)), 10.0)
    main = bj.nautical_to_grid(Dp)
    D = np.cos((sp.dir - main) / 2) ** (2 * sc)
    assert np.allclose(Sd / Sd.max(), D / D.max(), rtol=1e-5, atol=1e-12)


def test_nautical_to_grid():
    assert bj.nautical_to_grid(270.0) == pytest.approx(0.0, abs=1e-8)
    assert bj.nautical_to_grid(260.0) == pytest.approx(math.radians(10.0), rel=1e-7)


def minstd_uniforms(seed, n):
    """libstdc++: default_random_engine = minstd_rand0 (x <- 16807 x mod
    2^31-1, seed 0 -> 1); uniform_real_distribution<double>(0, 1) =
    generate_canonical with two draws accumulated in double:
    sum = d1 + d2*R, result sum / R^2 (R = 2^31 - 2, d = draw - 1)."""
    m, a = 2147483647, 16807
    x = seed % m
    if x == 0:
        x = 1
    R = np.float64(2147483646.0)
    out = []
    for _ in range(n):
        x = (a * x) % m
        d1 = np.float64(x - 1)
        x = (a * x) % m
        d2 = np.float64(x - 1)
        s = d1 * np.float64(1.0)
        s = s + d2 * R
        v = s / np.float64(float(2147483646 * 2147483646))
        out.append(v if v < 1.0 else np.nextafter(1.0, 0.0))
    return np.array(out)


@pytest.mark.parametrize("seed", [0, 1, 12345])
def test_random_draws_are_the_reference_engine(seed):
    got = bj.uniforms(seed, 200)
    assert np.array_equal(got, minstd_uniforms(seed, 200))


@pytest.mark.parametrize("m", [64, 1000, 3602])
def test_hilbert_is_the_padded_analytic_signal(m):
    rng = np.random.default_rng(m)
    x = rng.normal(size=m)
    n = 1 << (m - 1).bit_length()
    want = scipy.signal.hilbert(np.pad(x, (0, n - m)))[:m]
    assert np.allclose(bj.hilbert(x), want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("N", [16, 3602, 4096])
def test_sparse_inverse_transform_equals_numpy(N):
    rng = np.random.default_rng(N)
    idx = np.sort(rng.choice(np.arange(1, N // 2), size=min(40, N // 2 - 1), replace=False))
    Gc = rng.normal(size=len(idx)) + 1j * rng.normal(size=len(idx))
    full = np.zeros(N, complex)
    full[idx] = Gc
    full[N - idx] = np.conj(Gc)
    want = np.real(np.fft.ifft(full) * N)
    got = bj.synth(N, idx, Gc)
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.fixture(scope="module")
def small_block():
    p = default_params(nx=200, ny=50, dx=5.0, thetamin=-60, thetamax=60, dtheta=10)
    p.rtlength = 600.0
    sp = bj.jonswap(1.5, 10.0, 270.0, 10.0, 3.3)
    return p, sp, bj.wave_groups(p, 8.0, sp, seed=0)


def test_wave_groups_records_and_trep(small_block):
    p, sp, blk = small_block
    nrec = math.ceil(p.rtlength / p.dtbc) + 1
    assert blk.Stfile.shape == (nrec, p.ntheta, p.ny) and blk.qfile.shape == (nrec, 4, p.ny)
    assert np.isfinite(blk.Stfile).all() and (blk.Stfile >= 0).all()
    assert (blk.qfile[:, 2:] == 0).all()
    # Trep = sum(Sf/f) / sum(Sf) over Sf >= 0.01 max (makjonswap.cpp:467-486)
    Sf, _, _, _ = moments(sp)
    sel = Sf >= Sf.max() * 0.01
    want = float(np.float32((Sf[sel] / np.maximum(sp.freq[sel], 0.001)).sum() / Sf[sel].sum()))
    assert blk.Trep == pytest.approx(want, rel=1e-6)


def test_wave_groups_energy_matches_the_spectrum(small_block):
    """The time-mean wave energy summed over the bins equals rho g Hm0^2 / 16
    of the components falling into the computational bins (the taper of the
    first 5 peak periods aside), the groups modulate it, and the directional
    distribution follows the spectrum's."""
    p, sp, blk = small_block
    E = blk.Stfile.sum(axis=1) * p.dtheta          # (nrec, ny)
    Etot = RHO * G * 1.5 ** 2 / 16
    steady = E[100:]                               # after the start taper
    assert 0.75 * Etot < steady.mean() < 1.1 * Etot
    cv = steady.std(axis=0) / steady.mean(axis=0)
    assert (cv > 0.2).all()                        # wave groups, not a constant
    # energy-weighted mean direction close to the spectral main direction (0)
    th = p.thetamin + (np.arange(p.ntheta) + 0.5) * p.dtheta
    mdir = (blk.Stfile[100:].mean(axis=(0, 2)) * th).sum() / blk.Stfile[100:].mean(axis=(0, 2)).sum()
    assert abs(mdir) < math.radians(8)


def test_bound_long_wave_flux(small_block):
    """Bound long waves: zero-mean flux at the group (difference) frequencies,
    below the peak frequency like the group envelope itself, tapered to zero
    at both ends of the block."""
    p, sp, blk = small_block
    qx = blk.qfile[:, 0]
    assert np.isfinite(qx).all() and np.abs(qx).max() > 0
    assert abs(qx[100:-100].mean()) < 0.2 * qx[100:-100].std()
    assert np.abs(qx[0]).max() == 0 and np.abs(qx[-1]).max() < 1e-12
    f = np.fft.rfftfreq(qx.shape[0] - 200, p.dtbc)
    Fq = np.abs(np.fft.rfft(qx[100:-100] - qx[100:-100].mean(0), axis=0)) ** 2
    E = blk.Stfile.sum(axis=1)[100:-100]
    FE = np.abs(np.fft.rfft(E - E.mean(0), axis=0)) ** 2
    fp = 1 / 10.0
    assert Fq[f < fp].sum() / Fq.sum() > 0.99
    assert FE[f < fp].sum() / FE.sum() > 0.99


def test_wave_groups_deterministic_in_the_seed(small_block):
    p, sp, blk = small_block
    again = bj.wave_groups(p, 8.0, sp, seed=0)
    other = bj.wave_groups(p, 8.0, sp, seed=7)
    assert np.array_equal(again.Stfile, blk.Stfile) and np.array_equal(again.qfile, blk.qfile)
    assert not np.allclose(other.Stfile, blk.Stfile)


def test_constant_waves():
    """GenCstWave: E = rho g Hs^2 / 16 spread by cos^2s over the bins."""
    p = default_params(nx=20, ny=6, dx=5.0, thetamin=-60, thetamax=60, dtheta=10)
    from xbeach_gpu_amd import theta_tables
    th, _, _ = theta_tables(p)
    blk = bj.constant_waves(p, [(1.0, 8.0, 270.0, 10.0), (2.0, 10.0, 260.0, 20.0)], th)
    for n, Hs in enumerate((1.0, 2.0)):
        assert (blk.Stfile[n].sum(axis=0) * p.dtheta) == pytest.approx(np.full(p.ny, RHO * G * Hs ** 2 / 16),
                                                                       rel=1e-12)
        assert np.allclose(blk.Stfile[n], blk.Stfile[n][:, :1])
    assert (blk.qfile == 0).all() and blk.Trep == 8.0
