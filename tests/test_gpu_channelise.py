"""GPU parity of the channeliser with the oracle: spectra against the FP64
DFT of the oracle's FP32 FIR (stated tolerance), history carried across
calls, fused Stokes-I/integration against the unfused stages, standalone
Stokes (I, Q, U, V) and integration bit-exact."""
import numpy as np
import pytest
import torch

from synth import noise_voltages

pytestmark = pytest.mark.gpu

# relative error bound of the FP32 FIR + FP32 FFT against FP64, per output,
# measured against the spectrum's RMS (|X - X64| <= TOL * rms(X64))
SPEC_TOL = 2e-5


def _coeffs(orc, n_taps, n_ch, window="kaiser"):
    return orc.ppf_coefficients(n_taps, n_ch, window)


@pytest.mark.parametrize("n_ch,n_taps", [(16, 8), (64, 8), (256, 8), (32, 4), (128, 16), (8, 8)])
def test_spectra_vs_fp64_dft(orc, ppl, n_ch, n_taps):
    n_sub, n_blk = 3, 48
    h = _coeffs(orc, n_taps, n_ch)
    x = noise_voltages(n_sub, n_blk * n_ch, 64.0, seed=n_ch)
    p = ppl.PPF(n_sub, n_ch, n_taps, h.astype(np.float32))
    out = p.spectra_i16(torch.from_numpy(x).cuda(), n_blk).cpu().numpy()
    z = out[..., 0] + 1j * out[..., 1].astype(np.float64)
    ref = orc.PPF(n_sub, 2, n_ch, n_taps, h).run_fir32(x, n_blk)
    rms = np.sqrt(np.mean(np.abs(ref) ** 2))
    err = np.max(np.abs(z - ref)) / rms
    assert err < SPEC_TOL, err


def test_history_carries_across_calls(orc, ppl):
    n_sub, n_ch, n_taps, n_blk = 2, 64, 8, 96
    h = _coeffs(orc, n_taps, n_ch)
    x = noise_voltages(n_sub, n_blk * n_ch, 64.0, seed=3)
    xd = torch.from_numpy(x).cuda()
    whole = ppl.PPF(n_sub, n_ch, n_taps, h.astype(np.float32)).spectra_i16(xd, n_blk)
    p = ppl.PPF(n_sub, n_ch, n_taps, h.astype(np.float32))
    parts = []
    for a, b in [(0, 5), (5, 40), (40, 96)]:  # includes a call shorter than nTaps
        parts.append(p.spectra_i16(xd[:, :, a * n_ch:b * n_ch].contiguous(), b - a))
    assert torch.equal(torch.cat(parts), whole)


def test_complex_float_input_matches_int16(orc, ppl):
    n_sub, n_ch, n_taps, n_blk = 2, 16, 8, 32
    h = _coeffs(orc, n_taps, n_ch).astype(np.float32)
    x = noise_voltages(n_sub, n_blk * n_ch, 64.0, seed=4)
    a = ppl.PPF(n_sub, n_ch, n_taps, h).spectra_i16(torch.from_numpy(x).cuda(), n_blk)
    xf = torch.from_numpy(x.astype(np.float32)).cuda()
    b = ppl.PPF(n_sub, n_ch, n_taps, h).spectra_c32(xf, n_blk)
    assert torch.equal(a, b)


def test_stokes_and_integrate_bit_exact(orc, ppl):
    rng = np.random.default_rng(8)
    n_blk, n_sub, n_ch = 64, 5, 32
    spec = (rng.normal(size=(n_blk, n_sub, 2, n_ch, 2)) * 300).astype(np.float32)
    d = torch.from_numpy(spec).cuda()
    I = ppl.stokes(d, 1).cpu().numpy()
    assert np.array_equal(I, orc.stokes_i_f32(spec))
    I4 = ppl.stokes(d, 4).cpu().numpy()
    assert np.array_equal(I4[:, :, 0:1], I)
    X, Y = spec[:, :, 0], spec[:, :, 1]
    Xr, Xi, Yr, Yi = X[..., 0], X[..., 1], Y[..., 0], Y[..., 1]
    # StokesGenerator.cpp:93-103, each product rounded to float32 before the sum
    power_x = Xr * Xr + Xi * Xi
    power_y = Yr * Yr + Yi * Yi
    assert np.array_equal(I4[:, :, 1], power_x - power_y)
    assert np.array_equal(I4[:, :, 2], np.float32(2.0) * (Xr * Yr + Xi * Yi))
    assert np.array_equal(I4[:, :, 3], np.float32(2.0) * (Xi * Yr - Xr * Yi))
    # a swapped sign or operand order in V, or U taken from X alone, would differ
    assert not np.array_equal(I4[:, :, 3], np.float32(2.0) * (Xr * Yi - Xi * Yr))
    assert not np.array_equal(I4[:, :, 2], np.float32(2.0) * (Xr * Xr + Xi * Xi))
    integ = ppl.integrate(torch.from_numpy(I).cuda(), 4).cpu().numpy()
    assert np.array_equal(integ, orc.integrate(I, 4))
    integ2 = ppl.integrate(torch.from_numpy(I).cuda(), 2, 4).cpu().numpy()
    assert np.array_equal(integ2, orc.integrate(I, 2, 4))


@pytest.mark.parametrize("n_ch,n_taps,n_sub,n_blk", [(16, 8, 4, 128), (64, 8, 4, 128), (256, 8, 3, 72),
                                                     (128, 16, 3, 100), (256, 16, 3, 72)])
def test_fused_stokes_matches_stages(orc, ppl, n_ch, n_taps, n_sub, n_blk):
    """Fused PPF -> Stokes-I -> x4 equals the unfused kernels bit for bit, and
    the FP64 chain within tolerance.  The 256- and 128-channel cases with
    8 and 16 taps take the pol-split form (a wave per polarisation); with
    72 blocks of 256 channels the runs are 32, 32 and 8 blocks, so a
    workgroup's two runs straddle subbands, the last run is short and the
    last workgroup's second pair has no run."""
    h = _coeffs(orc, n_taps, n_ch)
    x = noise_voltages(n_sub, n_blk * n_ch, 64.0, seed=9)
    xd = torch.from_numpy(x).cuda()
    fused = ppl.PPF(n_sub, n_ch, n_taps, h.astype(np.float32), integrate=4).stokes_i16(xd, n_blk)
    spec = ppl.PPF(n_sub, n_ch, n_taps, h.astype(np.float32)).spectra_i16(xd, n_blk)
    staged = ppl.integrate(ppl.stokes(spec, 1), 4)
    ref = orc.PPF(n_sub, 2, n_ch, n_taps, h).run_f64(x, n_blk)
    I64 = orc.integrate(orc.stokes_i_f64(ref), 4)
    err = np.max(np.abs(fused.cpu().numpy() - I64)) / np.mean(I64)
    assert torch.equal(fused, staged), (err, float((fused - staged).abs().max()))
    assert err < 1e-4, err
