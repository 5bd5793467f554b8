"""The split-precision channeliser (DFT on the matrix cores: BF16x3 or FP16x2
products, FP32 accumulation) against the oracle's FP64 DFT of the FP32 FIR and
against the all-FP64 chain, with the error per output stated in units of the
spectrum's RMS; history across calls, the fused Stokes-I/integration and the
shapes it rejects."""
import numpy as np
import pytest
import torch

from synth import noise_voltages

pytestmark = pytest.mark.gpu

# |X - X64| <= TOL * rms(X64) for every output (max over outputs), and the
# RMS of the error <= TOL_RMS * rms(X64)
# (measured on MI355X, noise voltages: BF16x3 max 1.2e-5 / 1.9e-5 and RMS
# 3.1e-6 / 4.4e-6 at 64 / 256 channels; FP16x2 max 4.0e-4 / 6.6e-4, RMS
# 8.3e-5 / 1.4e-4; FP32 max 5.9e-7)
TOL = {"bf16x3": 5e-5, "fp16x2": 1.5e-3}
TOL_RMS = {"bf16x3": 1e-5, "fp16x2": 3e-4}


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x2"])
@pytest.mark.parametrize("n_ch,n_taps", [(64, 8), (256, 8), (64, 4), (256, 16)])
def test_split_precision_spectra_error(orc, ppl, prec, n_ch, n_taps):
    n_sub, n_blk = 3, 88  # 88 blocks: runs end in a partial tile
    h = orc.ppf_coefficients(n_taps, n_ch, "kaiser")
    x = noise_voltages(n_sub, n_blk * n_ch, 64.0, seed=n_ch + n_taps)
    p = ppl.PPF(n_sub, n_ch, n_taps, h.astype(np.float32), precision=prec)
    out = p.spectra_i16(torch.from_numpy(x).cuda(), n_blk).cpu().numpy()
    z = out[..., 0] + 1j * out[..., 1].astype(np.float64)
    op = orc.PPF(n_sub, 2, n_ch, n_taps, h)
    ref32 = op.run_fir32(x, n_blk)  # FP32 FIR, FP64 DFT
    ref64 = orc.PPF(n_sub, 2, n_ch, n_taps, h).run_f64(x, n_blk)  # all FP64
    rms = np.sqrt(np.mean(np.abs(ref64) ** 2))
    for ref in (ref32, ref64):
        e = np.abs(z - ref)
        assert e.max() < TOL[prec] * rms, (prec, e.max() / rms)
        assert np.sqrt(np.mean(e ** 2)) < TOL_RMS[prec] * rms, (prec, np.sqrt(np.mean(e ** 2)) / rms)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x2"])
def test_split_precision_history_and_fused_stokes(orc, ppl, prec):
    n_sub, n_ch, n_taps, n_blk = 2, 256, 8, 160
    h = orc.ppf_coefficients(n_taps, n_ch, "hann").astype(np.float32)
    x = noise_voltages(n_sub, n_blk * n_ch, 64.0, seed=5)
    xd = torch.from_numpy(x).cuda()
    whole = ppl.PPF(n_sub, n_ch, n_taps, h, precision=prec).spectra_i16(xd, n_blk)
    p = ppl.PPF(n_sub, n_ch, n_taps, h, precision=prec)
    parts = [p.spectra_i16(xd[:, :, a * n_ch:b * n_ch].contiguous(), b - a) for a, b in [(0, 4), (4, 100), (100, 160)]]
    assert torch.equal(torch.cat(parts), whole)
    # fused PPF -> Stokes-I -> x4 equals Stokes and integration of the same spectra
    f = ppl.PPF(n_sub, n_ch, n_taps, h, precision=prec, integrate=4).stokes_i16(xd, n_blk)
    ref = ppl.integrate(ppl.stokes(whole, 1), 4)
    assert torch.equal(f, ref)


def test_split_precision_rejects_shapes(ppl, orc):
    h = orc.ppf_coefficients(8, 128, "kaiser").astype(np.float32)
    with pytest.raises(ppl.PPLError):
        ppl.PPF(2, 128, 8, h, precision="bf16x3")
    h = orc.ppf_coefficients(8, 64, "kaiser").astype(np.float32)
    with pytest.raises(ppl.PPLError):
        ppl.PPF(2, 64, 8, h, precision="fp16x2", integrate=3)
