"""BandPassRecorder on the GPU: the integrated spectrum
(ppl_bandpass_accumulate, src/lib/src/BandPassRecorder.cpp:85-97) equals the
oracle's sequential float sums bit for bit over ragged blob sequences, and the
whole recorder (device integration + host fit + BandPassOutput, through
pelican-lofar_amd/ppl_bandpass) gives the oracle's bandpass."""
import json
import os
import subprocess

import numpy as np
import pytest

import bandpass_recorder as bpr
from conftest import ROOT

pytestmark = pytest.mark.gpu
F32 = np.float32
EXE = os.path.join(ROOT, "pelican-lofar_amd", "ppl_bandpass")


def _stokes(rng, n_t, n_sub, n_pol, n_ch):
    # bandpass-shaped positive powers with a spread of magnitudes
    shape = 1000.0 + 300.0 * np.cos(np.linspace(0, 3, n_sub * n_ch)).reshape(n_sub, n_ch)
    x = rng.exponential(1.0, (n_t, n_sub, n_pol, n_ch)).astype(F32) * shape[None, :, None, :].astype(F32)
    return x.astype(F32)


@pytest.mark.parametrize("n_sub,n_pol,n_ch,blobs", [
    (3, 1, 16, [5, 0, 17, 1, 33]),
    (8, 4, 64, [16, 15, 1]),
    (62, 1, 256, [1024, 1000]),
])
def test_accumulate_bit_exact(ppl, n_sub, n_pol, n_ch, blobs):
    import torch
    rng = np.random.default_rng(n_sub * 100 + n_ch)
    host = [_stokes(rng, n, n_sub, n_pol, n_ch) for n in blobs]
    d_sum = torch.zeros(n_sub * n_ch, dtype=torch.float32, device="cuda")
    for h in host:
        ppl.bandpass_accumulate(torch.from_numpy(h).cuda(), d_sum)
    torch.cuda.synchronize()
    want = bpr.accumulate(host)
    got = d_sum.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_accumulate_sequential_order(ppl):
    import torch
    # 2^24 then ones: only the time-ordered float sum stays at 2^24
    blob = np.ones((129, 2, 1, 64), F32)
    blob[0] = 2.0 ** 24
    d_sum = torch.zeros(128, dtype=torch.float32, device="cuda")
    ppl.bandpass_accumulate(torch.from_numpy(blob).cuda(), d_sum)
    assert np.all(d_sum.cpu().numpy() == F32(2.0 ** 24))


def test_accumulate_rejects_bad_shapes(ppl):
    import torch
    d = torch.zeros(16, dtype=torch.float32, device="cuda")
    L = ppl.lib()
    assert L.ppl_bandpass_accumulate(ppl._ptr(d), 1, 0, 1, 16, ppl._ptr(d), None) != 0
    assert L.ppl_bandpass_accumulate(ppl._ptr(d), -1, 1, 1, 16, ppl._ptr(d), None) != 0
    assert L.ppl_bandpass_accumulate(None, 1, 1, 1, 16, ppl._ptr(d), None) != 0


def test_accumulate_rate(ppl):
    """Throughput of the integration at a LOFAR-beam shape (62 x 256 bins,
    4096 spectra = 260 MB read), for the record: printed, not asserted."""
    import torch
    n_t, n_sub, n_ch = 4096, 62, 256
    x = torch.rand((n_t, n_sub, 1, n_ch), dtype=torch.float32, device="cuda")
    d_sum = torch.zeros(n_sub * n_ch, dtype=torch.float32, device="cuda")
    ppl.bandpass_accumulate(x, d_sum)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    reps = 10
    for _ in range(reps):
        ppl.bandpass_accumulate(x, d_sum)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print("bandpass_accumulate %d x %d bins x %d spectra: %.3f ms, %.0f GB/s"
          % (n_sub, n_ch, n_t, ms, x.numel() * 4 / ms / 1e6))


CONFIG = """<config>
  <BandPassRecorder>
    <requiredSamples value="{req}"/>
    <fitParmaters coefficients="3"/>
    <Band startFrequency="{f0}" endFrequency="{f1}"/>
  </BandPassRecorder>
  <BandPassOutput><file name="{out}"/></BandPassOutput>
</config>
"""


def test_recorder_end_to_end(tmp_path):
    assert os.path.exists(EXE), "build the library first (make -C pelican-lofar_amd)"
    n_t, n_sub, n_pol, n_ch, chunk, req = 20000, 8, 1, 62, 4096, 20000
    f0, f1 = 138.0, 146.0
    rng = np.random.default_rng(77)
    n = n_sub * n_ch
    f = bpr.bin_freqs(n, f0, f1)[0].astype(np.float64)
    level = (2000.0 - 3000.0 * ((f - 142.0) / 8.0) ** 2).astype(F32)
    x = (rng.normal(0, 50.0, (n_t, n_sub, n_pol, n_ch)).astype(F32)
         + level.reshape(n_sub, n_ch)[None, :, None, :]).astype(F32)
    rfi = rng.choice(n, 12, replace=False)
    x[:, rfi // n_ch, 0, rfi % n_ch] += F32(400.0)
    path = tmp_path / "stokes.f32"
    x.tofile(path)
    out = tmp_path / "rec.bp"
    cfg = tmp_path / "bp.xml"
    cfg.write_text(CONFIG.format(req=req, f0=f0, f1=f1, out=out))
    p = subprocess.run([EXE, str(cfg), "stream", str(path), str(n_t), str(n_sub), str(n_pol), str(n_ch), str(chunk)],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    r = json.loads(p.stdout)
    assert r["done"] and r["blobs"] == 5 and r["spectra"] == req
    blobs = [x[t:t + chunk] for t in range(0, n_t, chunk)]
    avg = bpr.average(bpr.accumulate(blobs), req)
    params, rms, clean, fits = bpr.fit_spectrum(avg, f0, f1, 3)
    assert r["fits"] == fits
    freq = bpr.bin_freqs(n, f0, f1)[0]
    got_model = bpr.evaluate(np.array(r["params"], F32), freq)
    want_model = bpr.evaluate(params, freq)
    assert np.max(np.abs(got_model - want_model)) <= 2e-5 * float(np.max(np.abs(want_model)))
    # the RFI channels were trimmed, and the model follows the injected level
    assert np.all(clean[rfi] < avg[rfi] - 200)
    assert np.max(np.abs(got_model - level)) < 5.0
    assert r["rms"] == pytest.approx(float(rms), rel=1e-3)
    assert out.read_text() == bpr.bp_file_text(n, f0, f1, r["rms"], r["median"], r["params"])
