"""The drop-in module chain (Pelican-style C++ modules over the C ABI, run
by the DedispersionPipeline glue in the ppl_chain driver) on the
CPU-runnable configuration: its event lists equal those of the C-ABI chain
bit for bit, it finds the injected pulse, and every Stokes blob returns to
the pool once no analysed or pending buffer holds its samples."""
import os
import subprocess

import numpy as np
import pytest

from test_gpu_chain import CHUNK, FCH1, FOFF, N_CH, N_CHUNKS, N_SUB, PULSE, TDMS, DM_STEP, _gpu_chain, _inputs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHAIN = os.path.join(ROOT, "pelican-lofar_amd", "ppl_chain")

XML = """<pipeline>
  <DedispersionPipeline><history value="{history}"/><events min="0" max="0"/></DedispersionPipeline>
  <PPFChanneliser>
    <outputChannelsPerSubband value="{nch}"/>
    <filter nTaps="8" filterWindow="hann"/>
  </PPFChanneliser>
  <StokesGenerator><numberOfStokes value="1"/></StokesGenerator>
  <StokesIntegrator><integrateTimeBins value="4"/><integrateFrequencyChannels value="1"/></StokesIntegrator>
  <RFI_Clipper active="true" channelRejectionRMS="10.0" spectrumRejectionRMS="4.0">
    <BandPassData file="{bp}"/>
    <History maximum="10000"/>
    <zeroDMing active="true"/>
    <Band matching="true"/>
  </RFI_Clipper>
  <DedispersionModule>
    <timeBinsPerBufferPow2 value="12"/>
    <dedispersionSamples value="{tdms}"/>
    <dedispersionStepSize value="{dmstep}"/>
    <dedispersionMinimum value="0"/>
    <frequencyChannel1 MHz="{fch1}"/>
    <channelBandwidth MHz="{foff}"/>
    <numberOfBuffers value="2"/>
  </DedispersionModule>
  <DedispersionAnalyser>
    <detectionThreshold in_sigma="6.0"/>
    <power2ForBinning value="6"/>
  </DedispersionAnalyser>
</pipeline>
"""


def _write_bp(path, bp):
    with open(path, "w") as f:
        f.write("# bandpass of the synthetic test band\n")
        for k, v in enumerate(bp):
            f.write(("%d\n" % int(v)) if k == 0 else ("%.17g\n" % v))


def _read_events(path):
    bufs, pool = [], None
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines):
        w = lines[i].split()
        i += 1
        if not w:
            continue
        if w[0] == "buffer":
            n = int(w[2])
            rows = [lines[i + k].split() for k in range(n)]
            i += n
            ev = np.array([[int(r[0]), int(r[1]), int(r[2])] for r in rows], np.int32).reshape(-1, 3)
            vals = np.array([float(r[3]) for r in rows], np.float32)
            bufs.append((ev, vals))
        elif w[0] == "pool":
            pool = (int(w[1]), int(w[2]))
    return bufs, pool


def test_module_chain_matches_c_abi_chain(orc, ppl, tmp_path):
    assert os.path.exists(CHAIN), "build the library first (make -C pelican-lofar_amd)"
    coeffs, x, bp = _inputs(orc)
    plan, raw, clipped, buffers, series, events = _gpu_chain(ppl, coeffs, x, bp)
    bp_path, xml_path = str(tmp_path / "band.bp"), str(tmp_path / "pipeline.xml")
    volt_path, ev_path = str(tmp_path / "volt.i16"), str(tmp_path / "events.txt")
    _write_bp(bp_path, bp)
    history = 64
    with open(xml_path, "w") as f:
        f.write(XML.format(history=history, nch=N_CH, bp=bp_path, tdms=TDMS, dmstep=repr(DM_STEP),
                           fch1=repr(FCH1), foff=repr(FOFF)))
    np.ascontiguousarray(x).tofile(volt_path)
    r = subprocess.run([CHAIN, xml_path, volt_path, str(N_SUB), str(CHUNK * N_CHUNKS), str(CHUNK),
                        repr(5.12e-6), ev_path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got, pool = _read_events(ev_path)
    assert len(got) == len(events) >= 4
    for b, ((ge, gv), (ce, cv)) in enumerate(zip(got, events)):
        assert np.array_equal(ge, ce) and np.array_equal(gv, cv), b
    # every blob is back in the pool except those holding samples of the
    # unfinished buffer (its overlap plus the samples added since)
    overlap = plan.maxshift + plan.rem
    total = CHUNK // N_CH // 4 * N_CHUNKS
    consumed = 4096 + (len(events) - 1) * (4096 - overlap)
    held = -(-(overlap + total - consumed) // (CHUNK // N_CH // 4))
    assert pool == (history - held, history), (pool, held)
    dm_idx = int(round(PULSE["dm"] / DM_STEP))
    found = any(abs(e[0] - dm_idx) <= 2 and abs(b * plan.nout + e[1] - PULSE["t0"]) <= 16
                for b, (ev, _) in enumerate(got) for e in ev[1:])
    assert found
