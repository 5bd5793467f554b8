"""The drop-in module chain (Pelican-style C++ modules over the C ABI, run
by the DedispersionPipeline glue in the ppl_chain driver) on the
CPU-runnable configuration: its event lists equal those of the C-ABI chain
bit for bit, it finds the injected pulse, every Stokes blob returns to
the pool once no analysed or pending buffer holds its samples, and the two
event writers on the pipeline's output stream report those events (text
rows, and the FRATS message of each buffer's best event)."""
import glob
import os
import re
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


OUTPUTS = """<output>
  <dataStreams>
    <stream name="DedispersionDataAnalysis" listeners="DedispersionDataAnalysisOutput"/>
    <stream name="TriggerInput" listeners="TriggerOutput"/>
    <stream name="SignalFoundSpectrum" listeners="SigprocStokesWriter"/>
  </dataStreams>
  <streamers>
    <DedispersionDataAnalysisOutput><file name="{prefix}"/></DedispersionDataAnalysisOutput>
    <TriggerOutput>
      <Trigger format="FRATS"/><Logfile name="{log}"/>
      <Telescope stationID="CS302" beamID="0"/><RADec rajd="1" decjd="2"/>
      <frequencyChannel1 MHz="{fch1}"/><Threshold snr="0" events="2"/>
    </TriggerOutput>
    <SigprocStokesWriter writeHeader="true">
      <file filepath="{fil}"/>
      <subbandsPerPacket value="{nsub}"/><outputChannelsPerSubband value="{nch}"/>
      <fch1 value="{fch1}"/><foff value="{foff}"/><tsamp value="{tsamp}"/>
      <RAJ value="053431.9"/><DecJ value="-220052.1"/>
    </SigprocStokesWriter>
  </streamers>
</output>
"""


def _check_filterbank(got, plan, clipped, fil_prefix):
    """The filterbank of every input blob of every written buffer, in order:
    the header once, then each blob's spectra with subbands and channels
    reversed (the clipped Stokes-I the C-ABI chain produced)."""
    import time
    import sigproc
    fil = glob.glob(fil_prefix + "_D*.fil")
    assert len(fil) == 1, fil
    data = open(fil[0], "rb").read()
    epoch = time.mktime((2011, 1, 1, 0, 0, 0, 0, 0, -1))
    hdr = sigproc.header(9, 0, "0534-2200", 53431.9, -220052.1, FCH1, FOFF, N_SUB * N_CH,
                         5.12e-6 * N_CH * 4, 32, (0.0 - epoch) / 86400 + 55562.0, 1)
    assert data[:len(hdr)] == hdr
    per_blob = CHUNK // N_CH // 4
    body = b""
    for b in range(len(got)):
        first, last = b * plan.nout // per_blob, (b * plan.nout + 4095) // per_blob
        for k in range(first, last + 1):
            body += sigproc.pack(clipped[k].reshape(per_blob, N_SUB, 1, N_CH))
    assert len(data) == len(hdr) + len(body)
    assert data[len(hdr):] == body


def _check_writers(got, plan, prefix, log):
    """The writers' text against the chain's own events: the dm of index i
    is float(i * double(float(step))), snr = value / (rms * sqrt(binFactor))
    with rms = sqrt(channels) as the analyser sets it, and event times are
    the integrated-sample times from the first chunk's time stamp (0)."""
    dat = glob.glob(prefix + "_D*.dat")
    assert len(dat) == 1, dat
    text = open(dat[0]).read().split("\n")
    assert text[:5] == ["# Generated by DedispersionDataAnalysisOutput", "# Events", "# ------",
                        "# Time|Dm|Amplitude|BinFactor", "# ------"]
    rms = np.float32(np.sqrt(np.float32(N_CH * N_SUB)))
    step = np.float64(np.float32(DM_STEP))
    tsamp = 5.12e-6 * N_CH * 4
    row = 5
    messages = []
    for b, (ev, vals) in enumerate(got):
        snr = [np.float32(np.float64(v) / (np.float64(rms) * np.sqrt(np.float64(e[2]))))
               for e, v in zip(ev[1:], vals[1:])]
        for (dm, t, bf), s in zip(ev[1:], snr):
            mjd, dmv, sv, bfv = [w.strip() for w in text[row].split(",")]
            row += 1
            assert dmv == "%.14g" % np.float32(step * dm) and int(bfv) == bf
            assert sv == "%.14g" % s
            assert abs(float(mjd) * 86400 - (b * plan.nout + t) * tsamp) <= 1e-9 * max(1.0, float(mjd) * 86400)
        m = re.fullmatch(r"# Written buffer :(\d+) \| MJDstart: (\S+) \| Best DM: (\S+) \| Max SNR: (\S+)  Done",
                         text[row])
        assert m and int(m.group(1)) == b + 1, text[row]
        row += 1
        assert abs(float(m.group(2)) * 86400 - b * plan.nout * tsamp) <= 1e-9 * max(1.0, b * plan.nout * tsamp)
        if snr and max(snr) > 0:
            k = int(np.argmax(snr))
            assert m.group(4) == "%.14g" % snr[k]
            assert m.group(3) == "%.14g" % np.float32(step * ev[1 + k][0])
            if len(ev) >= 2:
                messages.append((b, ev[1 + k], snr[k]))
    assert row == len(text) - 1 and text[row] == ""
    lines = open(log).read().split("\n")
    assert lines[-1] == "" and len(lines) - 1 == len(messages) > 0
    for n, (line, (b, (dm, t, bf), s)) in enumerate(zip(lines, messages)):
        m = re.fullmatch(r"artemis:CS302##ID:\d{12}-00(\d{4})##cFreqMHz:(.{10})##beamRA:0000000001"
                         r"##beamDec:000000002##eTime:(.{20})##eDM:(.{9})##eSNR:(.{6})", line)
        assert m, line
        assert int(m.group(1)) == n + 1
        assert m.group(4) == ("%.3f" % np.float32(step * dm)).rjust(9, "0")
        assert m.group(5) == ("%.2f" % s).rjust(6, "0")
        assert abs(float(m.group(3)) - (b * plan.nout + t) * tsamp) <= 1e-8


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
    out_path, prefix, log = str(tmp_path / "outputs.xml"), str(tmp_path / "events"), str(tmp_path / "trigger.log")
    fil = str(tmp_path / "found")
    with open(out_path, "w") as f:
        f.write(OUTPUTS.format(prefix=prefix, log=log, fch1=repr(FCH1), fil=fil, nsub=N_SUB, nch=N_CH,
                               foff=repr(FOFF), tsamp=repr(5.12e-6 * N_CH * 4)))
    r = subprocess.run([CHAIN, xml_path, volt_path, str(N_SUB), str(CHUNK * N_CHUNKS), str(CHUNK),
                        repr(5.12e-6), ev_path, out_path], capture_output=True, text=True, timeout=300)
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
    _check_writers(got, plan, prefix, log)
    _check_filterbank(got, plan, clipped, fil)
