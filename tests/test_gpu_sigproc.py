"""SigprocStokesWriter on the GPU: ppl_sigproc_pack gives the oracle's bytes
(src/lib/src/SigprocStokesWriter.cpp:184-240, 285-290) bit for bit in 32 and
8 bits, including the crop edges, rint ties, infinities and NaN, and the
writer's file (through pelican-lofar_amd/ppl_outputs) is the header followed
by every blob's bytes."""
import os
import re
import subprocess
import time

import numpy as np
import pytest

import sigproc
from conftest import ROOT

pytestmark = pytest.mark.gpu
F32 = np.float32
EXE = os.path.join(ROOT, "pelican-lofar_amd", "ppl_outputs")


def _blob(rng, n_t, n_sub, n_pol, n_ch):
    x = rng.normal(128.0, 60.0, (n_t, n_sub, n_pol, n_ch)).astype(F32)
    flat = x.reshape(-1)
    k = flat.size
    # edges: crop limits, rint ties, specials
    picks = rng.choice(k, 64, replace=False)
    specials = np.array([0.0, 255.0, -1e-7, 255.00002, 0.5, 1.5, 2.5, 127.5, 128.5, np.inf, -np.inf, np.nan,
                         -0.0, 1e30, -1e30, 254.5], F32)
    flat[picks[:len(specials)]] = specials
    flat[picks[len(specials):]] = np.round(flat[picks[len(specials):]]) + F32(0.5)
    return x


@pytest.mark.parametrize("n_bits,crop", [(32, (0.0, 255.0)), (8, (0.0, 255.0)), (8, (20.0, 230.5)),
                                         (8, (5.0, 5.0))])
@pytest.mark.parametrize("shape,n_pw", [((37, 3, 1, 16), 1), ((8, 31, 4, 64), 2), ((512, 62, 1, 256), 1)])
def test_pack_bit_exact(ppl, n_bits, crop, shape, n_pw):
    import torch
    rng = np.random.default_rng(shape[0] + n_bits)
    x = _blob(rng, *shape)
    out = ppl.sigproc_pack(torch.from_numpy(x).cuda(), n_pols_write=n_pw, n_bits=n_bits, crop_min=crop[0],
                           crop_max=crop[1])
    got = out.cpu().numpy().tobytes()
    want = sigproc.pack(x, n_pw, n_bits, crop[0], crop[1])
    assert got == want


def test_pack_rejects(ppl):
    import torch
    L = ppl.lib()
    x = torch.zeros((2, 2, 1, 4), dtype=torch.float32, device="cuda")
    o = torch.zeros(64, dtype=torch.float32, device="cuda")
    assert L.ppl_sigproc_pack(ppl._ptr(x), 2, 2, 1, 4, 1, 16, 0.0, 1.0, ppl._ptr(o), None) != 0
    assert L.ppl_sigproc_pack(ppl._ptr(x), 2, 2, 1, 4, 2, 32, 0.0, 1.0, ppl._ptr(o), None) != 0  # 2 > nPol
    # an empty blob writes nothing, whatever the pointers
    assert L.ppl_sigproc_pack(None, 0, 2, 1, 4, 1, 32, 0.0, 1.0, None, None) == 0


XML = """<output>
  <dataStreams><stream name="SignalFoundSpectrum" listeners="SigprocStokesWriter"/></dataStreams>
  <streamers>
    <SigprocStokesWriter writeHeader="true">
      <file filepath="{path}"/>
      <subbandsPerPacket value="{nsub}"/><outputChannelsPerSubband value="{nch}"/>
      <dataBits value="{bits}"/><scale min="{cmin}" max="{cmax}"/>
      <params nPolsToWrite="{npw}"/>
      <fch1 value="150.0"/><foff value="-0.1"/><tsamp value="0.001"/>
    </SigprocStokesWriter>
  </streamers>
</output>"""


@pytest.mark.parametrize("bits,npw", [(32, 1), (8, 2)])
def test_writer_file(tmp_path, bits, npw):
    assert os.path.exists(EXE), "build the library first (make -C pelican-lofar_amd)"
    rng = np.random.default_rng(bits)
    n_sub, n_pol, n_ch = 4, 2, 32
    blobs = [_blob(rng, n, n_sub, n_pol, n_ch) for n in (64, 17, 128)]
    cfg = tmp_path / "out.xml"
    cfg.write_text(XML.format(path=tmp_path / "beam", nsub=n_sub, nch=n_ch, bits=bits, cmin=10.0, cmax=250.0,
                              npw=npw))
    script = ""
    stamp = 1.4e9
    for k, b in enumerate(blobs):
        path = tmp_path / ("b%d.f32" % k)
        b.tofile(path)
        script += "stokes %s %d %d %d %d %r 0.001\nsendstokes SignalFoundSpectrum\n" % (
            path, b.shape[0], n_sub, n_pol, n_ch, stamp + k)
    p = subprocess.run([EXE, str(cfg)], input=script, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    name = re.search(r"^fil (.*)$", p.stdout, re.M).group(1)
    epoch = time.mktime((2011, 1, 1, 0, 0, 0, 0, 0, -1))
    # default RAJ/DecJ "000000.0": dec 0 is not > 0, so the name is
    # "0000" + the first five characters of the dec string
    hdr = sigproc.header(9, 0, "000000000", 0.0, 0.0, 150.0, -0.1, n_sub * n_ch, 0.001, bits,
                         (stamp - epoch) / 86400 + 55562.0, npw)
    body = b"".join(sigproc.pack(b, npw, bits, 10.0, 250.0) for b in blobs)
    data = open(name, "rb").read()
    assert data == hdr + body
