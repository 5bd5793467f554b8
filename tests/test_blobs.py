"""Host blobs against the reference's own known answers, on the CPU:
DedispersionSpectraTest::test_dmIndex
(src/lib/test/src/DedispersionSpectraTest.cpp:32-55) and BinMapTest::test_bin
(src/lib/test/src/BinMapTest.cpp:48-62), through pelican-lofar_amd/ppl_blobs
(host/lofar/DataSets.h)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "pelican-lofar_amd", "ppl_blobs")


def run(script):
    if not os.path.exists(EXE):
        pytest.fail("pelican-lofar_amd/ppl_blobs is not built (make -C pelican-lofar_amd)")
    r = subprocess.run([EXE], input=script, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.split("\n")[:-1]


def test_dedispersion_spectra_dm_index():
    """DedispersionSpectraTest.cpp:32-55: 100 time bins x 100 DM bins from
    0.1 in steps of 0.2; the last element holds 1001.1."""
    timebins, dm_bins, start, width, amp = 100, 100, 0.1, 0.2, 1001.1
    dm_max = np.float32(start) + np.float32(width) * (dm_bins - 1)
    out = run("\n".join([
        "resize %d %d %r %r" % (timebins, dm_bins, start, width),
        "size",
        "set %d %r" % (timebins * dm_bins - 1, amp),
        "dmMax", "dmBins", "dmIndex %r" % float(dm_max), "timeSamples",
        "ampf %d %r" % (timebins - 1, float(dm_max)),
        "ampi %d %d" % (timebins - 1, dm_bins - 1),
        "dm 0", "dm 1", "dm 99", "dmIndex 0.1", "dmIndex 0.3", "dmIndex 0.19", "dmIndex 0.21",
    ]) + "\n")
    assert int(out[0]) == timebins * dm_bins
    assert abs(float(out[1]) - dm_max) <= 1e-4           # CPPUNIT_ASSERT_DOUBLES_EQUAL(max, dmMax(), 0.0001)
    assert int(out[2]) == dm_bins
    assert int(out[3]) == dm_bins - 1                    # dmIndex(max) == 99
    assert int(out[4]) == timebins
    assert np.float32(out[5]) == np.float32(amp)          # dmAmplitude(99, max) == 1001.1
    assert np.float32(out[6]) == np.float32(amp)          # dmAmplitude(99, 99) == 1001.1
    # dm(i) = start + width*i in double, narrowed to float; dmIndex rounds half up
    assert [np.float32(v) for v in out[7:10]] == [np.float32(np.float64(np.float32(start)) + np.float64(np.float32(width)) * i)
                                                  for i in (0, 1, 99)]
    assert [int(v) for v in out[10:14]] == [0, 1, 0, 1]


def test_dedispersion_spectra_layout():
    """[dm][t]: element t + timeBins*dm is time slice t of DM bin dm."""
    lines = ["resize 7 5 0 1"]
    for dm in range(5):
        for t in range(7):
            lines.append("set %d %d" % (t + 7 * dm, 100 * dm + t))
    q = [(t, dm) for dm in range(5) for t in range(7)]
    lines += ["ampi %d %d" % (t, dm) for t, dm in q]
    lines += ["ampf %d %r" % (t, float(dm)) for t, dm in q]
    out = run("\n".join(lines) + "\n")
    vals = [float(v) for v in out]
    assert vals == [100 * dm + t for t, dm in q] * 2


def test_bin_map_known_answers():
    """BinMapTest.cpp:48-62: 8192 bins from 142.779541, width 0.006104."""
    start, width = 142.779541, 0.006104
    out = run("\n".join([
        "map 8192 %r %r" % (start, width),
        "mapNumber 0", "mapStart 0", "mapEnd 0", "mapNumber 1", "mapNumber 2",
        "mapIndex %r" % start, "mapIndex %r" % (start + width),
    ]) + "\n")
    v = [float(x) for x in out[:5]]
    assert abs(v[0] - start) < 1e-5
    assert abs(v[1] - (start - width / 2)) < 1e-5
    assert abs(v[2] - (start + width / 2)) < 1e-5
    assert abs(v[3] - (start + width)) < 1e-5
    assert abs(v[4] - (start + 2 * width)) < 1e-5
    assert [int(x) for x in out[5:7]] == [0, 1]
