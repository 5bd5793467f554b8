"""GPU parity of brute-force dedispersion with the oracle: bit-exact FP32
(channel-order sums), bit-exact FP16-input, exact int32 for u8 input."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SMALL = dict(n_chan=128, nsamp_buf=4096, tdms=128, dm_low=0.0, dm_step=50 / 128,
             tsamp=5.12e-6 * 16 * 4, fch1=141.5625, foff=-0.01220703125)


def _oracle(orc, x, cfg, dtype="f32"):
    p = orc.dedisp_plan(cfg["n_chan"], cfg["nsamp_buf"], cfg["tdms"], cfg["dm_low"], cfg["dm_step"],
                        cfg["tsamp"], cfg["fch1"], cfg["foff"])
    assert p["rc"] == 0
    d = orc.dedisp_delays(p["shift"], cfg["tdms"], p["mstartdm"], p["mdmstep"])
    return p, d


def test_delay_table_matches_oracle(orc, ppl):
    plan = ppl.DedispPlan(**SMALL)
    p, d = _oracle(orc, None, SMALL)
    assert (plan.maxshift, plan.rem, plan.nout) == (p["maxshift"], p["rem"], p["nout"])
    assert np.array_equal(plan.delays(), d)


@pytest.mark.parametrize("tile", [None, (32, 256), (16, 512), (8, 512), (16, 256), (64, 256),
                                  (16, 512, 0, 16), (32, 512, 0, 16), (32, 512, 1, 16), (16, 256, 0, 16),
                                  (32, 256, 0, 16), (64, 512, 0, 16), (64, 256, 0, 16)])
def test_f32_bit_exact_small(orc, ppl, tile):
    rng = np.random.default_rng(11)
    x = rng.normal(size=(SMALL["n_chan"], SMALL["nsamp_buf"])).astype(np.float32)
    plan = ppl.DedispPlan(**SMALL)
    if tile:
        plan.set_tile(*tile)
    out = plan.run(torch.from_numpy(x).cuda()).cpu().numpy()
    p, d = _oracle(orc, x, SMALL)
    exp = orc.dedisperse(x, d, p["nout"])
    assert out.shape == exp.shape
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("reads", [4, 16])
@pytest.mark.parametrize("shape", [(37, 2048, 100, 0.7), (300, 8192, 200, 0.1), (1, 1024, 3, 2.0),
                                   (129, 4096, 41, 0.3)])
def test_f32_bit_exact_ragged(orc, ppl, shape, reads):
    """Ragged shapes with one-sample (4-byte) and four-sample (16-byte) LDS
    reads; the last channel's staged rows run past the end of the input."""
    n_chan, ns, tdms, step = shape
    cfg = dict(n_chan=n_chan, nsamp_buf=ns, tdms=tdms, dm_low=0.0, dm_step=step,
               tsamp=0.00032768, fch1=150.0, foff=-6.0 / 2048)
    rng = np.random.default_rng(n_chan)
    x = (rng.normal(size=(n_chan, ns)) * 3 + 1).astype(np.float32)
    plan = ppl.DedispPlan(**cfg)
    if reads == 16:
        plan.set_tile(32, 512, 0, 16)
        assert plan.tile == (32, 512, 0, 16)
    out = plan.run(torch.from_numpy(x).cuda()).cpu().numpy()
    p, d = _oracle(orc, x, cfg)
    assert np.array_equal(out, orc.dedisperse(x, d, p["nout"]))


def test_known_answer_dm10(ppl):
    """DedispersionModuleTest.cpp:167-168: dmAmplitude(0, 10) == nSub*nCh."""
    nsub, nch = 32, 64
    n = nsub * nch
    fch1, foff, tsamp = 150.0, -6.0 / n, 0.00032768
    gen = [int(10.0 * (4148.741601 * ((1.0 / (fch1 + foff * c) / (fch1 + foff * c)) -
                                      (1.0 / fch1 / fch1)) / tsamp)) for c in range(n)]
    x = np.zeros((n, 32768), np.float32)
    for c in range(n):
        x[c, gen[c]:gen[c] + 10] = 1.0
    plan = ppl.DedispPlan(n, 32768, 200, 0.0, 0.1, tsamp, fch1, foff)
    out = plan.run(torch.from_numpy(x).cuda()).cpu().numpy()
    assert out[100, 0] == 2048.0


def test_f16_and_u8_exact(orc, ppl):
    rng = np.random.default_rng(12)
    x = rng.normal(size=(SMALL["n_chan"], SMALL["nsamp_buf"])).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    # FP16 input
    h = torch.empty(x.shape, dtype=torch.float16, device="cuda")
    ppl.convert_f16(xd, h)
    hb = h.cpu().numpy().view(np.uint16)
    assert np.array_equal(hb, orc.f32_to_f16_bits(x))
    plan16 = ppl.DedispPlan(**SMALL, dtype="f16")
    out16 = plan16.run(h).cpu().numpy()
    p, d = _oracle(orc, x, SMALL)
    assert np.array_equal(out16, orc.dedisperse(hb, d, p["nout"]))
    # offset-binary u8 input, int32 sums
    inv = np.float32(127.0 / 10.0)
    q = torch.empty(x.shape, dtype=torch.uint8, device="cuda")
    ppl.quantise_u8(xd, q, float(inv))
    qh = q.cpu().numpy()
    assert np.array_equal(qh, orc.quantise_u8(x, inv))
    plan8 = ppl.DedispPlan(**SMALL, dtype="u8")
    out8 = plan8.run(q).cpu().numpy()
    assert out8.dtype == np.int32
    assert np.array_equal(out8, orc.dedisperse(qh, d, p["nout"]))


def _tile_spread(d, td=32):
    """Largest delay spread inside any tile of td consecutive trials."""
    return max(int((d[min(k + td, len(d)) - 1] - d[k]).max()) for k in range(0, len(d), td))


def test_reference_c_seam(orc, ppl):
    """cacheDedisperseLoop's own arguments, for tdms not a multiple of 40.
    The 32-trial tiles' delay spread (171 samples) fits the ring kernel's
    staged row, so the seam runs the ring kernel with the table it builds
    from its device delays."""
    rng = np.random.default_rng(13)
    x = rng.normal(size=(SMALL["n_chan"], SMALL["nsamp_buf"])).astype(np.float32)
    p, d = _oracle(orc, x, SMALL)
    assert _tile_spread(d) // 4 + 512 // 4 <= 192  # the ring path
    shift = torch.from_numpy(p["shift"]).cuda()
    out = torch.full((SMALL["tdms"], p["nout"] + 5), -1.0, device="cuda")
    ppl.cache_dedisperse_loop(out, torch.from_numpy(x).cuda(), float(p["mstartdm"]), float(p["mdmstep"]),
                              SMALL["tdms"], SMALL["nsamp_buf"], shift, p["maxshift"] + p["rem"],
                              SMALL["n_chan"])
    flat = out.cpu().numpy().ravel()
    n = SMALL["tdms"] * p["nout"]
    assert np.array_equal(flat[:n].reshape(SMALL["tdms"], p["nout"]), orc.dedisperse(x, d, p["nout"]))
    assert np.all(flat[n:] == 0.0)  # the rest of outbufSize is cleared, as cudaMemset did


def test_reference_c_seam_wide_spread(orc, ppl):
    """The seam with delay spreads too wide for the ring kernel's row (the
    one-sample kernel path), tdms = 100."""
    cfg = dict(SMALL, dm_step=300 / 128, nsamp_buf=16384, tdms=100)
    rng = np.random.default_rng(14)
    x = rng.normal(size=(cfg["n_chan"], cfg["nsamp_buf"])).astype(np.float32)
    p, d = _oracle(orc, x, cfg)
    assert _tile_spread(d) // 4 + 512 // 4 > 192  # the one-sample path
    out = torch.empty((cfg["tdms"], p["nout"]), device="cuda")
    ppl.cache_dedisperse_loop(out, torch.from_numpy(x).cuda(), float(p["mstartdm"]), float(p["mdmstep"]),
                              cfg["tdms"], cfg["nsamp_buf"], torch.from_numpy(p["shift"]).cuda(),
                              p["maxshift"] + p["rem"], cfg["n_chan"])
    assert np.array_equal(out.cpu().numpy(), orc.dedisperse(x, d, p["nout"]))


def test_reference_c_seam_lofar_equals_plan(orc, ppl):
    """LOFAR-beam shape (3968 channels, 2048 trials, 65536 samples) through
    the seam (ring kernel, table built on the device) and through a plan
    (table built on the host from the oracle-checked delays): the same
    sums, bit for bit."""
    cfg = dict(n_chan=3968, nsamp_buf=65536, tdms=2048, dm_low=0.0, dm_step=500 / 2048,
               tsamp=5.12e-6 * 64 * 4, fch1=150.0, foff=-0.1953125 / 64)
    p, d = _oracle(orc, None, cfg)
    assert _tile_spread(d) // 4 + 512 // 4 <= 192  # the ring path
    plan = ppl.DedispPlan(**cfg)
    assert np.array_equal(plan.delays(), d)
    g = torch.Generator(device="cuda").manual_seed(21)
    x = torch.randn((cfg["n_chan"], cfg["nsamp_buf"]), device="cuda", generator=g)
    want = plan.run(x)
    out = torch.empty((cfg["tdms"], p["nout"]), device="cuda")
    ppl.cache_dedisperse_loop(out, x, float(p["mstartdm"]), float(p["mdmstep"]), cfg["tdms"], cfg["nsamp_buf"],
                              torch.from_numpy(p["shift"]).cuda(), p["maxshift"] + p["rem"], cfg["n_chan"])
    assert torch.equal(out, want)
    del x, out, want


def test_default_fp32_tile_is_the_ring_and_batches_fall_back(orc, ppl):
    """An FP32 plan starts on the ring tile (32 x 512, 16-byte reads) where
    its delay spread fits, and on one-sample reads where it does not; a batch
    under job 0's own tile falls back to one-sample reads when another job's
    spread does not fit the ring.  Every output equals the oracle's."""
    wide = dict(SMALL, dm_step=300 / 128, nsamp_buf=16384, tdms=100)
    rng = np.random.default_rng(15)
    xs, exps, plans = [], [], []
    for cfg, ring in ((SMALL, True), (wide, False)):
        x = rng.normal(size=(cfg["n_chan"], cfg["nsamp_buf"])).astype(np.float32)
        p, d = _oracle(orc, x, cfg)
        plan = ppl.DedispPlan(**cfg)
        assert (plan.tile == (32, 512, 0, 16)) == ring, plan.tile
        assert np.array_equal(plan.run(torch.from_numpy(x).cuda()).cpu().numpy(), orc.dedisperse(x, d, p["nout"]))
        xs.append(torch.from_numpy(x).cuda())
        exps.append(orc.dedisperse(x, d, p["nout"]))
        plans.append(plan)
    outs = ppl.dedisperse_batch(plans, xs)
    for o, e in zip(outs, exps):
        assert np.array_equal(o.cpu().numpy(), e)


def test_plan_counts_fused_delay_differences(orc, ppl):
    """The plan reports how many delays a fused multiply-add trial value
    (DedispersionKernel.cu:29 under contraction) would change: none with
    dmLow = 0, the oracle's count with dmLow > 0."""
    lofar = dict(n_chan=3968, nsamp_buf=65536, tdms=1024, dm_low=0.0, dm_step=500 / 2048,
                 tsamp=5.12e-6 * 64 * 4, fch1=150.0, foff=-0.1953125 / 64)
    assert ppl.DedispPlan(**lofar).fused_delay_diffs == 0
    lofar["dm_low"] = 37.3
    plan = ppl.DedispPlan(**lofar)
    p, d = _oracle(orc, None, lofar)
    assert np.array_equal(plan.delays(), d)
    fused = orc.dedisp_delays_fused(p["shift"], lofar["tdms"], p["mstartdm"], p["mdmstep"])
    assert plan.fused_delay_diffs == int(np.count_nonzero(fused != d)) > 0


def test_negative_dm_steps_and_shifts_are_refused(orc, ppl):
    """Delays that decrease with the trial index are refused (PPL_E_CONFIG):
    a negative dmStep in the plan, a negative dmShift, mstartdm or mdmstep
    in the C seam."""
    cfg = dict(SMALL)
    cfg["dm_step"] = -0.5
    with pytest.raises(ppl.PPLError, match="dmStep"):
        ppl.DedispPlan(**cfg)
    x = torch.zeros((SMALL["n_chan"], SMALL["nsamp_buf"]), device="cuda")
    p, d = _oracle(orc, None, SMALL)
    out = torch.zeros((SMALL["tdms"], p["nout"]), device="cuda")
    args = (SMALL["tdms"], SMALL["nsamp_buf"])
    bad = p["shift"].copy()
    bad[7] = -bad[7] - 1.0
    with pytest.raises(ppl.PPLError, match="negative"):
        ppl.cache_dedisperse_loop(out, x, float(p["mstartdm"]), float(p["mdmstep"]), *args,
                                  torch.from_numpy(bad).cuda(), p["maxshift"] + p["rem"], SMALL["n_chan"])
    shift = torch.from_numpy(p["shift"]).cuda()
    with pytest.raises(ppl.PPLError, match="negative"):
        ppl.cache_dedisperse_loop(out, x, 0.0, -float(p["mdmstep"]), *args, shift, p["maxshift"] + p["rem"],
                                  SMALL["n_chan"])


def test_plan_rejects_oversized_maxshift(ppl):
    with pytest.raises(ppl.PPLError):
        ppl.DedispPlan(3968, 4096, 2048, 0.0, 500 / 2048, 5.12e-6 * 64 * 4, 150.0, -0.1953125 / 64)


def test_linearity_full_size(ppl):
    """LOFAR-beam shape (3968 channels, 2048 trials, 65536 samples):
    dedisperse(a + b) == dedisperse(a) + dedisperse(b) for integer-valued
    inputs (exact in FP32), and an all-ones input sums to nChan."""
    cfg = dict(n_chan=3968, nsamp_buf=65536, tdms=2048, dm_low=0.0, dm_step=500 / 2048,
               tsamp=5.12e-6 * 64 * 4, fch1=150.0, foff=-0.1953125 / 64)
    plan = ppl.DedispPlan(**cfg)
    g = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randint(-8, 8, (3968, 65536), device="cuda", generator=g).float()
    b = torch.randint(-8, 8, (3968, 65536), device="cuda", generator=g).float()
    oa = plan.run(a)
    ob = plan.run(b)
    oab = plan.run(a + b)
    assert torch.equal(oab, oa + ob)
    ones = torch.ones((3968, 65536), device="cuda")
    assert torch.all(plan.run(ones) == 3968.0)
    # a single hot channel sample lands at t = s - delay[dm][c] for every DM
    z = torch.zeros((3968, 65536), device="cuda")
    c, s = 3000, 30000
    z[c, s] = 1.0
    out = plan.run(z)
    d = torch.from_numpy(plan.delays()[:, c].astype(np.int64)).cuda()
    hit = out.nonzero()
    assert hit.shape[0] == 2048
    assert torch.equal(hit[:, 1], s - d)


@pytest.mark.parametrize("dtype", ["f32", "f16", "u8"])
@pytest.mark.parametrize("tile", [(32, 512, 2), (64, 256, 4), (16, 512, 8), (32, 256, 16), (8, 512, 0),
                                  (64, 128, 3), (16, 512, 3, 16), (32, 512, 1, 16), (32, 512, 0, 16),
                                  (16, 256, 0, 16), (32, 256, 0, 16), (64, 512, 0, 16), (64, 256, 0, 16)])
def test_all_inputs_all_tiles_exact(orc, ppl, dtype, tile):
    """Every instantiated tile, channel block and read width, odd channel
    counts and odd delay parities (the FP16 form reads pairs from a shifted
    copy, the 16-byte FP32 form from four copies shifted by 0..3, staged
    through its three-stage ring: 53 channels end part way through a round
    of stages)."""
    # 18 samples of delay per trial in the lowest channel; a third of that for
    # the 16-byte form, whose ring stage holds a spread of up to 256 samples
    # at 512-sample tiles
    cfg = dict(n_chan=53, nsamp_buf=3000, tdms=77, dm_low=0.0, dm_step=0.3 if len(tile) == 4 else 0.9,
               tsamp=0.00032768, fch1=150.0, foff=-0.05)
    rng = np.random.default_rng(hash(tile) % 1000)
    x = rng.normal(size=(cfg["n_chan"], cfg["nsamp_buf"])).astype(np.float32)
    p, d = _oracle(orc, x, cfg)
    if dtype == "f32":
        dev, host = torch.from_numpy(x).cuda(), x
    elif dtype == "f16":
        host = orc.f32_to_f16_bits(x)
        dev = torch.from_numpy(host.view(np.float16)).cuda()
    else:
        host = orc.quantise_u8(x, np.float32(12.7))
        dev = torch.from_numpy(host).cuda()
    plan = ppl.DedispPlan(**cfg, dtype=dtype)
    if len(tile) == 4 and (dtype == "f16" or (dtype == "u8" and tile[0] > 32)):
        with pytest.raises(ppl.PPLError):  # the ring kernels take FP32, and u8 up to 32 trials
            plan.set_tile(*tile)
        return
    plan.set_tile(*tile)
    if dtype == "u8" and tile[1] % 256:
        with pytest.raises(ppl.PPLError):  # the u8 form needs 4 samples per lane
            plan.run(dev)
        return
    out = plan.run(dev).cpu().numpy()
    assert np.array_equal(out, orc.dedisperse(host, d, p["nout"]))


@pytest.mark.parametrize("dtype,tile", [("u8", (32, 512, 0, 4)), ("u8", (16, 256, 3, 4)), ("f16", (32, 512, 0, 4)),
                                        ("f16", (16, 512, 2, 4)), ("f32", (32, 512, 0, 16))])
def test_last_partial_dword_of_the_input(orc, ppl, dtype, tile):
    """A buffer whose byte count is not a multiple of four, with the delays
    of the top trial reaching the last sample of the last channel: the
    kernels that stage whole dwords read the input's final, partial dword
    byte by byte."""
    # dm_step 0.31: maxshift 391, and the top trial's delay in the last channel is 391 too
    base = dict(n_chan=53, tdms=64, dm_low=0.0, dm_step=0.31, tsamp=0.00032768, fch1=150.0, foff=-0.05)
    cfg = None
    for nsamp in range(2001, 3000, 2):  # odd lengths: 53 * nsamp is odd
        c = dict(base, nsamp_buf=nsamp)
        p, d = _oracle(orc, None, c)
        if p["nout"] - 1 + int(d.max()) == nsamp - 1:
            cfg = c
            break
    assert cfg is not None
    rng = np.random.default_rng(nsamp)
    x = rng.normal(size=(cfg["n_chan"], cfg["nsamp_buf"])).astype(np.float32)
    x[-1, -1] = 9.0  # the last sample, read by the top trial's last output
    if dtype == "f32":
        dev, host = torch.from_numpy(x).cuda(), x
    elif dtype == "f16":
        host = orc.f32_to_f16_bits(x)
        dev = torch.from_numpy(host.view(np.float16)).cuda()
    else:
        host = orc.quantise_u8(x, np.float32(12.7))
        dev = torch.from_numpy(host).cuda()
    plan = ppl.DedispPlan(**cfg, dtype=dtype)
    plan.set_tile(*tile)
    out = plan.run(dev).cpu().numpy()
    assert np.array_equal(out, orc.dedisperse(host, d, p["nout"]))


def test_u8_ring_refuses_a_ragged_input(ppl):
    """The u8 ring kernel stages whole dwords, so it is refused when the
    input's byte count is not a multiple of 4 (the one-word u8 kernel takes
    such inputs); the autotuner then keeps another tile."""
    cfg = dict(n_chan=53, nsamp_buf=3001, tdms=40, dm_low=0.0, dm_step=0.3, tsamp=0.00032768, fch1=150.0,
               foff=-0.05)
    plan = ppl.DedispPlan(**cfg, dtype="u8")
    plan.set_tile(32, 512, 0, 16)
    with pytest.raises(ppl.PPLError, match="multiple of 4"):
        plan.run(torch.zeros((53, 3001), dtype=torch.uint8, device="cuda"))


def test_wide_reads_reject_a_spread_too_wide(ppl):
    """A DM tile whose delay spread (about 4600 samples over 32 trials) needs
    a staged row longer than a ring stage holds (768 samples a copy; 1024
    for 64 trials) is refused, not truncated; 16-byte tiles that are not
    instantiated are refused when set."""
    cfg = dict(n_chan=64, nsamp_buf=16384, tdms=64, dm_low=0.0, dm_step=6.0, tsamp=0.00032768,
               fch1=150.0, foff=-0.05)
    plan = ppl.DedispPlan(**cfg)
    for tile in [(32, 512, 0, 16), (64, 512, 0, 16)]:
        plan.set_tile(*tile)
        with pytest.raises(ppl.PPLError):
            plan.run(torch.zeros((64, 16384), device="cuda"))
    for tile in [(128, 512), (16, 1024), (8, 512)]:
        with pytest.raises(ppl.PPLError, match="not instantiated"):
            plan.set_tile(*tile, 0, 16)


def test_autotuner_keeps_results_exact(orc, ppl):
    cfg = dict(SMALL)
    rng = np.random.default_rng(17)
    x = rng.normal(size=(cfg["n_chan"], cfg["nsamp_buf"])).astype(np.float32)
    plan = ppl.DedispPlan(**cfg)
    xd = torch.from_numpy(x).cuda()
    out = torch.empty((cfg["tdms"], plan.nout), device="cuda")
    ms = plan.autotune(xd, out, reps=1)
    assert ms > 0 and plan.tile[0] % 4 == 0
    p, d = _oracle(orc, x, cfg)
    assert np.array_equal(plan.run(xd).cpu().numpy(), orc.dedisperse(x, d, p["nout"]))


def _beam_plan(ppl, n_sub, n_ch, nsamp, tdms, dm_hi, dtype, n_int=4):
    n_chan = n_sub * n_ch
    foff = -0.1953125 / n_ch  # 195.3125 kHz subbands
    tsamp = 5.12e-6 * n_ch * n_int
    return ppl.DedispPlan(n_chan, nsamp, tdms, 0.0, dm_hi / tdms, tsamp, 150.0, foff, dtype)


@pytest.mark.parametrize("dtype", ["f32", "f16", "u8"])
def test_batch_equals_separate_runs(ppl, dtype):
    """Beams with different channel counts, buffer lengths, DM counts and
    ranges in one launch give, bit for bit, the outputs of separate runs,
    under the plans' own tile and under a forced one."""
    import torch
    shapes = [(8, 16, 4096, 128, 50.0), (31, 16, 8192, 512, 100.0), (12, 64, 8192, 300, 200.0),
              (4, 256, 4096, 64, 20.0)]
    rng = np.random.default_rng(9)
    plans, ins = [], []
    for n_sub, n_ch, nsamp, tdms, dm_hi in shapes:
        p = _beam_plan(ppl, n_sub, n_ch, nsamp, tdms, dm_hi, dtype)
        n_chan = n_sub * n_ch
        if dtype == "u8":
            x = torch.from_numpy(rng.integers(0, 256, size=(n_chan, nsamp), dtype=np.uint8))
        else:
            x = torch.from_numpy(rng.normal(size=(n_chan, nsamp)).astype(np.float32))
            if dtype == "f16":
                x = x.half()
        plans.append(p)
        ins.append(x.cuda())
    single = [p.run(x) for p, x in zip(plans, ins)]
    tiles = [(0, 0, 0), (16, 256, 4), (32, 512, 2)]
    if dtype == "f32":
        # the 31x16-channel beam spreads 292 samples over 32 trials: more than
        # a 512-sample tile's ring stage holds, so that tile is refused
        tiles += [(16, 512, 0, 16), (16, 256, 0, 16), (32, 256, 0, 16)]
        with pytest.raises(ppl.PPLError, match="too large"):
            ppl.dedisperse_batch(plans, ins, tile=(32, 512, 2, 16))
    elif dtype == "u8":
        # the u8 ring stages a row of up to 1020 bytes: every ring tile fits
        tiles += [(32, 512, 2, 16), (16, 512, 0, 16), (16, 256, 0, 16), (32, 256, 0, 16)]
    else:
        with pytest.raises(ppl.PPLError):
            ppl.dedisperse_batch(plans, ins, tile=(32, 512, 2, 16))
    for tile in tiles:
        outs = ppl.dedisperse_batch(plans, ins, tile=tile)
        for k in range(len(plans)):
            assert torch.equal(outs[k], single[k]), (tile, k)
    tile, ms = ppl.autotune_batch(plans, ins, outs)
    assert ms > 0 and tile[0] % 4 == 0 and tile[1] % 64 == 0
    outs = ppl.dedisperse_batch(plans, ins, tile=tile)
    for k in range(len(plans)):
        assert torch.equal(outs[k], single[k]), ("tuned", tile, k)


def test_batch_rejects_mixed_dtypes_and_too_many_jobs(ppl):
    import torch
    a = _beam_plan(ppl, 8, 16, 4096, 128, 50.0, "f32")
    b = _beam_plan(ppl, 8, 16, 4096, 128, 50.0, "f16")
    xa = torch.zeros((128, 4096), device="cuda")
    xb = torch.zeros((128, 4096), device="cuda", dtype=torch.float16)
    with pytest.raises(ppl.PPLError):
        ppl.dedisperse_batch([a, b], [xa, xb])
    with pytest.raises(ppl.PPLError):
        ppl.dedisperse_batch([a] * 49, [xa] * 49)


def test_tile_setters_of_the_c_abi(orc, ppl):
    """ppl_dedisp_set_tile and ppl_dedisp_set_shape (the 4-byte-read forms of
    ppl_dedisp_set_kernel) set the tile, keep the sums exact, and refuse
    tiles that are not instantiated."""
    cfg = dict(SMALL)
    rng = np.random.default_rng(23)
    x = rng.normal(size=(cfg["n_chan"], cfg["nsamp_buf"])).astype(np.float32)
    p, d = _oracle(orc, x, cfg)
    exp = orc.dedisperse(x, d, p["nout"])
    plan = ppl.DedispPlan(**cfg)
    L = ppl.lib()
    assert L.ppl_dedisp_set_tile(plan.h, 32, 256) == 0
    plan._refresh()
    assert plan.tile[:2] == (32, 256) and plan.tile[3] == 4
    assert np.array_equal(plan.run(torch.from_numpy(x).cuda()).cpu().numpy(), exp)
    assert L.ppl_dedisp_set_shape(plan.h, 16, 512, 3) == 0
    plan._refresh()
    assert plan.tile == (16, 512, 3, 4)
    assert np.array_equal(plan.run(torch.from_numpy(x).cuda()).cpu().numpy(), exp)
    assert L.ppl_dedisp_set_tile(plan.h, 12, 512) != 0      # 12 DMs: not instantiated
    assert L.ppl_dedisp_set_shape(plan.h, 32, 512, 17) != 0  # channel block above 16
    plan._refresh()
    assert plan.tile == (16, 512, 3, 4)                     # a refused call changes nothing


def test_u8_series_matches_oracle_and_restores_rejected_rows(orc, ppl):
    """ppl_dedisp_u8_series on a buffer with rejected spectra: bit for bit the
    oracle's series, and within the quantisation bound of the FP32 sums of
    the unquantised buffer everywhere -- also where a rejected spectrum lands,
    which plain dequantisation misses by thousands.  (The bound holds while a
    rejected spectrum's spread about its middle channel fits the 8-bit range,
    here 5 units against 10: the clipper's rejected spectra are
    -0.2 - bandpass - dataModel, flat for a flat bandpass model.)"""
    cfg = dict(SMALL)
    inv = np.float32(127.0 / 10.0)
    rng = np.random.default_rng(71)
    n, ns = cfg["n_chan"], cfg["nsamp_buf"]
    x = rng.normal(size=(n, ns)).astype(np.float32)
    rej = rng.choice(ns, size=20, replace=False)
    x[:, rej] = (-2500.0 - 5.0 * rng.random(size=(n, rej.size))).astype(np.float32)
    q, off = orc.quantise_u8_rows(x, inv)
    assert np.count_nonzero(off) == rej.size
    p, d = _oracle(orc, x, cfg)
    plan = ppl.DedispPlan(**cfg, dtype="u8")
    sums = plan.run(torch.from_numpy(q).cuda())
    assert np.array_equal(sums.cpu().numpy(), orc.dedisperse(q, d, p["nout"]))
    series = plan.u8_series(sums, torch.from_numpy(off).cuda(), float(1.0 / inv)).cpu().numpy()
    want = orc.u8_series(sums.cpu().numpy(), off, d, p["nout"], np.float32(1.0 / inv))
    assert np.array_equal(series, want)
    ref = orc.dedisperse(x, d, p["nout"])
    bound = n * 0.5 / float(inv) * 1.001 + 1e-5 * np.abs(ref)
    assert np.all(np.abs(series - ref) <= bound)
    plain = (sums.cpu().numpy() - 128 * n).astype(np.float32) * np.float32(1.0 / inv)
    assert np.max(np.abs(plain - ref)) > 1000.0
