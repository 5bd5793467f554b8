#!/usr/bin/env python3
"""Benchmark of the MI355X pulsar-search chain on one GPU.

Workload (the LOFAR-scale single beam): 62 subbands x 2 polarisations of
complex int16 voltages at 5.12 us, 138-150 MHz; 64-channel 8-tap polyphase
channeliser (3968 channels), Stokes-I integrated x4 (tsamp 1.31072 ms), RFI
clipper (thresholds 10/4, history 10000, zero-DM, flat .bp model), brute-force
dedispersion of 2048 DM trials 0-500 pc/cm^3 over 65536-sample buffers, and
dyadic boxcar detection at 6 sigma.  The synthetic voltages carry Gaussian
noise (sigma 64), 10 narrowband sinusoid channels, broadband impulses in 0.5%
of the integrated samples at 20 sigma and 5 dispersed pulses per buffer (DM
50-450, 1-8 ms, 6-12 sigma per channel).

One step is one pass of the chain over one buffer of new data: nOut = 52640
integrated spectra (13,475,840 input samples per subband and polarisation),
channelised, clipped, transposed into the resident dedispersion buffer,
dedispersed over all 2048 trials and searched; the buffer's last
maxshift+remainder samples carry to the next step.  Inputs are generated on
the device before timing (a ring of distinct chunks).

Prints one JSON line.  value = DM trials x output samples per second of the
whole chain (tdms * nOut * steps / wall), with the real-time factor, per-stage
times, the dedispersion roofline and the CPU baseline alongside.  The steps
are timed one after the other on one stream, and in the overlapped schedule a
streaming deployment runs (step k's clipper walk, one workgroup, on CU slot 0
of every XCD beside step k-1's dedispersion on the other slots), timed in its
steady state over the same number of steps; --two-stream and --cu-split N
time two more pipelined schedules.  value is the fastest schedule's, and
every timed schedule is reported.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DM-trials×samples/s + real-time factor, 1 GPU, mixed vs plain at stated error"

# workload: LOFAR-scale single beam
N_SUB, N_CH, N_TAPS, INTEG = 62, 64, 8, 4
TS_IN = 5.12e-6
TSAMP = TS_IN * N_CH * INTEG
FCH1, FOFF = 150.0, -0.1953125 / N_CH
TDMS, DM_STEP, NSAMP_BUF = 2048, 500.0 / 2048, 65536
SIGMA = 64.0
N_CHAN = N_SUB * N_CH

PEAK_TFLOPS_FP32 = 157.3   # MI355X FP32 (vector/MFMA) peak
PEAK_HBM_GBS = 8000.0


def log(msg):
    """Progress on stderr (the JSON line is the only stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dedisp-dtype", default="f32", choices=["f32", "f16", "u8"])
    ap.add_argument("--chunks", type=int, default=2, help="distinct input chunks kept in HBM")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP16 / u8 comparison chains")
    ap.add_argument("--tile", default="", help="DM,T,chanblock[,readbytes] tile for dedispersion (default: autotune)")
    ap.add_argument("--cpu-spectra", type=int, default=8192,
                    help="integrated spectra in the timed CPU sample")
    ap.add_argument("--cpu-passes", type=int, default=5,
                    help="timed passes of the CPU baseline over its sample (the stream continues)")
    ap.add_argument("--beams", type=int, default=24, help="beams in the batched multi-beam run")
    ap.add_argument("--mb-steps", type=int, default=2, help="timed steps of the multi-beam run")
    ap.add_argument("--no-multibeam", action="store_true", help="skip the batched multi-beam run")
    ap.add_argument("--no-service", action="store_true", help="skip the streaming-service run")
    ap.add_argument("--two-stream", action="store_true",
                    help="also time the two-stream schedule (front stages of step k beside the "
                         "dedispersion of step k-1); off by default, since it is slower on one MI355X")
    ap.add_argument("--cu-split", type=int, default=0,
                    help="also time the CU-split schedule with this many CU slots per XCD for the clipper, "
                         "buffer fill and carry (the rest run the channeliser, dedispersion and detection); "
                         "0: off")
    ap.add_argument("--no-overlapped", action="store_true",
                    help="skip the overlapped schedule (clipper walk beside the previous step's dedispersion)")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 256-channel split-precision run")
    ap.add_argument("--prec-steps", type=int, default=2, help="timed steps per chain of the 256-channel run")
    ap.add_argument("--svc-seconds", type=float, default=30.0, help="wall seconds of the streaming run")
    ap.add_argument("--svc-accel", type=float, default=10.0,
                    help="data seconds per wall second in the streaming run (arrival schedule speed-up)")
    ap.add_argument("--svc-dtype", default="f16", choices=["f32", "f16", "u8"])
    ap.add_argument("--svc-recall-sessions", type=int, default=16,
                    help="sessions replayed through plain FP32 chains for the event recall")
    return ap.parse_args()


def gen_chunk(torch, ppl_coeffs, n_samp, sample0, seed, pulses, rfi, impulses=True):
    """int16 voltages [sub][pol][n_samp][2] on the device.  sample0 is the
    chunk's first sample in the stream (tone phases); a pulse (dm, t0, width,
    amp) starts t0 integrated spectra after the chunk's start in the top
    channel and arrives later in lower channels by its dispersion delay; what
    falls past the chunk's end is dropped."""
    from tests.synth import shift_table, tone_amplitude
    g = torch.Generator(device="cuda").manual_seed(seed)
    out = torch.empty((N_SUB, 2, n_samp, 2), dtype=torch.int16, device="cuda")
    shift = torch.from_numpy(shift_table(N_CHAN, FCH1, FOFF)).cuda()
    spi = N_CH * INTEG  # input samples per integrated sample
    n_int = n_samp // spi
    imp = torch.rand(n_int, generator=g, device="cuda") < (0.005 if impulses else 0.0)  # broadband impulses
    imp_idx = torch.nonzero(imp).flatten()
    for s in range(N_SUB):
        x = torch.randn((2, n_samp, 2), generator=g, device="cuda") * SIGMA
        if imp_idx.numel():
            seg = (imp_idx[:, None] * spi + torch.arange(spi, device="cuda")[None, :]).flatten()
            x[:, seg, :] += torch.randn((2, seg.numel(), 2), generator=g, device="cuda") * (20.0 * SIGMA)
        m = torch.arange(n_samp, device="cuda", dtype=torch.float64) + sample0
        for (rs, rk, ramp) in rfi:
            if rs != s:
                continue
            nu = (rk - N_CH / 2.0) / N_CH
            ph = 2 * np.pi * nu * m
            x[:, :, 0] += (ramp * torch.cos(ph)).float()[None, :]
            x[:, :, 1] += (ramp * torch.sin(ph)).float()[None, :]
        for (dm, t0, width, amp) in pulses:
            d = shift[s * N_CH:(s + 1) * N_CH] * dm / TSAMP
            a = torch.round((t0 + d) * spi).long()
            L = int(round(width * spi))
            idx = a[:, None] + torch.arange(L, device="cuda")[None, :]
            ok = (idx >= 0) & (idx < n_samp)
            k = torch.arange(N_CH, device="cuda", dtype=torch.float64)
            nu = ((k - N_CH / 2.0) / N_CH)[:, None]
            ph = 2 * np.pi * nu * (idx.double() + sample0)
            for p in range(2):
                re = (amp * torch.cos(ph + p)).float()[ok]
                im = (amp * torch.sin(ph + p)).float()[ok]
                x[p, :, 0].index_add_(0, idx[ok], re)
                x[p, :, 1].index_add_(0, idx[ok], im)
        out[s] = torch.clamp(torch.round(x), -32768, 32767).to(torch.int16)
    return out


def traffic_from_profile(dtype):
    """HBM bytes per dedispersion launch from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a copy of known size),
    or None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "dedisperse_lofar_beam_counters.md")
    try:
        with open(path) as f:
            for ln in f:
                w = ln.split()
                if len(w) == 3 and w[0] == "traffic_bytes_per_launch" and w[1] == dtype:
                    return float(w[2])
    except OSError:
        pass
    return None


def run_steps(chain, chunks, n, timing=None, graph=False):
    """n steps of a chain over the chunk ring; returns per-step event lists."""
    import torch
    evs = []
    for k in range(n):
        ev = None
        if timing is not None:
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(7)]
            ev[0].record()
        if graph:
            chain.replay()
            if ev is not None:
                for e in ev[1:6]:
                    e.record()
        else:
            chain.step(chunks[k % len(chunks)], None if ev is None else ev[1:6])
        ints, vals, _ = chain.detect()
        if timing is not None:
            ev[6].record()
            timing.append(ev)
        evs.append((ints, vals))
    return evs


def event_keys(ints):
    return {tuple(r) for r in ints[1:]}  # without the dummy event


def last_step_keeping_input(chain, volt):
    """One step of a BeamChain by its parts, returning a copy of the
    dedispersion input (the buffer, before its carry) of this step."""
    chain.ppf.stokes_i16(volt, chain.nout * chain.integrate, d_out=chain.stokes)
    chain.clip.run(chain.stokes)
    chain.buf.add(chain.stokes.view(chain.nout, chain.n_chan))
    keep = chain.buf.tensor().clone()
    chain._dedisperse(chain.buf.tensor(), chain.buf.row_offsets(), chain.out, chain.series)
    chain.buf.rotate()
    chain.parity ^= 1
    return keep


def fp64_sample(torch, buf, plan, n_trials):
    """FP64 dedispersion of a [nChan][nsamp] buffer (any dtype, taken as its
    values; u8 codes are not meant here) over n_trials trials spread across
    the range: (trial indices, [n_trials][nOut] float64)."""
    delays = torch.from_numpy(plan.delays()).cuda().long()
    tdms, n_chan = delays.shape
    trials = torch.linspace(0, tdms - 1, n_trials, device="cuda").round().long()
    x = buf.double()
    nout = plan.nout
    out = torch.zeros((n_trials, nout), dtype=torch.float64, device="cuda")
    t = torch.arange(nout, device="cuda")
    for k, dm in enumerate(trials.tolist()):
        for c0 in range(0, n_chan, 512):
            c1 = min(c0 + 512, n_chan)
            idx = t[None, :] + delays[dm, c0:c1, None]
            out[k] += x[c0:c1].gather(1, idx).sum(0)
    return trials, out


def count_recovered(step_events, pulses, step0, n_chunks, nout, overlap):
    """Injected pulses found in the steps' event lists: a pulse of chunk c at
    integrated spectrum t0 reaches the series of the step that reads c at
    index overlap + t0 (channel 0 has no delay), or the next step's at
    overlap + t0 - nOut; it counts as found if an event lies within 10
    pc/cm^3 of its DM and within its width plus one 64-sample bin of that
    index.  Returns "found/injected"."""
    found = total = 0
    for k, (ints, _) in enumerate(step_events):
        ci = (step0 + k) % n_chunks
        for (dm, t0, width, _amp) in pulses[ci]:
            t = overlap + int(round(t0))
            step_t = (k, t) if t < nout else (k + 1, t - nout)
            if step_t[0] >= len(step_events):
                continue
            total += 1
            ev = step_events[step_t[0]][0][1:]
            if len(ev) == 0:
                continue
            ok = (np.abs(ev[:, 0] * DM_STEP - dm) <= 10.0) & (np.abs(ev[:, 1] - step_t[1]) <= width + 64)
            found += bool(np.any(ok))
    return "%d/%d" % (found, total)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the hot path is measured on one MI355X, and a "
              "multi-GPU layout would place beams or DM ranges one per GPU; it is "
              "not launched here.  Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as ge
    from tests.synth import tone_amplitude

    ppl = ge.load_package()
    ppl.init(0)

    coeffs = ppl.ppf_coefficients(N_TAPS, N_CH, "hann")
    e_h = float(np.sum(coeffs * coeffs))
    level = 2 * INTEG * 2.0 * SIGMA * SIGMA * e_h          # mean integrated Stokes-I
    noise_std = np.sqrt(2 * INTEG) * 2.0 * SIGMA * SIGMA * e_h
    bp = np.array([N_CHAN, FCH1, FCH1 + FOFF * (N_CHAN - 1), FOFF, noise_std, level], np.float64)
    inv_scale = 127.0 / 10.0  # u8: the clipper's normalised range (crFactor = 10 sigma) over 127 steps

    def make_chain(dtype):
        tile = [int(v) for v in args.tile.split(",")] if args.tile and dtype == args.dedisp_dtype else None
        return ppl.BeamChain(N_SUB, N_CH, N_TAPS, coeffs, INTEG, bp, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP,
                             FCH1, FOFF, dtype=dtype, inv_scale=inv_scale, tile=tile)

    dd_dtype = args.dedisp_dtype
    chain = make_chain(dd_dtype)
    nout, overlap = chain.nout, chain.overlap
    n_samp = chain.samples_per_step()
    def make_chunks(seed, impulses, with_pulses=False):
        rng = np.random.default_rng(seed)
        out, all_pulses = [], []
        for ci in range(args.chunks):
            pulses = []
            for _ in range(5):
                dm = rng.uniform(50, 450)
                width = rng.uniform(1e-3, 8e-3) / TSAMP
                amp_sig = rng.uniform(6, 12)
                t0 = rng.uniform(0.05, 0.8) * nout
                pulses.append((dm, t0, width, tone_amplitude(coeffs, N_CH, SIGMA, amp_sig, INTEG)))
            rfi = [(int(rng.integers(N_SUB)), int(rng.integers(N_CH)),
                    tone_amplitude(coeffs, N_CH, SIGMA, 50.0, INTEG)) for _ in range(10)]
            out.append(gen_chunk(torch, coeffs, n_samp, ci * n_samp, seed + ci, pulses, rfi, impulses))
            all_pulses.append(pulses)
        return (out, all_pulses) if with_pulses else out

    log("generating inputs")
    t_gen = time.time()
    chunks = make_chunks(2024, True)
    prime = chunks[-1][:, :, :overlap * INTEG * N_CH].contiguous()
    torch.cuda.synchronize()
    t_gen = time.time() - t_gen

    log("autotune + prime")
    tune_ms = None if args.tile else chain.autotune()
    chain.prime(prime)
    run_steps(chain, chunks, args.warmup)
    torch.cuda.synchronize()

    # ---- timed: the plain (FP32) chain, eager
    log("timed steps")
    timing = []
    t0 = time.perf_counter()
    plain_events = run_steps(chain, chunks, args.steps, timing)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    last_plain = chain.out.clone() if dd_dtype == "f32" else None
    stage_names = ["ppf_stokes", "clipper", "buffer", "dedisperse", "carry", "detect"]
    stages = {n: sum(ev[j].elapsed_time(ev[j + 1]) for ev in timing) / args.steps
              for j, n in enumerate(stage_names)}

    ms_per_step = 1e3 * wall / args.steps
    value = TDMS * nout * args.steps / wall
    rtf = nout * TSAMP * args.steps / wall

    # ---- the same steps replayed from HIP graphs (detection stays eager)
    log("graph replay")
    chain.capture_steps([chunks[(args.warmup + args.steps + k) % len(chunks)] for k in range(2)])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_steps(chain, chunks, args.steps, graph=True)
    torch.cuda.synchronize()
    graph_ms = 1e3 * (time.perf_counter() - t0) / args.steps

    # ---- pipelined schedules (a streaming deployment's order: step k's
    # front work beside step k-1's dedispersion and detection), timed in
    # their steady state: one run over warm + K + 1 chunks, the window from
    # the detection of step warm-1 to that of step warm+K-1, with a
    # synchronisation at both ends.  The window holds exactly K steps'
    # work: K PPF/clipper/buffer stages (steps warm+1 .. warm+K; step warm's
    # finished at the first synchronisation) and K dedispersions and
    # detections (steps warm .. warm+K-1).
    def pipelined(run, base):
        warm = max(2, args.warmup)
        n = warm + args.steps + 1
        res, t0, t1 = [], None, None
        for r in run(chunks[(base + k) % len(chunks)] for k in range(n)):
            res.append(r)
            if len(res) == warm:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            elif len(res) == warm + args.steps:
                torch.cuda.synchronize()
                t1 = time.perf_counter()
        w = t1 - t0
        return {"ms_per_step": 1e3 * w / args.steps, "value": TDMS * nout * args.steps / w,
                "real_time_factor": nout * TSAMP * args.steps / w,
                "events_per_step": [len(e[0]) for e in res[warm:warm + args.steps]]}

    # the two-stream schedule: PPF, clipper, buffer fill and carry on one
    # stream, dedispersion and detection on another
    pipe = None
    if args.two_stream:
        log("timed steps, two-stream schedule")
        pipe = pipelined(chain.run_pipelined, args.warmup + args.steps)
    # the CU-split schedule: as above, with the clipper, buffer fill and
    # carry on their own CU slots of every XCD
    split = None
    if args.cu_split:
        log("timed steps, CU-split schedule (%d CU slots per XCD)" % args.cu_split)
        split = pipelined(lambda it: chain.run_split(it, small_cus=args.cu_split), args.warmup + 2 * args.steps)
        split["cu_slots_per_xcd"] = args.cu_split
    # the overlapped schedule: step k's clipper walk (one workgroup) on CU
    # slot 0 of every XCD beside step k-1's dedispersion on the other slots;
    # PPF, the clipper's precompute and write passes and the buffer work on
    # the whole GPU
    over = None
    if not args.no_overlapped:
        log("timed steps, overlapped schedule")
        over = pipelined(chain.run_overlapped, args.warmup + 3 * args.steps)
    sequential = {"ms_per_step": ms_per_step, "value": value, "real_time_factor": rtf}
    schedule = "one-stream"
    for name, alt in (("two-stream", pipe), ("cu-split", split), ("overlapped", over)):
        if alt is not None and alt["value"] > value:
            ms_per_step, value, rtf = alt["ms_per_step"], alt["value"], alt["real_time_factor"]
            schedule = name

    # ---- roofline of the dedispersion kernel: algorithmic adds / kernel time
    adds = float(TDMS) * N_CHAN * nout
    elem = {"f32": 4, "f16": 2, "u8": 1}[dd_dtype]
    alg_bytes = N_CHAN * NSAMP_BUF * elem + TDMS * nout * 4
    t_dd = stages["dedisperse"] * 1e-3
    achieved = adds / t_dd / 1e12
    kname = {"f32": "dedisperse_kernel", "f16": "dedisperse_f16_kernel", "u8": "dedisperse_u8_kernel"}[dd_dtype]
    if dd_dtype == "f32" and chain.plan.tile[3] == 16:
        kname = "dedisperse_ring_kernel"
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_TFLOPS_FP32, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_TFLOPS_FP32, 4), "traffic": traffic_from_profile(dd_dtype),
                "traffic_source": "profiles/dedisperse_lofar_beam_counters.md (rocprofv3 --pmc FETCH_SIZE, "
                                  "WRITE_SIZE per dispatch, scaled by a copy of known size; not measured in this run)",
                "kernel": kname, "ops_per_launch": adds, "alg_bytes_per_launch": alg_bytes,
                "hbm_gbs_algorithmic": round(alg_bytes / t_dd / 1e9, 1),
                "note": "ops are channel-order FP32 adds (no FMA, no packed FP32), priced against the FP32 peak: "
                        "one v_add_f32 per add plus one address add per eight (72 VALU instructions per 64 "
                        "adds); counters show 0.64 VALU wave-instructions per CU-clock against the 1.19 the "
                        "channeliser sustains, so LDS binds: the 16-byte reads plus the LDS-DMA staging of "
                        "four shifted copies of every row (a variant staging one copy ran 17 % faster)"}

    # ---- mixed precision: the same inputs through chains with FP16 / u8
    # dedispersion input (FP32 / int32 sums), against a plain FP32 chain.
    # On the workload above the rejected impulse spectra (-0.2 minus the
    # bandpass, never normalised, as in the reference) pull every
    # dedispersed series far below the threshold, so the comparison runs on
    # the same stream without the broadband impulses, where the pulses are
    # found.
    mixed = {}
    pulses_recovered = None
    if dd_dtype == "f32" and not args.no_mixed:
        pchunks, ppulses = make_chunks(4048, False, with_pulses=True)
        pprime = pchunks[-1][:, :, :overlap * INTEG * N_CH].contiguous()
        nrm = float(np.sqrt(N_CHAN))  # series noise in the clipper's normalised units
        ref_out, ref_events = None, None
        fp64 = None  # (trial indices, FP64 sums of the plain chain's last buffer)
        for mdt in ("f32", "f16", "u8"):
            log("mixed-precision chain " + mdt)
            mc = make_chain(mdt)
            if mdt == "f32":
                mc.plan.set_tile(*chain.plan.tile)
                m_tune = None
            else:
                m_tune = mc.autotune()
            mc.prime(pprime)
            run_steps(mc, pchunks, args.warmup)
            torch.cuda.synchronize()
            mt = []
            t0 = time.perf_counter()
            mev = run_steps(mc, pchunks, args.steps, mt)
            torch.cuda.synchronize()
            mwall = time.perf_counter() - t0
            rec = {"value": TDMS * nout * args.steps / mwall, "ms_per_step": 1e3 * mwall / args.steps,
                   "dedisperse_ms": round(sum(ev[3].elapsed_time(ev[4]) for ev in mt) / args.steps, 3),
                   "tile_dm_t_chanblock_readbytes": list(mc.plan.tile),
                   "autotune_ms": None if m_tune is None else round(m_tune, 3),
                   "events": sum(len(e[0]) - 1 for e in mev)}
            # one more step by hand, keeping the dedispersion input: its series
            # against the FP64 sums of the plain chain's input on a DM sample
            keep = last_step_keeping_input(mc, pchunks[args.steps % len(pchunks)])
            series = (mc.series if mdt == "u8" else mc.out).double()
            if mdt == "f32":
                fp64 = fp64_sample(torch, keep, mc.plan, 16)
                pulses_recovered = count_recovered(mev, ppulses, 0, len(pchunks), nout, overlap)
            del keep
            trials, ref64 = fp64
            e64 = (series[trials] - ref64).abs()
            rec.update({"series_vs_fp64_max_abs_err_sigma": float(e64.max()) / nrm,
                        "series_vs_fp64_rms_err_sigma": float(e64.pow(2).mean().sqrt()) / nrm})
            if mdt == "f32":
                ref_out, ref_events = series.clone(), mev
            else:
                err = (series - ref_out).abs()
                n_ref = n_got = inter = 0
                for (pe, _), (me, _) in zip(ref_events, mev):
                    rk, gk = event_keys(pe), event_keys(me)
                    n_ref += len(rk)
                    n_got += len(gk)
                    inter += len(rk & gk)
                rec.update({"series_max_abs_err_sigma": float(err.max()) / nrm,
                            "series_rms_err_sigma": float(err.pow(2).mean().sqrt()) / nrm,
                            "event_recall": inter / max(n_ref, 1), "event_precision": inter / max(n_got, 1)})
            mixed[mdt] = rec
            del mc, series
            torch.cuda.empty_cache()
        del fp64
        mixed["stream"] = "the workload without broadband impulses (pulses and narrowband RFI only)"
        mixed["error_unit"] = "sigma of a dedispersed sample (sqrt(nChan) in normalised units)"
        mixed["fp64_reference"] = ("FP64 sums, over 16 trials spread across the DM range and every output "
                                   "sample, of the plain chain's last dedispersion input (the clipper's FP32 "
                                   "output, which every chain here shares); the series_* errors without "
                                   "vs_fp64 compare against the plain FP32 chain's series")
        mixed["note"] = ("f16 rounds each clipped sample to 11 bits; u8 quantises to 1/12.7 over +-10 sigma, "
                         "and spectra most of whose samples leave that range (the clipper's rejected spectra, "
                         "-0.2 minus the bandpass) are quantised relative to their middle channel and restored "
                         "at their own level in the series (ppl_dedisp_u8_series)")

    mbeam = None
    if not args.no_multibeam:
        mbeam = multibeam(args, torch, ppl)

    prec = None
    if not args.no_precision:
        prec = precision(args, torch, ppl)

    svc = None
    if not args.no_service:
        svc = service(args, torch, ppl, coeffs, bp)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(args.cpu_spectra, coeffs, bp, nout, args.cpu_passes)

    # a device copy of known size for scaling the HBM byte counters
    a = torch.empty(1 << 28, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DM-trials*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dd_dtype,
        "data": "synthetic (seeded int16 voltages: noise, narrowband RFI, impulses, 5 pulses/buffer)",
        "config": {"workload": "lofar-beam: 62 sub x 2 pol x 64 ch (3968 ch), 8-tap PPF, Stokes-I x4, "
                               "RFI clipper, 2048 DMs 0-500 on 65536-sample buffers, 6-sigma detection",
                   "nOut_per_step": nout, "input_samples_per_stream_per_step": n_samp,
                   "dedisp_input": dd_dtype, "dft": "fp32", "tile_dm_t_chanblock_readbytes": list(chain.plan.tile),
                   "autotune_ms": None if tune_ms is None else round(tune_ms, 3)},
        "real_time_factor": rtf,
        "schedule": schedule,
        "one_stream": sequential,
        "two_stream": pipe,
        "cu_split": split,
        "overlapped": over,
        "stages_ms": {k: round(v, 3) for k, v in stages.items()},
        "graph_ms_per_step": round(graph_ms, 3),
        "events_per_step": [len(e[0]) for e in plain_events],
        "events_note": ("on this stream the clipper rejects the 0.5 % impulse spectra and leaves them at -0.2 "
                        "minus the bandpass, as the reference does (RFI_Clipper.cpp:127-150,488); nearly every "
                        "DM trial's series then lands on one and stays far below the threshold, so a step's "
                        "event list is the dummy event alone; pulses_recovered counts the injected pulses "
                        "found on the same stream without the impulses (the mixed run)"),
        "pulses_recovered": pulses_recovered,
        "clipper_walk": chain.clip.counters(),
        "mixed": mixed,
        "precision_256ch": prec,
        "multibeam": mbeam,
        "service": svc,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gen_seconds": round(t_gen, 1),
    }
    print(json.dumps(line), flush=True)


def ppf_fp64(volt, coeffs, n_ch, n_blk):
    """FP64 polyphase channeliser of int16 voltages [sub][pol][n][2]: FIR with
    the oldest block on tap 0 and zero history, then the unnormalised forward
    DFT; returns [n_blk][sub][pol][n_ch] complex."""
    x = volt[..., 0].astype(np.float64) + 1j * volt[..., 1].astype(np.float64)
    n_sub, n_pol = x.shape[:2]
    taps = coeffs.size // n_ch
    h = np.asarray(coeffs, np.float64).reshape(taps, n_ch)
    blocks = x[:, :, :n_blk * n_ch].reshape(n_sub, n_pol, n_blk, n_ch)
    pad = np.concatenate([np.zeros((n_sub, n_pol, taps - 1, n_ch)), blocks], axis=2)
    f = np.zeros_like(blocks)
    for t in range(taps):
        f += h[t][None, None, None, :] * pad[:, :, t:t + n_blk, :]
    return np.fft.fft(f, axis=-1).transpose(2, 0, 1, 3)


def precision(args, torch, ppl):
    """The LOFAR beam at 256 channels per subband (15872 channels, tsamp
    5.24 ms): the plain chain (FP32 DFT, FP32 dedispersion input) against
    split-precision DFTs on the matrix cores (BF16x3, FP16x2) with FP16 and
    8-bit dedispersion input, on the same voltages (pulses and narrowband
    RFI, no broadband impulses).  Reports per chain the step time and stage
    times, the channeliser's error against an FP64 channeliser on a sample,
    the dedispersed series' error against the plain chain and the event
    recall/precision against the plain chain's events."""
    from tests.synth import beam_bandpass, device_voltages, tone_amplitude
    n_ch = 256
    tsamp = TS_IN * n_ch * INTEG
    foff = -0.1953125 / n_ch
    n_chan = N_SUB * n_ch
    coeffs = ppl.ppf_coefficients(N_TAPS, n_ch, "hann")
    bp = beam_bandpass(coeffs, N_SUB, n_ch, INTEG, FCH1, foff, SIGMA)
    inv_scale = 127.0 / 10.0
    chains = [("fp32", "f32"), ("bf16x3", "f16"), ("fp16x2", "f16"), ("bf16x3", "u8")]
    # f32: autotuned; f16 and u8: the tiles the autotuner picks on the LOFAR beam (the u8 ring kernel)
    tiles = {"f32": None, "f16": (32, 512, 4), "u8": (32, 512, 0, 16)}
    # channeliser error on a sample: 2 subbands x 96 blocks against FP64
    rng = np.random.default_rng(77)
    sample = np.clip(np.rint(rng.normal(0, SIGMA, size=(2, 2, 96 * n_ch, 2))), -32768, 32767).astype(np.int16)
    z64 = ppf_fp64(sample, coeffs, n_ch, 96)
    rms64 = float(np.sqrt(np.mean(np.abs(z64) ** 2)))
    dft_err = {}
    for prec in ("fp32", "bf16x3", "fp16x2"):
        pp = ppl.PPF(2, n_ch, N_TAPS, coeffs.astype(np.float32), precision=prec)
        o = pp.spectra_i16(torch.from_numpy(sample).cuda(), 96).cpu().numpy().astype(np.float64)
        e = np.abs(o[..., 0] + 1j * o[..., 1] - z64)
        dft_err[prec] = {"max_err_rms": float(e.max() / rms64), "rms_err_rms": float(np.sqrt(np.mean(e ** 2)) / rms64)}
        del pp
    log("precision: 256-channel beam, generating one step of voltages")
    probe = ppl.DedispPlan(n_chan, NSAMP_BUF, TDMS, 0.0, DM_STEP, tsamp, FCH1, foff)
    nout, overlap = probe.nout, probe.maxshift + probe.rem
    del probe
    n_samp = nout * INTEG * n_ch
    pulses = []
    for _ in range(5):
        pulses.append((float(rng.uniform(50, 450)), float(rng.uniform(0.05, 0.8) * nout),
                       float(rng.uniform(1e-3, 8e-3) / tsamp), tone_amplitude(coeffs, n_ch, SIGMA, float(rng.uniform(6, 12)), INTEG)))
    rfi = [(int(rng.integers(N_SUB)), int(rng.integers(n_ch)), tone_amplitude(coeffs, n_ch, SIGMA, 50.0, INTEG))
           for _ in range(10)]
    volt = device_voltages(torch, N_SUB, n_ch, INTEG, n_samp, 909, FCH1, foff, SIGMA, pulses=pulses, rfi=rfi)
    prime = volt[:, :, :overlap * INTEG * n_ch].contiguous()
    torch.cuda.synchronize()
    out = {"workload": "62 sub x 256 ch (15872 ch), 8-tap PPF, Stokes-I x4 (tsamp %.3f ms), clipper, 2048 DMs 0-500 "
                       "on 65536-sample buffers, pulses + narrowband RFI" % (tsamp * 1e3),
           "nOut_per_step": nout, "data_s_per_step": round(nout * tsamp, 1), "dft_error_vs_fp64": dft_err}
    ref_series, ref_events = None, None
    fp64 = None
    nrm = float(np.sqrt(n_chan))
    for dft, dd in chains:
        log("precision: chain dft=%s dedisp=%s" % (dft, dd))
        ch = ppl.BeamChain(N_SUB, n_ch, N_TAPS, coeffs, INTEG, bp, NSAMP_BUF, TDMS, 0.0, DM_STEP, tsamp, FCH1, foff,
                           dtype=dd, inv_scale=inv_scale, tile=tiles[dd], dft=dft)
        if tiles[dd] is None:
            ch.autotune()
        ch.prime(prime)
        for _ in range(args.warmup):
            ch.step(volt)
            ch.detect()
        torch.cuda.synchronize()
        timing, events = [], []
        t0 = time.perf_counter()
        for _ in range(args.prec_steps):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(7)]
            ev[0].record()
            ch.step(volt, ev[1:6])
            ints, _, _ = ch.detect()
            ev[6].record()
            timing.append(ev)
            events.append(ints)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.prec_steps
        names = ["ppf_stokes", "clipper", "buffer", "dedisperse", "carry", "detect"]
        stages = {nm: round(sum(e[j].elapsed_time(e[j + 1]) for e in timing) / args.prec_steps, 3)
                  for j, nm in enumerate(names)}
        # one more step by hand, keeping its dedispersion input: the plain
        # chain's gives the FP64 reference sums on a DM sample
        keep = last_step_keeping_input(ch, volt)
        if fp64 is None:
            fp64 = fp64_sample(torch, keep, ch.plan, 16)
        del keep
        series = (ch.series if dd == "u8" else ch.out).double()
        e64 = (series[fp64[0]] - fp64[1]).abs()
        rec = {"dft": dft, "dedisp_input": dd, "tile_dm_t_chanblock_readbytes": list(ch.plan.tile),
               "ms_per_step": round(1e3 * wall, 3),
               "value": TDMS * nout / wall, "real_time_factor": nout * tsamp / wall, "stages_ms": stages,
               "events": sum(len(e) - 1 for e in events),
               "series_vs_fp64_max_abs_err_sigma": float(e64.max()) / nrm,
               "series_vs_fp64_rms_err_sigma": float(e64.pow(2).mean().sqrt()) / nrm}
        if ref_series is None:
            ref_series, ref_events = series.clone(), events
        else:
            err = (series - ref_series).abs()
            n_ref = n_got = inter = 0
            for er, eg in zip(ref_events, events):
                rk, gk = event_keys(er), event_keys(eg)
                n_ref, n_got, inter = n_ref + len(rk), n_got + len(gk), inter + len(rk & gk)
            rec.update({"series_max_abs_err_sigma": float(err.max()) / nrm,
                        "series_rms_err_sigma": float(err.pow(2).mean().sqrt()) / nrm,
                        "event_recall_vs_plain": inter / max(n_ref, 1),
                        "event_precision_vs_plain": inter / max(n_got, 1)})
        out["%s+%s" % (dft, dd)] = rec
        del ch, series
        torch.cuda.empty_cache()
    del volt, prime
    torch.cuda.empty_cache()
    out["error_unit"] = "channeliser: max / RMS error over the spectrum's RMS; series: sigma of a dedispersed sample"
    out["fp64_reference"] = ("series_vs_fp64: against FP64 sums, over 16 trials spread across the DM range and "
                             "every output sample, of the fp32+f32 chain's last dedispersion input (its FP32 "
                             "channeliser and clipper output); the other chains' errors include their "
                             "channeliser's.  series_* without vs_fp64 compare with the fp32+f32 chain's series")
    return out


def draw_beams(seed, n):
    """The heterogeneous multi-beam workload (tests/synth.py draw_multibeam)."""
    from tests.synth import draw_multibeam
    return draw_multibeam(seed, n, TS_IN, INTEG, FCH1)


def multibeam(args, torch, ppl):
    """24 beams of different shapes through the batched chain, plain (FP32
    DFT and dedispersion input) and mixed (BF16x3 DFT on the matrix cores for
    the 64- and 256-channel beams, FP16 or u8 dedispersion input): one clipper
    launch per stage for all beams, one dedispersion launch for all beams
    (or per group of beams sharing a tile, when the autotuner finds that
    faster), HIP-graph replayed steps.  Returns the report object."""
    from tests.synth import beam_bandpass, device_voltages, tone_amplitude
    beams = draw_beams(2024, args.beams)
    specs = []
    for bm in beams:
        h = ppl.ppf_coefficients(N_TAPS, bm["n_ch"], "hann")
        foff = -0.1953125 / bm["n_ch"]
        bp = beam_bandpass(h, bm["n_sub"], bm["n_ch"], INTEG, FCH1, foff, SIGMA)
        specs.append(ppl.BeamSpec(bm["n_sub"], bm["n_ch"], N_TAPS, h, INTEG, bp, bm["nsamp"], bm["tdms"], 0.0,
                                  bm["dm_hi"] / bm["tdms"], FCH1, foff))
    cands = [(64, 512, 1, 16), (32, 512, 1, 16), (16, 512, 1, 16), (32, 256, 1, 16), (16, 512, 4), (32, 512, 4),
             (16, 256, 4), (32, 256, 8), (16, 512, 8), (8, 512, 8)]
    report = {"beams": [{k: bm[k] for k in ("n_sub", "n_ch", "dm_hi", "tdms", "nsamp")} for bm in beams]}
    ref_events = None
    # the plain form (FP32 DFT, FP32 input) and the mixed forms (BF16x3 DFT
    # where built, FP16 or u8 input); the autotuner keeps one batch tile or
    # per-beam tiles, whichever is faster on the whole batch
    for dtype in ("f32", "f16", "u8"):
        log("multi-beam %d beams, %s" % (len(specs), dtype))
        mb = ppl.MultiBeam(specs, dtype=dtype, dft="fp32" if dtype == "f32" else "bf16x3")
        volts = []
        for b, (bm, sp) in enumerate(zip(beams, specs)):
            n_samp = mb.chunk_samples(b)
            m = mb.chunk_spectra[b]
            pulses = [(dm, t0 * m, w, tone_amplitude(sp.coeffs, sp.n_ch, SIGMA, a, INTEG))
                      for (dm, t0, w, a) in bm["pulses"]]
            volts.append(device_voltages(torch, sp.n_sub, sp.n_ch, INTEG, n_samp, bm["seed"], FCH1, sp.foff,
                                         SIGMA, pulses=pulses))
        torch.cuda.synchronize()
        mem_gb = torch.cuda.memory_allocated() / 1e9
        log("multi-beam prime (%.1f GB resident)" % mem_gb)
        mb.prime(volts)
        log("multi-beam autotune")
        t0 = time.perf_counter()
        tune_ms = mb.autotune(candidates=cands, per_beam=True)
        tune_s = time.perf_counter() - t0
        def run(n, graph=False, timing=None):
            evs = []
            for _ in range(n):
                ev = None
                if timing is not None:
                    ev = [torch.cuda.Event(enable_timing=True) for _ in range(7)]
                    ev[0].record()
                if graph:
                    mb.replay()
                else:
                    mb.step(volts, None if ev is None else ev[1:6])
                evs.append([mb.detect(b)[0] for b in range(len(specs))])
                if ev is not None:
                    ev[6].record()
                    timing.append(ev)
            return evs
        # warm up until every beam's clipper has left its learning phase
        # (history 10000 plus the 999 rejected warm-up spectra)
        need = max(-(-max(0, 11000 - ov) // n) for ov, n in zip(mb.overlap, mb.nout))
        warm = max(args.warmup, need)
        log("multi-beam steps (%d warm-up)" % warm)
        run(warm)
        torch.cuda.synchronize()
        cnt0 = [c.counters() for c in mb.clips]
        timing = []
        t0 = time.perf_counter()
        events = run(args.mb_steps, timing=timing)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.mb_steps
        names = ["ppf_stokes", "clipper", "buffer", "dedisperse", "carry", "detect"]
        stages = {nm: round(sum(ev[j].elapsed_time(ev[j + 1]) for ev in timing) / args.mb_steps, 3)
                  for j, nm in enumerate(names)}
        cnt1 = [c.counters() for c in mb.clips]
        walk = {k: sum(c1[k] - c0[k] for c0, c1 in zip(cnt0, cnt1)) for k in cnt0[0]}
        mb.capture_steps(volts)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(args.mb_steps, graph=True)
        torch.cuda.synchronize()
        gwall = (time.perf_counter() - t0) / args.mb_steps
        secs = mb.data_seconds()
        rec = {"value": mb.trials_samples() / wall, "ms_per_step": round(1e3 * wall, 3),
               "graph_ms_per_step": round(1e3 * gwall, 3),
               "graph_value": mb.trials_samples() / gwall,
               "real_time_factor_aggregate": sum(secs) / wall, "real_time_factor_min_beam": min(secs) / wall,
               "stages_ms": stages,
               "tile_dm_t_chanblock_readbytes": list(mb.tile) if mb.tile != "own" else [list(p.tile) for p in mb.plans],
               "dft": "fp32" if dtype == "f32" else "bf16x3 (64, 256 ch), fp32 (16 ch)",
               "autotune_ms": round(tune_ms, 3),
               "autotune_s": round(tune_s, 1), "resident_gb": round(mem_gb, 1), "warmup_steps": warm,
               "clipper_walk": walk,
               "dedisperse_adds_per_step": float(sum(sp.tdms * n * sp.n_sub * sp.n_ch
                                                     for sp, n in zip(specs, mb.nout))),
               "events": sum(len(e) - 1 for step in events for e in step)}
        rec["dedisperse_tadds_per_s"] = round(rec["dedisperse_adds_per_step"] / (stages["dedisperse"] * 1e-3) / 1e12, 3)
        if dtype == "f32":
            ref_events = events
        else:
            n_ref = n_got = inter = 0
            for se_r, se_g in zip(ref_events, events):
                for er, eg in zip(se_r, se_g):
                    rk, gk = event_keys(er), event_keys(eg)
                    n_ref, n_got, inter = n_ref + len(rk), n_got + len(gk), inter + len(rk & gk)
            rec["event_recall_vs_f32"] = inter / max(n_ref, 1)
            rec["event_precision_vs_f32"] = inter / max(n_got, 1)
        report[dtype] = rec
        report["data_seconds_per_step"] = round(sum(secs), 1)
        report["trials_samples_per_step"] = mb.trials_samples()
        del mb, volts
        torch.cuda.empty_cache()
    return report


def service(args, torch, ppl, coeffs, bp):
    """48 beams of the LOFAR-beam shape as a continuous chunk stream: beam
    sessions start at Poisson times (mean 2/s) and last 30-300 s, each beam
    admitted while its device footprint fits the HBM budget, one pulse per
    10.7 s chunk per beam (DM 0-500, 6-15 sigma, 1-8 ms) plus narrowband RFI.
    The schedule runs args.svc_accel times faster than real time, so the
    deadline of a chunk is its data duration / accel.  Chunk latency is from
    the moment its last sample is due to the end of its round (including the
    dedispersion and detection of a buffer it completes).  Returns the report
    object; event recall compares the service's buffers with a plain FP32
    per-beam run over the same chunk sequence for a sample of sessions."""
    from tests.synth import device_voltages, tone_amplitude
    rng = np.random.default_rng(4096)
    n_beams, chunk_spec, ring = 48, 8192, 6
    accel = args.svc_accel
    spec = ppl.BeamSpec(N_SUB, N_CH, N_TAPS, coeffs, INTEG, bp, NSAMP_BUF, TDMS, 0.0, DM_STEP, FCH1, FOFF)
    chunk_s = chunk_spec * TSAMP             # data seconds per chunk
    n_samp = chunk_spec * INTEG * N_CH
    log("service: %d-chunk input ring" % ring)
    chunks = []
    for c in range(ring):
        pulses = [(float(rng.uniform(0, 500)), float(rng.uniform(0.05, 0.95) * chunk_spec),
                   float(rng.uniform(1e-3, 8e-3) / TSAMP),
                   tone_amplitude(coeffs, N_CH, SIGMA, float(rng.uniform(6, 15)), INTEG))]
        rfi = [(int(rng.integers(N_SUB)), int(rng.integers(N_CH)), tone_amplitude(coeffs, N_CH, SIGMA, 50.0, INTEG))
               for _ in range(10)]
        chunks.append(device_voltages(torch, N_SUB, N_CH, INTEG, n_samp, 600 + c, FCH1, FOFF, SIGMA,
                                      pulses=pulses, rfi=rfi, wrap=False))
    torch.cuda.synchronize()
    free, total = torch.cuda.mem_get_info()
    budget = int(0.85 * free)
    svc = ppl.StreamService(spec, chunk_spec, budget, dtype=args.svc_dtype, max_beams=n_beams)
    fp_gb = svc.footprint() / 1e9
    # arrival schedule over the run, in data seconds
    horizon = args.svc_seconds * accel
    t, arrivals = 0.0, []
    while t < horizon:
        t += rng.exponential(0.5)
        arrivals.append((t, float(rng.uniform(30.0, 300.0)), int(rng.integers(1 << 30))))
    idle = list(range(n_beams))
    sessions = {}          # beam -> dict(start, end, next chunk index, sequence, key)
    log_seq = {}           # session key -> list of ring indices fed
    events = {}            # session key -> {buffer index: event set}
    lat, missed, rounds, n_chunks = [], 0, 0, 0
    rejected = dropped = admitted = 0
    max_active = 0
    ai = 0
    log("service: %.0f s of data at x%.0f, budget %.1f GB, footprint %.2f GB per beam" %
        (horizon, accel, budget / 1e9, svc.footprint() / 1e9))
    w0 = time.perf_counter()
    last_log = w0
    while True:
        now = (time.perf_counter() - w0) * accel
        if now >= horizon:
            break
        while ai < len(arrivals) and arrivals[ai][0] <= now:
            t_a, dur, seed = arrivals[ai]
            ai += 1
            if not idle:
                dropped += 1
                continue
            b = idle.pop(int(seed % len(idle)))
            key = (b, t_a)
            if not svc.admit(key):
                rejected += 1
                idle.append(b)
                continue
            admitted += 1
            sessions[b] = {"start": t_a, "end": t_a + dur, "k": 0, "key": key, "seed": seed}
            log_seq[key] = []
            events[key] = {}
        max_active = max(max_active, len(sessions))
        ready, due = [], []
        for b, ss in list(sessions.items()):
            t_due = ss["start"] + (ss["k"] + 1) * chunk_s
            if t_due > ss["end"]:
                svc.release(ss["key"])
                del sessions[b]
                idle.append(b)
                continue
            if t_due <= now:
                ci = (ss["seed"] + ss["k"]) % ring
                ready.append((ss["key"], chunks[ci]))
                due.append(t_due)
                log_seq[ss["key"]].append(ci)
                ss["k"] += 1
        if not ready:
            nxt = min([ss["start"] + (ss["k"] + 1) * chunk_s for ss in sessions.values()] +
                      ([arrivals[ai][0]] if ai < len(arrivals) else [horizon]))
            time.sleep(max(0.0, min(0.05, (nxt - now) / accel)))
            continue
        for key, idx, ints, vals in svc.process(ready):
            events[key][idx] = event_keys(ints)
        torch.cuda.synchronize()
        done = (time.perf_counter() - w0) * accel
        for t_due in due:
            l_wall = (done - t_due) / accel
            lat.append(l_wall)
            missed += l_wall > chunk_s / accel
        rounds += 1
        n_chunks += len(ready)
        if time.perf_counter() - last_log > 10.0:
            last_log = time.perf_counter()
            log("service: %.0f s data, %d active, %d chunks" % (now, len(sessions), n_chunks))
    wall = time.perf_counter() - w0
    for b, ss in list(sessions.items()):
        svc.release(ss["key"])
    del svc
    torch.cuda.empty_cache()
    lat_ms = np.array(lat) * 1e3
    # recall against plain FP32 per-beam runs of the same chunk sequences
    log("service: recall check")
    # a sample of the sessions that produced events, spread over the run
    cand = sorted([k for k in events if events[k]], key=lambda k: k[1])
    step = max(1, len(cand) // max(args.svc_recall_sessions, 1))
    keys = cand[::step][:args.svc_recall_sessions]
    n_ref = n_got = inter = 0
    for key in keys:
        plain = ppl.StreamService(spec, chunk_spec, 1 << 40, dtype="f32", max_beams=1)
        plain.admit(key)
        ref = {}
        for ci in log_seq[key]:
            for _, idx, ints, _ in plain.process([(key, chunks[ci])]):
                ref[idx] = event_keys(ints)
        for idx, ev in events[key].items():
            r = ref.get(idx, set())
            n_ref, n_got, inter = n_ref + len(r), n_got + len(ev), inter + len(r & ev)
        del plain
    torch.cuda.empty_cache()
    data_s = n_chunks * chunk_s
    return {"beams": n_beams, "arrival_rate_per_s": 2.0, "session_s": [30, 300], "accel": accel,
            "dtype": args.svc_dtype, "chunk_spectra": chunk_spec, "chunk_data_s": round(chunk_s, 3),
            "deadline_ms": round(1e3 * chunk_s / accel, 1), "deadline_ms_real_time": round(1e3 * chunk_s, 1),
            "latency_ms_p50": round(float(np.percentile(lat_ms, 50)), 2) if lat else None,
            "latency_ms_p99": round(float(np.percentile(lat_ms, 99)), 2) if lat else None,
            "latency_ms_max": round(float(lat_ms.max()), 2) if lat else None,
            "missed_deadline": int(missed), "chunks": n_chunks, "rounds": rounds, "wall_s": round(wall, 1),
            "sessions_admitted": admitted, "rejected_by_hbm": rejected, "dropped_no_idle_beam": dropped,
            "max_active": max_active, "footprint_gb_per_beam": round(fp_gb, 3),
            "budget_gb": round(budget / 1e9, 1),
            "throughput_data_s_per_s": round(data_s / wall, 1),
            "throughput_dm_trials_samples_per_s": data_s / TSAMP * TDMS / wall,
            "event_recall_vs_plain_f32": inter / max(n_ref, 1), "event_precision_vs_plain_f32": inter / max(n_got, 1),
            "recall_sessions": len(keys), "recall_sessions_with_events": len(cand), "recall_events_ref": n_ref}


def cpu_baseline(n_spec, coeffs, bp, nout, passes=1):
    """Times the oracle (the C++ restatement of the reference CPU path, OpenMP
    where the reference uses it) on a bounded sample and scales to the metric:
    `passes` passes over n_spec integrated spectra, the channeliser and
    clipper continuing their streams from pass to pass."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as orc
    from tests.synth import noise_voltages
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n_blk = n_spec * INTEG
    x = noise_voltages(N_SUB, n_blk * N_CH, SIGMA, seed=7)
    p = orc.PPF(N_SUB, 2, N_CH, N_TAPS, coeffs)
    clip = orc.Clipper(bp, 10.0, 4.0, 10000, True)
    plan = orc.dedisp_plan(N_CHAN, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP, FCH1, FOFF)
    delays = orc.dedisp_delays(plan["shift"], TDMS, plan["mstartdm"], plan["mdmstep"])
    overlap = plan["maxshift"] + plan["rem"]
    buf_in = np.random.default_rng(8).normal(size=(N_CHAN, overlap + n_spec)).astype(np.float32)
    t = {k: 0.0 for k in ("ppf", "stokes_integrate", "clipper", "buffer", "dedisperse", "detect")}
    primed = False
    for _ in range(passes):
        t0 = time.perf_counter()
        spec = p.run_f32(x, n_blk, threads)
        t["ppf"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        I = orc.stokes_i_f32(spec)
        I4 = orc.integrate(I, INTEG)
        t["stokes_integrate"] += time.perf_counter() - t0
        del spec, I
        if not primed:
            # the clipper's first `history` spectra are its learning phase (band
            # minima on every spectrum); prime it past that off the clock so the
            # timed spectra take its steady-state path, as a long-running stream does
            reps = -(-10000 // n_spec)
            clip.run(np.ascontiguousarray(np.concatenate([I4] * reps)[:10000]))
            primed = True
        t0 = time.perf_counter()
        clip.run(I4)
        t["clipper"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        buf = orc.DDBuffer(N_CHAN, overlap + n_spec, overlap)
        buf.add(I4.reshape(n_spec, N_CHAN), 1)
        t["buffer"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        out = orc.dedisperse(buf_in, delays, n_spec, threads)
        t["dedisperse"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.detect(out, N_CH, N_SUB, 6.0, 6)
        t["detect"] += time.perf_counter() - t0
    total = sum(t.values())
    per_spectrum = total / (n_spec * passes)
    import platform
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": TDMS / per_spectrum, "unit": "DM-trials*samples/s", "cores": threads,
            "kind": "port",
            "sample": "oracle chain (FP32 FIR+FFT PPF, Stokes-I, x4, clipper past its learning phase, "
                      "buffer, OpenMP brute-force dedispersion of 2048 DMs, detection), %d passes over %d "
                      "integrated spectra of the same workload, %.1f s; scaled per output sample"
                      % (passes, n_spec, total),
            "real_time_factor": TSAMP / per_spectrum,
            "stages_s": {k: round(v, 3) for k, v in t.items()},
            "cpu_model": model, "nproc": os.cpu_count()}


if __name__ == "__main__":
    main()
