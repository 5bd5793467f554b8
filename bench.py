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
times, the dedispersion roofline and the CPU baseline alongside.
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


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dedisp-dtype", default="f32", choices=["f32", "f16", "u8"])
    ap.add_argument("--chunks", type=int, default=2, help="distinct input chunks kept in HBM")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile", default="", help="DM,T,chanblock tile for dedispersion (default: autotune)")
    ap.add_argument("--cpu-spectra", type=int, default=8192,
                    help="integrated spectra in the timed CPU sample")
    return ap.parse_args()


def gen_chunk(torch, ppl_coeffs, n_samp, sample0, seed, pulses, rfi):
    """int16 voltages [sub][pol][n_samp][2] on the device."""
    from tests.synth import shift_table, tone_amplitude
    g = torch.Generator(device="cuda").manual_seed(seed)
    out = torch.empty((N_SUB, 2, n_samp, 2), dtype=torch.int16, device="cuda")
    shift = torch.from_numpy(shift_table(N_CHAN, FCH1, FOFF)).cuda()
    spi = N_CH * INTEG  # input samples per integrated sample
    n_int = n_samp // spi
    imp = torch.rand(n_int, generator=g, device="cuda") < 0.005  # broadband impulses
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
            a = torch.round((t0 + d) * spi).long() - sample0
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


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the hot path is measured on one MI355X, and a "
              "multi-GPU layout (beams or DM ranges per GPU) is described in DESIGN.md, "
              "not launched here.  Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as ge
    from tests.synth import tone_amplitude

    ppl = ge.load_package()
    ppl.init(0)
    dev = torch.device("cuda:0")

    coeffs = ppl.ppf_coefficients(N_TAPS, N_CH, "hann")
    e_h = float(np.sum(coeffs * coeffs))
    level = 2 * INTEG * 2.0 * SIGMA * SIGMA * e_h          # mean integrated Stokes-I
    noise_std = np.sqrt(2 * INTEG) * 2.0 * SIGMA * SIGMA * e_h
    bp = np.array([N_CHAN, FCH1, FCH1 + FOFF * (N_CHAN - 1), FOFF, noise_std, level], np.float64)

    dd_dtype = args.dedisp_dtype
    inv_scale = 127.0 / 10.0  # u8: clipper's normalised range (crFactor = 10 sigma) over 127 steps
    plan = ppl.DedispPlan(N_CHAN, NSAMP_BUF, TDMS, 0.0, DM_STEP, TSAMP, FCH1, FOFF, dtype=dd_dtype)
    nout, overlap = plan.nout, plan.maxshift + plan.rem
    ppf = ppl.PPF(N_SUB, N_CH, N_TAPS, coeffs.astype(np.float32), integrate=INTEG)
    clip = ppl.Clipper(bp, 10.0, 4.0, 10000, True)
    ddbuf = ppl.DDBuffer(N_CHAN, NSAMP_BUF, overlap, dd_dtype, inv_scale)

    blk_step = nout * INTEG
    n_samp = blk_step * N_CH
    rng = np.random.default_rng(2024)
    chunks = []
    t_gen = time.time()
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
        chunks.append(gen_chunk(torch, coeffs, n_samp, ci * n_samp, 100 + ci, pulses, rfi))
    prime_samples = overlap * INTEG * N_CH
    prime = chunks[-1][:, :, :prime_samples].contiguous()
    torch.cuda.synchronize()
    t_gen = time.time() - t_gen

    stokes = torch.empty((nout, N_SUB, 1, N_CH), dtype=torch.float32, device=dev)
    dd_out = torch.empty((TDMS, nout), dtype=plan.out_dtype(), device=dev)
    ev_cap = 1 << 22

    # pick the DM x time tile for this shape (the project's autotuner)
    tune_in = torch.zeros((N_CHAN, NSAMP_BUF), dtype={"f32": torch.float32, "f16": torch.float16,
                                                      "u8": torch.uint8}[dd_dtype], device=dev)
    if args.tile:
        plan.set_tile(*[int(v) for v in args.tile.split(",")])
        tune_ms = None
    else:
        tune_ms = plan.autotune(tune_in, dd_out, reps=1)
    del tune_in

    # prime the buffer with the overlap so that every step consumes nOut spectra
    st = ppf.stokes_i16(prime, overlap * INTEG)
    clip.run(st)
    consumed, full = ddbuf.add(st.view(overlap, N_CHAN))
    assert consumed == overlap and not full

    stage_names = ["ppf_stokes", "clipper", "buffer", "dedisperse", "detect"]

    def step(chunk, timing):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(stage_names) + 1)]
        ev[0].record()
        ppf.stokes_i16(chunk, blk_step, d_out=stokes)
        ev[1].record()
        clip.run(stokes)
        ev[2].record()
        c, full = ddbuf.add(stokes.view(nout, N_CHAN))
        assert c == nout and full
        ev[3].record()
        plan.run(ddbuf.tensor(), d_out=dd_out)
        ev[4].record()
        if dd_dtype == "u8":
            _, _, n_ev = ppl.detect_u8(dd_out, N_CHAN, 1.0 / inv_scale, N_CH, N_SUB, 6.0, 6, ev_cap)
        else:
            _, _, n_ev = ppl.detect(dd_out, N_CH, N_SUB, 6.0, 6, ev_cap)
        ev[5].record()
        ddbuf.rotate()
        if timing is not None:
            timing.append((ev, n_ev))
        return n_ev

    for w in range(args.warmup):
        step(chunks[w % len(chunks)], None)
    torch.cuda.synchronize()
    timing = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(chunks[k % len(chunks)], timing)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    stages = {n: 0.0 for n in stage_names}
    for ev, _ in timing:
        for j, n in enumerate(stage_names):
            stages[n] += ev[j].elapsed_time(ev[j + 1])
    stages = {n: v / args.steps for n, v in stages.items()}
    events = [n for _, n in timing]

    ms_per_step = 1e3 * wall / args.steps
    value = TDMS * nout * args.steps / wall
    rtf = nout * TSAMP * args.steps / wall
    # roofline of the dedispersion kernel: algorithmic adds / kernel time
    adds = float(TDMS) * N_CHAN * nout
    elem = {"f32": 4, "f16": 2, "u8": 1}[dd_dtype]
    alg_bytes = N_CHAN * NSAMP_BUF * elem + TDMS * nout * 4
    t_dd = stages["dedisperse"] * 1e-3
    achieved = adds / t_dd / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_TFLOPS_FP32,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_TFLOPS_FP32, 4), "traffic": None,
                "kernel": "dedisperse_kernel", "ops_per_launch": adds,
                "alg_bytes_per_launch": alg_bytes,
                "hbm_gbs_algorithmic": round(alg_bytes / t_dd / 1e9, 1)}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_spectra, coeffs, bp, nout)

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
                   "dedisp_input": dd_dtype, "dft": "fp32", "tile_dm_t_chanblock": list(plan.tile),
                   "autotune_ms": None if tune_ms is None else round(tune_ms, 3)},
        "real_time_factor": rtf,
        "stages_ms": {k: round(v, 3) for k, v in stages.items()},
        "events_per_step": events,
        "clipper_walk": clip.counters(),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gen_seconds": round(t_gen, 1),
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(n_spec, coeffs, bp, nout):
    """Times the oracle (the C++ restatement of the reference CPU path, OpenMP
    where the reference uses it) on a bounded sample and scales to the metric."""
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
    t = {}
    t0 = time.perf_counter()
    spec = p.run_f32(x, n_blk, threads)
    t["ppf"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    I = orc.stokes_i_f32(spec)
    I4 = orc.integrate(I, INTEG)
    t["stokes_integrate"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    clip.run(I4)
    t["clipper"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    buf = orc.DDBuffer(N_CHAN, overlap + n_spec, overlap)
    buf.add(I4.reshape(n_spec, N_CHAN), 1)
    t["buffer"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = orc.dedisperse(buf_in, delays, n_spec, threads)
    t["dedisperse"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.detect(out, N_CH, N_SUB, 6.0, 6)
    t["detect"] = time.perf_counter() - t0
    total = sum(t.values())
    per_spectrum = total / n_spec
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
            "sample": "oracle chain (FP32 FIR+FFT PPF, Stokes-I, x4, clipper, buffer, OpenMP brute-force "
                      "dedispersion of 2048 DMs, detection) over %d integrated spectra of the same "
                      "workload, %.1f s; scaled per output sample" % (n_spec, total),
            "real_time_factor": TSAMP / per_spectrum,
            "stages_s": {k: round(v, 3) for k, v in t.items()},
            "cpu_model": model, "nproc": os.cpu_count()}


if __name__ == "__main__":
    main()
