"""Seeded synthetic inputs for the pulsar-search chain (numpy, CPU).

Conventions (build-defined, the reference leaves them open):
- voltages are complex int16, layout [sub][pol][t][re,im] (TimeSeriesDataSetC32
  order with block length nCh), Gaussian noise of standard deviation sigma
  per component, rounded and saturated to int16;
- buffer channel c' = sub*nCh + k is at frequency fch1 + foff*c' with foff < 0
  (descending order, the reference's invert = 0 path);
- a dispersed pulse is a set of channel-centre tone bursts: channel c' carries
  a tone at baseband frequency (k - nCh/2)/nCh cycles/sample (the bin the
  fftshifted channeliser maps it to) during [t0 + delay(c'), t0 + delay(c') +
  width) integrated samples, delay = shift[c']*DM/tsamp;
- the pulse amplitude is given as the mean excess of one channel's
  integrated Stokes-I over its noise standard deviation.
"""
import numpy as np

KDM = 4148.741601


def shift_table(n_chan, fch1, foff):
    c = np.arange(n_chan, dtype=np.float64)
    f = fch1 + foff * c
    return KDM * ((1.0 / f / f) - (1.0 / fch1 / fch1))


def noise_voltages(n_sub, n_samples, sigma, seed, n_pol=2):
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, sigma, size=(n_sub, n_pol, n_samples, 2))
    return np.clip(np.rint(x), -32768, 32767).astype(np.int16)


def tone_amplitude(coeffs, n_ch, sigma, excess_sigma, integrate):
    """Tone amplitude giving `excess_sigma` standard deviations of excess
    integrated Stokes-I in its channel (both polarisations carry the tone)."""
    h = np.asarray(coeffs, np.float64)
    e_h = np.sum(h * h)
    s = abs(np.sum(h * ((-1.0) ** (np.arange(h.size) % n_ch))))
    m = 2.0 * sigma * sigma * e_h                # mean |X|^2 of one pol, one block
    n_terms = 2 * integrate                      # pols x integrated blocks
    std = np.sqrt(n_terms) * m
    return float(np.sqrt(excess_sigma * std / (n_terms * s * s)))


def inject_pulse_samples(volt, n_ch, fch1, foff, tsamp, samples_per_int, dm, t0, width, amp,
                         sample0=0, seed=0):
    """Tone-burst pulse: see module docstring.  t0, width in integrated samples."""
    n_sub, n_pol, n_samp, _ = volt.shape
    n_chan = n_sub * n_ch
    shift = shift_table(n_chan, fch1, foff)
    rng = np.random.default_rng(seed)
    acc = volt.astype(np.float64)
    for cp in range(n_chan):
        s, k = divmod(cp, n_ch)
        d = shift[cp] * dm / tsamp
        a = int(round((t0 + d) * samples_per_int)) - sample0
        b = a + int(round(width * samples_per_int))
        a0, b0 = max(a, 0), min(b, n_samp)
        if a0 >= b0:
            continue
        nu = (k - n_ch / 2.0) / n_ch
        m = np.arange(a0, b0) + sample0
        for p in range(n_pol):
            ph = rng.uniform(0, 2 * np.pi)
            z = amp * np.exp(1j * (2 * np.pi * nu * m + ph))
            acc[s, p, a0:b0, 0] += z.real
            acc[s, p, a0:b0, 1] += z.imag
    volt[...] = np.clip(np.rint(acc), -32768, 32767).astype(np.int16)
    return volt


def bandpass_values(n_bins, start, width, median, rms, poly=()):
    """Numbers of a .bp file: nChan, start, end, width, rms, median, poly..."""
    end = start + width * (n_bins - 1)
    return np.array([n_bins, start, end, width, rms, median] + list(poly), np.float64)


def stokes_noise(n_t, n_sub, n_ch, level, seed, dof=16, n_pol=1):
    """Chi-square-like Stokes-I spectra with mean `level` per bin
    ([nT][nSub][nPol][nCh] float32), optional per-bin level array."""
    rng = np.random.default_rng(seed)
    lv = np.broadcast_to(np.asarray(level, np.float64), (n_sub * n_ch,)).reshape(n_sub, n_ch)
    x = rng.chisquare(dof, size=(n_t, n_sub, n_pol, n_ch)) / dof
    return (x * lv[None, :, None, :]).astype(np.float32)


def device_voltages(torch, n_sub, n_ch, integrate, n_samp, seed, fch1, foff, sigma=64.0, pulses=(),
                    rfi=(), impulse_frac=0.0, impulse_sigma=20.0, wrap=True):
    """int16 voltages [sub][pol][n_samp][2] generated on the device: Gaussian
    noise, narrowband tones rfi = [(sub, ch, amp)], broadband impulses in a
    fraction of the integrated samples, and dispersed tone-burst pulses
    [(dm, t0, width, amp)] (t0, width in integrated samples).  With wrap the
    chunk is periodic: a pulse's samples past the end re-enter at the start,
    so the chunk can be fed repeatedly as a continuous stream."""
    n_chan = n_sub * n_ch
    tsamp = 5.12e-6 * n_ch * integrate
    g = torch.Generator(device="cuda").manual_seed(seed)
    out = torch.empty((n_sub, 2, n_samp, 2), dtype=torch.int16, device="cuda")
    shift = torch.from_numpy(shift_table(n_chan, fch1, foff)).cuda()
    spi = n_ch * integrate
    n_int = n_samp // spi
    imp_idx = torch.nonzero(torch.rand(n_int, generator=g, device="cuda") < impulse_frac).flatten()
    k = torch.arange(n_ch, device="cuda", dtype=torch.float64)
    nu = ((k - n_ch / 2.0) / n_ch)[:, None]
    for s in range(n_sub):
        x = torch.randn((2, n_samp, 2), generator=g, device="cuda") * sigma
        if imp_idx.numel():
            seg = (imp_idx[:, None] * spi + torch.arange(spi, device="cuda")[None, :]).flatten()
            x[:, seg, :] += torch.randn((2, seg.numel(), 2), generator=g, device="cuda") * (impulse_sigma * sigma)
        m = torch.arange(n_samp, device="cuda", dtype=torch.float64)
        for (rs, rk, ramp) in rfi:
            if rs != s:
                continue
            ph = 2 * np.pi * ((rk - n_ch / 2.0) / n_ch) * m
            x[:, :, 0] += (ramp * torch.cos(ph)).float()[None, :]
            x[:, :, 1] += (ramp * torch.sin(ph)).float()[None, :]
        for (dm, t0, width, amp) in pulses:
            d = shift[s * n_ch:(s + 1) * n_ch] * dm / tsamp
            a = torch.round((t0 + d) * spi).long()
            L = max(1, int(round(width * spi)))
            idx = a[:, None] + torch.arange(L, device="cuda")[None, :]
            ph = 2 * np.pi * nu * idx.double()
            if wrap:
                idx = idx % n_samp
                ok = torch.ones_like(idx, dtype=torch.bool)
            else:
                ok = (idx >= 0) & (idx < n_samp)
            for p in range(2):
                x[p, :, 0].index_add_(0, idx[ok], (amp * torch.cos(ph + p)).float()[ok])
                x[p, :, 1].index_add_(0, idx[ok], (amp * torch.sin(ph + p)).float()[ok])
        out[s] = torch.clamp(torch.round(x), -32768, 32767).to(torch.int16)
    return out


def beam_bandpass(coeffs, n_sub, n_ch, integrate, fch1, foff, sigma=64.0):
    """.bp values of a flat bandpass at the mean integrated Stokes-I of noise
    of this sigma through this filter, with its standard deviation as RMS."""
    e_h = float(np.sum(np.asarray(coeffs, np.float64) ** 2))
    level = 2 * integrate * 2.0 * sigma * sigma * e_h
    noise_std = np.sqrt(2 * integrate) * 2.0 * sigma * sigma * e_h
    n_chan = n_sub * n_ch
    return np.array([n_chan, fch1, fch1 + foff * (n_chan - 1), foff, noise_std, level], np.float64)


def generator_blob(dm, n_samples=6400, n_sub=32, n_ch=64, start_bin=0, width=10, fch1=150.0, tsamp=0.00032768):
    """The reference's DedispersionDataGenerator::generate for one blob
    (src/lib/test/src/DedispersionDataGenerator.cpp:29-37,47-81): every
    channel is 1.0 for `width` samples from its FP64 dispersion index,
    else 0.  Returns the dedispersion-buffer layout [channel][sample]."""
    n = n_sub * n_ch
    foff = -6.0 / n
    out = np.zeros((n, n_samples), np.float32)
    for c in range(n):
        f = fch1 + foff * c
        idx = int(dm * (4148.741601 * ((1.0 / f / f) - (1.0 / fch1 / fch1)) / tsamp)) + start_bin
        a, b = max(idx, 0), min(idx + width, n_samples)
        if a < b:
            out[c, a:b] = 1.0
    return out


def draw_multibeam(seed, n, ts_in=5.12e-6, integrate=4, fch1=150.0):
    """The heterogeneous multi-beam workload: per beam 31-62 subbands of
    195.3125 kHz below fch1, 16/64/256 channels per subband, DM 0 to
    100-1000 pc/cm^3 over 512-4096 trials, a 2^14-2^17-sample dedispersion
    buffer, raised to the next power of two while it holds fewer than
    1.25 x maxshift samples (a shorter buffer cannot hold the beam's DM
    sweep), and 1-3 pulses (DM, start as a fraction of a chunk, width in
    integrated samples for 1-8 ms, 6-12 sigma)."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n):
        n_sub = int(rng.integers(31, 63))
        n_ch = int(rng.choice([16, 64, 256]))
        dm_hi = float(rng.choice([100, 200, 300, 500, 750, 1000]))
        tdms = int(rng.choice([512, 1024, 2048, 4096]))
        nsamp = 1 << int(rng.integers(14, 18))
        tsamp = ts_in * n_ch * integrate
        foff = -0.1953125 / n_ch
        fmin = fch1 + foff * (n_sub * n_ch - 1)
        maxshift = KDM * (1.0 / fmin ** 2 - 1.0 / fch1 ** 2) * dm_hi / tsamp
        while nsamp < 1.25 * maxshift:
            nsamp *= 2
        n_pulse = int(rng.integers(1, 4))
        pulses = [(float(rng.uniform(0.1, 0.9) * dm_hi), float(rng.uniform(0.0, 1.0)),
                   float(rng.uniform(1e-3, 8e-3) / tsamp), float(rng.uniform(6.0, 12.0))) for _ in range(n_pulse)]
        out.append(dict(n_sub=n_sub, n_ch=n_ch, dm_hi=dm_hi, tdms=tdms, nsamp=nsamp, seed=seed + 17 * b,
                        pulses=pulses))
    return out
