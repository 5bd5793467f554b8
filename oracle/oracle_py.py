"""ctypes access to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker (or the timed CPU
baseline).  The product never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

WIN = {"kaiser": 0, "gaussian": 1, "blackman": 2, "hamming": 3, "hann": 4}

_lib = None


def build(force=False):
    """Compile liboracle.so with make (g++ -fopenmp -ffp-contract=off)."""
    cmd = ["make", "-s", "-C", HERE] + (["-B"] if force else [])
    subprocess.check_call(cmd)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _declare(L):
    vp, u, i, f, d = C.c_void_p, C.c_uint, C.c_int, C.c_float, C.c_double
    sig = {
        "orc_ppf_coefficients": (i, [u, u, i, vp]),
        "orc_fir2_unwindowed": (i, [u, u, vp]),
        "orc_assign_threads": (i, [vp, vp, u, u, u]),
        "orc_ppf_create": (vp, [u, u, u, u, vp, vp]),
        "orc_ppf_destroy": (None, [vp]),
        "orc_ppf_run_fir32": (None, [vp, vp, u, vp]),
        "orc_ppf_run_f64": (None, [vp, vp, u, vp]),
        "orc_ppf_run_f32": (None, [vp, vp, u, vp, i]),
        "orc_ppf_fir_only_f32": (None, [vp, vp, u, vp]),
        "orc_stokes_i_f32": (None, [vp, u, u, u, vp]),
        "orc_stokes_i_f64": (None, [vp, u, u, u, vp]),
        "orc_integrate_f32": (None, [vp, u, u, u, u, u, u, vp]),
        "orc_integrate_f64": (None, [vp, u, u, u, u, u, u, vp]),
        "orc_clipper_create": (vp, [vp, i, f, f, i, i, i, i]),
        "orc_clipper_create_kill": (vp, [vp, i, vp, i, f, f, i, i, i, i]),
        "orc_clipper_destroy": (None, [vp]),
        "orc_clipper_run": (i, [vp, vp, vp, u, u, u, u]),
        "orc_clipper_last_flags": (i, [vp, vp, vp, u]),
        "orc_clipper_state": (None, [vp, vp, vp]),
        "orc_clipper_bandpass": (i, [vp, vp, u]),
        "orc_bandpass_probe": (i, [u, d, d, vp, i, f, f, u, d, d, vp]),
        "orc_dedisp_plan": (i, [u, u, u, f, f, d, d, d, vp, vp, vp]),
        "orc_dedisp_delays": (None, [vp, u, u, f, f, vp]),
        "orc_dedisp_delays_fused": (None, [vp, u, u, f, f, vp]),
        "orc_dedisperse_f32": (None, [vp, u, u, u, vp, u, vp, i]),
        "orc_dedisperse_f16": (None, [vp, u, u, u, vp, u, vp, i]),
        "orc_dedisperse_u8": (None, [vp, u, u, u, vp, u, vp, i]),
        "orc_f32_to_f16": (None, [vp, vp, C.c_size_t]),
        "orc_quantise_u8": (None, [vp, vp, C.c_size_t, f]),
        "orc_quantise_u8_rows": (None, [vp, u, u, f, vp, vp]),
        "orc_u8_series": (None, [vp, vp, vp, u, u, u, u, f, vp]),
        "orc_ddbuf_create": (vp, [u, u, u]),
        "orc_ddbuf_destroy": (None, [vp]),
        "orc_ddbuf_add": (i, [vp, vp, u, vp, i]),
        "orc_detect": (i, [vp, i, i, i, i, f, u, vp, vp, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


# ------------------------------------------------------------------ PPF
def ppf_coefficients(n_taps, n_ch, window="kaiser"):
    out = np.zeros(n_taps * n_ch, np.float64)
    rc = lib().orc_ppf_coefficients(n_taps, n_ch, WIN[window], _p(out))
    if rc != 0:
        raise ValueError("unknown window %r" % window)
    return out


def fir2_unwindowed(n_taps, n_ch):
    out = np.zeros(n_taps * n_ch, np.float64)
    lib().orc_fir2_unwindowed(n_taps, n_ch, _p(out))
    return out


def assign_threads(n_values, n_threads, thread_id):
    s, e = C.c_uint(0), C.c_uint(0)
    rc = lib().orc_assign_threads(C.byref(s), C.byref(e), n_values, n_threads, thread_id)
    if rc != 0:
        raise ValueError("threadId out of range")
    return s.value, e.value


class PPF:
    """Stateful PPF restatement; input int16 (re,im) [sub][pol][t]."""

    def __init__(self, n_sub, n_pol, n_ch, n_taps, coeffs64):
        self.shape = (n_sub, n_pol, n_ch, n_taps)
        self.c64 = np.ascontiguousarray(coeffs64, np.float64)
        self.c32 = self.c64.astype(np.float32)
        self.h = lib().orc_ppf_create(n_sub, n_pol, n_ch, n_taps, _p(self.c32), _p(self.c64))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ppf_destroy(self.h)
            self.h = None

    def _run(self, fn, x, n_blk, dtype, comps):
        n_sub, n_pol, n_ch, _ = self.shape
        x = np.ascontiguousarray(x, np.int16)
        assert x.size == n_sub * n_pol * n_blk * n_ch * 2
        out = np.zeros((n_blk, n_sub, n_pol, n_ch, comps), dtype)
        fn(self.h, _p(x), n_blk, _p(out))
        return out

    def run_fir32(self, x, n_blk):
        o = self._run(lib().orc_ppf_run_fir32, x, n_blk, np.float64, 2)
        return o[..., 0] + 1j * o[..., 1]

    def run_f64(self, x, n_blk):
        o = self._run(lib().orc_ppf_run_f64, x, n_blk, np.float64, 2)
        return o[..., 0] + 1j * o[..., 1]

    def run_f32(self, x, n_blk, nthreads=0):
        """FP32 FIR + FP32 FFT (CPU baseline path); float32 [blk][sub][pol][nCh][2]."""
        n_sub, n_pol, n_ch, _ = self.shape
        x = np.ascontiguousarray(x, np.int16)
        out = np.zeros((n_blk, n_sub, n_pol, n_ch, 2), np.float32)
        lib().orc_ppf_run_f32(self.h, _p(x), n_blk, _p(out), nthreads)
        return out

    def fir_only_f32(self, x, n_blk):
        return self._run(lib().orc_ppf_fir_only_f32, x, n_blk, np.float32, 2)


def stokes_i_f32(spec):
    """spec: float32 [blk][sub][2][nch][2] (re,im)."""
    spec = np.ascontiguousarray(spec, np.float32)
    n_blk, n_sub, _, n_ch, _ = spec.shape
    out = np.zeros((n_blk, n_sub, 1, n_ch), np.float32)
    lib().orc_stokes_i_f32(_p(spec), n_blk, n_sub, n_ch, _p(out))
    return out


def stokes_i_f64(spec):
    """spec: complex128 [blk][sub][2][nch]."""
    s = np.ascontiguousarray(np.stack([spec.real, spec.imag], -1), np.float64)
    n_blk, n_sub, _, n_ch, _ = s.shape
    out = np.zeros((n_blk, n_sub, 1, n_ch), np.float64)
    lib().orc_stokes_i_f64(_p(s), n_blk, n_sub, n_ch, _p(out))
    return out


def integrate(x, window, bin_ch=1):
    dt = np.float32 if x.dtype == np.float32 else np.float64
    x = np.ascontiguousarray(x, dt)
    n_t, n_sub, n_pol, n_ch = x.shape
    w = min(window, n_t)
    out = np.zeros((n_t // w, n_sub, n_pol, n_ch // bin_ch), dt)
    fn = lib().orc_integrate_f32 if dt == np.float32 else lib().orc_integrate_f64
    fn(_p(x), n_t, n_sub, n_pol, n_ch, window, bin_ch, _p(out))
    return out


# -------------------------------------------------------------- clipper
class Clipper:
    def __init__(self, bp_values, cr=10.0, sr=4.0, history=10000, zero_dm=True,
                 band_matching=True, mean_over_rms=True, kill=()):
        v = np.ascontiguousarray(bp_values, np.float64)
        k = np.ascontiguousarray(list(kill), np.uint32)
        self.h = lib().orc_clipper_create_kill(_p(v), v.size, _p(k) if k.size else None, k.size, cr, sr,
                                               history, int(zero_dm), int(band_matching), int(mean_over_rms))
        if not self.h:
            raise ValueError("bad bandpass values or kill indices")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_clipper_destroy(self.h)
            self.h = None

    def run(self, I, W=None):
        """In place on float32 I [nT][nSub][nPol][nCh]."""
        assert I.dtype == np.float32 and I.flags.c_contiguous
        n_t, n_sub, n_pol, n_ch = I.shape
        wp = _p(W) if W is not None else None
        rc = lib().orc_clipper_run(self.h, _p(I), wp, n_t, n_sub, n_pol, n_ch)
        if rc == -2:
            raise ValueError("a killed band reaches outside the run map")
        if rc != 0:
            raise ValueError("nBins must be divisible by 8")
        return I

    def last_flags(self, n_t):
        good = np.zeros(n_t, np.uint8)
        nclip = np.zeros(n_t, np.int32)
        lib().orc_clipper_last_flags(self.h, _p(good), _p(nclip), n_t)
        return good, nclip

    def state(self):
        o = np.zeros(5, np.float32)
        s = np.zeros(2, np.int32)
        lib().orc_clipper_state(self.h, _p(o), _p(s))
        return o, s

    def bandpass(self, n_bins):
        o = np.zeros(n_bins, np.float32)
        if lib().orc_clipper_bandpass(self.h, _p(o), n_bins) != 0:
            raise ValueError("bin count mismatch")
        return o


def bandpass_probe(n_bins, start, width, params, rms, median=float("nan"), rebin=None):
    """BandPass set up on (n_bins, start, width) with params, rms and median,
    optionally rebinned to rebin = (n_bins2, start2, width2); returns
    (median, rms, intensityOfBin(0), params[0])."""
    p = np.ascontiguousarray(params, np.float32)
    o = np.zeros(4, np.float32)
    n2, s2, w2 = rebin if rebin is not None else (0, 0.0, 0.0)
    if lib().orc_bandpass_probe(n_bins, start, width, _p(p), p.size, rms, median, n2, s2, w2, _p(o)) != 0:
        raise ValueError("bad bandpass")
    return tuple(float(v) for v in o)


def read_bp(path, with_kills=False):
    """Numbers of a .bp bandpass file in order (BandPassAdapter::deserialise,
    src/lib/src/BandPassAdapter.cpp:30-100): '#' lines are comments, one
    number per line.  A line starting with 'K' holds an index and a
    frequency (:58-71): at least two fields, the first an unsigned integer;
    with_kills returns (numbers, sorted unique kill indices), otherwise a
    file with kill lines is refused."""
    vals, kills = [], set()
    with open(path) as fh:
        for n, raw in enumerate(fh, 1):
            if raw.startswith("#"):
                continue
            if raw.startswith("K"):
                nums = raw[1:].split()
                if len(nums) < 2:
                    raise ValueError("line %d: expecting two numbers on Kill line: index frequency" % n)
                if not nums[0].isdigit() or int(nums[0]) > 0xFFFFFFFF:
                    raise ValueError("line %d: cannot convert index to unsigned integer" % n)
                kills.add(int(nums[0]))
                continue
            line = raw.strip()
            if not line:
                continue
            vals.append(float(line))
    if len(vals) < 6:
        raise ValueError("expecting at least 6 parameters")
    if kills and not with_kills:
        raise ValueError("the file has kill lines: read it with with_kills=True")
    v = np.array(vals, np.float64)
    return (v, sorted(kills)) if with_kills else v


# --------------------------------------------------------- dedispersion
def dedisp_plan(n_chan, nsamp_buf, tdms, dm_low, dm_step, tsamp, fch1, foff):
    shift = np.zeros(n_chan, np.float32)
    ms = np.zeros(2, np.float32)
    ints = np.zeros(3, np.int32)
    rc = lib().orc_dedisp_plan(n_chan, nsamp_buf, tdms, dm_low, dm_step, tsamp, fch1, foff,
                               _p(shift), _p(ms), _p(ints))
    return {"rc": rc, "shift": shift, "mstartdm": ms[0], "mdmstep": ms[1],
            "maxshift": int(ints[0]), "rem": int(ints[1]), "nout": int(ints[2])}


def dedisp_delays(shift, tdms, mstartdm, mdmstep):
    shift = np.ascontiguousarray(shift, np.float32)
    out = np.zeros((tdms, shift.size), np.int32)
    lib().orc_dedisp_delays(_p(shift), shift.size, tdms, mstartdm, mdmstep, _p(out))
    return out


def dedisp_delays_fused(shift, tdms, mstartdm, mdmstep):
    """The delay table with the trial value as fmaf(dm, mdmstep, mstartdm)."""
    shift = np.ascontiguousarray(shift, np.float32)
    out = np.zeros((tdms, shift.size), np.int32)
    lib().orc_dedisp_delays_fused(_p(shift), shift.size, tdms, mstartdm, mdmstep, _p(out))
    return out


def dedisperse(inp, delays, n_out, nthreads=0):
    inp = np.ascontiguousarray(inp)
    n_chan, nsamp = inp.shape
    tdms = delays.shape[0]
    d = np.ascontiguousarray(delays, np.int32)
    if inp.dtype == np.float32:
        out = np.zeros((tdms, n_out), np.float32)
        lib().orc_dedisperse_f32(_p(inp), n_chan, nsamp, tdms, _p(d), n_out, _p(out), nthreads)
    elif inp.dtype == np.uint16:
        out = np.zeros((tdms, n_out), np.float32)
        lib().orc_dedisperse_f16(_p(inp), n_chan, nsamp, tdms, _p(d), n_out, _p(out), nthreads)
    elif inp.dtype == np.uint8:
        out = np.zeros((tdms, n_out), np.int32)
        lib().orc_dedisperse_u8(_p(inp), n_chan, nsamp, tdms, _p(d), n_out, _p(out), nthreads)
    else:
        raise TypeError(inp.dtype)
    return out


def f32_to_f16_bits(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.uint16)
    lib().orc_f32_to_f16(_p(x), _p(out), x.size)
    return out


def quantise_u8_rows(x, inv_scale):
    """The u8 buffer [nChan][nsamp] with exceptional rows: (codes, row offsets)."""
    x = np.ascontiguousarray(x, np.float32)
    n_chan, nsamp = x.shape
    q = np.zeros(x.shape, np.uint8)
    off = np.zeros(nsamp, np.float32)
    lib().orc_quantise_u8_rows(_p(x), n_chan, nsamp, inv_scale, _p(q), _p(off))
    return q, off


def u8_series(sums, off, delays, n_out, scale):
    """Dedispersed values of a u8 buffer's int32 sums with its row offsets."""
    sums = np.ascontiguousarray(sums, np.int32)
    off = np.ascontiguousarray(off, np.float32)
    d = np.ascontiguousarray(delays, np.int32)
    tdms, n_chan = d.shape
    out = np.zeros((tdms, n_out), np.float32)
    lib().orc_u8_series(_p(sums), _p(off), _p(d), n_chan, off.size, tdms, n_out, scale, _p(out))
    return out


def quantise_u8(x, inv_scale):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.uint8)
    lib().orc_quantise_u8(_p(x), _p(out), x.size, inv_scale)
    return out


class DDBuffer:
    def __init__(self, n_chan, nsamp_buf, overlap):
        self.n_chan, self.nsamp = n_chan, nsamp_buf
        self.h = lib().orc_ddbuf_create(n_chan, nsamp_buf, overlap)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ddbuf_destroy(self.h)
            self.h = None

    def add(self, spectra, max_out=8):
        s = np.ascontiguousarray(spectra, np.float32).reshape(spectra.shape[0], -1)
        out = np.zeros((max_out, self.n_chan, self.nsamp), np.float32)
        n = lib().orc_ddbuf_add(self.h, _p(s), s.shape[0], _p(out), max_out)
        assert n <= max_out
        return out[:n]


def detect(data, n_channels, n_subbands, threshold=6.0, bin_pow2=6, max_events=1 << 20):
    data = np.ascontiguousarray(data, np.float32)
    tdms, nsamp = data.shape
    ints = np.zeros((max_events, 3), np.int32)
    vals = np.zeros(max_events, np.float32)
    n = lib().orc_detect(_p(data), tdms, nsamp, n_channels, n_subbands, threshold, bin_pow2,
                         _p(ints), _p(vals), max_events)
    if n < 0:
        return detect(data, n_channels, n_subbands, threshold, bin_pow2, -n)
    return ints[:n].copy(), vals[:n].copy()
