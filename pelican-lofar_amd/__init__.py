"""Python access to the MI355X pulsar-search chain (libppl_amd.so).

The compute lives in the C-ABI library built from csrc/ (HIP for gfx950);
this module only binds it with ctypes and hands it device pointers and
streams owned by PyTorch.  There is no CPU fallback: when the library or a
GPU is missing every entry point raises.

The directory name is not a Python identifier; load it with
``load_package()`` from ``__graft_entry__`` (or importlib) as
``pelican_lofar_amd``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libppl_amd.so")

F32, F16, U8 = 0, 1, 2
DFT_FP32, DFT_BF16X3, DFT_FP16X2 = 0, 1, 2
DTYPE_NAMES = {"f32": F32, "f16": F16, "u8": U8}
PRECISION_NAMES = {"fp32": DFT_FP32, "bf16x3": DFT_BF16X3, "fp16x2": DFT_FP16X2}

_lib = None


class PPLError(RuntimeError):
    pass


def build(force=False, jobs=8):
    """Compile libppl_amd.so (hipcc --offload-arch=gfx950)."""
    cmd = ["make", "-s", "-j%d" % jobs, "-C", HERE]
    if force:
        cmd.append("-B")
    subprocess.check_call(cmd)


class DedispConfig(C.Structure):
    _fields_ = [("nChan", C.c_uint), ("nsampBuf", C.c_uint), ("tdms", C.c_uint),
                ("dmLow", C.c_float), ("dmStep", C.c_float), ("tsamp", C.c_double),
                ("fch1", C.c_double), ("foff", C.c_double), ("dtype", C.c_int)]


class DedispInfo(C.Structure):
    _fields_ = [("maxshift", C.c_int), ("remainingSamples", C.c_int), ("nOut", C.c_int),
                ("mstartdm", C.c_float), ("mdmstep", C.c_float), ("tileDM", C.c_int),
                ("tileT", C.c_int), ("chanBlock", C.c_int), ("readBytes", C.c_int),
                ("fusedDelayDiffs", C.c_longlong)]


class PPFConfig(C.Structure):
    _fields_ = [("nSub", C.c_uint), ("nPol", C.c_uint), ("nCh", C.c_uint), ("nTaps", C.c_uint),
                ("coeffs", C.c_void_p), ("precision", C.c_int), ("integrate", C.c_uint)]


class PacketSeq(C.Structure):
    """Stream position and counts of ppl_packet_sequence (zero for a new stream)."""
    _fields_ = [("startTime", C.c_uint), ("startBlockid", C.c_uint), ("accepted", C.c_ulonglong),
                ("rejected", C.c_ulonglong), ("lost", C.c_ulonglong)]


class ClipperConfig(C.Structure):
    _fields_ = [("bp_values", C.c_void_p), ("n_values", C.c_int), ("crFactor", C.c_float),
                ("srFactor", C.c_float), ("maxHistory", C.c_int), ("zeroDM", C.c_int),
                ("bandMatching", C.c_int), ("useMeanOverRMS", C.c_int), ("kill_index", C.c_void_p),
                ("n_kill", C.c_int)]


# name -> (restype, argtypes); every symbol declared in include/ppl_amd.h
vp, u, i, f, d, sz = C.c_void_p, C.c_uint, C.c_int, C.c_float, C.c_double, C.c_size_t
SIGNATURES = {
    "ppl_init": (i, [i]),
    "ppl_shutdown": (i, []),
    "ppl_last_error": (C.c_char_p, []),
    "ppl_version": (i, []),
    "ppl_device_alloc": (i, [vp, sz]),
    "ppl_device_free": (i, [vp]),
    "ppl_host_alloc": (i, [vp, sz]),
    "ppl_host_free": (i, [vp]),
    "ppl_memcpy_h2d": (i, [vp, vp, sz, vp]),
    "ppl_memcpy_d2h": (i, [vp, vp, sz, vp]),
    "ppl_memset_async": (i, [vp, i, sz, vp]),
    "ppl_stream_create": (i, [vp]),
    "ppl_stream_destroy": (i, [vp]),
    "ppl_stream_create_cu_range": (i, [i, i, vp]),
    "ppl_stream_sync": (i, [vp]),
    "ppl_stream_callback": (i, [vp, vp, vp]),
    "ppl_dedisp_plan_create": (i, [vp, vp]),
    "ppl_dedisp_plan_destroy": (i, [vp]),
    "ppl_dedisp_plan_info": (i, [vp, vp]),
    "ppl_dedisp_set_tile": (i, [vp, i, i]),
    "ppl_dedisp_set_shape": (i, [vp, i, i, i]),
    "ppl_dedisp_set_kernel": (i, [vp, i, i, i, i]),
    "ppl_dedisp_autotune": (i, [vp, vp, vp, i, vp, vp]),
    "ppl_dedisp_delays": (i, [vp, vp]),
    "ppl_dedisperse": (i, [vp, vp, vp, vp]),
    "ppl_dedisperse_batch": (i, [vp, i, i, i, i, i, vp]),
    "ppl_dedisp_autotune_batch": (i, [vp, i, i, vp, vp, vp, vp, vp, vp]),
    "ppl_cache_dedisperse_loop": (i, [vp, C.c_long, vp, f, f, i, i, vp, i, i, vp]),
    "ppl_convert_f16": (i, [vp, vp, sz, vp]),
    "ppl_quantise_u8": (i, [vp, vp, sz, f, vp]),
    "ppl_ddbuf_create": (i, [vp, u, u, u, i, f]),
    "ppl_ddbuf_destroy": (i, [vp]),
    "ppl_ddbuf_add": (i, [vp, vp, u, u, vp, vp, vp]),
    "ppl_ddbuf_add_blob": (i, [vp, vp, u, u, u, u, u, vp, vp, vp]),
    "ppl_ddbuf_rotate": (i, [vp, vp]),
    "ppl_ddbuf_data": (vp, [vp]),
    "ppl_ddbuf_fill": (u, [vp]),
    "ppl_detect": (i, [vp, i, i, i, i, f, u, vp, vp, i, vp, vp]),
    "ppl_dedisp_u8_series": (i, [vp, vp, vp, f, vp, vp]),
    "ppl_ddbuf_row_offsets": (vp, [vp]),
    "ppl_detect_noalloc": (i, [vp, i, i, i, i, f, u, vp, vp, i, vp, vp]),
    "ppl_detect_u8": (i, [vp, i, i, i, f, i, i, f, u, vp, vp, i, vp, vp]),
    "ppl_dequantise_sums": (i, [vp, vp, sz, u, f, vp]),
    "ppl_ingest_c32": (i, [vp, sz, u, u, u, u, u, i, vp, vp]),
    "ppl_ingest_i16": (i, [vp, sz, u, u, u, u, u, i, vp, vp]),
    "ppl_packet_sequence": (i, [vp, vp, sz, u, u, i, vp, u, vp, vp]),
    "ppl_ingest_c32_slots": (i, [vp, sz, u, u, vp, u, u, u, u, i, vp, vp]),
    "ppl_ingest_i16_slots": (i, [vp, sz, u, u, vp, u, u, u, u, i, vp, vp]),
    "ppl_bandpass_accumulate": (i, [vp, i, i, i, i, vp, vp]),
    "ppl_sigproc_pack": (i, [vp, u, u, u, u, u, i, f, f, vp, vp]),
    "ppl_ppf_coefficients": (i, [u, u, i, vp]),
    "ppl_ppf_plan_create": (i, [vp, vp]),
    "ppl_ppf_plan_destroy": (i, [vp]),
    "ppl_ppf_reset": (i, [vp, vp]),
    "ppl_ppf_run_i16": (i, [vp, vp, u, vp, vp]),
    "ppl_ppf_run_c32": (i, [vp, vp, u, vp, vp]),
    "ppl_ppf_stokes_i16": (i, [vp, vp, u, vp, vp]),
    "ppl_ppf_stokes_i16_batch": (i, [vp, i, vp]),
    "ppl_stokes": (i, [vp, u, u, u, u, vp, vp]),
    "ppl_integrate": (i, [vp, u, u, u, u, u, u, vp, vp]),
    "ppl_clipper_create": (i, [vp, vp]),
    "ppl_clipper_destroy": (i, [vp]),
    "ppl_clipper_run": (i, [vp, vp, vp, u, u, u, u, vp]),
    "ppl_clipper_run_into": (i, [vp, vp, u, u, u, vp, u, vp, vp, vp]),
    "ppl_clipper_stats": (i, [vp, vp, vp]),
    "ppl_clipper_flags": (i, [vp, vp, u]),
    "ppl_clipper_model": (i, [vp, vp, vp, vp, vp, u]),
    "ppl_clipper_counters": (i, [vp, vp, i]),
    "ppl_clipper_run_batch": (i, [vp, vp, i, vp]),
    "ppl_clipper_run_front": (i, [vp, vp, vp, u, u, u, u, vp]),
    "ppl_clipper_run_walk": (i, [vp, vp]),
    "ppl_clipper_run_back": (i, [vp, vp]),
}


def lib():
    """Load libppl_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PPLError("libppl_amd.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise PPLError("ppl error %d: %s" % (rc, lib().ppl_last_error().decode()))
    return rc


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class CuRangeStream:
    """A stream whose kernels run only on CU slots [first, first + count) of
    every XCD (ppl_stream_create_cu_range); .stream is the torch view."""

    def __init__(self, first, count):
        import torch
        h = C.c_void_p()
        check(lib().ppl_stream_create_cu_range(first, count, C.byref(h)))
        self.h = h
        self.stream = torch.cuda.ExternalStream(h.value)

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            self.stream.synchronize()
            _lib.ppl_stream_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def init(device=0):
    import torch
    if not torch.cuda.is_available():
        raise PPLError("no GPU visible: the MI355X kernels need a device")
    check(lib().ppl_init(device))


# ----------------------------------------------------------- dedispersion
class DedispPlan:
    """DedispersionModule's kernel configuration (resize arithmetic + delay table)."""

    def __init__(self, n_chan, nsamp_buf, tdms, dm_low, dm_step, tsamp, fch1, foff, dtype="f32"):
        cfg = DedispConfig(n_chan, nsamp_buf, tdms, dm_low, dm_step, tsamp, fch1, foff,
                           DTYPE_NAMES[dtype] if isinstance(dtype, str) else dtype)
        h = C.c_void_p()
        check(lib().ppl_dedisp_plan_create(C.byref(h), C.byref(cfg)))
        self.h = h
        self.cfg = cfg
        self.dtype = cfg.dtype
        info = DedispInfo()
        check(lib().ppl_dedisp_plan_info(h, C.byref(info)))
        self.maxshift, self.rem, self.nout = info.maxshift, info.remainingSamples, info.nOut
        self.mstartdm, self.mdmstep = info.mstartdm, info.mdmstep
        self.tile = (info.tileDM, info.tileT, info.chanBlock, info.readBytes)
        # delays that a contracted (fused multiply-add) trial value would change
        self.fused_delay_diffs = info.fusedDelayDiffs

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ppl_dedisp_plan_destroy(self.h)
            self.h = None

    def _refresh(self):
        info = DedispInfo()
        check(lib().ppl_dedisp_plan_info(self.h, C.byref(info)))
        self.tile = (info.tileDM, info.tileT, info.chanBlock, info.readBytes)

    def set_tile(self, tile_dm, tile_t, chan_block=0, read_bytes=4):
        """Tile (DM trials x samples), channels staged per block (0: as many
        as fit) and LDS read width per lane (4, or 16 for FP32 input)."""
        check(lib().ppl_dedisp_set_kernel(self.h, tile_dm, tile_t, chan_block, read_bytes))
        self._refresh()

    def autotune(self, d_in, d_out, reps=2, stream=None):
        """Time every tile shape / channel block on these buffers and keep the fastest."""
        best = C.c_float(0)
        check(lib().ppl_dedisp_autotune(self.h, _ptr(d_in), _ptr(d_out), reps, _stream(stream),
                                        C.byref(best)))
        self._refresh()
        return best.value

    def delays(self):
        out = np.zeros((self.cfg.tdms, self.cfg.nChan), np.int32)
        check(lib().ppl_dedisp_delays(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def out_dtype(self):
        import torch
        return torch.int32 if self.dtype == U8 else torch.float32

    def run(self, d_in, d_out=None, stream=None):
        import torch
        if d_out is None:
            d_out = torch.empty((self.cfg.tdms, self.nout), dtype=self.out_dtype(), device=d_in.device)
        check(lib().ppl_dedisperse(self.h, _ptr(d_in), _ptr(d_out), _stream(stream)))
        return d_out

    def u8_series(self, d_sums, d_row_offsets, scale, d_out=None, stream=None):
        """Float series of a u8 plan's int32 sums: dequantised, with the
        buffer's exceptional rows (d_row_offsets, DDBuffer.row_offsets())
        restored at their own level (ppl_dedisp_u8_series)."""
        import torch
        if d_out is None:
            d_out = torch.empty((self.cfg.tdms, self.nout), dtype=torch.float32, device=d_sums.device)
        check(lib().ppl_dedisp_u8_series(self.h, _ptr(d_sums), _ptr(d_row_offsets), scale, _ptr(d_out),
                                         _stream(stream)))
        return d_out


class DedispJob(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("d_in", C.c_void_p), ("d_out", C.c_void_p)]


MAX_DEDISP_JOBS = 48


def _dedisp_jobs(plans, d_ins, d_outs):
    n = len(plans)
    if not (len(d_ins) == n == len(d_outs)):
        raise PPLError("plans, inputs and outputs differ in number")
    arr = (DedispJob * n)()
    for k in range(n):
        arr[k] = DedispJob(plans[k].h, _ptr(d_ins[k]), _ptr(d_outs[k]))
    return arr, n


def _tile4(tile):
    """(tileDM, tileT, chanBlock[, readBytes]) with the read width filled in (4 bytes)."""
    t = tuple(int(v) for v in tile)
    if len(t) == 3:
        t = t + (4,)
    if len(t) != 4:
        raise PPLError("tile must be (tileDM, tileT, chanBlock[, readBytes])")
    return t


def dedisperse_batch(plans, d_ins, d_outs=None, tile=(0, 0, 0, 4), stream=None):
    """One launch dedispersing every (plan, input) pair under one tile; outputs
    as separate DedispPlan.run calls would give them.  tile="own": one launch
    per group of plans that share their own (autotuned) tile."""
    import torch
    if d_outs is None:
        d_outs = [torch.empty((p.cfg.tdms, p.nout), dtype=p.out_dtype(), device=x.device)
                  for p, x in zip(plans, d_ins)]
    if isinstance(tile, str):
        if tile != "own":
            raise PPLError("tile must be a (tileDM, tileT, chanBlock[, readBytes]) tuple or 'own'")
        groups = {}
        for k, p in enumerate(plans):
            groups.setdefault(tuple(p.tile), []).append(k)
        for t, ks in groups.items():
            dedisperse_batch([plans[k] for k in ks], [d_ins[k] for k in ks], [d_outs[k] for k in ks], tile=t,
                             stream=stream)
        return d_outs
    arr, n = _dedisp_jobs(plans, d_ins, d_outs)
    t = _tile4(tile)
    check(lib().ppl_dedisperse_batch(arr, n, t[0], t[1], t[2], t[3], _stream(stream)))
    return d_outs


def autotune_batch(plans, d_ins, d_outs, reps=1, stream=None, candidates=None):
    """Fastest (tileDM, tileT, chanBlock, readBytes) for the batch, and its time
    in ms: every instantiated tile, channel block and read width, or only the given candidates
    (each timed over reps launches after one untimed launch)."""
    if candidates is not None:
        import torch
        best = None
        for cand in candidates:
            try:
                dedisperse_batch(plans, d_ins, d_outs, tile=cand, stream=stream)
            except PPLError:
                continue  # does not fit every job
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                dedisperse_batch(plans, d_ins, d_outs, tile=cand, stream=stream)
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1) / reps
            if best is None or ms < best[1]:
                best = (_tile4(cand), ms)
        if best is None:
            raise PPLError("autotune: no candidate tile fits every job")
        return best
    arr, n = _dedisp_jobs(plans, d_ins, d_outs)
    td, tt, cb, rb, best = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_float()
    check(lib().ppl_dedisp_autotune_batch(arr, n, reps, _stream(stream), C.byref(td), C.byref(tt),
                                          C.byref(cb), C.byref(rb), C.byref(best)))
    return (td.value, tt.value, cb.value, rb.value), best.value


def cache_dedisperse_loop(outbuff, buff, mstartdm, mdmstep, tdms, num_samples, dm_shift, maxshift,
                          nchans, stream=None):
    check(lib().ppl_cache_dedisperse_loop(_ptr(outbuff), outbuff.numel() * 4, _ptr(buff), mstartdm,
                                          mdmstep, tdms, num_samples, _ptr(dm_shift), maxshift,
                                          nchans, _stream(stream)))


def convert_f16(d_in, d_out, stream=None):
    check(lib().ppl_convert_f16(_ptr(d_in), _ptr(d_out), d_in.numel(), _stream(stream)))


def quantise_u8(d_in, d_out, inv_scale, stream=None):
    check(lib().ppl_quantise_u8(_ptr(d_in), _ptr(d_out), d_in.numel(), inv_scale, _stream(stream)))


class DDBuffer:
    """Device-resident dedispersion buffer with overlap carry."""

    def __init__(self, n_chan, nsamp_buf, overlap, dtype="f32", inv_scale=1.0):
        h = C.c_void_p()
        dt = DTYPE_NAMES[dtype] if isinstance(dtype, str) else dtype
        check(lib().ppl_ddbuf_create(C.byref(h), n_chan, nsamp_buf, overlap, dt, inv_scale))
        self.h, self.n_chan, self.nsamp, self.dtype = h, n_chan, nsamp_buf, dt

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ppl_ddbuf_destroy(self.h)
            self.h = None

    def add(self, d_spectra, first=0, stream=None):
        consumed, full = C.c_uint(0), C.c_int(0)
        n_t = d_spectra.shape[0]
        check(lib().ppl_ddbuf_add(self.h, _ptr(d_spectra), n_t, first, C.byref(consumed),
                                  C.byref(full), _stream(stream)))
        return consumed.value, bool(full.value)

    def add_blob(self, d_blob, first=0, stream=None):
        """addSamples from a [nT][nSub][nPol][nCh] blob (polarisation 0)."""
        consumed, full = C.c_uint(0), C.c_int(0)
        n_t, n_sub, n_pol, n_ch = d_blob.shape
        check(lib().ppl_ddbuf_add_blob(self.h, _ptr(d_blob), n_t, first, n_sub, n_pol, n_ch,
                                       C.byref(consumed), C.byref(full), _stream(stream)))
        return consumed.value, bool(full.value)

    def rotate(self, stream=None):
        check(lib().ppl_ddbuf_rotate(self.h, _stream(stream)))

    def data_ptr(self):
        return lib().ppl_ddbuf_data(self.h)

    def fill(self):
        return lib().ppl_ddbuf_fill(self.h)

    def tensor(self):
        """A torch view of the current buffer (shares device memory)."""
        import torch
        dt = {F32: torch.float32, F16: torch.float16, U8: torch.uint8}[self.dtype]
        return _device_view(self.data_ptr(), (self.n_chan, self.nsamp), dt)

    def row_offsets(self):
        """u8 buffers: a torch view of the current buffer's row offsets
        ([nsampBuf] float32; nonzero for rows quantised relative to their
        middle channel), else None."""
        import torch
        ptr = lib().ppl_ddbuf_row_offsets(self.h)
        return None if not ptr else _device_view(ptr, (self.nsamp,), torch.float32)


def _device_view(ptr, shape, dtype):
    import torch
    n = int(np.prod(shape))
    elt = torch.empty((), dtype=dtype).element_size()

    class _Holder:
        __cuda_array_interface__ = {
            "shape": (n,), "typestr": {torch.float32: "<f4", torch.float16: "<f2",
                                       torch.uint8: "|u1", torch.int32: "<i4"}[dtype],
            "data": (ptr, False), "version": 3, "strides": (elt,)}
    return torch.as_tensor(_Holder(), device="cuda").view(shape)


# --------------------------------------------------------------- detection
def detect(d_data, n_channels, n_subbands, threshold=6.0, bin_pow2=6, max_events=1 << 20,
           stream=None):
    """DedispersionAnalyser::analyse on a device [tdms][nsamp] float tensor."""
    tdms, nsamp = d_data.shape
    # uninitialised: only the pages of the events written are ever touched,
    # and the results are views of these arrays (no second copy)
    ints = np.empty((max(max_events, 1), 3), np.int32)
    vals = np.empty(max(max_events, 1), np.float32)
    count = C.c_int(0)
    check(lib().ppl_detect(_ptr(d_data), tdms, nsamp, n_channels, n_subbands, threshold, bin_pow2,
                           ints.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p),
                           max_events, C.byref(count), _stream(stream)))
    n = min(count.value, max_events)
    return ints[:n], vals[:n], count.value


DETECT_NOALLOC_MAX_SEGMENTS = 1 << 23


def detect_noalloc(d_data, n_channels, n_subbands, d_ints, d_vals, d_count, threshold=6.0, bin_pow2=6,
                   stream=None):
    """ppl_detect_noalloc: detection into device tensors d_ints [max][3]
    int32, d_vals [max] float32 and d_count [1] int32, with the library's
    static scratch (nothing allocated, nothing synchronised)."""
    tdms, nsamp = d_data.shape
    check(lib().ppl_detect_noalloc(_ptr(d_data), tdms, nsamp, n_channels, n_subbands, threshold, bin_pow2,
                                   _ptr(d_ints), _ptr(d_vals), d_vals.shape[0], _ptr(d_count), _stream(stream)))


def detect_u8(d_data, n_chan, scale, n_channels, n_subbands, threshold=6.0, bin_pow2=6,
              max_events=1 << 20, stream=None):
    tdms, nsamp = d_data.shape
    ints = np.empty((max(max_events, 1), 3), np.int32)  # as in detect()
    vals = np.empty(max(max_events, 1), np.float32)
    count = C.c_int(0)
    check(lib().ppl_detect_u8(_ptr(d_data), tdms, nsamp, n_chan, scale, n_channels, n_subbands,
                              threshold, bin_pow2, ints.ctypes.data_as(C.c_void_p),
                              vals.ctypes.data_as(C.c_void_p), max_events, C.byref(count),
                              _stream(stream)))
    n = min(count.value, max_events)
    return ints[:n], vals[:n], count.value


# -------------------------------------------------------------- channeliser
WINDOWS = {"kaiser": 0, "gaussian": 1, "blackman": 2, "hamming": 3, "hann": 4}


def sequence_packets(state, h_packets, n_spp, n_slots, clock_mhz=200):
    """Place received packets (a host uint8 array [packet][bytes], LOFAR
    headers first) in n_slots chunk slots as LofarDataSplittingChunker does;
    returns (slot sources: int32 array, -1 = lost packet, packets used).
    state: a PacketSeq, updated."""
    raw = np.ascontiguousarray(h_packets, np.uint8)
    slots = np.full(n_slots, -2, np.int32)
    used, filled = C.c_uint(0), C.c_uint(0)
    check(lib().ppl_packet_sequence(C.byref(state), raw.ctypes.data_as(C.c_void_p), raw.shape[1], raw.shape[0],
                                    n_spp, clock_mhz, slots.ctypes.data_as(C.c_void_p), n_slots, C.byref(used),
                                    C.byref(filled)))
    return slots[:filled.value], used.value


def ingest(d_packets, n_packets, n_sub, n_spp, n_pol=2, sample_bits=16, header_bytes=0, packet_bytes=None,
           out="c32", stream=None, slots=None):
    """Packet payloads -> [sub][pol][t] samples (complex float or int16 pairs)
    on the device (AdapterTimeSeriesDataSet::_readData); with `slots`, chunk
    slot s holds packet slots[s], or zeros where it is -1."""
    import torch
    payload = n_sub * n_spp * n_pol * 2 * (sample_bits // 8)
    pb = packet_bytes if packet_bytes is not None else header_bytes + payload
    if slots is not None:
        # slots: a device int32 tensor of packet indices, -1 for a lost packet
        n_slots = slots.numel()
        n_samp = n_spp * n_slots
        if out == "c32":
            o = torch.empty((n_sub, n_pol, n_samp, 2), dtype=torch.float32, device=d_packets.device)
            check(lib().ppl_ingest_c32_slots(_ptr(d_packets), pb, header_bytes, n_packets, _ptr(slots), n_slots,
                                             n_sub, n_spp, n_pol, sample_bits, _ptr(o), _stream(stream)))
        else:
            o = torch.empty((n_sub, n_pol, n_samp, 2), dtype=torch.int16, device=d_packets.device)
            check(lib().ppl_ingest_i16_slots(_ptr(d_packets), pb, header_bytes, n_packets, _ptr(slots), n_slots,
                                             n_sub, n_spp, n_pol, sample_bits, _ptr(o), _stream(stream)))
        return o
    n_samp = n_spp * n_packets
    if out == "c32":
        o = torch.empty((n_sub, n_pol, n_samp, 2), dtype=torch.float32, device=d_packets.device)
        check(lib().ppl_ingest_c32(_ptr(d_packets), pb, header_bytes, n_packets, n_sub, n_spp, n_pol,
                                   sample_bits, _ptr(o), _stream(stream)))
    else:
        o = torch.empty((n_sub, n_pol, n_samp, 2), dtype=torch.int16, device=d_packets.device)
        check(lib().ppl_ingest_i16(_ptr(d_packets), pb, header_bytes, n_packets, n_sub, n_spp, n_pol,
                                   sample_bits, _ptr(o), _stream(stream)))
    return o


def bandpass_accumulate(stokes, d_sum, stream=None):
    """d_sum[s*nCh + c] += stokes[t, s, 0, c] over t in order, float
    (BandPassRecorder::run's integrated spectrum).  stokes is a float32
    [t][s][pol][c] device tensor, d_sum a float32 device tensor of s*nCh."""
    import torch
    n_t, n_sub, n_pol, n_ch = stokes.shape
    if stokes.dtype != torch.float32 or not stokes.is_contiguous():
        raise ValueError("stokes must be a contiguous float32 tensor [t][s][pol][c]")
    if d_sum.numel() != n_sub * n_ch or d_sum.dtype != stokes.dtype:
        raise ValueError("d_sum must hold nSubbands*nChannels float32 values")
    check(lib().ppl_bandpass_accumulate(_ptr(stokes), n_t, n_sub, n_pol, n_ch, _ptr(d_sum), _stream(stream)))
    return d_sum


def sigproc_pack(stokes, n_pols_write=1, n_bits=32, crop_min=0.0, crop_max=255.0, stream=None):
    """The filterbank bytes of a Stokes blob [t][s][pol][c] as
    SigprocStokesWriter writes them (subbands and channels reversed; 32-bit
    floats or 8-bit scaled bytes); returns a device tensor (float32 or uint8)
    of shape [t][n_pols_write][s][c]."""
    import torch
    n_t, n_sub, n_pol, n_ch = stokes.shape
    if stokes.dtype != torch.float32 or not stokes.is_contiguous():
        raise ValueError("stokes must be a contiguous float32 tensor [t][s][pol][c]")
    dt = torch.float32 if n_bits == 32 else torch.uint8
    out = torch.empty((n_t, n_pols_write, n_sub, n_ch), dtype=dt, device=stokes.device)
    check(lib().ppl_sigproc_pack(_ptr(stokes), n_t, n_sub, n_pol, n_ch, n_pols_write, n_bits, crop_min, crop_max,
                                 _ptr(out), _stream(stream)))
    return out


def ppf_coefficients(n_taps, n_ch, window="kaiser"):
    """FP64 FIR coefficients (PolyphaseCoefficients::genereateFilter), host code."""
    out = np.zeros(n_taps * n_ch, np.float64)
    check(lib().ppl_ppf_coefficients(n_taps, n_ch, WINDOWS[window], out.ctypes.data_as(C.c_void_p)))
    return out


class PPF:
    """PPFChanneliser with device-resident FIR history."""

    def __init__(self, n_sub, n_ch, n_taps, coeffs, precision="fp32", integrate=1, n_pol=2):
        c = np.ascontiguousarray(coeffs, np.float32)
        assert c.size == n_taps * n_ch
        prec = PRECISION_NAMES[precision] if isinstance(precision, str) else precision
        cfg = PPFConfig(n_sub, n_pol, n_ch, n_taps, c.ctypes.data_as(C.c_void_p), prec, integrate)
        h = C.c_void_p()
        check(lib().ppl_ppf_plan_create(C.byref(h), C.byref(cfg)))
        self.h = h
        self.n_sub, self.n_ch, self.n_taps, self.integrate = n_sub, n_ch, n_taps, integrate

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ppl_ppf_plan_destroy(self.h)
            self.h = None

    def reset(self, stream=None):
        check(lib().ppl_ppf_reset(self.h, _stream(stream)))

    def spectra_i16(self, d_in, n_blk, d_out=None, stream=None):
        import torch
        if d_out is None:
            d_out = torch.empty((n_blk, self.n_sub, 2, self.n_ch, 2), dtype=torch.float32,
                                device=d_in.device)
        check(lib().ppl_ppf_run_i16(self.h, _ptr(d_in), n_blk, _ptr(d_out), _stream(stream)))
        return d_out

    def spectra_c32(self, d_in, n_blk, d_out=None, stream=None):
        import torch
        if d_out is None:
            d_out = torch.empty((n_blk, self.n_sub, 2, self.n_ch, 2), dtype=torch.float32,
                                device=d_in.device)
        check(lib().ppl_ppf_run_c32(self.h, _ptr(d_in), n_blk, _ptr(d_out), _stream(stream)))
        return d_out

    def stokes_i16(self, d_in, n_blk, d_out=None, stream=None):
        import torch
        if d_out is None:
            d_out = torch.empty((n_blk // self.integrate, self.n_sub, 1, self.n_ch),
                                dtype=torch.float32, device=d_in.device)
        check(lib().ppl_ppf_stokes_i16(self.h, _ptr(d_in), n_blk, _ptr(d_out), _stream(stream)))
        return d_out


class PPFJob(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("d_in", C.c_void_p), ("nBlk", C.c_uint), ("strideBlk", C.c_uint),
                ("d_stokes", C.c_void_p)]


MAX_PPF_JOBS = 48


def ppf_stokes_i16_batch(jobs, stream=None):
    """PPF.stokes_i16 of many plans in one launch per (channel count, taps,
    DFT method) and one history launch: jobs = [(PPF, int16 voltages
    [sub][pol][samples][2], n_blk, d_out)].  The first n_blk blocks of each
    (sub, pol) stream are read, so a chunk can be a prefix of a longer
    voltage buffer; d_out is [n_blk / integrate][sub][1][nCh] (contiguous)."""
    n = len(jobs)
    if not 1 <= n <= MAX_PPF_JOBS:
        raise PPLError("%d channeliser jobs (1..%d)" % (n, MAX_PPF_JOBS))
    arr = (PPFJob * n)()
    for k, (ppf, d_in, n_blk, d_out) in enumerate(jobs):
        if not (d_in.is_contiguous() and d_out.is_contiguous()):
            raise PPLError("job %d: voltages and output must be contiguous" % k)
        if d_in.shape[0] != ppf.n_sub or d_in.shape[1] != 2 or d_in.shape[2] % ppf.n_ch:
            raise PPLError("job %d: voltages of shape %s for %d subbands x %d channels"
                           % (k, tuple(d_in.shape), ppf.n_sub, ppf.n_ch))
        stride = d_in.shape[2] // ppf.n_ch
        if n_blk > stride:
            raise PPLError("job %d: %d blocks from a buffer of %d" % (k, n_blk, stride))
        if d_out.numel() < (n_blk // ppf.integrate) * ppf.n_sub * ppf.n_ch:
            raise PPLError("job %d: output too small" % k)
        arr[k] = PPFJob(ppf.h, _ptr(d_in), n_blk, stride, _ptr(d_out))
    check(lib().ppl_ppf_stokes_i16_batch(arr, n, _stream(stream)))


def stokes(d_spectra, n_stokes=1, d_out=None, stream=None):
    import torch
    n_blk, n_sub, _, n_ch, _ = d_spectra.shape
    if d_out is None:
        d_out = torch.empty((n_blk, n_sub, n_stokes, n_ch), dtype=torch.float32,
                            device=d_spectra.device)
    check(lib().ppl_stokes(_ptr(d_spectra), n_blk, n_sub, n_ch, n_stokes, _ptr(d_out),
                           _stream(stream)))
    return d_out


def integrate(d_in, window, bin_ch=1, d_out=None, stream=None):
    import torch
    n_t, n_sub, n_pol, n_ch = d_in.shape
    w = min(window, n_t)
    if d_out is None:
        d_out = torch.empty((n_t // w, n_sub, n_pol, n_ch // bin_ch), dtype=torch.float32,
                            device=d_in.device)
    check(lib().ppl_integrate(_ptr(d_in), n_t, n_sub, n_pol, n_ch, window, bin_ch, _ptr(d_out),
                              _stream(stream)))
    return d_out


# ------------------------------------------------------------------ clipper
class ClipJob(C.Structure):
    _fields_ = [("d_I", C.c_void_p), ("d_W", C.c_void_p), ("nT", C.c_uint), ("nSub", C.c_uint),
                ("nPol", C.c_uint), ("nCh", C.c_uint)]


def clipper_run_batch(clippers, d_spectra, d_weights=None, stream=None):
    """RFI_Clipper::run on several beams at once (one launch per stage for all
    beams); d_spectra[k] is beam k's [nT][nSub][nPol][nCh] float tensor."""
    n = len(clippers)
    handles = (C.c_void_p * n)(*[c.h for c in clippers])
    jobs = (ClipJob * n)()
    for k, x in enumerate(d_spectra):
        n_t, n_sub, n_pol, n_ch = x.shape
        w = None if d_weights is None else d_weights[k]
        jobs[k] = ClipJob(x.data_ptr(), None if w is None else w.data_ptr(), n_t, n_sub, n_pol, n_ch)
    check(lib().ppl_clipper_run_batch(handles, jobs, n, _stream(stream)))


class Clipper:
    """RFI_Clipper with its running state on the device."""

    def __init__(self, bp_values, cr=10.0, sr=4.0, history=10000, zero_dm=True,
                 band_matching=True, mean_over_rms=True, kill=()):
        """kill: the indices of the .bp file's K lines (killed channels)."""
        self._bp = np.ascontiguousarray(bp_values, np.float64)
        self._kill = np.ascontiguousarray(list(kill), np.uint32)
        cfg = ClipperConfig(self._bp.ctypes.data_as(C.c_void_p), self._bp.size, cr, sr, history,
                            int(zero_dm), int(band_matching), int(mean_over_rms),
                            self._kill.ctypes.data_as(C.c_void_p) if self._kill.size else None,
                            int(self._kill.size))
        h = C.c_void_p()
        check(lib().ppl_clipper_create(C.byref(h), C.byref(cfg)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ppl_clipper_destroy(self.h)
            self.h = None

    def run(self, d_I, d_W=None, stream=None):
        n_t, n_sub, n_pol, n_ch = d_I.shape
        check(lib().ppl_clipper_run(self.h, _ptr(d_I), _ptr(d_W), n_t, n_sub, n_pol, n_ch,
                                    _stream(stream)))
        return d_I

    def run_front(self, d_I, d_W=None, stream=None):
        """First of the three calls of a run split over streams: the run's
        start and the precompute of all its spectra."""
        n_t, n_sub, n_pol, n_ch = d_I.shape
        check(lib().ppl_clipper_run_front(self.h, _ptr(d_I), _ptr(d_W), n_t, n_sub, n_pol, n_ch,
                                          _stream(stream)))

    def run_walk(self, stream=None):
        """The sequential walk of the run started by run_front."""
        check(lib().ppl_clipper_run_walk(self.h, _stream(stream)))

    def run_back(self, stream=None):
        """The statistics and write passes of the run; the spectra then equal
        run()'s."""
        check(lib().ppl_clipper_run_back(self.h, _stream(stream)))

    def run_into(self, d_I, ddbuf, first=0, stream=None):
        """Clip d_I [nT][nSub][1][nCh] in place, then add spectra [first, nT)
        to ddbuf; returns (consumed, full) as DDBuffer.add does."""
        n_t, n_sub, n_pol, n_ch = d_I.shape
        if n_pol != 1:
            raise PPLError("run_into takes Stokes-I spectra (nPol = 1)")
        consumed, full = C.c_uint(0), C.c_int(0)
        check(lib().ppl_clipper_run_into(self.h, _ptr(d_I), n_t, n_sub, n_ch, ddbuf.h, first,
                                         C.byref(consumed), C.byref(full), _stream(stream)))
        return consumed.value, bool(full.value)

    def stats(self):
        m, r = C.c_float(0), C.c_float(0)
        check(lib().ppl_clipper_stats(self.h, C.byref(m), C.byref(r)))
        return m.value, r.value

    COUNTER_NAMES = ("fast", "slow", "learn", "box", "undecided", "fill", "ver", "stop", "resolved")

    def counters(self):
        """Walk counters since creation: spectra decided from the precomputed
        mask ("fast"), spectra on the full path ("slow") and the reason they
        took it, and sub-chunks ended early ("stop")."""
        out = (C.c_longlong * len(self.COUNTER_NAMES))()
        check(lib().ppl_clipper_counters(self.h, out, len(self.COUNTER_NAMES)))
        return dict(zip(self.COUNTER_NAMES, (int(v) for v in out)))

    def flags(self, n_t):
        out = np.zeros(n_t, np.uint8)
        check(lib().ppl_clipper_flags(self.h, out.ctypes.data_as(C.c_void_p), n_t))
        return out

    def model(self, n_bins=None):
        """The live bandpass model (RFI_Clipper::bandPass()): (params[0],
        mean, rms) and, with n_bins (the run map's bins), its values."""
        p0, m, r = C.c_float(), C.c_float(), C.c_float()
        vals = None if n_bins is None else np.zeros(n_bins, np.float32)
        check(lib().ppl_clipper_model(self.h, C.byref(p0), C.byref(m), C.byref(r),
                                      None if vals is None else vals.ctypes.data_as(C.c_void_p),
                                      0 if vals is None else n_bins))
        return (p0.value, m.value, r.value) if vals is None else (p0.value, m.value, r.value, vals)


# ------------------------------------------------------------- beam chain
class BeamChain:
    """One beam's device-resident chain, stepped one dedispersion buffer at a
    time: fused PPF -> Stokes-I -> integration, RFI clipper, transposition
    into the dedispersion buffer, dedispersion, overlap carry (the
    DedispersionPipeline stage order).  A step consumes exactly nOut new
    integrated spectra, i.e. nOut*integrate*nCh voltage samples per subband
    and polarisation, after prime() has filled the buffer's overlap.

    Every launch goes to the current stream, so a step can be captured in a
    HIP graph (capture_steps): the channeliser's FIR history and the
    dedispersion buffer alternate between two device copies, so two graphs
    (even and odd steps) replay the chain.  Detection synchronises and stays
    outside the graphs."""

    def __init__(self, n_sub, n_ch, n_taps, coeffs, integrate, bp_values, nsamp_buf, tdms, dm_low,
                 dm_step, tsamp, fch1, foff, dtype="f32", inv_scale=127.0 / 10.0, cr=10.0, sr=4.0,
                 history=10000, zero_dm=True, tile=None, dft="fp32"):
        import torch
        self.n_sub, self.n_ch, self.integrate = n_sub, n_ch, integrate
        self.n_chan = n_sub * n_ch
        self.dtype, self.inv_scale = dtype, inv_scale
        self.plan = DedispPlan(self.n_chan, nsamp_buf, tdms, dm_low, dm_step, tsamp, fch1, foff, dtype=dtype)
        self.nout, self.overlap = self.plan.nout, self.plan.maxshift + self.plan.rem
        self.ppf = PPF(n_sub, n_ch, n_taps, np.asarray(coeffs, np.float32), precision=dft, integrate=integrate)
        self.clip = Clipper(bp_values, cr, sr, history, zero_dm)
        self.buf = DDBuffer(self.n_chan, nsamp_buf, self.overlap, dtype, inv_scale)
        self.stokes = torch.empty((self.nout, n_sub, 1, n_ch), dtype=torch.float32, device="cuda")
        self.out = torch.empty((tdms, self.nout), dtype=self.plan.out_dtype(), device="cuda")
        # u8: the float series the detector reads (dequantised sums with the
        # buffer's exceptional rows restored); self.out holds the int32 sums
        self.series = (torch.empty((tdms, self.nout), dtype=torch.float32, device="cuda")
                       if dtype == "u8" else None)
        self.graphs = None
        self.parity = 0
        if tile is not None:
            self.plan.set_tile(*tile)

    def samples_per_step(self):
        return self.nout * self.integrate * self.n_ch

    def autotune(self, reps=1):
        import torch
        tdt = {"f32": torch.float32, "f16": torch.float16, "u8": torch.uint8}[self.dtype]
        probe = torch.zeros((self.n_chan, self.plan.cfg.nsampBuf), dtype=tdt, device="cuda")
        ms = self.plan.autotune(probe, self.out, reps=reps)
        del probe
        return ms

    def prime(self, d_volt):
        """Fill the buffer's overlap from overlap*integrate*nCh samples."""
        st = self.ppf.stokes_i16(d_volt, self.overlap * self.integrate)
        self.clip.run(st)
        used, full = self.buf.add(st.view(self.overlap, self.n_chan))
        if used != self.overlap or full:
            raise PPLError("prime: expected %d spectra to fill the overlap" % self.overlap)

    def step(self, d_volt, events=None):
        """One buffer: the dedispersed series land in self.out.  events: five
        torch events recorded after each stage (PPF+Stokes, clipper, buffer,
        dedispersion, carry), or None."""
        def mark(i):
            if events is not None:
                events[i].record()
        self.ppf.stokes_i16(d_volt, self.nout * self.integrate, d_out=self.stokes)
        mark(0)
        self.clip.run(self.stokes)
        mark(1)
        used, full = self.buf.add(self.stokes.view(self.nout, self.n_chan))
        if used != self.nout or not full:
            raise PPLError("step: the buffer did not fill with %d spectra" % self.nout)
        mark(2)
        self._dedisperse(self.buf.tensor(), self.buf.row_offsets(), self.out, self.series)
        mark(3)
        self.buf.rotate()
        mark(4)
        self.parity ^= 1

    def _dedisperse(self, d_in, row_offsets, out, series, stream=None):
        """Dedispersion of a full buffer into out; for u8 also the float
        series (before the buffer's rotate, which moves its row offsets).
        Returns the tensor the detector reads."""
        self.plan.run(d_in, d_out=out, stream=stream)
        if self.dtype != "u8":
            return out
        return self.plan.u8_series(out, row_offsets, 1.0 / self.inv_scale, d_out=series, stream=stream)

    def capture_steps(self, d_volt):
        """Capture two consecutive steps as two graphs (even and odd steps);
        d_volt is one input tensor, or a pair read by the graph of the
        current parity and the next.  The chain's device state is not
        advanced by the capture."""
        import torch
        if self.graphs is not None:
            return
        ins = list(d_volt) if isinstance(d_volt, (list, tuple)) else [d_volt, d_volt]
        gs = {}
        for k in range(2):
            g = torch.cuda.CUDAGraph()
            p = self.parity
            with torch.cuda.graph(g):
                self.step(ins[k])
            gs[p] = g
        self.graphs = gs  # host state went round twice: back to this parity

    def replay(self):
        """One step by graph, reading the input captured for this parity."""
        self.graphs[self.parity].replay()
        self.parity ^= 1

    def detect(self, threshold=6.0, bin_pow2=6, max_events=1 << 22, d_out=None, stream=None):
        """Events of the last step's series (or of d_out, a float series)."""
        d = (self.series if self.dtype == "u8" else self.out) if d_out is None else d_out
        return detect(d, self.n_ch, self.n_sub, threshold, bin_pow2, max_events, stream=stream)

    def run_pipelined(self, volts, threshold=6.0, bin_pow2=6, max_events=1 << 22):
        """Steps over the voltage chunks in `volts` on two streams, yielding
        each step's (ints, vals, count) detection in order.

        The front stream runs PPF+Stokes, the clipper, the buffer fill and
        the overlap carry; the back stream runs dedispersion and detection.
        Step k's front work overlaps step k-1's dedispersion, as the
        reference's asynchronous DedispersionModule overlaps the next blobs
        with the GPU job.  Ordering: dedispersion k waits for buffer k to be
        filled; the carry of step k (which writes the buffer copy
        dedispersion k-1 reads) waits for dedispersion k-1; dedispersion k
        writes one of two output tensors, the other holding step k-1's series
        until its detection has run.  The results equal step()'s."""
        import torch
        front, back = torch.cuda.Stream(), torch.cuda.Stream()
        outs = [self.out, torch.empty_like(self.out)]
        series = [self.series, torch.empty_like(self.series)] if self.series is not None else [None, None]
        cur = torch.cuda.current_stream()
        front.wait_stream(cur)
        back.wait_stream(cur)
        done = None
        k = 0
        pending = None
        try:
            for d_volt in volts:
                with torch.cuda.stream(front):
                    self.ppf.stokes_i16(d_volt, self.nout * self.integrate, d_out=self.stokes)
                    self.clip.run(self.stokes)
                    used, full = self.buf.add(self.stokes.view(self.nout, self.n_chan))
                    if used != self.nout or not full:
                        raise PPLError("step: the buffer did not fill with %d spectra" % self.nout)
                    d_in, r_in = self.buf.tensor(), self.buf.row_offsets()
                    filled = torch.cuda.Event()
                    filled.record(front)
                    if done is not None:
                        front.wait_event(done)
                    self.buf.rotate()
                result = None
                with torch.cuda.stream(back):
                    if pending is not None:
                        result = self.detect(threshold, bin_pow2, max_events, d_out=pending, stream=back)
                    back.wait_event(filled)
                    # the event is recorded after the u8 series too: it reads
                    # the buffer's row offsets, which the carry then moves
                    seen = self._dedisperse(d_in, r_in, outs[k & 1], series[k & 1])
                    done = torch.cuda.Event()
                    done.record(back)
                if result is not None:
                    yield result
                pending = seen
                self.parity ^= 1
                k += 1
            if pending is not None:
                with torch.cuda.stream(back):
                    result = self.detect(threshold, bin_pow2, max_events, d_out=pending, stream=back)
                yield result
        finally:
            # on every exit (the consumer stopping early, an error) the
            # current stream waits for both streams' work on self.stokes,
            # the buffer and the outputs
            cur.wait_stream(front)
            cur.wait_stream(back)
            if k and not k & 1:  # leave the last series in self.out, as step() does
                self.out.copy_(outs[1])
                if self.series is not None:
                    self.series.copy_(series[1])

    def run_overlapped(self, volts, threshold=6.0, bin_pow2=6, max_events=1 << 22):
        """Steps over the voltage chunks in `volts` with step k's clipper walk
        beside step k-1's dedispersion, yielding each step's (ints, vals,
        count) detection in order.

        The clipper runs split in three (Clipper.run_front / run_walk /
        run_back): the precompute of the whole chunk, the sequential walk
        (one workgroup per beam), then the statistics and write passes.
        Three streams: the current one runs PPF+Stokes, the clipper's front
        and back, the buffer fill and the carry; a walk stream on CU slot 0
        of every XCD runs the walk; a dedispersion stream on the other slots
        runs dedispersion and detection.  The two CU ranges do not overlap:
        a kernel whose range overlaps the running dedispersion grid's is not
        dispatched before that grid's last workgroup is (measured: the walk
        then ends after the dedispersion instead of beside it).  Ordering
        per step k: the walk and dedispersion k-1 wait
        for the clipper's front (so they do not share the GPU with PPF k and
        the precompute); the back waits for the walk; the carry that starts
        buffer k+1 overwrites the copy dedispersion k-1 read, so it waits for
        that dedispersion.  The results equal step()'s."""
        import torch
        per_xcd = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count // 8
        dd_s, walk_s = CuRangeStream(1, per_xcd - 1), CuRangeStream(0, 1)
        dd, walk = dd_s.stream, walk_s.stream
        main = torch.cuda.current_stream()
        outs = [self.out, torch.empty_like(self.out)]
        series = [self.series, torch.empty_like(self.series)] if self.series is not None else [None, None]
        dd.wait_stream(main)
        walk.wait_stream(main)
        prev = None  # (buffer, row offsets, filled event) of step k-1
        k = 0
        try:
            for d_volt in volts:
                self.ppf.stokes_i16(d_volt, self.nout * self.integrate, d_out=self.stokes)
                self.clip.run_front(self.stokes)
                fronted = torch.cuda.Event()
                fronted.record(main)
                with torch.cuda.stream(walk):
                    walk.wait_event(fronted)
                    self.clip.run_walk(stream=walk)
                    walked = torch.cuda.Event()
                    walked.record(walk)
                seen, done = None, None
                if prev is not None:
                    with torch.cuda.stream(dd):
                        dd.wait_event(fronted)
                        dd.wait_event(prev[2])
                        seen = self._dedisperse(prev[0], prev[1], outs[(k - 1) & 1], series[(k - 1) & 1],
                                                stream=dd)
                        done = torch.cuda.Event()
                        done.record(dd)
                main.wait_event(walked)
                self.clip.run_back()
                used, full = self.buf.add(self.stokes.view(self.nout, self.n_chan))
                if used != self.nout or not full:
                    raise PPLError("step: the buffer did not fill with %d spectra" % self.nout)
                d_in, r_in = self.buf.tensor(), self.buf.row_offsets()
                filled = torch.cuda.Event()
                filled.record(main)
                if done is not None:
                    main.wait_event(done)
                self.buf.rotate()
                if seen is not None:
                    with torch.cuda.stream(dd):
                        result = self.detect(threshold, bin_pow2, max_events, d_out=seen, stream=dd)
                    yield result
                prev = (d_in, r_in, filled)
                self.parity ^= 1
                k += 1
            if prev is not None:
                with torch.cuda.stream(dd):
                    dd.wait_event(prev[2])
                    seen = self._dedisperse(prev[0], prev[1], outs[(k - 1) & 1], series[(k - 1) & 1], stream=dd)
                    result = self.detect(threshold, bin_pow2, max_events, d_out=seen, stream=dd)
                yield result
        finally:
            main.wait_stream(walk)
            main.wait_stream(dd)
            torch.cuda.synchronize()
            # leave the last step's series in out, as step() does
            if k and not k & 1:
                self.out.copy_(outs[1])
                if self.series is not None:
                    self.series.copy_(series[1])
            dd_s.close()
            walk_s.close()

    def run_split(self, volts, small_cus=4, threshold=6.0, bin_pow2=6, max_events=1 << 22):
        """run_pipelined with the GPU split between the two streams by CU
        range, the same slots on every XCD: the big stream (slots small_cus
        and up) runs PPF+Stokes, dedispersion and detection; the small stream
        (slots below small_cus) runs the clipper, the buffer fill and the
        carry.  The clipper's walk is one workgroup per beam, so on its own
        CUs it runs beside step k-1's dedispersion instead of after it, and
        no stream's kernels wait for CU slots held by the other's.
        Ordering: the clipper of step k waits for PPF k; dedispersion k-1
        (issued after PPF k) waits for buffer k-1 to be filled; the carry of
        step k waits for dedispersion k-1; PPF k waits for the clipper of
        step k-2, whose Stokes slot it reuses (two slots).  The results
        equal step()'s, and the chain carries on with step() afterwards."""
        import torch
        per_xcd = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count // 8
        if not 0 < small_cus < per_xcd:
            raise PPLError("small_cus must be in 1..%d" % (per_xcd - 1))
        big_s, small_s = CuRangeStream(small_cus, per_xcd - small_cus), CuRangeStream(0, small_cus)
        big, small = big_s.stream, small_s.stream
        stokes = [self.stokes, torch.empty_like(self.stokes)]
        outs = [self.out, torch.empty_like(self.out)]
        series = [self.series, torch.empty_like(self.series)] if self.series is not None else [None, None]
        cur = torch.cuda.current_stream()
        big.wait_stream(cur)
        small.wait_stream(cur)
        filled = [None, None]  # per Stokes slot: its step's clipper and fill are done
        prev = None            # (buffer, filled event) of step k-1
        k = 0
        try:
            for d_volt in volts:
                sk = stokes[k & 1]
                with torch.cuda.stream(big):
                    if filled[k & 1] is not None:
                        big.wait_event(filled[k & 1])
                    self.ppf.stokes_i16(d_volt, self.nout * self.integrate, d_out=sk)
                    ppf_done = torch.cuda.Event()
                    ppf_done.record(big)
                with torch.cuda.stream(small):
                    small.wait_event(ppf_done)
                    self.clip.run(sk)
                    used, full = self.buf.add(sk.view(self.nout, self.n_chan))
                    if used != self.nout or not full:
                        raise PPLError("step: the buffer did not fill with %d spectra" % self.nout)
                    d_in, r_in = self.buf.tensor(), self.buf.row_offsets()
                    ev = torch.cuda.Event()
                    ev.record(small)
                    filled[k & 1] = ev
                result, dd_done = None, None
                if prev is not None:
                    with torch.cuda.stream(big):
                        big.wait_event(prev[2])
                        seen = self._dedisperse(prev[0], prev[1], outs[(k - 1) & 1], series[(k - 1) & 1])
                        dd_done = torch.cuda.Event()
                        dd_done.record(big)
                        result = self.detect(threshold, bin_pow2, max_events, d_out=seen, stream=big)
                with torch.cuda.stream(small):
                    if dd_done is not None:
                        small.wait_event(dd_done)
                    self.buf.rotate()
                if result is not None:
                    yield result
                prev = (d_in, r_in, ev)
                self.parity ^= 1
                k += 1
            if prev is not None:
                with torch.cuda.stream(big):
                    big.wait_event(prev[2])
                    seen = self._dedisperse(prev[0], prev[1], outs[(k - 1) & 1], series[(k - 1) & 1])
                    result = self.detect(threshold, bin_pow2, max_events, d_out=seen, stream=big)
                yield result
        finally:
            cur.wait_stream(big)
            cur.wait_stream(small)
            torch.cuda.synchronize()
            # leave the last step's Stokes and series where step() leaves them
            if k and not k & 1:
                self.stokes.copy_(stokes[1])
                self.out.copy_(outs[1])
                if self.series is not None:
                    self.series.copy_(series[1])
            big_s.close()
            small_s.close()


# ------------------------------------------------------------- many beams
class BeamSpec:
    """One beam's shape for MultiBeam: nSub x nCh channels, the FIR, the
    clipper's bandpass, the dedispersion buffer and DM range."""

    def __init__(self, n_sub, n_ch, n_taps, coeffs, integrate, bp_values, nsamp_buf, tdms, dm_low, dm_step,
                 fch1, foff, tsamp=None, cr=10.0, sr=4.0, history=10000, zero_dm=True):
        self.n_sub, self.n_ch, self.n_taps, self.integrate = n_sub, n_ch, n_taps, integrate
        self.coeffs = np.asarray(coeffs, np.float32)
        self.bp_values = np.asarray(bp_values, np.float64)
        self.nsamp_buf, self.tdms, self.dm_low, self.dm_step = nsamp_buf, tdms, dm_low, dm_step
        self.fch1, self.foff = fch1, foff
        self.tsamp = tsamp if tsamp is not None else 5.12e-6 * n_ch * integrate
        self.cr, self.sr, self.history, self.zero_dm = cr, sr, history, zero_dm


class MultiBeam:
    """Many beams with different shapes stepped together on one device, one
    dedispersion buffer of new data per beam per step: per beam the fused
    PPF -> Stokes-I -> integration over its input chunk, then one batched
    clipper run for every beam, the transposition into each beam's buffer,
    one batched dedispersion launch for every beam and the overlap carries.

    Each beam reads a voltage chunk of chunk_spectra[b] integrated spectra,
    fed repeatedly as a continuous stream: n spectra are channelised in
    chunks of chunk_spectra[b] read from the chunk's start, the last one
    shorter (chunk_lengths).  The chunk sizes give every beam the same number
    of chunks per step (at most max_chunk_samples input samples per
    subband and polarisation in the largest), and each round of chunks, one
    per beam, is one batched channeliser launch per (channel count, DFT
    method) plus one history launch.  Every launch goes to the current
    stream, so steps can be captured in two HIP graphs (even and odd steps:
    the FIR histories and buffers alternate)."""

    def __init__(self, specs, dtype="f32", inv_scale=127.0 / 10.0, max_chunk_samples=1 << 22, tile=None,
                 dft="fp32"):
        import torch
        if len(specs) > min(MAX_DEDISP_JOBS, 48):
            raise PPLError("at most 48 beams")
        self.specs, self.dtype, self.inv_scale = list(specs), dtype, inv_scale
        self.plans, self.ppfs, self.clips, self.bufs, self.stokes, self.outs = [], [], [], [], [], []
        self.series = []
        self.nout, self.overlap, self.chunk_spectra = [], [], []
        for sp in self.specs:
            n_chan = sp.n_sub * sp.n_ch
            plan = DedispPlan(n_chan, sp.nsamp_buf, sp.tdms, sp.dm_low, sp.dm_step, sp.tsamp, sp.fch1, sp.foff,
                              dtype=dtype)
            self.plans.append(plan)
            self.nout.append(plan.nout)
            self.overlap.append(plan.maxshift + plan.rem)
            # the split-precision DFT where it is built (64 and 256 channels)
            prec = dft if (dft != "fp32" and sp.n_ch in (64, 256)) else "fp32"
            self.ppfs.append(PPF(sp.n_sub, sp.n_ch, sp.n_taps, sp.coeffs, precision=prec, integrate=sp.integrate))
            self.clips.append(Clipper(sp.bp_values, sp.cr, sp.sr, sp.history, sp.zero_dm))
            self.bufs.append(DDBuffer(n_chan, sp.nsamp_buf, plan.maxshift + plan.rem, dtype, inv_scale))
            self.stokes.append(torch.empty((plan.nout, sp.n_sub, 1, sp.n_ch), dtype=torch.float32, device="cuda"))
            self.outs.append(torch.empty((sp.tdms, plan.nout), dtype=plan.out_dtype(), device="cuda"))
            # u8: the float series the detector reads (see BeamChain)
            self.series.append(torch.empty((sp.tdms, plan.nout), dtype=torch.float32, device="cuda")
                               if dtype == "u8" else None)
        # chunks per step: the largest beam's input in chunks of at most
        # max_chunk_samples samples; every beam's step in that many chunks
        per = [sp.integrate * sp.n_ch for sp in self.specs]  # input samples per integrated spectrum
        rounds = max(-(-n * q // max_chunk_samples) for n, q in zip(self.nout, per))
        self.chunk_spectra = [max(1, -(-n // rounds)) for n in self.nout]
        self.tile = _tile4(tile) if tile else (0, 0, 0, 4)
        self.graphs = None
        self.parity = 0

    def chunk_samples(self, b):
        """Input samples per subband and polarisation in beam b's chunk."""
        sp = self.specs[b]
        return self.chunk_spectra[b] * sp.integrate * sp.n_ch

    def data_seconds(self):
        """Seconds of telescope data per step, per beam."""
        return [n * sp.tsamp for n, sp in zip(self.nout, self.specs)]

    def trials_samples(self):
        return sum(sp.tdms * n for sp, n in zip(self.specs, self.nout))

    def autotune(self, reps=1, candidates=None, per_beam=False):
        """Picks the tile on the beams' own buffers (their contents do not
        change the time); candidates: (tileDM, tileT, chanBlock[, readBytes]) tuples to try
        instead of every instantiated shape.  One tile for the whole batch,
        or with per_beam each beam's own fastest tile (one launch per group
        of beams sharing a tile) when that is faster on the whole batch.
        Returns the tuned dedispersion time per step in ms."""
        import torch
        ins = [bf.tensor() for bf in self.bufs]
        tile, ms = autotune_batch(self.plans, ins, self.outs, reps=reps, candidates=candidates)
        self.tile = tile
        if not per_beam:
            return ms
        # per-beam tiles (each tuned alone), kept only if the grouped launches
        # beat the single batch tile on the whole batch
        old = [p.tile for p in self.plans]
        for p, x, o in zip(self.plans, ins, self.outs):
            t, _ = autotune_batch([p], [x], [o], reps=reps, candidates=candidates)
            p.set_tile(*t)
        dedisperse_batch(self.plans, ins, self.outs, tile="own")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            dedisperse_batch(self.plans, ins, self.outs, tile="own")
        e1.record()
        e1.synchronize()
        own_ms = e0.elapsed_time(e1) / reps
        if own_ms < ms:
            self.tile = "own"
            return own_ms
        for p, t in zip(self.plans, old):
            p.set_tile(*t)
        return ms

    def chunk_lengths(self, b, n_spec):
        """The chunks (integrated spectra each) in which n_spec spectra of
        beam b are channelised, each read from the start of its voltage chunk."""
        m = self.chunk_spectra[b]
        return [min(m, n_spec - d) for d in range(0, n_spec, m)]

    def _channelise(self, volts, n_specs, d_outs):
        """n_specs[b] integrated spectra of every beam b into d_outs[b]: round
        r channelises chunk r of every beam that has one, in one batched call."""
        sched = [self.chunk_lengths(b, n) for b, n in enumerate(n_specs)]
        done = [0] * len(sched)
        for r in range(max(len(x) for x in sched)):
            jobs = []
            for b, lens in enumerate(sched):
                if r < len(lens):
                    k, sp = lens[r], self.specs[b]
                    jobs.append((self.ppfs[b], volts[b], k * sp.integrate, d_outs[b][done[b]:done[b] + k]))
                    done[b] += k
            ppf_stokes_i16_batch(jobs)

    def prime(self, volts):
        """Fill every beam's overlap (untimed set-up)."""
        import torch
        sts = [torch.empty((ov, sp.n_sub, 1, sp.n_ch), dtype=torch.float32, device="cuda")
               for ov, sp in zip(self.overlap, self.specs)]
        self._channelise(volts, self.overlap, sts)
        for b, sp in enumerate(self.specs):
            ov, st = self.overlap[b], sts[b]
            self.clips[b].run(st)
            used, full = self.bufs[b].add(st.view(ov, sp.n_sub * sp.n_ch))
            if used != ov or full:
                raise PPLError("prime: beam %d did not take its overlap" % b)

    def step(self, volts, events=None):
        """One buffer for every beam; events: five torch events recorded after
        the channeliser, clipper, buffer, dedispersion and carry stages."""
        def mark(i):
            if events is not None:
                events[i].record()
        self._channelise(volts, self.nout, self.stokes)
        mark(0)
        clipper_run_batch(self.clips, self.stokes)
        mark(1)
        for b, sp in enumerate(self.specs):
            used, full = self.bufs[b].add(self.stokes[b].view(self.nout[b], sp.n_sub * sp.n_ch))
            if used != self.nout[b] or not full:
                raise PPLError("step: beam %d's buffer did not fill" % b)
        mark(2)
        dedisperse_batch(self.plans, [bf.tensor() for bf in self.bufs], self.outs, tile=self.tile)
        if self.dtype == "u8":
            for b, bf in enumerate(self.bufs):
                self.plans[b].u8_series(self.outs[b], bf.row_offsets(), 1.0 / self.inv_scale, d_out=self.series[b])
        mark(3)
        for bf in self.bufs:
            bf.rotate()
        mark(4)
        self.parity ^= 1

    def capture_steps(self, volts):
        import torch
        if self.graphs is not None:
            return
        gs = {}
        for _ in range(2):
            g = torch.cuda.CUDAGraph()
            p = self.parity
            with torch.cuda.graph(g):
                self.step(volts)
            gs[p] = g
        self.graphs = gs

    def replay(self):
        self.graphs[self.parity].replay()
        self.parity ^= 1

    def detect(self, b, threshold=6.0, bin_pow2=6, max_events=1 << 22):
        sp = self.specs[b]
        d = self.series[b] if self.dtype == "u8" else self.outs[b]
        return detect(d, sp.n_ch, sp.n_sub, threshold, bin_pow2, max_events)


# ------------------------------------------------------------ streaming
class StreamService:
    """A streaming loop over many beams of one shape, each arriving and
    leaving on its own: beams are admitted while their device footprint fits
    the HBM budget, and their FIR history, clipper state and dedispersion
    overlap stay resident across chunks.  Chunks from all beams that are
    ready are processed in rounds: per beam the fused channeliser over its
    chunk, one batched clipper run for every beam in the round, the
    transposition into each beam's buffer, one batched dedispersion launch
    for every beam whose buffer filled, its detection, then the overlap carry
    and the rest of the chunk into the next buffer (DedispersionBuffer's
    split, src/lib/src/DedispersionBuffer.cpp:98-179).  All beams share one
    dedispersion plan (one delay table)."""

    class Beam:
        def __init__(self, svc, key):
            import torch
            sp = svc.spec
            self.key = key
            self.ppf = PPF(sp.n_sub, sp.n_ch, sp.n_taps, sp.coeffs, integrate=sp.integrate)
            self.clip = Clipper(sp.bp_values, sp.cr, sp.sr, sp.history, sp.zero_dm)
            self.buf = DDBuffer(svc.n_chan, sp.nsamp_buf, svc.overlap, svc.dtype, svc.inv_scale)
            self.stokes = torch.empty((svc.chunk_spectra, sp.n_sub, 1, sp.n_ch), dtype=torch.float32,
                                      device="cuda")
            self.out = torch.empty((sp.tdms, svc.plan.nout), dtype=svc.plan.out_dtype(), device="cuda")
            self.series = (torch.empty((sp.tdms, svc.plan.nout), dtype=torch.float32, device="cuda")
                           if svc.dtype == "u8" else None)
            self.buffers = 0  # dedispersion buffers completed

    def __init__(self, spec, chunk_spectra, budget_bytes, dtype="f32", inv_scale=127.0 / 10.0, tile=None,
                 max_beams=48):
        self.spec, self.dtype, self.inv_scale = spec, dtype, inv_scale
        self.n_chan = spec.n_sub * spec.n_ch
        self.plan = DedispPlan(self.n_chan, spec.nsamp_buf, spec.tdms, spec.dm_low, spec.dm_step, spec.tsamp,
                               spec.fch1, spec.foff, dtype=dtype)
        if tile is not None:
            self.plan.set_tile(*tile)
        self.overlap = self.plan.maxshift + self.plan.rem
        self.chunk_spectra = int(chunk_spectra)
        self.budget = int(budget_bytes)
        self.max_beams = min(int(max_beams), MAX_DEDISP_JOBS)
        self.beams = {}
        self.used = 0

    def footprint(self):
        """Device bytes one admitted beam holds: the two alternating
        dedispersion buffers, the chunk's spectra, the dedispersed output (for
        u8 the int32 sums and the float series), the channeliser history and
        the clipper's state and work areas."""
        sp = self.spec
        elem = {"f32": 4, "f16": 2, "u8": 1}[self.dtype]
        buf = 2 * self.n_chan * sp.nsamp_buf * elem
        stokes = self.chunk_spectra * self.n_chan * 4
        out = sp.tdms * self.plan.nout * 4 * (2 if self.dtype == "u8" else 1)  # u8: sums and series
        ppf = 2 * sp.n_sub * 2 * (sp.n_taps - 1) * sp.n_ch * 8
        clip = 64 * self.n_chan * 4 + self.chunk_spectra * 256 + (16 << 20)
        return buf + stokes + out + ppf + clip

    def admit(self, key):
        """Admits beam `key` if its footprint fits the budget; returns whether
        it did."""
        if key in self.beams:
            raise PPLError("beam %r already admitted" % (key,))
        fp = self.footprint()
        if len(self.beams) >= self.max_beams or self.used + fp > self.budget:
            return False
        self.beams[key] = StreamService.Beam(self, key)
        self.used += fp
        return True

    def release(self, key):
        b = self.beams.pop(key)
        self.used -= self.footprint()
        del b

    def process(self, ready, threshold=6.0, bin_pow2=6):
        """One round over ready = [(key, int16 voltages [sub][pol][chunk*integrate*nCh][2])],
        at most one chunk per beam.  Returns [(key, buffer index, event ints,
        event values)] for the buffers completed in this round."""
        sp = self.spec
        m = self.chunk_spectra
        beams = [self.beams[k] for k, _ in ready]
        if len({id(b) for b in beams}) != len(beams):
            raise PPLError("one chunk per beam per round")
        for b, (_, v) in zip(beams, ready):
            b.ppf.stokes_i16(v, m * sp.integrate, d_out=b.stokes)
        if beams:
            clipper_run_batch([b.clip for b in beams], [b.stokes for b in beams])
        used = []
        full = []
        for b in beams:
            u, f = b.buf.add(b.stokes.view(m, self.n_chan))
            used.append(u)
            if f:
                full.append(b)
        results = []
        if full:
            dedisperse_batch([self.plan] * len(full), [b.buf.tensor() for b in full], [b.out for b in full])
            for b in full:
                if self.dtype == "u8":
                    self.plan.u8_series(b.out, b.buf.row_offsets(), 1.0 / self.inv_scale, d_out=b.series)
                    ints, vals, _ = detect(b.series, sp.n_ch, sp.n_sub, threshold, bin_pow2)
                else:
                    ints, vals, _ = detect(b.out, sp.n_ch, sp.n_sub, threshold, bin_pow2)
                results.append((b.key, b.buffers, ints, vals))
                b.buffers += 1
                b.buf.rotate()
        for b, u in zip(beams, used):
            if u < m:
                u2, f2 = b.buf.add(b.stokes.view(m, self.n_chan), first=u)
                if u + u2 != m or f2:
                    raise PPLError("chunk larger than one buffer of new samples")
        return results
