/*
 * ppl_amd.h -- C ABI of the MI355X-native pulsar-search chain.
 *
 * Plain C: device pointers, sizes and a hipStream_t passed as void*.  No
 * exception crosses this boundary; every entry point returns PPL_OK (0) or
 * a negative PPL_E* code and leaves a message in ppl_last_error().
 *
 * Each entry point names the pelican-lofar interface it replaces
 * (file:line relative to the pelican-lofar source tree).
 */
#ifndef PPL_AMD_H
#define PPL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PPL_OK = 0,
    PPL_E_ARG = -1,      /* invalid argument or shape */
    PPL_E_HIP = -2,      /* HIP runtime error */
    PPL_E_NOMEM = -3,    /* allocation failed */
    PPL_E_CONFIG = -4,   /* configuration rejected (e.g. maxshift >= buffer) */
    PPL_E_NODEV = -5     /* no GPU / not initialised */
};

/* Element types of dedispersion input buffers. */
enum { PPL_F32 = 0, PPL_F16 = 1, PPL_U8 = 2 };

/* Channeliser DFT precision. */
enum {
    PPL_DFT_FP32 = 0,    /* FP32 FIR + FP32 FFT (the plain method) */
    PPL_DFT_BF16X3 = 1,  /* DFT as three BF16 MFMA GEMMs (hi*hi + hi*lo + lo*hi), FP32 accumulate */
    PPL_DFT_FP16X2 = 2   /* DFT as two FP16 MFMA GEMMs (hi*hi + lo*hi, data split), FP32 accumulate */
};

/* --------------------------------------------------------------- runtime */

/* Replaces GPU_NVidia::initialiseResources (src/lib/src/GPU_NVidia.cpp:71-77):
 * selects the device and prepares per-device state.  Idempotent. */
int ppl_init(int device);
int ppl_shutdown(void);
const char* ppl_last_error(void);
int ppl_version(void);

/* Device / pinned memory and streams (replaces GPU_Param cudaMalloc/Memcpy,
 * src/lib/src/GPU_Param.cpp:18,41,49).  Streams are created at the default
 * priority. */
int ppl_device_alloc(void** ptr, size_t bytes);
int ppl_device_free(void* ptr);
int ppl_host_alloc(void** ptr, size_t bytes);   /* pinned */
int ppl_host_free(void* ptr);
int ppl_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int ppl_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int ppl_memset_async(void* dst, int value, size_t bytes, void* stream);
int ppl_stream_create(void** stream);
int ppl_stream_destroy(void* stream);
int ppl_stream_sync(void* stream);
/* A stream restricted to CU slots [first, first + count) of every XCD
 * (hipExtStreamCreateWithCUMask with the same range on each XCD); two
 * disjoint ranges split the GPU between two streams.  No counterpart in the
 * reference, where a GPU job holds a whole GPU resource while it runs
 * (src/lib/src/GPU_Manager.cpp:48-86). */
int ppl_stream_create_cu_range(int first, int count, void** stream);
/* Enqueue a host callback after the work queued so far on the stream
 * (the GPU_Job callback of src/lib/src/GPU_Manager.cpp:83-85). */
int ppl_stream_callback(void* stream, void (*fn)(void* user, int status), void* user);

/* ------------------------------------------------------------ dedispersion */

typedef struct ppl_dedisp_plan ppl_dedisp_plan;

typedef struct {
    unsigned nChan;      /* channels in buffer order (nSubbands*nChannels) */
    unsigned nsampBuf;   /* samples per buffer (2^timeBinsPerBufferPow2) */
    unsigned tdms;       /* number of DM trials (dedispersionSamples) */
    float dmLow;         /* dedispersionMinimum (clamped at 0) */
    float dmStep;        /* dedispersionStepSize */
    double tsamp;        /* block rate of the incoming Stokes blob, s */
    double fch1;         /* frequencyChannel1, MHz (highest channel) */
    double foff;         /* channelBandwidth, MHz (< 0: descending order) */
    int dtype;           /* PPL_F32, PPL_F16 or PPL_U8 input buffer */
} ppl_dedisp_config;

typedef struct {
    int maxshift;           /* DedispersionModule::maxshift() */
    int remainingSamples;   /* (nsampBuf - maxshift) % 224 */
    int nOut;               /* nsampBuf - maxshift - remainingSamples */
    float mstartdm;         /* dmLow / tsamp (float) */
    float mdmstep;          /* dmStep / tsamp (float) */
    int tileDM, tileT;      /* DM x time tile chosen for this plan */
    int chanBlock;          /* channels staged per block (0: as many as fit in LDS) */
    int readBytes;          /* LDS read width per lane of the chosen kernel: 4 or 16 */
    long long fusedDelayDiffs; /* delay entries that would differ were the trial value
                                  mstartdm + dm*mdmstep one fused multiply-add, as a
                                  contracting build of DedispersionKernel.cu:29 may form
                                  it; always 0 when dmLow == 0 (the sum is then exact) */
} ppl_dedisp_info;

/* DedispersionModule::resize (src/lib/src/DedispersionModule.cpp:109-190):
 * builds the FP32 delay table on the host from the FP64 shift table exactly
 * as the reference (:152-162, DedispersionKernel.cu:29,36), uploads it and
 * picks the tile shape: for FP32 input the ring kernel at 32 trials x 512
 * samples where a tile's delay spread fits its staged row, else (and for
 * f16 and u8) the one-word kernel at 32 (or fewer) trials x 512.  Fails with PPL_E_CONFIG when nsampBuf <= maxshift
 * (the module's throw at :175-178), when dmStep < 0 (delays would decrease
 * with the trial index) or when a delayed read would leave the channel row. */
int ppl_dedisp_plan_create(ppl_dedisp_plan** plan, const ppl_dedisp_config* cfg);
int ppl_dedisp_plan_destroy(ppl_dedisp_plan* plan);
int ppl_dedisp_plan_info(const ppl_dedisp_plan* plan, ppl_dedisp_info* info);
/* Override the DM x time tile (the autotuner's hook); 0 keeps the default. */
int ppl_dedisp_set_tile(ppl_dedisp_plan* plan, int tileDM, int tileT);
/* Same with the number of channels staged per LDS block (0: as many as fit). */
int ppl_dedisp_set_shape(ppl_dedisp_plan* plan, int tileDM, int tileT, int chanBlock);
/* Same with the kernel family: readBytes 4 for the one-word kernels, or 16
 * for the ring kernels (FP32 input: four consecutive samples per 16-byte
 * read, tileDM 16, 32 or 64; u8 input: four samples per dword, tileDM 16 or
 * 32, byte count of the input a multiple of 4; tileT 256 or 512).  All give
 * the same sums bit for bit. */
int ppl_dedisp_set_kernel(ppl_dedisp_plan* plan, int tileDM, int tileT, int chanBlock, int readBytes);
/* Autotuner: times every instantiated DM x time tile, channel block and read width on
 * the given buffers (reps launches each, HIP events on the stream), keeps the
 * fastest for the plan and returns its time per launch in *best_ms. */
int ppl_dedisp_autotune(ppl_dedisp_plan* plan, const void* d_in, void* d_out, int reps, void* stream,
                        float* best_ms);
/* Copy the int32 delay table [tdms][nChan] to host memory (for checks). */
int ppl_dedisp_delays(const ppl_dedisp_plan* plan, int32_t* host_out);

/* cacheDedisperseLoop (src/lib/DedispersionKernel.cu:53-86; declared at
 * src/lib/src/DedispersionModule.cpp:21-24): out[dm][t] = sum over channels,
 * in channel order with FP32 adds, of in[c][t + delay[dm][c]], t < nOut.
 * in: [nChan][nsampBuf] of the plan's dtype; out: [tdms][nOut] float for
 * F32/F16 input, int32 for U8 input (exact integer sums). */
int ppl_dedisperse(const ppl_dedisp_plan* plan, const void* d_in, void* d_out, void* stream);

/* Many beams' dedispersion in one launch (no reference counterpart: the
 * reference dedisperses one beam per GPU job, DedispersionModule.cpp:263-297).
 * Each job is a plan with its own input buffer and output; plans may differ
 * in channels, buffer length, DM count and range, but share the dtype.  All
 * jobs run under one DM x time tile, channel block and read width (tileDM =
 * tileT = 0: the first plan's; when that is a ring tile another job's delay
 * spread does not fit, the same tile with one-word reads), each with its own
 * LDS row length.  Results equal n separate ppl_dedisperse calls, bit for
 * bit. */
typedef struct {
    const ppl_dedisp_plan* plan;
    const void* d_in;
    void* d_out;
} ppl_dedisp_job;
enum { PPL_DEDISP_MAX_JOBS = 48 };
int ppl_dedisperse_batch(const ppl_dedisp_job* jobs, int n, int tileDM, int tileT, int chanBlock,
                         int readBytes, void* stream);
/* Autotuner for a batch: times every instantiated tile, channel block and
 * read width that fits every job (reps launches each) and returns the fastest. */
int ppl_dedisp_autotune_batch(const ppl_dedisp_job* jobs, int n, int reps, void* stream, int* tileDM,
                              int* tileT, int* chanBlock, int* readBytes, float* best_ms);

/* The reference's C seam itself, with the reference's arguments (device
 * pointers; maxshift includes the remainder; outbufSize in bytes), computed
 * for every tdms (the reference's launch covers all rows only when
 * tdms % 40 == 0).  Builds its delay table from dmShift on the device, and
 * from it the FP32 ring kernel's table where a 32-trial tile's delay spread
 * fits the ring (the one-word kernel otherwise).
 * PPL_E_CONFIG for a negative dmShift, mstartdm or mdmstep (delays that
 * decrease with the trial index) or a delayed read outside the buffer. */
int ppl_cache_dedisperse_loop(float* outbuff, long outbufSize, float* buff, float mstartdm,
                              float mdmstep, int tdms, int numSamples, const float* dmShift,
                              int maxshift, int i_nchans, void* stream);

/* FP32 -> FP16 and FP32 -> offset-binary u8 conversion of a buffer
 * (q = clamp(rint(x*inv_scale), -128, 127) + 128). */
int ppl_convert_f16(const float* d_in, uint16_t* d_out, size_t n, void* stream);
int ppl_quantise_u8(const float* d_in, uint8_t* d_out, size_t n, float inv_scale, void* stream);

/* ---------------------------------------------------- dedispersion buffer */

typedef struct ppl_ddbuf ppl_ddbuf;
/* DedispersionBuffer (src/lib/src/DedispersionBuffer.cpp): a device-resident
 * [nChan][nsampBuf] buffer of dtype, filled from Stokes-I spectra
 * [t][nChan] (float) by transposition, whose last `overlap` samples
 * (maxshift + remainingSamples) start the next buffer. */
int ppl_ddbuf_create(ppl_ddbuf** buf, unsigned nChan, unsigned nsampBuf, unsigned overlap,
                     int dtype, float inv_scale);
int ppl_ddbuf_destroy(ppl_ddbuf* buf);
/* addSamples (:98-179): appends spectra[first .. first+n) and returns in
 * *consumed how many were taken before the buffer filled; *full is set to 1
 * when the buffer holds nsampBuf samples. */
int ppl_ddbuf_add(ppl_ddbuf* buf, const float* d_spectra, unsigned nT, unsigned first,
                  unsigned* consumed, int* full, void* stream);
/* addSamples on a whole SpectrumDataSet blob [nT][nSub][nPol][nCh]: takes
 * polarisation 0 ("dedispersion on total power only", :116) of every
 * spectrum; nSub*nCh must equal the buffer's nChan. */
int ppl_ddbuf_add_blob(ppl_ddbuf* buf, const float* d_spectra, unsigned nT, unsigned first,
                       unsigned nSub, unsigned nPol, unsigned nCh, unsigned* consumed, int* full,
                       void* stream);
/* copy (:65-96): carries the overlap into the start and empties the rest. */
int ppl_ddbuf_rotate(ppl_ddbuf* buf, void* stream);
void* ppl_ddbuf_data(ppl_ddbuf* buf);
/* u8 buffers only (NULL otherwise; no reference counterpart: the reference
 * has no quantised mode): the device array [nsampBuf] of the offsets the
 * current buffer's rows were quantised relative to.  A row (spectrum) more
 * than half of whose samples fall outside the 8-bit range -- here the RFI
 * clipper's rejected spectra, left at -0.2 - bandpass - dataModel, far below
 * the normalised data -- is stored as rint((x - off) * inv_scale) with off =
 * its middle channel's value (channel nChan/2); every other row has off = 0
 * and is stored as before.  The offsets are carried with the overlap. */
const float* ppl_ddbuf_row_offsets(const ppl_ddbuf* buf);
unsigned ppl_ddbuf_fill(const ppl_ddbuf* buf);

/* -------------------------------------------------------------- detection */

/* DedispersionAnalyser::analyse (src/lib/src/DedispersionAnalyser.cpp:30-203):
 * dyadic boxcar threshold detection over data [tdms][nsamp].  Writes the
 * events (dm, t, binFactor) as int32 triples plus their values, in the
 * reference's order (dummy (0,0,1,0) first; then by DM, 2^binPow2 segment,
 * level, position).  *count receives the total number of events; when it
 * exceeds max_events only the first max_events are written.  Event buffers
 * are host pointers; the call synchronises the stream. */
int ppl_detect(const float* d_data, int tdms, int nsamp, int nChannels, int nSubbands,
               float threshold, unsigned binPow2, int32_t* h_ev_ints, float* h_ev_vals,
               int max_events, int* count, void* stream);
/* No counterpart in the reference: ppl_detect for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same events in the same order, written to device
 * memory: d_ev_ints [max_events][3], d_ev_vals [max_events], and the total
 * number of events to the device int *d_count.  The per-call scratch (a hit
 * count per segment, the workgroup offsets) is a static device array of the
 * library, about 8.5 MB per device, held until the process ends; nothing is
 * allocated and nothing synchronises (the results are ready when the stream
 * reaches them).  tdms * (nsamp >> binPow2) may not exceed
 * PPL_DETECT_NOALLOC_MAX_SEGMENTS (PPL_E_ARG).  Every call on a device shares
 * that device's scratch: calls that overlap on one device (other streams,
 * other threads) race. */
enum { PPL_DETECT_NOALLOC_MAX_SEGMENTS = 1 << 23 };
int ppl_detect_noalloc(const float* d_data, int tdms, int nsamp, int nChannels, int nSubbands,
                       float threshold, unsigned binPow2, int32_t* d_ev_ints, float* d_ev_vals,
                       int max_events, int* d_count, void* stream);
/* Same over int32 u8-mode sums, dequantised as (S - 128*nChan) * scale. */
int ppl_detect_u8(const int32_t* d_data, int tdms, int nsamp, int nChan, float scale,
                  int nChannels, int nSubbands, float threshold, unsigned binPow2,
                  int32_t* h_ev_ints, float* h_ev_vals, int max_events, int* count,
                  void* stream);

/* u8 series (no reference counterpart): the dedispersed values of a u8
 * buffer, out[dm][t] = float(S[dm][t] - 128*nChan) * scale, plus for every
 * row e of the buffer with a nonzero offset (ppl_ddbuf_row_offsets), taken
 * in increasing e, off[e] times the number of channels whose delay at
 * trial dm is e - t (product rounded, then added).  So a rejected spectrum
 * enters the series at its own level, as in the FP32 chain, instead of at
 * the bottom of the 8-bit range.  d_sums [tdms][nOut] int32 from
 * ppl_dedisperse; d_rowOffsets the offsets of the buffer those sums were
 * taken from (before its rotate); d_out [tdms][nOut] float. */
int ppl_dedisp_u8_series(const ppl_dedisp_plan* plan, const int32_t* d_sums, const float* d_rowOffsets,
                         float scale, float* d_out, void* stream);
/* u8-mode sums as dedispersed values: out = float(S - 128*nChan) * scale,
 * the values ppl_detect_u8 compares (no reference counterpart: the
 * reference has no quantised mode). */
int ppl_dequantise_sums(const int32_t* d_in, float* d_out, size_t n, unsigned nChan, float scale,
                        void* stream);

/* ---------------------------------------------------------------- ingest */

/* AdapterTimeSeriesDataSet::_readData (src/lib/src/AdapterTimeSeriesDataSet.cpp:
 * 217-293), 8- and 16-bit branches: nPackets packets, each packetBytes long
 * with its payload after headerBytes, payload samples [sub][t][pol] complex
 * integers (re, im) of sampleBits bits, nSamplesPerPacket samples per packet.
 * ppl_ingest_c32 writes the TimeSeriesDataSetC32 blob [sub][pol][block][t]
 * (complex float, nPackets*nSamplesPerPacket samples per (sub, pol)); the
 * int16 variant writes the fused channeliser's [sub][pol][t] complex int16
 * input (exact for both sample widths). */
int ppl_ingest_c32(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nPackets,
                   unsigned nSub, unsigned nSamplesPerPacket, unsigned nPol, int sampleBits,
                   float* d_blob, void* stream);
int ppl_ingest_i16(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nPackets,
                   unsigned nSub, unsigned nSamplesPerPacket, unsigned nPol, int sampleBits,
                   int16_t* d_volt, void* stream);

/* LofarDataSplittingChunker's packet sequencing (src/lib/src/
 * LofarDataSplittingChunker.cpp:186-325), host code: packets received in
 * arrival order, each starting with the 16-byte LOFAR UDP header
 * (src/lib/LofarUdpHeader.h:25-36), are placed in the slots of a chunk.
 * h_slotSource[s] receives the packet index for slot s, or -1 for a lost
 * packet, which the chunker writes as an empty (zero) packet.  As the
 * chunker does: a packet whose timestamp is 0xFFFFFFFF or more than 10 s
 * ahead is rejected, as is one less than nSamplesPerPacket blocks on
 * (duplicates; the very first packet of a stream is always one); a gap of
 * whole packets becomes empty slots, with blocks per second 156250 (160 MHz
 * clock) or alternately 195313 and 195312 (200 MHz); a packet after gap
 * slots that filled the chunk is dropped.  Stops when nSlots slots are
 * filled or the packets run out (*slotsFilled, *packetsUsed); the state
 * carries the stream position to the next call.  Zero-initialise it for a
 * new stream. */
typedef struct ppl_packet_seq {
    unsigned startTime;    /* timestamp of the last slot filled (0 before any) */
    unsigned startBlockid; /* its block sequence number */
    unsigned long long accepted, rejected, lost;
} ppl_packet_seq;
int ppl_packet_sequence(ppl_packet_seq* state, const void* h_packets, size_t packetBytes, unsigned nPackets,
                        unsigned nSamplesPerPacket, int clockMHz, int* h_slotSource, unsigned nSlots,
                        unsigned* packetsUsed, unsigned* slotsFilled);
/* ppl_ingest_c32 / _i16 over the slots of a chunk: slot s takes packet
 * d_slotSource[s] (device int array) of the nPackets given, or zeros where
 * it is negative (or not below nPackets). */
int ppl_ingest_c32_slots(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nPackets,
                         const int* d_slotSource, unsigned nSlots, unsigned nSub, unsigned nSamplesPerPacket,
                         unsigned nPol, int sampleBits, float* d_blob, void* stream);
int ppl_ingest_i16_slots(const void* d_packets, size_t packetBytes, unsigned headerBytes, unsigned nPackets,
                         const int* d_slotSource, unsigned nSlots, unsigned nSub, unsigned nSamplesPerPacket,
                         unsigned nPol, int sampleBits, int16_t* d_volt, void* stream);

/* --------------------------------------------------------------- bandpass */

/* BandPassRecorder::run, the integrated spectrum (src/lib/src/BandPassRecorder.cpp:
 * 85-97): for each bin b = s*nChannels + c of a Stokes blob [t][s][pol][c],
 * d_sum[b] += I[t][s][0][c] over t = 0..nSamples-1 in order, in float (the
 * reference's sequential sums, bit for bit, across any split into blobs).
 * d_sum holds nSubbands*nChannels floats; clear it before the first blob.
 * An empty blob (nSamples 0, d_stokes may be null) adds nothing. */
int ppl_bandpass_accumulate(const float* d_stokes, int nSamples, int nSubbands, int nPolarisations,
                            int nChannels, float* d_sum, void* stream);

/* ---------------------------------------------------------------- sigproc */

/* SigprocStokesWriter::sendStream, the data written per Stokes blob
 * (src/lib/src/SigprocStokesWriter.cpp:184-240, _float2int :285-290): d_out
 * receives nSamples*nPolsToWrite*nSubbands*nChannels values of nBits (32:
 * float, 8: byte) in file order -- t, p < nPolsToWrite, subbands last to
 * first, channels last to first.  8 bits: x = min(v, cropMax), byte =
 * x < cropMin ? 0 : low byte of (int)rint((x - cropMin) * 255 / (cropMax -
 * cropMin)), float arithmetic.  Other nBits are rejected, as the reference
 * throws for them. */
int ppl_sigproc_pack(const float* d_stokes, unsigned nSamples, unsigned nSubbands, unsigned nPolarisations,
                     unsigned nChannels, unsigned nPolsToWrite, int nBits, float cropMin, float cropMax,
                     void* d_out, void* stream);

/* ------------------------------------------------------------ channeliser */

typedef struct ppl_ppf_plan ppl_ppf_plan;

/* Window codes of ppl_ppf_coefficients. */
enum { PPL_WIN_KAISER = 0, PPL_WIN_GAUSSIAN = 1, PPL_WIN_BLACKMAN = 2, PPL_WIN_HAMMING = 3,
       PPL_WIN_HANN = 4 };

/* PolyphaseCoefficients::genereateFilter (src/lib/src/PolyphaseCoefficients.cpp:
 * 308-364), host code in FP64: nTaps*nCh coefficients, index t*nCh+c, with
 * the 1/nCh scale and the (-1)^c sign.  PPL_WIN_HANN is this library's
 * addition (0.5 - 0.5 cos(2 pi i/(n-1))); the reference has no Hann window. */
int ppl_ppf_coefficients(unsigned nTaps, unsigned nCh, int window, double* out);

typedef struct {
    unsigned nSub, nPol, nCh, nTaps;   /* nPol must be 2 for Stokes */
    const float* coeffs;               /* host, nTaps*nCh, index t*nCh+c */
    int precision;                     /* PPL_DFT_* */
    unsigned integrate;                /* integrateTimeBins (>= 1) */
} ppl_ppf_config;

/* PPFChanneliser (src/lib/src/PPFChanneliser.cpp:29-56,81-187): keeps the
 * FIR history of every (subband, pol) resident on the device. */
int ppl_ppf_plan_create(ppl_ppf_plan** plan, const ppl_ppf_config* cfg);
int ppl_ppf_plan_destroy(ppl_ppf_plan* plan);
int ppl_ppf_reset(ppl_ppf_plan* plan, void* stream);
/* PPFChanneliser::run: complex int16 (re,im) voltages [sub][pol][nBlk*nCh]
 * -> complex float spectra [nBlk][sub][pol][nCh]. */
int ppl_ppf_run_i16(ppl_ppf_plan* plan, const int16_t* d_in, unsigned nBlk, float* d_spectra,
                    void* stream);
/* Same from complex float input (the TimeSeriesDataSetC32 blob layout). */
int ppl_ppf_run_c32(ppl_ppf_plan* plan, const float* d_in, unsigned nBlk, float* d_spectra,
                    void* stream);
/* Fused PPF -> Stokes-I (StokesGenerator.cpp:88-99) -> time integration
 * (StokesIntegrator.cpp:97-114): int16 voltages [sub][pol][nBlk*nCh] ->
 * float [nBlk/integrate][sub][1][nCh].  nBlk must be a multiple of
 * integrate. */
int ppl_ppf_stokes_i16(ppl_ppf_plan* plan, const int16_t* d_in, unsigned nBlk, float* d_stokes,
                       void* stream);

/* One plan's chunk in ppl_ppf_stokes_i16_batch. */
enum { PPL_PPF_MAX_JOBS = 48 };
typedef struct {
    ppl_ppf_plan* plan;
    const int16_t* d_in;   /* int16 voltages [sub][pol][strideBlk*nCh]; the first nBlk blocks of
                              every stream are read */
    unsigned nBlk;         /* blocks to channelise, a multiple of the plan's integrate; 0 skips the job */
    unsigned strideBlk;    /* blocks per (sub, pol) stream of d_in, >= nBlk (0: nBlk) */
    float* d_stokes;       /* [nBlk/integrate][sub][1][nCh] */
} ppl_ppf_job;
/* ppl_ppf_stokes_i16 of up to PPL_PPF_MAX_JOBS plans (beams of any channel
 * count, tap count and DFT method) in one launch per (channel count, taps,
 * method) group, and one launch that saves every plan's FIR history.  Each
 * job's output equals ppl_ppf_stokes_i16 on the same plan and chunk, bit for
 * bit.  A plan appears at most once per call (its chunks run in order).  No
 * counterpart in the reference, which channelises one beam per module
 * (src/lib/src/PPFChanneliser.cpp:81-187); for callers that step many beams
 * together. */
int ppl_ppf_stokes_i16_batch(const ppl_ppf_job* jobs, int n, void* stream);

/* StokesGenerator::run (src/lib/src/StokesGenerator.cpp:40-108):
 * spectra [nBlk][nSub][2][nCh] complex -> [nBlk][nSub][nStokes][nCh]
 * (nStokes 1: I; 4: I, Q, U, V). */
int ppl_stokes(const float* d_spectra, unsigned nBlk, unsigned nSub, unsigned nCh,
               unsigned nStokes, float* d_out, void* stream);
/* StokesIntegrator::run (src/lib/src/StokesIntegrator.cpp:32-123). */
int ppl_integrate(const float* d_in, unsigned nT, unsigned nSub, unsigned nPol, unsigned nCh,
                  unsigned window, unsigned binCh, float* d_out, void* stream);

/* ---------------------------------------------------------------- clipper */

typedef struct ppl_clipper ppl_clipper;

typedef struct {
    const double* bp_values;  /* .bp file numbers: nChan, start, end, width, rms, median, poly... */
    int n_values;
    float crFactor;           /* channelRejectionRMS (default 10) */
    float srFactor;           /* spectrumRejectionRMS (default 4) */
    int maxHistory;           /* History maximum */
    int zeroDM;               /* zeroDMing active */
    int bandMatching;         /* Band matching (must be 1) */
    int useMeanOverRMS;       /* computeRMSfromMean active (default 1) */
    const unsigned* kill_index; /* indices of the .bp file's K lines (BandPassAdapter.cpp:58-71), or NULL */
    int n_kill;               /* entries in kill_index; 0: no killed bands */
} ppl_clipper_config;

/* RFI_Clipper (src/lib/src/RFI_Clipper.cpp:22-104): bandpass model from the
 * .bp values (BandPassAdapter.cpp:30-107), running state on the device.
 * Killed channels (K lines) are checked as the adapter checks them
 * (binIndex(index) == index on the file's map, else PPL_E_ARG with the
 * adapter's message) and hold 0.0 in the bandpass on every map
 * (BandPass.cpp:101-157). */
int ppl_clipper_create(ppl_clipper** clip, const ppl_clipper_config* cfg);
int ppl_clipper_destroy(ppl_clipper* clip);
/* RFI_Clipper::run (:154-573) in place on I [nT][nSub][nPol][nCh]; W (may be
 * NULL) receives the weights (always 1, :143,332). */
int ppl_clipper_run(ppl_clipper* clip, float* d_I, float* d_W, unsigned nT, unsigned nSub,
                    unsigned nPol, unsigned nCh, void* stream);
/* One beam's part of a batched run: spectra (and weights, may be NULL) of
 * nT spectra [nT][nSub][nPol][nCh]. */
typedef struct {
    float* d_I;
    float* d_W;
    unsigned nT, nSub, nPol, nCh;
} ppl_clip_job;
enum { PPL_CLIP_MAX_BEAMS = 48 };
/* RFI_Clipper::run on several independent clippers (beams) at once, each
 * on its own spectra and shape: every stage is one launch for all beams and
 * the sequential walk runs one workgroup per beam, so the beams' recurrences
 * proceed side by side.  Results equal nBeams separate ppl_clipper_run calls.
 * At most PPL_CLIP_MAX_BEAMS beams; a clipper may appear once per batch. */
int ppl_clipper_run_batch(ppl_clipper* const* clips, const ppl_clip_job* jobs, int nBeams, void* stream);
/* Same (Stokes-I only, nPol = 1, weights not written), then
 * DedispersionBuffer::addSamples of the clipped spectra [first, nT) into a
 * dedispersion buffer, both on the given stream: one call for a caller that
 * feeds a buffer (the two steps are separate kernels).  nSub*nCh must equal
 * the buffer's nChan (checked before the clipper runs).  As ppl_ddbuf_add,
 * *consumed receives how many spectra the buffer took and *full whether it
 * is full: when it filled part way, the caller rotates it and adds the
 * remaining clipped spectra (from first + *consumed) with ppl_ddbuf_add. */
int ppl_clipper_run_into(ppl_clipper* clip, float* d_I, unsigned nT, unsigned nSub, unsigned nCh,
                         ppl_ddbuf* buf, unsigned first, unsigned* consumed, int* full, void* stream);
/* RFI_Clipper::run (as ppl_clipper_run) in three calls, so that the
 * sequential walk can run on a stream of its own beside other work:
 *   _front: the run's start and the precompute of all its spectra (the
 *           whole chunk is one sub-chunk; a wider box than ppl_clipper_run's);
 *   _walk:  the walk over the chunk, one workgroup (then the precompute and
 *           walk of whatever follows a point where the state left the box);
 *   _back:  the statistics and write passes.
 * The caller orders them: the walk after the front, the back after the walk
 * (events between streams).  The results equal ppl_clipper_run's bit for
 * bit.  Calls out of this order fail with PPL_E_ARG.  No counterpart in the
 * reference, whose clipper runs inside its module (RFI_Clipper.cpp:154-573);
 * for schedules that overlap the walk with the previous buffer's
 * dedispersion. */
int ppl_clipper_run_front(ppl_clipper* clip, float* d_I, float* d_W, unsigned nT, unsigned nSub,
                          unsigned nPol, unsigned nCh, void* stream);
int ppl_clipper_run_walk(ppl_clipper* clip, void* stream);
int ppl_clipper_run_back(ppl_clipper* clip, void* stream);
/* Scalars after the last run: meanRunAve, rmsRunAve (WeightedSpectrumDataSet
 * setMean/setRMS, :569-570); synchronises. */
int ppl_clipper_stats(ppl_clipper* clip, float* mean, float* rms);
/* The live bandpass model, RFI_Clipper::bandPass() (RFI_Clipper.h:38, kept
 * current by setMean/setRMS after every spectrum, RFI_Clipper.cpp:562-563):
 * its constant coefficient, mean and RMS on the current map and, when
 * values is not NULL, its nValues (= the run's bins) values on the run map
 * (killed bins 0.0).  Before the first run: the file's model, no values. */
int ppl_clipper_model(ppl_clipper* clip, float* params0, float* mean, float* rms, float* values,
                      unsigned nValues);
/* Per-spectrum flags of the last run (1 = spectrum kept), host buffer. */
int ppl_clipper_flags(ppl_clipper* clip, uint8_t* h_good, unsigned nT);
/* Walk counters since creation (indices below); writes min(n, PPL_CLIP_NCOUNTERS)
 * values and synchronises.  FAST: spectra decided from the precomputed
 * channel mask; SLOW: spectra walked by the full workgroup, split by reason:
 * LEARN (history not yet full), BOX (state outside the speculation box, or
 * beyond the precomputed range), UNDECIDED (channels the box does not
 * decide: more than one, or one where the fast path could not run), FILL
 * (last-good slots to fill), VER (last-good versions used up); STOP:
 * sub-chunks ended early to re-centre the box; RESOLVED: FAST spectra with
 * exactly one undecided channel, decided at their own state. */
enum { PPL_CLIP_FAST = 0, PPL_CLIP_SLOW = 1, PPL_CLIP_LEARN = 2, PPL_CLIP_BOX = 3,
       PPL_CLIP_UNDECIDED = 4, PPL_CLIP_FILL = 5, PPL_CLIP_VER = 6, PPL_CLIP_STOP = 7,
       PPL_CLIP_RESOLVED = 8, PPL_CLIP_NCOUNTERS = 9 };
int ppl_clipper_counters(ppl_clipper* clip, long long* out, int n);

#ifdef __cplusplus
}
#endif
#endif /* PPL_AMD_H */
