/*
 * oracle.h -- CPU restatement of pelican-lofar's pulsar-search chain.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this code: it is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, and only as the checker (or the timed CPU
 * baseline), never as the thing measured on the GPU.
 *
 * Every function restates one reference function, following its FP32
 * expression forms and loop order (cited file:line are relative to the
 * pelican-lofar source tree).  All files are compiled with
 * -ffp-contract=off so that a*b+c is never fused, as in the reference's
 * x86-64 build.
 *
 * Parity status per function (what pins it) is listed beside each entry
 * point; "parity unpinned" marks what nothing but this restatement pins.
 */
#ifndef PPL_ORACLE_H
#define PPL_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- PPF */

/* Window codes for orc_ppf_coefficients. */
enum { ORC_WIN_KAISER = 0, ORC_WIN_GAUSSIAN = 1, ORC_WIN_BLACKMAN = 2,
       ORC_WIN_HAMMING = 3, ORC_WIN_HANN = 4 };

/* PolyphaseCoefficients::genereateFilter, src/lib/src/PolyphaseCoefficients.cpp:308-364
 * (windows :77-163, fir2 :216-304).  Writes nTaps*nCh doubles, index t*nCh+c,
 * with the (-1)^c sign and the 1/nCh scale.  The inverse transform is an FP64
 * DFT (the reference uses an FFTW float plan: bit parity with it is unpinned,
 * pinned instead by the Dirichlet closed form, symmetry and normalisation).
 * Returns 0 on success, -1 for an unknown window. */
int orc_ppf_coefficients(unsigned nTaps, unsigned nCh, int window, double* out);

/* The oversampled fir2 response before windowing, for closed-form tests:
 * writes the n+1 = nTaps*nCh taps of _generateFirFilter with a unit window. */
int orc_fir2_unwindowed(unsigned nTaps, unsigned nCh, double* out);

/* PPFChanneliser::_assign_threads, src/lib/src/PPFChanneliser.cpp:298-319. */
int orc_assign_threads(unsigned* start, unsigned* end, unsigned nValues,
                       unsigned nThreads, unsigned threadId);

/* PPF state: per (subband, pol) FIR history of nTaps blocks, initialised to 0
 * (PPFChanneliser.cpp:327-342). */
typedef struct orc_ppf orc_ppf;
orc_ppf* orc_ppf_create(unsigned nSub, unsigned nPol, unsigned nCh,
                        unsigned nTaps, const float* coeffs_f32,
                        const double* coeffs_f64);
void orc_ppf_destroy(orc_ppf* p);

/* PPFChanneliser::run, src/lib/src/PPFChanneliser.cpp:81-187.
 * in : complex int16 voltages, interleaved (re,im), layout [sub][pol][t]
 *      (the TimeSeriesDataSetC32 layout of TimeSeriesDataSet.h:145-151 with
 *      block length nCh), nBlk*nCh samples per (sub,pol).
 * out: complex spectra [blk][sub][pol][nCh] (SpectrumDataSet.h:229-235).
 * fir32: FIR in FP32 in the reference's order (_filter, :266-290); the
 *        N-point forward DFT is then evaluated in FP64 (the reference's FFTW
 *        float FFT is parity unpinned; its closed form is the DFT).
 * Writes double complex (re,im) to out. */
void orc_ppf_run_fir32(orc_ppf* p, const int16_t* in, unsigned nBlk, double* out);
/* FP32 FIR + FP32 radix-2 FFT, OpenMP over streams: the timed CPU baseline
 * (complex float out [blk][sub][pol][nCh]; nCh a power of two). */
void orc_ppf_run_f32(orc_ppf* p, const int16_t* in, unsigned nBlk, float* out, int nthreads);
/* FP64 twin: FP64 coefficients, FP64 FIR and FP64 DFT. */
void orc_ppf_run_f64(orc_ppf* p, const int16_t* in, unsigned nBlk, double* out);
/* The FP32 FIR stage alone (bit-exact target), complex float out [blk][sub][pol][nCh]. */
void orc_ppf_fir_only_f32(orc_ppf* p, const int16_t* in, unsigned nBlk, float* out);

/* ---------------------------------------------------- Stokes / integrate */

/* StokesGenerator::run, src/lib/src/StokesGenerator.cpp:40-108, Stokes-I only:
 * I = (Xr*Xr + Xi*Xi) + (Yr*Yr + Yi*Yi) in FP32.  in: complex float
 * [blk][sub][2][nCh]; out: float [blk][sub][1][nCh]. */
void orc_stokes_i_f32(const float* in, unsigned nBlk, unsigned nSub,
                      unsigned nCh, float* out);
/* Same from double spectra, in double (the FP64 twin). */
void orc_stokes_i_f64(const double* in, unsigned nBlk, unsigned nSub,
                      unsigned nCh, double* out);

/* StokesIntegrator::run, src/lib/src/StokesIntegrator.cpp:32-123:
 * out[u] = sum over window (time-major, then channel bins) starting at 0.0f.
 * in [nT][nSub][nPol][nCh] -> out [nT/window][nSub][nPol][nCh/binCh]. */
void orc_integrate_f32(const float* in, unsigned nT, unsigned nSub,
                       unsigned nPol, unsigned nCh, unsigned window,
                       unsigned binCh, float* out);
void orc_integrate_f64(const double* in, unsigned nT, unsigned nSub,
                       unsigned nPol, unsigned nCh, unsigned window,
                       unsigned binCh, double* out);

/* -------------------------------------------------------------- clipper */

/* RFI_Clipper (src/lib/src/RFI_Clipper.cpp:22-573) with BandPass
 * (src/lib/src/BandPass.cpp:26-166), BandPassAdapter::deserialise
 * (src/lib/src/BandPassAdapter.cpp:30-107) and BinMap
 * (src/lib/src/BinMap.cpp:49-70).
 * bp_values: the numbers of a .bp file in order (nChan, start, end, width,
 * rms, median, poly...), n_values >= 6. */
typedef struct orc_clipper orc_clipper;
orc_clipper* orc_clipper_create(const double* bp_values, int n_values,
                                float crFactor, float srFactor, int maxHistory,
                                int zeroDM, int bandMatching,
                                int useMeanOverRMS);
/* The same with the indices of the .bp file's K (kill) lines
 * (BandPassAdapter.cpp:58-71,101-106, BandPass.cpp:111-157); nullptr when an
 * index fails the adapter's frequency check or reaches outside the map. */
orc_clipper* orc_clipper_create_kill(const double* bp_values, int n_values,
                                     const unsigned* kill_index, int n_kill,
                                     float crFactor, float srFactor, int maxHistory,
                                     int zeroDM, int bandMatching, int useMeanOverRMS);
void orc_clipper_destroy(orc_clipper* c);
/* In-place run over one blob [nT][nSub][nPol][nCh]; weights likewise
 * (may be NULL).  Returns 0, or -1 if nSub*nCh is not divisible by 8. */
int orc_clipper_run(orc_clipper* c, float* I, float* W, unsigned nT,
                    unsigned nSub, unsigned nPol, unsigned nCh);
/* Per-spectrum record of the last run (for stage-wise checks):
 * good[t] (1 = spectrum kept), nclip[t] channels clipped. */
int orc_clipper_last_flags(orc_clipper* c, uint8_t* good, int32_t* nclip,
                           unsigned nT);
/* Scalar state after the last spectrum: meanRunAve, rmsRunAve, params0,
 * bandpass mean, meanOverRMS; history sizes. */
void orc_clipper_state(orc_clipper* c, float* out5, int* sizes2);
/* BandPass alone (src/lib/src/BandPass.cpp:26-108): setData over the map
 * (nBins, start, width) with params, setRMS(rms), setMedian(median) unless
 * median is NaN, then reBin to (nBins2, start2, width2) when nBins2 > 0.
 * out4 = median(), rms(), intensityOfBin(0), params[0] of the result.
 * Pinned by the known answers of BandPassTest.cpp:35-101. */
int orc_bandpass_probe(unsigned nBins, double start, double width, const float* params,
                       int nParams, float rms, float median, unsigned nBins2,
                       double start2, double width2, float* out4);
/* The bandpass values of the currently-binned map (nBins floats). */
int orc_clipper_bandpass(orc_clipper* c, float* out, unsigned nBins);

/* ---------------------------------------------------------- dedispersion */

/* DedispersionModule::resize arithmetic, src/lib/src/DedispersionModule.cpp:
 * 152-162, kernel members :363-364 (DedispersionModule.h:50-52).
 * shift[nChan] (float), scalars out: mstartdm, mdmstep (float);
 * ints out: maxshift, remainingSamples, nOut.  Returns 0 or -1 when
 * nsampBuf <= maxshift (the module's throw, :175-178). */
int orc_dedisp_plan(unsigned nChan, unsigned nsampBuf, unsigned tdms,
                    float dmLow, float dmStep, double tsamp, double fch1,
                    double foff, float* shift, float* mstart_mstep,
                    int* maxshift_rem_nout);
/* Delay table delay[dm][c] = trunc(shift[c] * (mstartdm + dm*mdmstep))
 * (src/lib/DedispersionKernel.cu:29,36). */
void orc_dedisp_delays(const float* shift, unsigned nChan, unsigned tdms,
                       float mstartdm, float mdmstep, int32_t* delay);
void orc_dedisp_delays_fused(const float* shift, unsigned nChan, unsigned tdms,
                             float mstartdm, float mdmstep, int32_t* delay);
/* Brute force, channel order, FP32 sums: out[dm][t], t < nOut
 * (DedispersionKernel.cu:15-50 arithmetic; the launch geometry is not
 * restated).  in [nChan][nsampBuf]. */
void orc_dedisperse_f32(const float* in, unsigned nChan, unsigned nsampBuf,
                        unsigned tdms, const int32_t* delay, unsigned nOut,
                        float* out, int nthreads);
/* FP16 input (IEEE binary16 bits), FP32 sums in channel order. */
void orc_dedisperse_f16(const uint16_t* in, unsigned nChan, unsigned nsampBuf,
                        unsigned tdms, const int32_t* delay, unsigned nOut,
                        float* out, int nthreads);
/* 8-bit offset-binary input, int32 sums. */
void orc_dedisperse_u8(const uint8_t* in, unsigned nChan, unsigned nsampBuf,
                       unsigned tdms, const int32_t* delay, unsigned nOut,
                       int32_t* out, int nthreads);
/* float -> binary16, round to nearest even (for the FP16 mode). */
void orc_f32_to_f16(const float* in, uint16_t* out, size_t n);
/* Offset-binary quantiser: q = clamp(rint(x*inv_scale), -128, 127) + 128. */
void orc_quantise_u8(const float* in, uint8_t* out, size_t n, float inv_scale);
void orc_quantise_u8_rows(const float* x, unsigned nChan, unsigned nsamp, float inv_scale, uint8_t* q,
                          float* off);
void orc_u8_series(const int32_t* sums, const float* off, const int32_t* delay, unsigned nChan,
                   unsigned nsamp, unsigned tdms, unsigned nOut, float scale, float* out);

/* DedispersionBuffer::addSamples + copy (src/lib/src/DedispersionBuffer.cpp:
 * 65-96,98-179) for Stokes-I blobs [t][sub][0][ch] with invert = 0.
 * A streaming assembler: feed blobs, get whole buffers [nChan][nsampBuf]. */
typedef struct orc_ddbuf orc_ddbuf;
orc_ddbuf* orc_ddbuf_create(unsigned nChan, unsigned nsampBuf, unsigned overlap);
void orc_ddbuf_destroy(orc_ddbuf* b);
/* Appends nT spectra (sampleSize nChan each, already [t][c]).  Each time the
 * buffer fills, it is copied to out + k*nChan*nsampBuf (k = 0,1,...), the
 * overlap is carried, and k is incremented.  Returns the number of full
 * buffers written (out may be NULL to count). */
int orc_ddbuf_add(orc_ddbuf* b, const float* spectra, unsigned nT, float* out,
                  int max_out);

/* ------------------------------------------------------------- detection */

/* DedispersionAnalyser::analyse, src/lib/src/DedispersionAnalyser.cpp:30-203.
 * data [tdms][nsamp].  events (dm, t, binFactor) int32 triples and values,
 * in the reference's order, starting with the dummy (0,0,1,0.0) event.
 * Returns the number of events, or -(needed) if max_events is too small. */
int orc_detect(const float* data, int tdms, int nsamp, int nChannels,
               int nSubbands, float threshold, unsigned binPow2,
               int32_t* ev_ints, float* ev_vals, int max_events);

#ifdef __cplusplus
}
#endif
#endif
