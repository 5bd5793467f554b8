// ppl_clipper.hip -- RFI clipper (src/lib/src/RFI_Clipper.cpp:154-573) on the
// device, bit-exact with the reference's FP32/FP64 expression forms.
//
// The clipper is a recurrence over spectra: every spectrum updates running
// means, ring buffers and the bandpass model that the next spectrum is
// judged against.  The work splits in two:
//
//  * walk  (one workgroup, sequential over spectra): everything that feeds
//    the recurrence -- the learning-phase band statistics, the channel mask,
//    the ordered sum of the kept channels, the spectrum decision, the ring
//    buffers, the last-good spectrum and the bandpass offset.  Elementwise
//    work is spread over the workgroup; sums whose rounding the reference
//    fixes by its loop order are taken in that order by one lane.  For each
//    spectrum it records the state the rest of the spectrum needs.
//  * apply (parallel over spectra): replaces clipped spectra, flattens by
//    the bandpass, takes the ordered mean/RMS of each spectrum (one lane per
//    spectrum, LDS-staged tiles) and normalises.  These values never feed
//    back into the recurrence.
//
// The bandpass values are scale * evaluate(v_i) with evaluate's float
// accumulation of double terms (BandPass.cpp:159-166); the terms
// params[j] * pow(v_i, j), j >= 1, do not change after construction and are
// tabulated on the host with the C library's pow, so only the offset
// params[0] moves on the device.
#include "ppl_internal.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

namespace {

using namespace ppl;

constexpr int kMaxParams = 8;
enum {
    CNT_FAST = PPL_CLIP_FAST, CNT_SLOW = PPL_CLIP_SLOW, CNT_LEARN = PPL_CLIP_LEARN, CNT_BOX = PPL_CLIP_BOX,
    CNT_UNDECIDED = PPL_CLIP_UNDECIDED, CNT_FILL = PPL_CLIP_FILL, CNT_VER = PPL_CLIP_VER, CNT_STOP = PPL_CLIP_STOP,
    CNT_RESOLVED = PPL_CLIP_RESOLVED
};
// Spectra per sub-chunk of the walk, and sub-chunk launches beyond
// nT / kSubLen that a run makes room for (each early end costs one).
constexpr int kSubLen = 4096;
constexpr int kExtraIter = 8;

// ---------------------------------------------------------------- host side:
// BandPass / BinMap restatement used to set the model up
// (src/lib/src/BandPass.cpp:26-108, src/lib/src/BinMap.cpp:49-70,
// src/lib/src/BandPassAdapter.cpp:76-100).
struct HostMap {
    unsigned nBins = 0;
    double lower = 0.0, width = 1.0;
    double binAssignmentNumber(int i) const { return lower + (width * i); }
    double endValue() const { return lower + width * nBins; }
    int binIndex(double v) const { return (int)(0.5 + ((v - lower) / width)); }       // BinMap.cpp:49-52
    double binStart(unsigned i) const { return binAssignmentNumber(i) - width / 2.0; }  // :72-75
    double binEnd(unsigned i) const { return binAssignmentNumber(i) + width / 2.0; }    // :77-80
};

struct HostBandPass {
    HostMap primary, current;
    std::vector<float> params;
    // values of the primary map and the map-scoped statistics
    float primaryMean = 0, primaryMedian = 0, primaryRms = 0;
    float curMean = 0, curMedian = 0, curRms = 0;
    bool currentIsPrimary = true;

    float evaluate(float v) const
    {
        float tot = 0.0f;
        for (size_t i = 0; i < params.size(); ++i) tot += (double)params[i] * std::pow((double)v, (double)(int)i);
        return tot;
    }
    std::vector<float> build(const HostMap& m, float scale) const
    {
        std::vector<float> d(m.nBins);
        for (unsigned i = 0; i < m.nBins; ++i) d[i] = scale * evaluate((float)m.binAssignmentNumber(i));
        return d;
    }
    // setData (:26-42): the primary map is the current map, QHash default 0
    void setData(const HostMap& m, const std::vector<float>& p)
    {
        primary = current = m;
        currentIsPrimary = true;
        params = p;
        // reBin(primary): scale 1, rms/median/mean = 0 * ...
        primaryRms = 0.0f * std::sqrt(1.0 / 1.0);
        primaryMedian = 0.0f;
        std::vector<float> d = build(m, 1.0f);
        float mean = 0;
        for (unsigned i = 0; i < m.nBins; ++i) mean += d[i];
        mean /= m.nBins;
        primaryMean = mean;
        std::vector<float> copy = d;
        std::nth_element(copy.begin(), copy.begin() + copy.size() / 2, copy.end());
        primaryMedian = copy[copy.size() / 2];
        curMean = primaryMean;
        curMedian = primaryMedian;
        curRms = primaryRms;
    }
    // setMedian (:48-61) while the current map is the primary map: the two
    // QHash entries are one, so the mean moves by delta twice.
    void setMedian(float median)
    {
        float delta = median - curMedian;
        if (std::fabs(delta) > 0.0f) {
            float scale = current.width / primary.width;
            curMedian = median;
            curMedian = median / scale;  // _median[primary] (same entry)
            params[0] += delta / scale;
            curMean += delta;
            curMean += delta / scale;
            primaryMedian = curMedian;
            primaryMean = curMean;
        }
    }
    void setRMS(float r)
    {
        curRms = r;
        if (currentIsPrimary) primaryRms = r;
    }
    // reBin to the run map (:81-99), a map distinct from the primary one
    void reBin(const HostMap& m)
    {
        current = m;
        currentIsPrimary = (m.nBins == primary.nBins && m.lower == primary.lower && m.width == primary.width);
        if (currentIsPrimary) return;
        double scale = m.width / primary.width;
        curRms = primaryRms * std::sqrt(1.0 / scale);
        curMedian = primaryMedian * scale;
        curMean = primaryMean * scale;
    }
};

// ------------------------------------------------------------ device state
struct ClipScalars {
    float meanRunAve, rmsRunAve, meanOverRMS, fractionBadChannels;
    float params0, bpMean, bpRms, spectrumRMS;
    int badSpectra, num, remainingZeros;
    int meanStart, meanSize, rmsStart, rmsSize;
    int initialised;
    // newest and second-newest ring entries, kept beside the rings so the
    // walk never reads them back from memory
    float meanBack, meanBack2, rmsBack, rmsBack2;
    // run control: next spectrum to walk, last-good version in force,
    // growth factor of the speculation box
    int tNext, ver;
    float boxGrow;
    // spectra decided from the precomputed mask / by the full workgroup, the
    // latter by reason, and sub-chunks ended early (ppl_clipper_counters)
    long long count[PPL_CLIP_NCOUNTERS];
};

struct ClipConst {
    int nBins, nSub, nPol, nCh, cap;
    int nTerms;              // params.size() - 1 tabulated terms
    float scale;             // float(current.width / primary.width)
    float crFactor, srFactor;
    float meanMinusMinimum;  // float(4 / sqrt(8.0))
    double sqrtNBins;        // sqrt((double)nBins)
    int zeroDM, useMeanOverRMS, maxHistory;
};

// Per-spectrum record written by the walk for the apply kernel.
struct SpecRec {
    float params0;     // bandpass offset used by this spectrum
    float dataModel;
    float margin;      // channel threshold of this spectrum
    float meanRunAve;  // for the non-zero-DM normalisation
    int flags;         // bit0: spectrum kept; bit1: channel replacement already applied
    int ver;           // last-good version in force for this spectrum
};

// Killed bins (BandPass::_zeroChannelsMap, BandPass.cpp:111-125) hold 0.0 in
// every model the clipper builds; kill is a bit mask over the run map's bins,
// or null.
__device__ __forceinline__ bool bin_killed(const unsigned* __restrict__ kill, int i)
{
    return kill != nullptr && ((kill[i >> 5] >> (i & 31)) & 1u);
}
__device__ __forceinline__ float bp_value(const double* __restrict__ terms, int nTerms, int nBins, int i,
                                          float p0, float scale, const unsigned* __restrict__ kill)
{
    if (bin_killed(kill, i)) return 0.0f;
    // evaluate(): tot = 0; tot += params[0]*1; tot += params[j]*pow(v,j)
    float tot = 0.0f;
    tot = (float)((double)tot + (double)p0 * 1.0);
    for (int j = 0; j < nTerms; ++j) tot = (float)((double)tot + terms[(size_t)j * nBins + i]);
    return scale * tot;
}

struct RingView {
    float* buf;
    double* prefix;  // prefix[k] = ((0 + e0) + ...) + e_{k-1} in double, valid while not full
    int cap;
};

__device__ __forceinline__ void ring_push(RingView r, int& start, int& size, float& back, float& back2, float v)
{
    if (r.cap == 0) return;
    if (size < r.cap) {
        r.buf[(start + size) % r.cap] = v;
        r.prefix[size + 1] = r.prefix[size] + (double)v;
        ++size;
    } else {
        r.buf[start] = v;
        start = (start + 1) % r.cap;
    }
    back2 = back;
    back = v;
}
// pop_back() then push_back(back()): the newest entry becomes a copy of the
// one before it (RFI_Clipper.cpp:432-437)
__device__ __forceinline__ void ring_repeat_previous(RingView r, int& start, int& size, float& back, float& back2)
{
    const float prev = back2;
    size -= 1;
    ring_push(r, start, size, back, back2, prev);
    back2 = prev;
}
__device__ __forceinline__ float ring_front(RingView r, int start) { return r.buf[start]; }

constexpr int kWalkThreads = 512;
constexpr int kStage = 512;      // spectra whose precomputed records are staged in LDS at a time
constexpr int kRowChunk = 8192;  // bins of one spectrum staged in LDS for the ordered sums
constexpr int kMaxBinsWalk = 32768;  // bins per spectrum the walk supports (kept-bin mask in LDS)
static_assert(kStage == kWalkThreads, "stage scans use one thread per staged spectrum");

// Per-sub-chunk speculation box: every spectrum of the sub-chunk whose
// bandpass offset lies in [p0lo, p0hi] and whose margin lies in [mlo, mhi]
// is judged by the channel mask the precompute pass fixed for that
// spectrum, provided no channel of it was undecided over the whole box.
struct Box {
    float p0lo, p0hi, mlo, mhi;
    int valid;
};
// Box half-widths around the sub-chunk-start state, before the growth
// factor: the bandpass offset moves the bandpass by +-kBoxP channel rms,
// the margin by a relative +-kBoxM.  In steady state both follow running
// averages over the history (10000 spectra by default), which move far less
// than this over a sub-chunk; a state that leaves the box ends the
// sub-chunk, and the next one starts with a box four times wider.
constexpr float kBoxP = 0.005f;
constexpr float kBoxM = 0.001f;
constexpr float kMaxBoxGrow = 64.0f;

// Precomputed per spectrum (steady state, dataModel = 0), for the mask that
// holds over the whole box: the spectrum mean of the kept channels (their
// ordered sum over their count), the increment of the bad-channel fraction
// and the bad-band count.
struct PreRec {
    float mean;
    float fracInc;
    int badBands;
    int flags;  // PRE_* bits
    // exactly one undecided channel (PRE_ONE): the same three for the mask
    // with that channel kept, the channel and its value
    float meanK, fracIncK;
    int badBandsK, undBin;
    float undVal;
};
enum {
    PRE_UND = 1,      // some channel undecided over the box
    PRE_FILL = 2,     // a kept (or undecided) non-zero channel whose last-good slot was empty
    PRE_REJ = 4,      // (walk) the mask in force rejects the spectrum for certain: 4 bad bands or more
    PRE_ONE = 8,       // exactly one channel undecided: its decision picks one of two records
    PRE_PREDKEPT = 16, // (PRE_ONE) the undecided channel is kept at the predicting state
    PRE_TOLREJ = 32    // (walk) the mask keeps it, but its mean exceeds the tolerance at the stage-start state
};

constexpr int kMaxVer = 64;  // last-good versions kept per run for the apply pass

// Row-tiled passes (precompute, statistics): a workgroup of 256 threads owns
// kTileRows spectra; bins stream through double-buffered LDS tiles of
// kTileBins, loaded 4 bins per thread; one lane per spectrum folds the tile
// in bin order while the other waves load and judge the next tile.
constexpr int kTileRows = 32;
constexpr int kTileBins = 256;
constexpr int kTileThreads = 256;
constexpr int kRowsPerWave = kTileRows / 4;
// The precompute and learning records: spectra per workgroup.  Each row's
// ordered sum is one lane's dependent chain, so fewer rows per workgroup
// spread a sub-chunk's rows over more CUs.
constexpr int kPreRows = 4;
constexpr int kPreRowsPerWave = kPreRows / 4;

// Loads of 4 consecutive bins (i % 4 == 0) of polarisation 0 of one spectrum.
// mode 0: rows of contiguous bins (nPol == 1), 16-byte aligned; mode 1:
// subbands of nCh (nCh % 4 == 0) interleaved with the other polarisations,
// aligned; mode 2: anything else, one bin at a time.
__device__ __forceinline__ long bin_offset(int i, int nCh, int nPol)
{
    return (long)(i / nCh) * nCh * nPol + i % nCh;
}
__device__ __forceinline__ float4 load_quad(const float* __restrict__ I, long rowOff, int i, int nCh, int nPol,
                                            int mode)
{
    if (mode == 0) return *reinterpret_cast<const float4*>(I + rowOff + i);
    if (mode == 1) return *reinterpret_cast<const float4*>(I + rowOff + bin_offset(i, nCh, nPol));
    float4 r;
    r.x = I[rowOff + bin_offset(i, nCh, nPol)];
    r.y = I[rowOff + bin_offset(i + 1, nCh, nPol)];
    r.z = I[rowOff + bin_offset(i + 2, nCh, nPol)];
    r.w = I[rowOff + bin_offset(i + 3, nCh, nPol)];
    return r;
}
__device__ __forceinline__ void store_quad(float* __restrict__ I, long rowOff, int i, int nCh, int nPol, int mode,
                                           float4 v)
{
    if (mode == 0) {
        *reinterpret_cast<float4*>(I + rowOff + i) = v;
    } else if (mode == 1) {
        *reinterpret_cast<float4*>(I + rowOff + bin_offset(i, nCh, nPol)) = v;
    } else {
        I[rowOff + bin_offset(i, nCh, nPol)] = v.x;
        I[rowOff + bin_offset(i + 1, nCh, nPol)] = v.y;
        I[rowOff + bin_offset(i + 2, nCh, nPol)] = v.z;
        I[rowOff + bin_offset(i + 3, nCh, nPol)] = v.w;
    }
}
__device__ __forceinline__ float& q_at(float4& q, int j) { return (&q.x)[j]; }
__device__ __forceinline__ float q_at(const float4& q, int j) { return (&q.x)[j]; }

// Polynomial terms of the bandpass for one bin held in registers (NT known
// at compile time), or read from memory (NT = -1).
template <int NT>
struct BinTerms {
    double t[NT > 0 ? NT : 1];
    bool killed;
    __device__ __forceinline__ void load(const double* __restrict__ terms, int, int nBins, int i,
                                         const unsigned* __restrict__ kill)
    {
#pragma unroll
        for (int j = 0; j < NT; ++j) t[j] = terms[(size_t)j * nBins + i];
        killed = bin_killed(kill, i);
    }
    __device__ __forceinline__ float bp(float p0, float scale) const
    {
        if (killed) return 0.0f;
        float tot = 0.0f;
        tot = (float)((double)tot + (double)p0 * 1.0);
#pragma unroll
        for (int j = 0; j < NT; ++j) tot = (float)((double)tot + t[j]);
        return scale * tot;
    }
};
template <>
struct BinTerms<-1> {
    const double* terms;
    const unsigned* kill;
    int nTerms, nBins, i;
    __device__ __forceinline__ void load(const double* __restrict__ tm, int nt, int nb, int bin,
                                         const unsigned* __restrict__ km)
    {
        terms = tm;
        nTerms = nt;
        nBins = nb;
        i = bin;
        kill = km;
    }
    __device__ __forceinline__ float bp(float p0, float scale) const
    {
        return bp_value(terms, nTerms, nBins, i, p0, scale, kill);
    }
};


__device__ __forceinline__ float block_min(float v, float* red)
{
    // NaN-ignoring min: the reference's `if (x < m) m = x` never takes a NaN
    const int tid = threadIdx.x;
    for (int off = 32; off >= 1; off >>= 1) {
        const float o = __shfl_down(v, off, 64);
        v = (o < v) ? o : v;
    }
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        float m = red[0];
        for (int w = 1; w < kWalkThreads / 64; ++w) m = (red[w] < m) ? red[w] : m;
        red[kWalkThreads / 64] = m;
    }
    __syncthreads();
    const float r = red[kWalkThreads / 64];
    __syncthreads();
    return r;
}

// Band bookkeeping of RFI_Clipper.cpp:345-372 from the kept-channel counts.
__device__ __forceinline__ void band_stats(const int* cnt, int nBins, int cpb, int& total, int& badBands,
                                           float& fracInc)
{
    total = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) total += cnt[b];
    unsigned bad = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        if ((double)(float)cnt[b] < 0.8 * (unsigned)cpb) ++bad;
        if (cnt[b] == 0) bad += 4;
    }
    if ((double)(unsigned)total < 0.8 * (unsigned)nBins) bad += 4;
    badBands = (int)bad;
    fracInc = (float)(unsigned)(nBins - total) / (float)(unsigned)nBins;
}

// Ordered folds of an LDS row of n floats (n % 8 == 0, n >= 8) by one lane,
// the reads of the next 8 values issued before the adds of the current 8.
__device__ __forceinline__ float fold_sum(const float* row, int n, float sum)
{
    float a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = row[j];
    for (int k = 8; k < n; k += 8) {
        float b[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = row[k + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) sum += a[j];
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = b[j];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) sum += a[j];
    return sum;
}

// The same fold for long rows (n % 32 == 0, 16-byte aligned): 32 values
// (eight 16-byte reads) in flight while the previous 32 are added, so the
// LDS latency hides behind the dependent adds.
__device__ __forceinline__ float fold_sum_deep(const float* row, int n, float sum)
{
    const float4* r4 = reinterpret_cast<const float4*>(row);
    float4 a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = r4[j];
    for (int k = 8; k < n / 4; k += 8) {
        float4 b[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = r4[k + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            sum += a[j].x;
            sum += a[j].y;
            sum += a[j].z;
            sum += a[j].w;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = b[j];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        sum += a[j].x;
        sum += a[j].y;
        sum += a[j].z;
        sum += a[j].w;
    }
    return sum;
}
__device__ __forceinline__ void fold_sum_sq(const float* row, int n, float& sum, float& sumSq)
{
    float a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = row[j];
    for (int k = 8; k < n; k += 8) {
        float b[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = row[k + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            sum += a[j];
            sumSq += a[j] * a[j];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = b[j];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        sum += a[j];
        sumSq += a[j] * a[j];
    }
}

// Band (0..7) of bin i: floor((i + 0.5) / cpb), exact in float while
// 0.5 / cpb exceeds the 2^-21 error of the product (cpb < 2^19).
__device__ __forceinline__ int band_of(int i, float invCpb) { return (int)(((float)i + 0.5f) * invCpb); }

// Data-only statistics of one learning-phase spectrum (RFI_Clipper.cpp:214-242
// and the kept-channel sum of :315-342), taken before the walk reaches it.
struct LearnRec {
    float bmin[8], bsum[8], bsq[8];  // per band: minimum, ordered sum and sum of squares
    float fullSum;   // ordered sum of every bin: the kept sum when nothing is clipped
    float predSum;   // ordered sum of the bins <= thr (the others as +0.0)
    float thr;       // predicted clipping level: mean + 10 standard deviations of the spectrum
    int predKept;    // bins <= thr
};

// Device-resident state of one clipper (one beam), fixed between runs.
struct BeamState {
    ClipConst K;
    ClipScalars* S;
    float *meanBuf, *rmsBuf, *lastGood, *lastGoodVer, *bpLo, *bpHi;
    double *meanPrefix, *rmsPrefix;
    const double* terms;
    const unsigned* kill;  // killed bins of the run map (bit mask), or null
    Box* box;
    PreRec* pre;
    LearnRec* learn;
    SpecRec* rec;
    float2* stats;
};

// What changes per run, for every beam of a batch (kernel argument): the
// beams' states, their spectra and weights, spectrum counts, sub-chunk
// lengths and access modes.  blockIdx.z selects the beam.
constexpr int kMaxBeams = PPL_CLIP_MAX_BEAMS;
struct BatchIO {
    const BeamState* st[kMaxBeams];
    float* I[kMaxBeams];
    float* W[kMaxBeams];
    int nT[kMaxBeams];
    int subLen[kMaxBeams];
    int ldMode[kMaxBeams];
};

// Start of a run: version 0 of the last-good spectrum is its current state.
__global__ void clip_begin_kernel(BatchIO io)
{
    const BeamState& bs = *io.st[blockIdx.z];
    ClipScalars* __restrict__ S = bs.S;
    const float* __restrict__ lastGood = bs.lastGood;
    float* __restrict__ lastGoodVer = bs.lastGoodVer;
    const int nBins = bs.K.nBins;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        S->tNext = 0;
        S->ver = 0;
    }
    if (i < nBins) lastGoodVer[i] = lastGood[i];
}

// The box around the sub-chunk-start state s and the bandpass at its two
// offsets: thread i0 (of a set of `stride` threads) writes the box when it
// is thread 0, and bins i0, i0 + stride, ...  Nothing is written once the
// run is done (s.tNext >= nT).
__device__ __forceinline__ void bounds_of_state(const BeamState& bs, const ClipScalars& s, int nT, int i0, int stride)
{
    const ClipConst& K = bs.K;
    const double* __restrict__ terms = bs.terms;
    Box* __restrict__ box = bs.box;
    float* __restrict__ bpLo = bs.bpLo;
    float* __restrict__ bpHi = bs.bpHi;
    if (s.tNext >= nT) return;
    const bool steady = s.rmsSize == K.cap && s.meanSize == K.cap;
    const float rms = s.rmsRunAve;
    const float g = s.boxGrow;
    const float dp = kBoxP * g * rms / fabsf(K.scale);
    const float p0lo = s.params0 - dp, p0hi = s.params0 + dp;
    const float margin = K.crFactor * rms;
    const bool valid = steady && isfinite(rms) && rms > 0.0f && isfinite(s.params0) && isfinite(dp) &&
                       p0lo < p0hi && margin > 0.0f && isfinite(margin);
    if (i0 == 0) {
        Box b;
        b.p0lo = p0lo;
        b.p0hi = p0hi;
        b.mlo = margin * (1.0f - kBoxM * g);
        b.mhi = margin * (1.0f + kBoxM * g);
        b.valid = valid ? 1 : 0;
        *box = b;
    }
    if (!valid) return;
    for (int i = i0; i < K.nBins; i += stride) {
        const float a = bp_value(terms, K.nTerms, K.nBins, i, p0lo, K.scale, bs.kill);
        const float b = bp_value(terms, K.nTerms, K.nBins, i, p0hi, K.scale, bs.kill);
        bpLo[i] = fminf(a, b);
        bpHi[i] = fmaxf(a, b);
    }
}

// The box of a run's first sub-chunk (the walk sets it up for the next ones
// as it ends).
__global__ void clip_bounds_kernel(BatchIO io)
{
    const int beam = blockIdx.z;
    const BeamState& bs = *io.st[beam];
    const ClipScalars s = *bs.S;
    bounds_of_state(bs, s, io.nT[beam], blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Precompute over the spectra of one sub-chunk: the kept-channel sum in bin
// order (the reference's order) for the mask that holds over the whole box,
// the kept count per band, and flags for undecided channels and last-good
// fill candidates.  A channel is kept for every state in the box when
// fl(I - bpLo) <= mlo and clipped for every state when fl(I - bpHi) > mhi
// (fl(I - bp) falls as bp rises; bp is monotone in the offset).  Clipped
// channels enter the fold as +0.0, which leaves a sum started at +0.0
// unchanged.
__device__ __forceinline__ void pre_rows(const BatchIO& io)
{
    const int beam = blockIdx.z;
    const BeamState& bs = *io.st[beam];
    const float* __restrict__ I = io.I[beam];
    const ClipConst K = bs.K;
    const ClipScalars* __restrict__ S = bs.S;
    const Box* __restrict__ box = bs.box;
    const float* __restrict__ bpLo = bs.bpLo;
    const float* __restrict__ bpHi = bs.bpHi;
    const float* __restrict__ lastGood = bs.lastGood;
    PreRec* __restrict__ pre = bs.pre;
    const int nT = io.nT[beam], subLen = io.subLen[beam], ldMode = io.ldMode[beam];
    __shared__ float vals[2][kPreRows][kTileBins + 1];
    // the same rows with the undecided channels kept, for rows with one
    __shared__ float valsK[2][kPreRows][kTileBins + 1];
    __shared__ int cnt[kPreRows][8];
    __shared__ int rowFlags[kPreRows];
    __shared__ int undCnt[kPreRows], undBinS[kPreRows];
    __shared__ float undValS[kPreRows];
    const int tStart = S->tNext;
    if (tStart >= nT) return;
    const Box bx = *box;
    if (!bx.valid) return;
    const int k0 = blockIdx.x * kPreRows;
    const int t0 = tStart + k0;
    const int nRows = min(min(kPreRows, nT - t0), subLen - k0);
    if (nRows <= 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nBins = K.nBins, nCh = K.nCh, nPol = K.nPol;
    const int cpb = nBins / 8;
    const float invCpb = 1.0f / (float)cpb;
    const long rowStride = (long)nBins * nPol;
    const int nTiles = (nBins + kTileBins - 1) / kTileBins;
    for (int i = tid; i < kPreRows * 8; i += kTileThreads) (&cnt[0][0])[i] = 0;
    if (tid < kPreRows) {
        rowFlags[tid] = 0;
        undCnt[tid] = 0;
    }
    __syncthreads();

    float4 raw[kPreRowsPerWave];
    auto issue = [&](int tile) {
        const int i = tile * kTileBins + lane * 4;
        if (i < nBins) {
#pragma unroll
            for (int q = 0; q < kPreRowsPerWave; ++q) {
                const int r = wave + 4 * q;
                if (r < nRows) raw[q] = load_quad(I, (long)(t0 + r) * rowStride, i, nCh, nPol, ldMode);
            }
        }
    };
    auto judge = [&](int tile, int buf) {
        const int tb0 = tile * kTileBins;
        const int i = tb0 + lane * 4;
        const bool inb = i < nBins;
        float4 lo4 = make_float4(0, 0, 0, 0), hi4 = lo4, lg4 = lo4;
        if (inb) {
            lo4 = *reinterpret_cast<const float4*>(bpLo + i);
            hi4 = *reinterpret_cast<const float4*>(bpHi + i);
            lg4 = *reinterpret_cast<const float4*>(lastGood + i);
        }
        int band[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) band[j] = band_of(i + j, invCpb);
        const int bandLo = band_of(tb0, invCpb);
        const int bandHi = band_of(min(tb0 + kTileBins, nBins) - 1, invCpb);
#pragma unroll
        for (int q = 0; q < kPreRowsPerWave; ++q) {
            const int r = wave + 4 * q;
            if (r >= nRows) break;  // uniform over the wave
            bool keep[4];
            bool und = false, fill = false;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float v = q_at(raw[q], j);
                const float eHi = v - q_at(lo4, j);  // largest fl(I - bp) over the box
                const float eLo = v - q_at(hi4, j);  // smallest
                keep[j] = inb && eHi <= bx.mlo;
                const bool uj = inb && !keep[j] && !(eLo > bx.mhi);
                und |= uj;
                fill |= (keep[j] || uj) && q_at(lg4, j) == 0.0f && v != 0.0f;
                if (inb) {
                    vals[buf][r][lane * 4 + j] = keep[j] ? v : 0.0f;
                    valsK[buf][r][lane * 4 + j] = (keep[j] || uj) ? v : 0.0f;
                }
                if (uj) {
                    // rare: with two or more the bin and value are not used
                    atomicAdd(&undCnt[r], 1);
                    undBinS[r] = i + j;
                    undValS[r] = v;
                }
            }
            for (int b = bandLo; b <= bandHi; ++b) {
                int c = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) c += __popcll(__ballot(keep[j] && band[j] == b));
                if (lane == 0) cnt[r][b] += c;
            }
            const bool anyUnd = __ballot(und) != 0, anyFill = __ballot(fill) != 0;
            if (lane == 0 && (anyUnd || anyFill)) rowFlags[r] |= (anyUnd ? 1 : 0) | (anyFill ? 2 : 0);
        }
    };

    // lanes 0..7 of wave 0 fold the rows as judged, lanes 8..15 the same
    // rows with the undecided channels kept (one fold instruction stream)
    float sum = 0.0f;
    const int foldRow = lane & (kPreRows - 1);
    const bool folder = wave == 0 && lane < 2 * kPreRows && foldRow < nRows;
    issue(0);
    judge(0, 0);
    __syncthreads();
    for (int tile = 0; tile < nTiles; ++tile) {
        const int buf = tile & 1;
        if (tile + 1 < nTiles) issue(tile + 1);
        if (folder) {
            const int nb = min(kTileBins, nBins - tile * kTileBins);
            sum = fold_sum(lane < kPreRows ? &vals[buf][foldRow][0] : &valsK[buf][foldRow][0], nb, sum);
        }
        if (tile + 1 < nTiles) judge(tile + 1, buf ^ 1);
        __syncthreads();
    }
    if (folder && lane < kPreRows) {
        int c[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) c[b] = cnt[lane][b];
        int total, badBands;
        float fracInc;
        band_stats(c, nBins, cpb, total, badBands, fracInc);
        PreRec r;
        r.mean = sum / (float)(unsigned)total;
        r.fracInc = fracInc;
        r.badBands = badBands;
        r.flags = rowFlags[lane];
        pre[k0 + lane].mean = r.mean;
        pre[k0 + lane].fracInc = r.fracInc;
        pre[k0 + lane].badBands = r.badBands;
        if (!(r.flags & PRE_UND) || undCnt[lane] != 1) pre[k0 + lane].flags = r.flags;
    } else if (folder && undCnt[foldRow] == 1) {
        // the record with the one undecided channel kept, and the decision
        // it gets at the sub-chunk-start state (steady state: dataModel 0,
        // margin crFactor * rmsRunAve, RFI_Clipper.cpp:305-311)
        const int row = foldRow, u = undBinS[row];
        int c[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) c[b] = cnt[row][b] + (band_of(u, invCpb) == b ? 1 : 0);
        int total, badBands;
        float fracInc;
        band_stats(c, nBins, cpb, total, badBands, fracInc);
        const float v = undValS[row];
        const float bpu = bp_value(bs.terms, K.nTerms, nBins, u, S->params0, K.scale, bs.kill);
        const bool predKept = !((v - 0.0f) - bpu > K.crFactor * S->rmsRunAve);
        PreRec& r = pre[k0 + row];
        r.meanK = sum / (float)(unsigned)total;
        r.fracIncK = fracInc;
        r.badBandsK = badBands;
        r.undBin = u;
        r.undVal = v;
        r.flags = rowFlags[row] | PRE_ONE | (predKept ? PRE_PREDKEPT : 0);
    }
}

// Learning-phase records (LearnRec) of the sub-chunk's spectra, one row per
// lane of wave 0 in bin order, tiles staged in LDS as in the precompute.
// Runs only while the clipper is learning at the sub-chunk's start.
__device__ __forceinline__ void learn_rows(const BatchIO& io)
{
    const int beam = blockIdx.z;
    const BeamState& bs = *io.st[beam];
    const float* __restrict__ I = io.I[beam];
    const ClipConst K = bs.K;
    const ClipScalars* __restrict__ S = bs.S;
    LearnRec* __restrict__ lr = bs.learn;
    const int nT = io.nT[beam], subLen = io.subLen[beam], ldMode = io.ldMode[beam];
    const int tStart = S->tNext;
    if (tStart >= nT || S->rmsSize >= K.cap) return;
    const int k0 = blockIdx.x * kPreRows;
    const int t0 = tStart + k0;
    const int nRows = min(min(kPreRows, nT - t0), subLen - k0);
    if (nRows <= 0) return;
    __shared__ float vals[kPreRows][kTileBins + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nBins = K.nBins, nCh = K.nCh, nPol = K.nPol;
    const int cpb = nBins / 8;
    const long rowStride = (long)nBins * nPol;
    const int nTiles = (nBins + kTileBins - 1) / kTileBins;
    const bool rowLane = wave == 0 && lane < nRows;
    LearnRec* out = lr + k0 + lane;
    float cmin = 1e6f, csum = 0.0f, csq = 0.0f, full = 0.0f, pred = 0.0f, thr = 0.0f;
    double ds = 0.0, dq = 0.0;
    int band = 0, bandEnd = cpb, kept = 0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int tile = 0; tile < nTiles; ++tile) {
            const int tb0 = tile * kTileBins;
            const int i = tb0 + lane * 4;
            if (i < nBins) {
#pragma unroll
                for (int q = 0; q < kPreRowsPerWave; ++q) {
                    const int r = wave + 4 * q;
                    if (r < nRows) {
                        const float4 v = load_quad(I, (long)(t0 + r) * rowStride, i, nCh, nPol, ldMode);
                        vals[r][lane * 4 + 0] = v.x;
                        vals[r][lane * 4 + 1] = v.y;
                        vals[r][lane * 4 + 2] = v.z;
                        vals[r][lane * 4 + 3] = v.w;
                    }
                }
            }
            __syncthreads();
            if (rowLane) {
                const int nb = min(kTileBins, nBins - tb0);
                if (pass == 0) {
                    for (int k = 0; k < nb; ++k) {
                        const float v = vals[lane][k];
                        cmin = (v < cmin) ? v : cmin;
                        csum += v;
                        csq += (v * v);
                        full += v;
                        ds += (double)v;
                        dq += (double)v * (double)v;
                        if (tb0 + k + 1 == bandEnd) {
                            out->bmin[band] = cmin;
                            out->bsum[band] = csum;
                            out->bsq[band] = csq;
                            cmin = 1e6f;
                            csum = 0.0f;
                            csq = 0.0f;
                            ++band;
                            bandEnd += cpb;
                        }
                    }
                } else {
                    for (int k = 0; k < nb; ++k) {
                        const float v = vals[lane][k];
                        const bool keep = v <= thr;
                        pred += keep ? v : 0.0f;
                        kept += keep ? 1 : 0;
                    }
                }
            }
            __syncthreads();
        }
        if (rowLane && pass == 0) {
            const double m = ds / nBins;
            const double var = dq / nBins - m * m;
            thr = (float)(m + 10.0 * sqrt(var > 0.0 ? var : 0.0));
        }
    }
    if (rowLane) {
        out->fullSum = full;
        out->predSum = pred;
        out->thr = thr;
        out->predKept = kept;
    }
}

// Staged per spectrum of the walk: the precomputed record, its mean and the
// two ring fronts, each divided by the (full) history length.
struct Staged {
    float mean, fracInc;
    int badBands, flags;
    float meanC, meanFC, rmsFC, pad;
};
// Staged for a spectrum with one undecided channel (PRE_ONE): the record
// with that channel kept, the channel, its value and its bandpass terms.
struct VarRec {
    float meanK, fracIncK;
    int badBandsK, undBin;
    float undVal;
    int killed;
    double terms[kMaxParams];
};
// The undecided channel's decision at a state (RFI_Clipper.cpp:319-322 with
// dataModel 0): kept unless (I - 0) - bandpass > margin.
__device__ __forceinline__ bool var_kept(const VarRec& v, int nTerms, float p0, float scale, float margin)
{
    float bp = 0.0f;
    if (!v.killed) {
        float tot = 0.0f;
        tot = (float)((double)tot + (double)p0 * 1.0);
        for (int j = 0; j < nTerms; ++j) tot = (float)((double)tot + v.terms[j]);
        bp = scale * tot;
    }
    return !((v.undVal - 0.0f) - bp > margin);
}

// A sub-chunk's records ahead of its walk: the precompute in the steady
// state (a valid box), the learning records while the rms ring fills; the
// two conditions exclude each other, so one of the bodies works.
__global__ __launch_bounds__(kTileThreads) void clip_prep_kernel(BatchIO io)
{
    pre_rows(io);
    learn_rows(io);
}

// The sequential walk over one sub-chunk of spectra.  Thread 0 runs ahead
// through spectra the precompute decided; the workgroup joins in for the
// others (the learning phase, undecided channels, a state outside the box,
// a last-good spectrum still to fill).  The walk starts at S->tNext and
// covers subLen spectra (all that remain when lastIter is set); a state
// that leaves the box ends it after that spectrum, so that the next
// sub-chunk speculates around the new state.
__global__ __launch_bounds__(kWalkThreads) void clip_walk_kernel(BatchIO io, int lastIter)
{
    const int beam = blockIdx.z;
    const BeamState& bst = *io.st[beam];
    float* __restrict__ I = io.I[beam];
    const ClipConst K = bst.K;
    ClipScalars* __restrict__ S = bst.S;
    float* __restrict__ meanBuf = bst.meanBuf;
    double* __restrict__ meanPrefix = bst.meanPrefix;
    float* __restrict__ rmsBuf = bst.rmsBuf;
    double* __restrict__ rmsPrefix = bst.rmsPrefix;
    float* __restrict__ lastGood = bst.lastGood;
    float* __restrict__ lastGoodVer = bst.lastGoodVer;
    const double* __restrict__ terms = bst.terms;
    const unsigned* __restrict__ kill = bst.kill;
    const Box* __restrict__ box = bst.box;
    const PreRec* __restrict__ pre = bst.pre;
    SpecRec* __restrict__ rec = bst.rec;
    const float* __restrict__ bpLo = bst.bpLo;
    const LearnRec* __restrict__ learn = bst.learn;
    const int nT = io.nT[beam], subLen = io.subLen[beam];
    __shared__ ClipScalars s;
    __shared__ float red[kWalkThreads / 64 + 1];
    __shared__ int goodCnt[8];
    __shared__ int fillCnt;
    __shared__ float sh_dataModel, sh_margin;
    __shared__ int sh_mode, sh_t, sh_fill, sh_frontsValid, sh_stop, sh_span, sh_why;
    // slow path: kept-bin mask, sums, and how the kept mean is obtained
    __shared__ unsigned long long keepMask[kMaxBinsWalk / 64];
    __shared__ double redD[2][kWalkThreads / 64];
    __shared__ int sh_mism, sh_mismL, sh_how, sh_total, sh_badBands;
    __shared__ float sh_fracInc, sh_meanArg;
    __shared__ float bandMin[8], modelMin[8];
    __shared__ float bandMean[8], bandMeanSq[8];
    __shared__ Staged stg[kStage];
    // the fold inputs of the staged spectra again, one array per field, so
    // that the single-lane folds read four spectra per LDS access
    __shared__ __align__(16) float soaMeanFC[kStage], soaMeanC[kStage], soaRmsFC[kStage], soaFracInc[kStage];
    // the speculative block's records: the precomputed mean (predicted
    // variant for one undecided channel) and whether its mask rejects it
    __shared__ __align__(16) float soaMean[kStage];
    __shared__ __align__(16) int soaRej[kStage];
    __shared__ int lastKept[kStage];  // last spectrum before k the mask keeps, or -1
    __shared__ VarRec var[kStage];
    // first spectrum at or after k the fast run must not take: undecided
    // channels (blkU), or undecided channels or a last-good fill (blkUF)
    __shared__ int blkU[kStage], blkUF[kStage];
    __shared__ __align__(16) float rowChunk[kRowChunk];

    const int tid = threadIdx.x;
    const int nBins = K.nBins, nCh = K.nCh, nSub = K.nSub, nPol = K.nPol;
    const int cpb = nBins / 8;
    RingView meanR{meanBuf, meanPrefix, K.cap}, rmsR{rmsBuf, rmsPrefix, K.cap};
    if (tid == 0) {
        s = *S;
        sh_frontsValid = 0;
        sh_stop = 0;
    }
    __syncthreads();
    const int tStart = s.tNext;
    if (tStart >= nT) return;
    const int tEnd = lastIter ? nT : min(nT, tStart + subLen);
    const int preN = min(subLen, nT - tStart);  // spectra the precompute covered
    const bool learnRun = s.rmsSize < K.cap;    // the learning records cover this sub-chunk
    const Box bx = *box;
    const float capf = (float)K.cap;

    auto idx = [&](int sb, int pol, int t) -> long { return (long)nCh * (nPol * ((long)nSub * t + sb)) + (long)nCh * pol; };

    int stageBase = tStart, stageN = 0;
    // Ring fronts of the staged spectra, divided by the (full) ring size:
    // with both rings full, the front of spectrum stageBase+k is entry
    // start+k (each spectrum evicts one front; a rejection rewrites only the
    // newest entry).
    auto rmsFrontDiv = [&](int t) {
        return sh_frontsValid ? stg[t - stageBase].rmsFC : ring_front(rmsR, s.rmsStart) / (float)s.rmsSize;
    };
    auto meanFrontDiv = [&](int t) {
        return sh_frontsValid ? stg[t - stageBase].meanFC : ring_front(meanR, s.meanStart) / (float)s.meanSize;
    };
    // Steady-state rms update and margin (RFI_Clipper.cpp:288-311), thread 0.
    auto phaseA_steady = [&](int t, float& dataModel, float& margin) {
        s.rmsRunAve -= rmsFrontDiv(t);
        s.rmsRunAve += (s.rmsBack / (float)K.cap);
        if (K.useMeanOverRMS) s.spectrumRMS = fabsf(s.meanRunAve / s.meanOverRMS);
        ring_push(rmsR, s.rmsStart, s.rmsSize, s.rmsBack, s.rmsBack2, s.spectrumRMS);
        dataModel = 0.0f;
        margin = K.crFactor * s.rmsRunAve;
    };

    // Decision and bookkeeping (:345-565) for spectrum t, thread 0, from the
    // kept-channel mean and band statistics.  Returns whether the last-good
    // spectrum must be filled from this spectrum.
    auto phaseC = [&](int t, float spectrumSum, float fracInc, int badBands, float dataModel, float margin,
                      int replaced) -> int {
        s.fractionBadChannels += fracInc;
        if (s.meanSize < K.cap) {
            ring_push(meanR, s.meanStart, s.meanSize, s.meanBack, s.meanBack2, spectrumSum);
            s.meanRunAve = (float)(meanPrefix[s.meanSize] / (double)s.meanSize);
        } else {
            s.meanRunAve -= meanFrontDiv(t);
            s.meanRunAve += s.meanBack / (float)s.meanSize;
            ring_push(meanR, s.meanStart, s.meanSize, s.meanBack, s.meanBack2, spectrumSum);
        }
        const float tol = (float)((double)(K.srFactor * s.bpRms) / K.sqrtNBins);
        int good = 1;
        if (s.meanSize < 1000) {
            good = 0;
        } else if (spectrumSum - s.meanRunAve > tol || badBands >= 4) {
            good = 0;
            ++s.badSpectra;
            ring_repeat_previous(meanR, s.meanStart, s.meanSize, s.meanBack, s.meanBack2);
            ring_repeat_previous(rmsR, s.rmsStart, s.rmsSize, s.rmsBack, s.rmsBack2);
        }
        SpecRec r;
        r.params0 = s.params0;
        r.dataModel = dataModel;
        r.margin = margin;
        r.flags = (good ? 1 : 0) | (replaced ? 2 : 0);
        r.ver = s.ver;
        const unsigned reportStatsEvery = 10u * (unsigned)K.maxHistory;
        if (s.num == 0) {
            const float fractionBadSpectra =
                (float)(100.0 * (double)(float)s.badSpectra / (double)(float)reportStatsEvery);
            if (fractionBadSpectra > 99.0) {
                s.rmsSize = 0;
                s.meanSize = 0;
                s.rmsStart = 0;
                s.meanStart = 0;
            }
            s.badSpectra = 0;
            s.fractionBadChannels = 0.0f;
        }
        r.meanRunAve = s.meanRunAve;
        rec[t] = r;
        return good && s.remainingZeros != 0;
    };
    // setMean (BandPass.cpp:63-74) and setRMS (:44-46) at the end of a spectrum.
    auto phaseD = [&]() {
        const float delta = s.meanRunAve - s.bpMean;
        if (fabsf(delta) > 0.0f) {
            s.params0 += delta / K.scale;
            s.bpMean = s.meanRunAve;
        }
        s.bpRms = s.rmsRunAve;
        s.num = (int)(((unsigned)s.num + 1u) % (10u * (unsigned)K.maxHistory));
    };

    // Why spectrum t (after its rms update) cannot take the precomputed mask,
    // or -1; sets sh_stop for a state that left the box.
    auto slow_reason = [&](int t, float margin) -> int {
        const int k = t - stageBase;
        const bool covered = t - tStart < preN;
        const bool inBox = bx.valid && covered && s.params0 >= bx.p0lo && s.params0 <= bx.p0hi &&
                           margin >= bx.mlo && margin <= bx.mhi;
        if (!inBox) {
            // leave the sub-chunk to re-centre the box (or to get one, when
            // the sub-chunk started in the learning phase)
            if (covered && !lastIter) sh_stop = bx.valid ? 1 : 2;
            return CNT_BOX;
        }
        // a kept spectrum fills empty last-good slots (:446-469); the
        // precompute flagged candidates against the sub-chunk-start slots,
        // a superset of the empty slots now
        if (stg[k].flags & 1) return CNT_UNDECIDED;
        if (s.remainingZeros != 0 && (stg[k].flags & 2)) return CNT_FILL;
        if (s.ver >= kMaxVer) return CNT_VER;
        return -1;
    };

    const int period = 10 * K.maxHistory;
    // ---- steady-state fast run (thread 0): spectra t, t+1, ... that the
    // precompute decided, with the scalars in registers.  Stops before the
    // first spectrum that is blocked (undecided channels, last-good fill,
    // outside the precomputed range) or at the end of the stage, or after
    // the rms update of a spectrum whose state left the box (FAST_BOX, the
    // margin in sh_margin), or after a statistics reset (FAST_RESET).
    // Division by a divisor d fixed over the run (history length, mean/rms
    // ratio, bandpass scale) is taken as RN_f32((double)x * RN_f64(1/d)):
    // for float x and a normal float d the exact quotient is never within
    // 2^-49 (relative) of a float rounding midpoint, and the double product
    // is within 2^-52 of it, so both round to the same float.  Results in
    // the subnormal range, where that bound does not hold, take the
    // division itself.  The spectrum tolerance RN_f32(RN_f64(a / sqrtN))
    // lies within one float step of RN_f32(RN_f64(a * RN_f64(1/sqrtN))),
    // so the division is needed only when the comparison falls in that step.
    enum { FAST_LIMIT = 0, FAST_BOX = 1, FAST_RESET = 2, FAST_MAXN = 3 };
    const double rcpCap = 1.0 / (double)capf;
    const double rcpScale = 1.0 / (double)K.scale;
    const double rcpSqrtN = 1.0 / K.sqrtNBins;
    const bool fastOK = K.cap >= 1000 && K.cap > kStage && isfinite(K.scale) && fabsf(K.scale) >= 0x1p-126f;
    auto fast_run = [&](int& t, int maxN) -> int {
        const int k0 = t - stageBase;
        const bool zerosLeft = s.remainingZeros != 0;
        const int kBlk = min(min(stageN, tEnd - stageBase),
                             min(preN - (stageBase - tStart), zerosLeft ? blkUF[k0] : blkU[k0]));
        const int kLim = min(kBlk, k0 + maxN);
        if (k0 >= kLim) return FAST_LIMIT;
        const float mOR = s.meanOverRMS;
        const bool useMOR = K.useMeanOverRMS != 0;
        if (useMOR && !(isfinite(mOR) && fabsf(mOR) >= 0x1p-126f)) return FAST_LIMIT;
        const double rcpMOR = 1.0 / (double)mOR;
        const int cap = K.cap;
        float mRA = s.meanRunAve, rA = s.rmsRunAve, p0 = s.params0, bpMean = s.bpMean, bpRms = s.bpRms;
        float sRMS = s.spectrumRMS, fbc = s.fractionBadChannels;
        float mBack = s.meanBack, rBack = s.rmsBack, mBack2 = s.meanBack2, rBack2 = s.rmsBack2;
        float mBackC = mBack / capf, rBackC = rBack / capf;
        int badSpectra = s.badSpectra, num = s.num, meanStart = s.meanStart, rmsStart = s.rmsStart;
        int status = FAST_LIMIT;
        int k = k0;
        int nResolved = 0;
        Staged cur = stg[k0];
        while (k < kLim) {
            const Staged nxt = stg[min(k + 1, kStage - 1)];  // consumed next iteration
            // phaseA (:288-311)
            const float rAn = (rA - cur.rmsFC) + rBackC;
            float sRMSn = sRMS;
            if (useMOR) sRMSn = fabsf((float)((double)mRA * rcpMOR));
            float rBCn = (float)((double)sRMSn * rcpCap);
            if (!(fabsf(rBCn) >= 0x1p-125f) || (useMOR && !(sRMSn >= 0x1p-125f))) {
                if (useMOR) sRMSn = fabsf(mRA / mOR);
                rBCn = sRMSn / capf;
            }
            const float margin = K.crFactor * rAn;
            if (!(p0 >= bx.p0lo && p0 <= bx.p0hi && margin >= bx.mlo && margin <= bx.mhi)) {
                // commit the rms update only; the workgroup takes it from here
                rmsBuf[rmsStart] = sRMSn;
                rmsStart = rmsStart + 1 == cap ? 0 : rmsStart + 1;
                rBack2 = rBack;
                rBack = sRMSn;
                rA = rAn;
                sRMS = sRMSn;
                sh_margin = margin;
                status = FAST_BOX;
                break;
            }
            // one undecided channel: its decision at this state picks the
            // precomputed record (:319-342)
            float cMean = cur.mean, cMeanC = cur.meanC, cFrac = cur.fracInc;
            bool cRej = (cur.flags & PRE_REJ) != 0;
            if (cur.flags & PRE_ONE) {
                const VarRec& v = var[k];
                if (var_kept(v, K.nTerms, p0, K.scale, margin)) {
                    cMean = v.meanK;
                    cMeanC = v.meanK / capf;
                    cFrac = v.fracIncK;
                    cRej = v.badBandsK >= 4;
                }
                ++nResolved;
            }
            // phaseC (:345-440)
            const float mRAn = (mRA - cur.meanFC) + mBackC;
            const float x = cMean - mRAn;
            const float a32 = K.srFactor * bpRms;
            const float tolA = (float)((double)a32 * rcpSqrtN);
            bool gt;
            if (tolA >= 0x1p-126f && tolA < 0x1p127f) {
                const bool gtHi = x > __int_as_float(__float_as_int(tolA) + 1);
                const bool gtLo = x > __int_as_float(__float_as_int(tolA) - 1);
                gt = gtHi;
                if (gtLo != gtHi) gt = x > (float)((double)a32 / K.sqrtNBins);
            } else {
                gt = x > (float)((double)a32 / K.sqrtNBins);
            }
            const bool good = !(gt || cRej);
            // push, or push then pop_back + push_back(back()): the slot ends up
            // with the new value or with the previous back
            rmsBuf[rmsStart] = good ? sRMSn : rBack;
            meanBuf[meanStart] = good ? cMean : mBack;
            rmsStart = rmsStart + 1 == cap ? 0 : rmsStart + 1;
            meanStart = meanStart + 1 == cap ? 0 : meanStart + 1;
            rBack2 = rBack;
            mBack2 = mBack;
            if (good) {
                rBack = sRMSn;
                rBackC = rBCn;
                mBack = cMean;
                mBackC = cMeanC;
            }
            badSpectra += good ? 0 : 1;
            fbc += cFrac;
            SpecRec rr;
            rr.params0 = p0;
            rr.dataModel = 0.0f;
            rr.margin = margin;
            rr.meanRunAve = mRAn;
            rr.flags = good ? 1 : 0;
            rr.ver = s.ver;
            rec[t] = rr;
            int reset = 0;
            if (num == 0) {
                const float fractionBadSpectra =
                    (float)(100.0 * (double)(float)badSpectra / (double)(float)(unsigned)period);
                if (fractionBadSpectra > 99.0) reset = 1;
                badSpectra = 0;
                fbc = 0.0f;
            }
            // phaseD
            const float delta = mRAn - bpMean;
            if (fabsf(delta) > 0.0f) {
                float qd = (float)((double)delta * rcpScale);
                if (!(fabsf(qd) >= 0x1p-125f)) qd = delta / K.scale;
                p0 += qd;
                bpMean = mRAn;
            }
            bpRms = rAn;
            num = num + 1 == period ? 0 : num + 1;
            rA = rAn;
            mRA = mRAn;
            sRMS = sRMSn;
            ++k;
            ++t;
            cur = nxt;
            if (reset) {
                status = FAST_RESET;
                break;
            }
        }
        if (status == FAST_LIMIT && kLim < kBlk) status = FAST_MAXN;
        s.count[CNT_FAST] += (t - stageBase) - k0;
        s.count[CNT_RESOLVED] += nResolved;
        s.meanRunAve = mRA;
        s.rmsRunAve = rA;
        s.params0 = p0;
        s.bpMean = bpMean;
        s.bpRms = bpRms;
        s.spectrumRMS = sRMS;
        s.fractionBadChannels = fbc;
        s.badSpectra = badSpectra;
        s.num = num;
        s.meanStart = meanStart;
        s.rmsStart = rmsStart;
        s.meanBack = mBack;
        s.rmsBack = rBack;
        s.meanBack2 = mBack2;
        s.rmsBack2 = rBack2;
        if (status == FAST_RESET) {
            s.rmsSize = 0;
            s.meanSize = 0;
            s.rmsStart = 0;
            s.meanStart = 0;
        }
        return status;
    };

    // ---- speculative steady-state block (whole workgroup).  Assumes every
    // spectrum from t on is kept and inside the box.  The four recurrences
    // (running mean, running rms, bandpass offset, bad-channel fraction) are
    // then folds over values known in advance; each is run by one lane with
    // exactly the operations of the scalar run, in the same order, so every
    // value is bit-identical to it.  The decisions are then checked in
    // parallel and the block is committed up to the first spectrum that is
    // not kept or leaves the box; that spectrum goes to the scalar run.
    // Returns the spectra committed.
    auto spec_span = [&](int t) -> int {
        const int k0 = t - stageBase;
        const bool zerosLeft = s.remainingZeros != 0;
        int kLim = min(min(stageN, tEnd - stageBase),
                       min(preN - (stageBase - tStart), zerosLeft ? blkUF[k0] : blkU[k0]));
        const float mOR = s.meanOverRMS;
        if (K.useMeanOverRMS && !(isfinite(mOR) && fabsf(mOR) >= 0x1p-126f)) return 0;
        if (s.num == 0) return 0;  // the statistics-reset check falls on spectrum t
        kLim = min(kLim, k0 + (period - s.num));
        return max(0, kLim - k0);
    };
    float* const spM = rowChunk;                                   // running mean after spectrum k
    float* const spR = rowChunk + kStage;                          // running rms after k
    float* const spS = rowChunk + 2 * kStage;                      // spectrumRMS pushed by k
    float* const spB = rowChunk + 3 * kStage;                      // that value / history length
    float* const spQ = rowChunk + 4 * kStage;                      // offset step of k
    int* const spD = reinterpret_cast<int*>(rowChunk + 5 * kStage); // k moves the offset
    float* const spP = rowChunk + 6 * kStage;                      // offset before k
    float* const spMBC = rowChunk + 7 * kStage;                    // newest mean entry before k / history
    float* const spRBC = rowChunk + 8 * kStage;                    // newest rms entry before k / history
    float* const spF = rowChunk + 9 * kStage;                      // bad-channel fraction after k
    float* const spMB = rowChunk + 10 * kStage;                    // newest mean-ring entry before k
    float* const spRB = rowChunk + 11 * kStage;                    // newest rms-ring entry before k
    static_assert(12 * kStage <= kRowChunk, "speculation arrays alias the row chunk");
    __shared__ int sh_fail, sh_nrej, sh_nres, sh_still;
    __shared__ float sh_pEnd;
    // spectrumRMS of a steady spectrum whose running mean before it is mPrev
    // (:288-311), and that value over the history length, as the scalar run
    // takes them
    auto spec_rms = [&](float mPrev, bool useMOR, float mOR, double rcpMOR, float& sR, float& rb) {
        sR = s.spectrumRMS;
        if (useMOR) sR = fabsf((float)((double)mPrev * rcpMOR));
        rb = (float)((double)sR * rcpCap);
        if (!(fabsf(rb) >= 0x1p-125f) || (useMOR && !(sR >= 0x1p-125f))) {
            if (useMOR) sR = fabsf(mPrev / mOR);
            rb = sR / capf;
        }
    };
    // ---- speculative steady-state block (whole workgroup).  Assumes that
    // every spectrum from t on stays inside the box, that each one with a
    // single undecided channel gets the decision it had at the stage-start
    // state, and that a spectrum is rejected exactly when its (so predicted)
    // mask has 4 bad bands or more.  The four recurrences (running mean,
    // running rms, bandpass offset, bad-channel fraction) are then folds over
    // values known in advance -- a rejection keeps the newest ring entry
    // (:432-435), so the entry each spectrum's running averages add is that
    // of the last kept spectrum before it; each fold is run by one lane with
    // exactly the operations of the scalar run, in the same order, so every
    // value is bit-identical to it.  The assumptions are then checked in
    // parallel and the block is committed up to the first spectrum that
    // breaks one; that spectrum goes to the scalar run.  Returns the spectra
    // committed.
    auto spec_block = [&](int t, int span) -> int {
        const int k0 = t - stageBase, kEnd = k0 + span;
        const bool useMOR = K.useMeanOverRMS != 0;
        const float mOR = s.meanOverRMS;
        const double rcpMOR = 1.0 / (double)mOR;
        // The folds read their inputs sixteen spectra at a time, ahead of the
        // dependent adds, so that LDS latency is paid once per sixteen; from
        // the first multiple of four on, four spectra per LDS access.
        constexpr int B = 16;
        auto ld4 = [](const float* a, int k) { return *reinterpret_cast<const float4*>(a + k); };
        auto ld4i = [](const int* a, int k) { return *reinterpret_cast<const int4*>(a + k); };
        auto st4 = [](float* a, int k, float x, float y, float z, float w) {
            *reinterpret_cast<float4*>(a + k) = make_float4(x, y, z, w);
        };
        // the mean-ring entry each spectrum's update adds: the last kept
        // spectrum's mean before it, or the state's newest entry
        if (tid == 0) sh_still = 0;
        {
            const int k = tid;
            if (k >= k0 && k < kEnd) {
                const int j = lastKept[k];
                spMB[k] = j >= k0 ? soaMean[j] : s.meanBack;
                spMBC[k] = j >= k0 ? soaMeanC[j] : s.meanBack / capf;
            }
        }
        __syncthreads();
        // a two-lane fold: (v - front/cap) + back/cap per spectrum
        auto fold2 = [&](float v, const float* fc, const float* bc, float* out) {
            int k = k0;
            for (; k < kEnd && (k & 3); ++k) {
                v = (v - fc[k]) + bc[k];
                out[k] = v;
            }
            for (; k + B <= kEnd; k += B) {
                float a[B], c[B];
#pragma unroll
                for (int j = 0; j < B; j += 4) {
                    const float4 x = ld4(fc, k + j), y = ld4(bc, k + j);
                    a[j] = x.x; a[j + 1] = x.y; a[j + 2] = x.z; a[j + 3] = x.w;
                    c[j] = y.x; c[j + 1] = y.y; c[j + 2] = y.z; c[j + 3] = y.w;
                }
#pragma unroll
                for (int j = 0; j < B; ++j) {
                    v = (v - a[j]) + c[j];
                    a[j] = v;
                }
#pragma unroll
                for (int j = 0; j < B; j += 4) st4(out, k + j, a[j], a[j + 1], a[j + 2], a[j + 3]);
            }
            for (; k < kEnd; ++k) {
                v = (v - fc[k]) + bc[k];
                out[k] = v;
            }
        };
        if (tid == 0) {
            fold2(s.meanRunAve, soaMeanFC, spMBC, spM);
        } else if (tid == 64) {
            float f = s.fractionBadChannels;
            int k = k0;
            for (; k < kEnd && (k & 3); ++k) {
                f += soaFracInc[k];
                spF[k] = f;
            }
            for (; k + B <= kEnd; k += B) {
                float fi[B];
#pragma unroll
                for (int j = 0; j < B; j += 4) {
                    const float4 a = ld4(soaFracInc, k + j);
                    fi[j] = a.x; fi[j + 1] = a.y; fi[j + 2] = a.z; fi[j + 3] = a.w;
                }
#pragma unroll
                for (int j = 0; j < B; ++j) {
                    f += fi[j];
                    fi[j] = f;
                }
#pragma unroll
                for (int j = 0; j < B; j += 4) st4(spF, k + j, fi[j], fi[j + 1], fi[j + 2], fi[j + 3]);
            }
            for (; k < kEnd; ++k) {
                f += soaFracInc[k];
                spF[k] = f;
            }
        }
        __syncthreads();
        {
            const int k = tid;
            if (k >= k0 && k < kEnd) {
                float sR, rb;
                spec_rms(k == k0 ? s.meanRunAve : spM[k - 1], useMOR, mOR, rcpMOR, sR, rb);
                spS[k] = sR;
                spB[k] = rb;
                // the rms-ring entry this spectrum's update adds: the last
                // kept spectrum's value before it (recomputed here, the
                // same operations as its own thread's), or the state's
                const int j = lastKept[k];
                if (j >= k0) {
                    float sRj, rbj;
                    spec_rms(j == k0 ? s.meanRunAve : spM[j - 1], useMOR, mOR, rcpMOR, sRj, rbj);
                    spRB[k] = sRj;
                    spRBC[k] = rbj;
                } else {
                    spRB[k] = s.rmsBack;
                    spRBC[k] = s.rmsBack / capf;
                }
                // the bandpass mean before k is the running mean after k-1 (an
                // unchanged one differs from it at most in the sign of a zero,
                // which neither the test nor the step below can see)
                const float bmPrev = k == k0 ? s.bpMean : spM[k - 1];
                const float delta = spM[k] - bmPrev;
                int d = 0;
                float qd = 0.0f;
                if (fabsf(delta) > 0.0f) {
                    d = 1;
                    qd = (float)((double)delta * rcpScale);
                    if (!(fabsf(qd) >= 0x1p-125f)) qd = delta / K.scale;
                }
                spQ[k] = qd;
                spD[k] = d;
                if (!d) sh_still = 1;
            }
        }
        __syncthreads();
        if (tid == 0) {
            fold2(s.rmsRunAve, soaRmsFC, spRBC, spR);
        } else if (tid == 64 && !sh_still) {
            // the bandpass offset when every spectrum moves it: p + q
            float p = s.params0;
            int k = k0;
            for (; k < kEnd && (k & 3); ++k) {
                spP[k] = p;
                p += spQ[k];
            }
            for (; k + B <= kEnd; k += B) {
                float q[B];
#pragma unroll
                for (int j = 0; j < B; j += 4) {
                    const float4 a = ld4(spQ, k + j);
                    q[j] = a.x; q[j + 1] = a.y; q[j + 2] = a.z; q[j + 3] = a.w;
                }
#pragma unroll
                for (int j = 0; j < B; ++j) {
                    const float pn = p + q[j];
                    q[j] = p;
                    p = pn;
                }
#pragma unroll
                for (int j = 0; j < B; j += 4) st4(spP, k + j, q[j], q[j + 1], q[j + 2], q[j + 3]);
            }
            for (; k < kEnd; ++k) {
                spP[k] = p;
                p += spQ[k];
            }
            sh_pEnd = p;
        } else if (tid == 64) {
            // the bandpass offset: p + q where the running mean moved
            float p = s.params0;
            const int* const spDi = spD;
            int k = k0;
            for (; k < kEnd && (k & 3); ++k) {
                spP[k] = p;
                if (spDi[k]) p += spQ[k];
            }
            for (; k + B <= kEnd; k += B) {
                float q[B], op[B];
                int d[B];
#pragma unroll
                for (int j = 0; j < B; j += 4) {
                    const float4 a = ld4(spQ, k + j);
                    const int4 e = ld4i(spDi, k + j);
                    q[j] = a.x; q[j + 1] = a.y; q[j + 2] = a.z; q[j + 3] = a.w;
                    d[j] = e.x; d[j + 1] = e.y; d[j + 2] = e.z; d[j + 3] = e.w;
                }
                // a select, not a branch: one add and one select per spectrum
#pragma unroll
                for (int j = 0; j < B; ++j) {
                    op[j] = p;
                    const float pn = p + q[j];
                    p = d[j] ? pn : p;
                }
#pragma unroll
                for (int j = 0; j < B; j += 4) st4(spP, k + j, op[j], op[j + 1], op[j + 2], op[j + 3]);
            }
            for (; k < kEnd; ++k) {
                spP[k] = p;
                if (spDi[k]) p += spQ[k];
            }
            sh_pEnd = p;
        } else if (tid == 128) {
            sh_fail = kEnd;
            sh_nrej = 0;
            sh_nres = 0;
        }
        __syncthreads();
        {
            const int k = tid;
            if (k >= k0 && k < kEnd) {
                const float margin = K.crFactor * spR[k];
                const float p0 = spP[k];
                bool ok = p0 >= bx.p0lo && p0 <= bx.p0hi && margin >= bx.mlo && margin <= bx.mhi;
                const int fl = stg[k].flags;
                if (fl & PRE_ONE)
                    ok = ok && var_kept(var[k], K.nTerms, p0, K.scale, margin) == ((fl & PRE_PREDKEPT) != 0);
                if (!soaRej[k] || (fl & PRE_TOLREJ)) {
                    // the tolerance test (:421-425) must keep the spectra
                    // predicted kept and reject those predicted rejected by it
                    const float x = soaMean[k] - spM[k];
                    const float a32 = K.srFactor * (k == k0 ? s.bpRms : spR[k - 1]);
                    const float tolA = (float)((double)a32 * rcpSqrtN);
                    bool gt;
                    if (tolA >= 0x1p-126f && tolA < 0x1p127f) {
                        const bool gtHi = x > __int_as_float(__float_as_int(tolA) + 1);
                        const bool gtLo = x > __int_as_float(__float_as_int(tolA) - 1);
                        gt = gtHi;
                        if (gtLo != gtHi) gt = x > (float)((double)a32 / K.sqrtNBins);
                    } else {
                        gt = x > (float)((double)a32 / K.sqrtNBins);
                    }
                    ok = ok && gt == ((fl & PRE_TOLREJ) != 0);
                }
                if (!ok) atomicMin(&sh_fail, k);
            }
        }
        __syncthreads();
        const int kf = sh_fail;
        {
            const int k = tid;
            if (k >= k0 && k < kf) {
                const bool rej = soaRej[k] != 0;
                SpecRec rr;
                rr.params0 = spP[k];
                rr.dataModel = 0.0f;
                rr.margin = K.crFactor * spR[k];
                rr.meanRunAve = spM[k];
                rr.flags = rej ? 0 : 1;
                rr.ver = s.ver;
                rec[stageBase + k] = rr;
                rmsBuf[(s.rmsStart + (k - k0)) % K.cap] = rej ? spRB[k] : spS[k];
                meanBuf[(s.meanStart + (k - k0)) % K.cap] = rej ? spMB[k] : soaMean[k];
                if (rej) atomicAdd(&sh_nrej, 1);
                if (stg[k].flags & PRE_ONE) atomicAdd(&sh_nres, 1);
            }
        }
        __syncthreads();
        const int n = kf - k0;
        if (tid == 0 && n > 0) {
            const int kl = kf - 1;
            const bool rejL = soaRej[kl] != 0;
            s.meanRunAve = spM[kl];
            s.rmsRunAve = spR[kl];
            s.params0 = kf < kEnd ? spP[kf] : sh_pEnd;
            // setMean moved the bandpass mean to the running mean at the
            // last spectrum whose running mean changed
            int jd = kl;
            while (jd >= k0 && !spD[jd]) --jd;
            if (jd >= k0) s.bpMean = spM[jd];
            s.bpRms = spR[kl];
            s.spectrumRMS = spS[kl];
            s.fractionBadChannels = spF[kl];
            s.num = (s.num + n) % period;
            s.meanStart = (s.meanStart + n) % K.cap;
            s.rmsStart = (s.rmsStart + n) % K.cap;
            s.meanBack2 = spMB[kl];
            s.meanBack = rejL ? spMB[kl] : soaMean[kl];
            s.rmsBack2 = spRB[kl];
            s.rmsBack = rejL ? spRB[kl] : spS[kl];
            s.badSpectra += sh_nrej;
            s.count[CNT_FAST] += n;
            s.count[CNT_RESOLVED] += sh_nres;
        }
        return n;
    };
    constexpr int kSpecMin = 24;  // shorter runs go to the scalar run

    enum { MODE_END = 0, MODE_STAGE = 1, MODE_FULL = 2, MODE_BC = 3, MODE_SPEC = 4 };
    int t = tStart;
    int scalarNext = 0;  // thread 0: spectra the scalar run takes before the next speculation
    for (;;) {
        if (tid == 0) {
            int mode = MODE_END;
            for (;;) {
                if (sh_stop || t >= tEnd) { mode = MODE_END; break; }
                if (t >= stageBase + stageN) { mode = MODE_STAGE; break; }
                if (s.rmsSize < K.cap) {  // learning phase
                    ++s.count[CNT_LEARN];
                    sh_why = CNT_LEARN;
                    mode = MODE_FULL;
                    break;
                }
                if (fastOK && bx.valid && sh_frontsValid && s.meanSize == K.cap && s.ver < kMaxVer) {
                    if (scalarNext == 0) {
                        const int span = spec_span(t);
                        if (span >= kSpecMin) {
                            sh_span = span;
                            mode = MODE_SPEC;
                            break;
                        }
                        // too short to speculate: the scalar run takes it and
                        // the spectrum that ends it, then speculation resumes
                        scalarNext = span + 1;
                    }
                    const int r = fast_run(t, scalarNext);
                    scalarNext = 0;
                    if (r == FAST_MAXN) continue;
                    if (r == FAST_BOX) {
                        // spectrum t had its rms update and left the box
                        const int why = slow_reason(t, sh_margin);
                        ++s.count[why < 0 ? CNT_BOX : why];
                        sh_why = why < 0 ? CNT_BOX : why;
                        sh_dataModel = 0.0f;
                        mode = MODE_BC;
                        break;
                    }
                    if (r == FAST_RESET || t >= tEnd || t >= stageBase + stageN) continue;
                    // otherwise spectrum t is blocked (undecided, fill, not
                    // precomputed): one spectrum by the general path below
                }
                float dataModel, margin;
                phaseA_steady(t, dataModel, margin);
                const int why = slow_reason(t, margin);
                if (why >= 0) {
                    ++s.count[why];
                    sh_why = why;
                    sh_dataModel = dataModel;
                    sh_margin = margin;
                    mode = MODE_BC;
                    break;
                }
                const int k = t - stageBase;
                phaseC(t, stg[k].mean, stg[k].fracInc, stg[k].badBands, dataModel, margin, 0);
                phaseD();
                ++s.count[CNT_FAST];
                ++t;
            }
            sh_mode = mode;
            sh_t = t;
        }
        __syncthreads();
        const int mode = sh_mode;
        t = sh_t;
        if (mode == MODE_END) break;
        if (mode == MODE_SPEC) {
            const int span = sh_span;
            const int n = spec_block(t, span);
            t += n;
            if (tid == 0) {
                if (n < span) scalarNext = 1;  // the spectrum that failed
                sh_t = t;
            }
            __syncthreads();
            continue;
        }
        if (mode == MODE_STAGE) {
            stageBase = t;
            stageN = min(kStage, tEnd - t);
            const int nPre = bx.valid ? min(stageN, preN - (t - tStart)) : 0;
            for (int k = tid; k < nPre; k += kWalkThreads) {
                const PreRec r = pre[t - tStart + k];
                const int fl = r.flags | (r.badBands >= 4 ? PRE_REJ : 0);
                stg[k].mean = r.mean;
                stg[k].fracInc = r.fracInc;
                stg[k].badBands = r.badBands;
                stg[k].flags = fl;
                stg[k].meanC = r.mean / capf;
                // the speculative block takes the record the one undecided
                // channel's decision at the stage-start state predicts
                bool useK = false;
                if (fl & PRE_ONE) {
                    VarRec& v = var[k];
                    v.meanK = r.meanK;
                    v.fracIncK = r.fracIncK;
                    v.badBandsK = r.badBandsK;
                    v.undBin = r.undBin;
                    v.undVal = r.undVal;
                    v.killed = bin_killed(kill, r.undBin) ? 1 : 0;
                    for (int j = 0; j < K.nTerms; ++j) v.terms[j] = terms[(size_t)j * nBins + r.undBin];
                    useK = var_kept(v, K.nTerms, s.params0, K.scale, K.crFactor * s.rmsRunAve);
                    stg[k].flags = useK ? (fl | PRE_PREDKEPT) : (fl & ~PRE_PREDKEPT);
                }
                soaMean[k] = useK ? r.meanK : r.mean;
                soaMeanC[k] = useK ? r.meanK / capf : stg[k].meanC;
                soaFracInc[k] = useK ? r.fracIncK : r.fracInc;
                // a spectrum the mask keeps is predicted rejected when its
                // mean exceeds the tolerance at the stage-start state
                // (:394-425): the speculative block then checks that it is
                const bool maskRej = (useK ? r.badBandsK : r.badBands) >= 4;
                const float tolS = (float)((double)(K.srFactor * s.bpRms) / K.sqrtNBins);
                const bool tolRej = !maskRej && soaMean[k] - s.meanRunAve > tolS;
                if (tolRej) stg[k].flags |= PRE_TOLREJ;
                soaRej[k] = (maskRej || tolRej) ? 1 : 0;
            }
            {
                // suffix minima over the stage (one thread per entry) of the
                // spectra neither the fast run nor a speculative block takes:
                // undecided channels other than one (blkU), or those or a
                // last-good fill (blkUF)
                const int k = tid;
                int bu = k, buf = k, lk = -1;
                if (k < nPre) {
                    const int fl = stg[k].flags;
                    const bool unres = (fl & PRE_UND) && !(fl & PRE_ONE);
                    bu = unres ? k : kStage;
                    buf = (unres || (fl & PRE_FILL)) ? k : kStage;
                    lk = soaRej[k] ? -1 : k;
                }
                blkU[k] = bu;
                blkUF[k] = buf;
                lastKept[k] = lk;
                __syncthreads();
                // and the prefix maxima of the kept indices (inclusive here,
                // shifted by one below)
                for (int off = 1; off < kStage; off <<= 1) {
                    if (k + off < kStage) {
                        bu = min(bu, blkU[k + off]);
                        buf = min(buf, blkUF[k + off]);
                    }
                    if (k >= off) lk = max(lk, lastKept[k - off]);
                    __syncthreads();
                    blkU[k] = bu;
                    blkUF[k] = buf;
                    lastKept[k] = lk;
                    __syncthreads();
                }
                lk = k > 0 ? lastKept[k - 1] : -1;
                __syncthreads();
                lastKept[k] = lk;
            }
            const int fv = s.rmsSize == K.cap && s.meanSize == K.cap && K.cap > kStage;
            if (fv) {
                for (int k = tid; k < stageN; k += kWalkThreads) {
                    stg[k].rmsFC = rmsBuf[(s.rmsStart + k) % K.cap] / capf;
                    stg[k].meanFC = meanBuf[(s.meanStart + k) % K.cap] / capf;
                    soaRmsFC[k] = stg[k].rmsFC;
                    soaMeanFC[k] = stg[k].meanFC;
                }
            }
            __syncthreads();
            if (tid == 0) sh_frontsValid = fv;
            __syncthreads();
            continue;
        }
        // ---- slow spectrum t, whole workgroup
        if (tid < 8) goodCnt[tid] = 0;
        if (tid == 0) fillCnt = 0;
        const bool useRec = mode == MODE_FULL && learnRun && (t - tStart) < preN;
        if (useRec) {
            // learning-phase band statistics (RFI_Clipper.cpp:214-242): the
            // data's from the record, the model's minimum per band by wave b
            const int b = tid >> 6;
            float mm = 1e6f;
            for (int i = b * cpb + (tid & 63); i < (b + 1) * cpb; i += 64) {
                const float bpv = bp_value(terms, K.nTerms, nBins, i, s.params0, K.scale, kill);
                mm = (bpv < mm) ? bpv : mm;
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const float o = __shfl_xor(mm, off, 64);
                mm = (o < mm) ? o : mm;
            }
            if ((tid & 63) == 0) {
                const LearnRec& R = learn[t - tStart];
                modelMin[b] = mm;
                bandMin[b] = R.bmin[b];
                bandMean[b] = R.bsum[b];
                bandMeanSq[b] = R.bsq[b];
            }
        }
        if (mode == MODE_FULL && !useRec) {
            // learning-phase band statistics (RFI_Clipper.cpp:214-242)
            for (int b = 0; b < 8; ++b) {
                float md = 1e6f, mm = 1e6f;
                for (int i = b * cpb + tid; i < (b + 1) * cpb; i += kWalkThreads) {
                    const int sb = i / nCh, c = i % nCh;
                    const float v = I[idx(sb, 0, t) + c];
                    const float bpv = bp_value(terms, K.nTerms, nBins, i, s.params0, K.scale, kill);
                    md = (v < md) ? v : md;
                    mm = (bpv < mm) ? bpv : mm;
                }
                md = block_min(md, red);
                mm = block_min(mm, red);
                if (tid == 0) {
                    bandMin[b] = md;
                    modelMin[b] = mm;
                }
            }
            // ordered band sums (band b by lane b, in bin order), rows staged in LDS
            float bsum = 0.0f, bsq = 0.0f;
            for (int x0 = 0; x0 < nBins; x0 += kRowChunk) {
                const int n = min(kRowChunk, nBins - x0);
                for (int k = tid; k < n; k += kWalkThreads) {
                    const int i = x0 + k, sb = i / nCh, c = i % nCh;
                    rowChunk[k] = I[idx(sb, 0, t) + c];
                }
                __syncthreads();
                if (tid < 8) {
                    const int a = max(tid * cpb, x0) - x0, b = min((tid + 1) * cpb, x0 + n) - x0;
                    for (int k = a; k < b; ++k) {
                        const float v = rowChunk[k];
                        bsum += v;
                        bsq += (v * v);
                    }
                }
                __syncthreads();
            }
            if (tid < 8) {
                bandMean[tid] = bsum;
                bandMeanSq[tid] = bsq;
            }
        }
        if (mode == MODE_FULL) {
            __syncthreads();
            if (tid == 0) {
                float dmm[8], sig[8];
                for (int b = 0; b < 8; ++b) {
                    dmm[b] = bandMin[b] - modelMin[b];
                    const float a = bandMeanSq[b] / (float)cpb;
                    const float m = bandMean[b] / (float)cpb;
                    sig[b] = (float)sqrt((double)a - (double)m * (double)m);
                }
                float spectrumRMS = sig[0];
                for (int b = 1; b < 8; ++b)
                    if (sig[b] < spectrumRMS) spectrumRMS = sig[b];  // std::min_element
                // nth_element(..., begin()+4): the 5th smallest of 8
                for (int a = 1; a < 8; ++a) {
                    const float v = dmm[a];
                    int j = a - 1;
                    while (j >= 0 && dmm[j] > v) {
                        dmm[j + 1] = dmm[j];
                        --j;
                    }
                    dmm[j + 1] = v;
                }
                float dataModel = dmm[4];
                s.spectrumRMS = spectrumRMS;
                ring_push(rmsR, s.rmsStart, s.rmsSize, s.rmsBack, s.rmsBack2, spectrumRMS);
                s.rmsRunAve = (float)(rmsPrefix[s.rmsSize] / (double)s.rmsSize);
                if (s.rmsSize == K.cap) s.meanOverRMS = s.meanRunAve / s.rmsRunAve;
                dataModel = dataModel + K.meanMinusMinimum * s.rmsRunAve;
                sh_dataModel = dataModel;
                sh_margin = K.crFactor * s.rmsRunAve;
            }
        }
        __syncthreads();
        const float dataModel = sh_dataModel, margin = sh_margin;
        // channel clipping (:315-342): the decisions (clipped bins take their
        // last-good value), the kept counts per band, a mask of the kept bins
        // and double-precision bounds of their sum.  The ordered sum of the
        // kept values (clipped bins as +0.0, which leaves a sum that starts at
        // +0.0 unchanged), taken by one lane, is needed only when
        //  * the spectrum may be kept (its mean enters the running mean), and
        //  * the mask differs from the precomputed one (whose sum is the
        //    precomputed mean): a spectrum that is rejected for certain
        //    leaves no trace of its mean, since the ring entry it pushes is
        //    overwritten by the rejection.
        float spectrumSum = 0.0f;
        const bool oneChunk = nBins <= kRowChunk;
        const float invCpbW = 1.0f / (float)cpb;
        const bool chkPre = sh_why == CNT_UNDECIDED && bx.valid;
        // learning phase: the recorded sums stand for the ordered sum when
        // nothing is clipped or the mask is the predicted one
        const bool chkLearn = useRec;
        const float lthr = chkLearn ? learn[t - tStart].thr : 0.0f;
        double sd = 0.0, ad = 0.0;
        int mism = 0, mismL = 0;
        if (tid == 0) {
            sh_mism = 0;
            sh_mismL = 0;
        }
        __syncthreads();
        // each thread's bins are read kG at a time, all in flight together,
        // before any of them is judged (the replacements write I); so are
        // their last-good values, which this pass only reads, so that the
        // replacements do not wait for one load after another
        constexpr int kG = 8;
        for (int kb0 = (tid & ~63); kb0 < nBins; kb0 += kWalkThreads * kG) {
            float vv[kG], bl[kG], lg[kG];
#pragma unroll
            for (int j = 0; j < kG; ++j) {
                const int i = kb0 + j * kWalkThreads + (tid & 63);
                vv[j] = 0.0f;
                bl[j] = 0.0f;
                lg[j] = 0.0f;
                if (i < nBins) {
                    const int sb = i / nCh, c = i % nCh;
                    vv[j] = I[idx(sb, 0, t) + c];
                    lg[j] = lastGood[i];
                    if (chkPre) bl[j] = bpLo[i];
                }
            }
#pragma unroll
            for (int j = 0; j < kG; ++j) {
                const int k0 = kb0 + j * kWalkThreads;  // uniform over the wave
                if (k0 >= nBins) break;
                const int i = k0 + (tid & 63);
                bool kept = false;
                int band = 0;
                if (i < nBins) {
                    const float v = vv[j];
                    const float bpv = bp_value(terms, K.nTerms, nBins, i, s.params0, K.scale, kill);
                    band = band_of(i, invCpbW);
                    if (v - dataModel - bpv > margin) {
                        const int sb = i / nCh, c = i % nCh;
                        const float g = lg[j];
                        for (int pol = 0; pol < nPol; ++pol) I[idx(sb, pol, t) + c] = g;
                    } else {
                        kept = true;
                        sd += (double)v;
                        ad += fabs((double)v);
                        // the precompute kept a bin only when v - bpLo <= mlo
                        if (chkPre && !((v - bl[j]) <= bx.mlo)) mism = 1;
                    }
                    if (chkLearn && kept != (v <= lthr)) mismL = 1;
                }
                const unsigned long long kb = __ballot(kept);
                if ((tid & 63) == 0) keepMask[k0 >> 6] = kb;
                const int bLo = band_of(k0, invCpbW);
                const int bHi = band_of(min(k0 + 63, nBins - 1), invCpbW);
                for (int b = bLo; b <= bHi; ++b) {
                    const int c = __popcll(__ballot(kept && band == b));
                    if ((tid & 63) == 0 && c) atomicAdd(&goodCnt[b], c);
                }
            }
        }
        if (__ballot(mism) != 0 && (tid & 63) == 0) atomicOr(&sh_mism, 1);
        if (__ballot(mismL) != 0 && (tid & 63) == 0) atomicOr(&sh_mismL, 1);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            sd += __shfl_xor(sd, off, 64);
            ad += __shfl_xor(ad, off, 64);
        }
        if ((tid & 63) == 0) {
            redD[0][tid >> 6] = sd;
            redD[1][tid >> 6] = ad;
        }
        __syncthreads();
        if (tid == 0) {
            int cnt[8];
            for (int b = 0; b < 8; ++b) cnt[b] = goodCnt[b];
            int total, badBands;
            float fracInc;
            band_stats(cnt, nBins, cpb, total, badBands, fracInc);
            sh_total = total;
            sh_badBands = badBands;
            sh_fracInc = fracInc;
            int how = 0;  // 0: fold; 1: the precomputed mean; 2: rejected for certain; 3: a recorded sum
            if (chkPre && !sh_mism) how = 1;
            if (chkLearn && total > 0) {
                const LearnRec& R = learn[t - tStart];
                if (total == nBins) {
                    how = 3;
                    sh_meanArg = R.fullSum / (float)(unsigned)total;
                } else if (!sh_mismL) {
                    how = 3;
                    sh_meanArg = R.predSum / (float)(unsigned)total;
                }
            }
            if (s.meanSize == K.cap && K.cap >= 1000) {
                if (badBands >= 4) {
                    how = 2;
                    sh_meanArg = 0.0f;
                } else if (how == 0 && total > 0) {
                    // the running mean and tolerance the decision will use (:394-425)
                    const float mPred = (s.meanRunAve - meanFrontDiv(t)) + s.meanBack / (float)s.meanSize;
                    const float tol = (float)((double)(K.srFactor * s.bpRms) / K.sqrtNBins);
                    double S = 0.0, A = 0.0;
                    for (int w = 0; w < kWalkThreads / 64; ++w) {
                        S += redD[0][w];
                        A += redD[1][w];
                    }
                    if (isfinite(mPred) && isfinite(tol) && tol > 0.0f && isfinite(S) && isfinite(A)) {
                        // |float fold - S| <= gamma_n * A (n additions, unit
                        // roundoff 2^-24), plus the double sums' own error;
                        // then the float division and subtraction, each
                        // within 2^-24 relative
                        const double n = (double)total;
                        const double err = A * (1.02 * n * 0x1p-24 + n * 0x1p-52) + 0x1p-1000;
                        double meanLo = (S - err) / n;
                        meanLo -= fabs(meanLo) * 0x1p-21;
                        double yLo = meanLo - (double)mPred;
                        yLo -= fabs(yLo) * 0x1p-21;
                        if (yLo > (double)tol * (1.0 + 0x1p-19)) {
                            how = 2;
                            sh_meanArg = INFINITY;  // the mean exceeds the tolerance for certain
                        }
                    }
                }
            }
            if (how == 1) sh_meanArg = stg[t - stageBase].mean;
            sh_how = how;
        }
        __syncthreads();
        if (sh_how == 0) {
            for (int x0 = 0; x0 < nBins; x0 += kRowChunk) {
                const int n = min(kRowChunk, nBins - x0);
                for (int k = tid; k < n; k += kWalkThreads) {
                    const int i = x0 + k;
                    const bool kept = (keepMask[i >> 6] >> (i & 63)) & 1ull;
                    float v = 0.0f;
                    if (kept) {
                        const int sb = i / nCh, c = i % nCh;
                        v = I[idx(sb, 0, t) + c];
                    }
                    rowChunk[k] = v;
                }
                __syncthreads();
                if (tid == 0) {
                    // a row of n % 32 == 0 takes the deep fold (n % 8 == 0 always: nBins is)
                    if ((n & 31) == 0) spectrumSum = fold_sum_deep(rowChunk, n, spectrumSum);
                    else spectrumSum = fold_sum(rowChunk, n, spectrumSum);
                }
                if (!oneChunk) __syncthreads();
            }
        }
        if (tid == 0) {
            const float meanArg = sh_how == 0 ? spectrumSum / (float)(unsigned)sh_total : sh_meanArg;
            sh_fill = phaseC(t, meanArg, sh_fracInc, sh_badBands, dataModel, margin, 1);
        }
        __syncthreads();
        // last-good spectrum (:441-469), read from the spectrum itself: a
        // clipped bin now holds its last-good value, so only kept non-zero
        // bins can fill an empty slot; every
        // change starts a new version for the spectra that follow (the apply
        // pass replaces from it)
        if (sh_fill) {
            for (int i = tid; i < nBins; i += kWalkThreads) {
                const int sb = i / nCh, c = i % nCh;
                const float v = I[idx(sb, 0, t) + c];
                if (lastGood[i] == 0.0f && v != 0.0f) {
                    lastGood[i] = v;
                    atomicAdd(&fillCnt, 1);
                }
            }
            __syncthreads();
            if (fillCnt > 0) {
                const int nv = s.ver + 1;
                if (nv < kMaxVer)
                    for (int i = tid; i < nBins; i += kWalkThreads) lastGoodVer[(size_t)nv * nBins + i] = lastGood[i];
                __syncthreads();
                if (tid == 0) s.ver = nv;  // >= kMaxVer: the rest of the run takes this path
            }
        }
        if (tid == 0) {
            s.remainingZeros -= fillCnt;
            phaseD();
            ++s.count[CNT_SLOW];
            sh_t = t + 1;
        }
        __syncthreads();
        t = sh_t;
    }
    if (tid == 0) {
        if (sh_stop == 1) {
            ++s.count[CNT_STOP];
            s.boxGrow = fminf(s.boxGrow * 4.0f, kMaxBoxGrow);
        } else if (sh_stop == 0) {
            s.boxGrow = fmaxf(s.boxGrow * 0.5f, 1.0f);
        }
        s.tNext = t;
        *S = s;
    }
    // the box for the next sub-chunk's precompute, around the state reached
    // (every thread is past its last read of the box and bandpass bounds)
    __syncthreads();
    bounds_of_state(bst, s, nT, tid, kWalkThreads);
}

// Replacement, clipSample and flattening of one bin of one spectrum
// (RFI_Clipper.cpp:319-335, 127-150, 478-485): returns the flattened value
// of polarisation 0; *set = 1 when every polarisation of the bin takes the
// value *g (replacement or clipSample), 0 when they keep their values.
__device__ __forceinline__ float flatten_bin(float v, float bpv, const SpecRec& r, const float* __restrict__ lgv,
                                             int nBins, int i, float* g, int* set)
{
    *set = 0;
    if ((r.flags & 3) == 1) {
        // kept spectrum the walk judged from the precomputed mask: same test,
        // replacement by the last-good value in force for it
        if (v - r.dataModel - bpv > r.margin) {
            v = lgv[(size_t)r.ver * nBins + i];
            *g = v;
            *set = 1;
        }
    }
    if (!(r.flags & 1)) {
        v = -0.2f;
        *g = v;
        *set = 1;
    }
    return v - (bpv + r.dataModel);
}

// Statistics pass over the flattened spectra: mean and rms of each spectrum
// from its ordered sums (RFI_Clipper.cpp:486-501), row-tiled as the
// precompute.  Writes stats[t] = (mean, rms).
template <int NT>
__global__ __launch_bounds__(kTileThreads) void clip_stats_kernel(BatchIO io)
{
    __shared__ float vals[2][kTileRows][kTileBins + 1];
    __shared__ SpecRec recs[kTileRows];
    const int beam = blockIdx.z;
    const BeamState& bs = *io.st[beam];
    const float* __restrict__ I = io.I[beam];
    const ClipConst K = bs.K;
    const double* __restrict__ terms = bs.terms;
    const float* __restrict__ lastGoodVer = bs.lastGoodVer;
    const SpecRec* __restrict__ rec = bs.rec;
    float2* __restrict__ stats = bs.stats;
    const int nT = io.nT[beam], ldMode = io.ldMode[beam];
    const int t0 = blockIdx.x * kTileRows;
    const int nRows = min(kTileRows, nT - t0);
    if (nRows <= 0) return;  // a shorter beam of the batch
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nBins = K.nBins, nCh = K.nCh, nPol = K.nPol;
    const long rowStride = (long)nBins * nPol;
    const int nTiles = (nBins + kTileBins - 1) / kTileBins;
    if (tid < nRows) recs[tid] = rec[t0 + tid];
    __syncthreads();

    float4 raw[kRowsPerWave];
    auto issue = [&](int tile) {
        const int i = tile * kTileBins + lane * 4;
        if (i < nBins) {
#pragma unroll
            for (int q = 0; q < kRowsPerWave; ++q) {
                const int r = wave + 4 * q;
                if (r < nRows) raw[q] = load_quad(I, (long)(t0 + r) * rowStride, i, nCh, nPol, ldMode);
            }
        }
    };
    auto flatten = [&](int tile, int buf) {
        const int i = tile * kTileBins + lane * 4;
        if (i >= nBins) return;
        BinTerms<NT> bt[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) bt[j].load(terms, K.nTerms, nBins, i + j, bs.kill);
#pragma unroll
        for (int q = 0; q < kRowsPerWave; ++q) {
            const int r = wave + 4 * q;
            if (r >= nRows) break;
            const SpecRec sr = recs[r];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float g;
                int set;
                const float bpv = bt[j].bp(sr.params0, K.scale);
                vals[buf][r][lane * 4 + j] = flatten_bin(q_at(raw[q], j), bpv, sr, lastGoodVer, nBins, i + j, &g, &set);
            }
        }
    };

    float sum = 0.0f, sumSq = 0.0f;
    issue(0);
    flatten(0, 0);
    __syncthreads();
    for (int tile = 0; tile < nTiles; ++tile) {
        const int buf = tile & 1;
        if (tile + 1 < nTiles) issue(tile + 1);
        if (wave == 0 && lane < nRows) {
            const int nb = min(kTileBins, nBins - tile * kTileBins);
            fold_sum_sq(&vals[buf][lane][0], nb, sum, sumSq);
        }
        if (tile + 1 < nTiles) flatten(tile + 1, buf ^ 1);
        __syncthreads();
    }
    if (wave == 0 && lane < nRows) {
        const float mean = sum / (float)(unsigned)nBins;
        float rms = (float)sqrt((double)(sumSq / (float)(unsigned)nBins) - (double)mean * (double)mean);
        if (rms == 0.0f) rms = 1.0f;
        stats[t0 + lane] = make_float2(mean, rms);
    }
}

constexpr int kWriteRows = 16;

// Elementwise pass: replacement / clipSample on every polarisation, the
// flattened polarisation 0 normalised for kept spectra (:502-532), weights.
// Threads own 4 consecutive bins; a workgroup walks kWriteRows spectra.
template <int NT>
__global__ __launch_bounds__(256) void clip_write_kernel(BatchIO io)
{
    __shared__ SpecRec recs[kWriteRows];
    __shared__ float2 st[kWriteRows];
    const int beam = blockIdx.z;
    const BeamState& bs = *io.st[beam];
    float* __restrict__ I = io.I[beam];
    float* __restrict__ W = io.W[beam];
    const ClipConst K = bs.K;
    const double* __restrict__ terms = bs.terms;
    const float* __restrict__ lastGoodVer = bs.lastGoodVer;
    const SpecRec* __restrict__ rec = bs.rec;
    const float2* __restrict__ stats = bs.stats;
    const int nT = io.nT[beam], ldMode = io.ldMode[beam];
    const int tid = threadIdx.x;
    const int t0 = blockIdx.y * kWriteRows;
    const int nRows = min(kWriteRows, nT - t0);
    if (nRows <= 0) return;  // a shorter beam of the batch
    if (tid < nRows) {
        recs[tid] = rec[t0 + tid];
        st[tid] = stats[t0 + tid];
    }
    __syncthreads();
    const int nBins = K.nBins, nCh = K.nCh, nPol = K.nPol;
    const int i = (blockIdx.x * 256 + tid) * 4;
    if (i >= nBins) return;
    const long rowStride = (long)nBins * nPol;
    BinTerms<NT> bt[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) bt[j].load(terms, K.nTerms, nBins, i + j, bs.kill);
    for (int r = 0; r < nRows; ++r) {
        const long rowOff = (long)(t0 + r) * rowStride;
        const SpecRec sr = recs[r];
        const float4 v4 = load_quad(I, rowOff, i, nCh, nPol, ldMode);
        float4 o, g4;
        int set[4];
        const bool kept = sr.flags & 1;
        const float sub = K.zeroDM == 1 ? (float)K.zeroDM * st[r].x : sr.meanRunAve;
        const float rms = st[r].y;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float bpv = bt[j].bp(sr.params0, K.scale);
            float v = flatten_bin(q_at(v4, j), bpv, sr, lastGoodVer, nBins, i + j, &q_at(g4, j), &set[j]);
            if (kept) {
                v -= sub;
                v /= rms;
                if (v > K.crFactor) v = 0.0f;
            }
            q_at(o, j) = v;
        }
        store_quad(I, rowOff, i, nCh, nPol, ldMode, o);
        if (nPol > 1 && (set[0] | set[1] | set[2] | set[3])) {
            for (int pol = 1; pol < nPol; ++pol)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (set[j]) I[rowOff + (long)pol * nCh + bin_offset(i + j, nCh, nPol)] = q_at(g4, j);
        }
        if (W) {
            const float4 one = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            for (int pol = 0; pol < nPol; ++pol) store_quad(W + (long)pol * nCh, rowOff, i, nCh, nPol, ldMode, one);
        }
    }
}

}  // namespace

// A run's launch set-up: the beams' descriptors and the grid sizes.
struct RunSetup {
    BatchIO io{};
    int nb = 0, maxBins = 0, maxNT = 0, maxSub = 0, nIter = 0, nTerms = -2;
    ppl_clipper* used[kMaxBeams];
};

struct ppl_clipper {
    ppl_clipper_config cfg;
    HostBandPass bp;
    float startFrequency = 0, endFrequency = 0;
    int nBins = 0;        // 0 until the first run fixes the map
    ClipConst K{};
    ClipScalars* d_S = nullptr;
    float *d_meanBuf = nullptr, *d_rmsBuf = nullptr, *d_lastGood = nullptr;
    double *d_meanPrefix = nullptr, *d_rmsPrefix = nullptr, *d_terms = nullptr;
    // killed bands of the .bp file, as Range<float> (min, max) pairs on the
    // file's map, and their bins on the run map
    std::vector<std::pair<float, float>> killed;
    unsigned* d_kill = nullptr;
    // host copies of the run map's polynomial terms and kill mask, for the
    // live model's values (ppl_clipper_model)
    std::vector<double> h_terms;
    std::vector<unsigned> h_kill;
    float2* d_stats = nullptr;
    SpecRec* d_rec = nullptr;
    PreRec* d_pre = nullptr;
    LearnRec* d_learn = nullptr;
    Box* d_box = nullptr;
    float *d_bpLo = nullptr, *d_bpHi = nullptr, *d_lastGoodVer = nullptr;
    BeamState* d_bs = nullptr;  // device descriptor of this clipper for the batched kernels
    int recCap = 0;
    std::vector<uint8_t> lastFlags;
    int lastNT = 0;
    // a run split over streams (ppl_clipper_run_front / _walk / _back): its
    // set-up and the phase reached (0: none, 1: front, 2: walks, 3 and 4:
    // front and walk of an empty run)
    RunSetup split;
    int phase = 0;
};

namespace {

void free_state(ppl_clipper* c)
{
    (void)hipFree(c->d_S);
    (void)hipFree(c->d_meanBuf);
    (void)hipFree(c->d_rmsBuf);
    (void)hipFree(c->d_lastGood);
    (void)hipFree(c->d_meanPrefix);
    (void)hipFree(c->d_rmsPrefix);
    (void)hipFree(c->d_terms);
    (void)hipFree(c->d_kill);
    c->d_kill = nullptr;
    (void)hipFree(c->d_stats);
    (void)hipFree(c->d_rec);
    (void)hipFree(c->d_pre);
    (void)hipFree(c->d_learn);
    (void)hipFree(c->d_box);
    (void)hipFree(c->d_bpLo);
    (void)hipFree(c->d_bpHi);
    (void)hipFree(c->d_lastGoodVer);
    (void)hipFree(c->d_bs);
    c->d_bs = nullptr;
    c->d_lastGoodVer = nullptr;
    c->d_pre = nullptr;
    c->d_learn = nullptr;
    c->d_box = nullptr;
    c->d_bpLo = c->d_bpHi = nullptr;
    c->d_S = nullptr;
    c->d_meanBuf = c->d_rmsBuf = c->d_lastGood = nullptr;
    c->d_meanPrefix = c->d_rmsPrefix = c->d_terms = nullptr;
    c->d_stats = nullptr;
    c->d_rec = nullptr;
}

// First run: fix the run map (RFI_Clipper.cpp:184-187), rebin the model,
// tabulate the polynomial terms and upload the initial state.
int setup_map(ppl_clipper* c, unsigned nSub, unsigned nPol, unsigned nCh)
{
    const unsigned nBins = nSub * nCh;
    HostMap m;
    m.nBins = nBins;
    m.lower = (double)c->startFrequency;
    m.width = ((double)c->endFrequency - m.lower) / nBins;  // BinMap::setEnd
    c->bp.reBin(m);
    // the killed ranges' bins on the run map (_zeroChannelsMap :111-125):
    // binIndex(min) to binIndex(max) inclusive; bins past the map's end are
    // dropped (the reference grows its vector to hold them), and a range
    // that starts before bin 0 or ends exactly at nBins is refused (the
    // reference writes outside its vector there)
    std::vector<unsigned> killMask;
    for (const auto& r : c->killed) {
        int lo = m.binIndex(r.first), hi = m.binIndex(r.second);
        if (hi < lo) std::swap(lo, hi);
        PPL_CHECK_ARG(lo >= 0 && hi != (int)nBins,
                      "RFI_Clipper: a killed band (%g to %g) reaches outside the %u-bin map", r.first, r.second,
                      nBins);
        if (killMask.empty()) killMask.assign((nBins + 31) / 32, 0u);
        for (int i = lo; i <= hi && i < (int)nBins; ++i) killMask[(size_t)i >> 5] |= 1u << (i & 31);
    }
    // _buildData and setMean both use the float width ratio
    const float scale = (float)(c->bp.current.width / c->bp.primary.width);
    const int nTerms = (int)c->bp.params.size() - 1;
    std::vector<double> terms((size_t)std::max(nTerms, 0) * nBins);
    for (int j = 1; j <= nTerms; ++j)
        for (unsigned i = 0; i < nBins; ++i) {
            const float v = (float)m.binAssignmentNumber(i);
            terms[(size_t)(j - 1) * nBins + i] = (double)c->bp.params[j] * std::pow((double)v, (double)j);
        }
    ClipConst& K = c->K;
    K.nBins = (int)nBins;
    K.nSub = (int)nSub;
    K.nPol = (int)nPol;
    K.nCh = (int)nCh;
    K.cap = c->cfg.maxHistory;
    K.nTerms = nTerms;
    K.scale = scale;
    K.crFactor = c->cfg.crFactor;
    K.srFactor = c->cfg.srFactor;
    {
        float k = 4;
        K.meanMinusMinimum = k / std::sqrt(2.0 * k);
    }
    K.sqrtNBins = std::sqrt((double)nBins);
    K.zeroDM = c->cfg.zeroDM ? 1 : 0;
    K.useMeanOverRMS = c->cfg.useMeanOverRMS ? 1 : 0;
    K.maxHistory = c->cfg.maxHistory;

    ClipScalars S{};
    S.params0 = c->bp.params[0];
    S.bpMean = c->bp.curMean;
    S.bpRms = c->bp.curRms;
    S.remainingZeros = (int)nBins;
    S.initialised = 1;
    S.boxGrow = 1.0f;

    const size_t cap = (size_t)std::max(K.cap, 1);
    PPL_HIP(hipMalloc(&c->d_S, sizeof(ClipScalars)));
    PPL_HIP(hipMalloc(&c->d_meanBuf, cap * sizeof(float)));
    PPL_HIP(hipMalloc(&c->d_rmsBuf, cap * sizeof(float)));
    PPL_HIP(hipMalloc(&c->d_meanPrefix, (cap + 1) * sizeof(double)));
    PPL_HIP(hipMalloc(&c->d_rmsPrefix, (cap + 1) * sizeof(double)));
    PPL_HIP(hipMalloc(&c->d_lastGood, nBins * sizeof(float)));
    PPL_HIP(hipMalloc(&c->d_box, sizeof(Box)));
    PPL_HIP(hipMalloc(&c->d_bpLo, nBins * sizeof(float)));
    PPL_HIP(hipMalloc(&c->d_bpHi, nBins * sizeof(float)));
    PPL_HIP(hipMalloc(&c->d_lastGoodVer, (size_t)kMaxVer * nBins * sizeof(float)));
    PPL_HIP(hipMalloc(&c->d_terms, std::max<size_t>(terms.size(), 1) * sizeof(double)));
    PPL_HIP(hipMemcpy(c->d_S, &S, sizeof S, hipMemcpyHostToDevice));
    PPL_HIP(hipMemset(c->d_meanPrefix, 0, (cap + 1) * sizeof(double)));
    PPL_HIP(hipMemset(c->d_rmsPrefix, 0, (cap + 1) * sizeof(double)));
    PPL_HIP(hipMemset(c->d_lastGood, 0, nBins * sizeof(float)));
    if (!terms.empty())
        PPL_HIP(hipMemcpy(c->d_terms, terms.data(), terms.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!killMask.empty()) {
        PPL_HIP(hipMalloc(&c->d_kill, killMask.size() * sizeof(unsigned)));
        PPL_HIP(hipMemcpy(c->d_kill, killMask.data(), killMask.size() * sizeof(unsigned), hipMemcpyHostToDevice));
    }
    c->h_terms = terms;
    c->h_kill = killMask;
    c->nBins = (int)nBins;
    return PPL_OK;
}

}  // namespace

extern "C" int ppl_clipper_create(ppl_clipper** out, const ppl_clipper_config* cfg)
{
    PPL_CHECK_ARG(out && cfg && cfg->bp_values, "null argument");
    PPL_CHECK_ARG(cfg->n_values >= 6,
                  "BandPassAdapter: data format wrong : got %d parameters, expecting at least 6", cfg->n_values);
    PPL_CHECK_ARG(cfg->n_values - 5 <= kMaxParams + 1, "too many bandpass polynomial terms (%d)", cfg->n_values - 5);
    PPL_CHECK_ARG(cfg->bandMatching,
                  "RFI_Clipper: <Band matching=\"true\"/> is required (without it the bin width is 0)");
    PPL_CHECK_ARG(cfg->maxHistory >= 1, "History maximum must be >= 1");
    ppl_clipper* c = new ppl_clipper;
    c->cfg = *cfg;
    c->cfg.bp_values = nullptr;
    // BandPassAdapter::deserialise :86-100, values through QVector<float>
    std::vector<float> params;
    for (int i = 0; i < cfg->n_values; ++i) params.push_back((float)cfg->bp_values[i]);
    const unsigned nChan = (unsigned)params[0];
    const float startFreq = params[1], deltaF = params[3], rms = params[4], median = params[5];
    params.erase(params.begin(), params.begin() + 5);
    HostMap m;
    m.nBins = nChan;
    m.lower = startFreq;
    m.width = deltaF;
    c->bp.setData(m, params);
    c->bp.setMedian(median);
    c->bp.setRMS(rms);
    // kill lines (BandPassAdapter.cpp:58-71,101-106): the line's first number
    // is both the index and the frequency; unique indices in ascending order
    // (a QMap), each checked against the file's map, then killChannel(index)
    // on the current (file's) map: the range [binStart, binEnd] as floats
    // (BandPass.cpp:141-157, Range<float>)
    if (cfg->n_kill < 0 || (cfg->n_kill > 0 && !cfg->kill_index)) {
        delete c;
        ::ppl::set_error("null kill list");
        return PPL_E_ARG;
    }
    std::vector<unsigned> kills(cfg->kill_index, cfg->kill_index + cfg->n_kill);
    std::sort(kills.begin(), kills.end());
    kills.erase(std::unique(kills.begin(), kills.end()), kills.end());
    for (unsigned index : kills) {
        const float freq = (float)(double)index;
        if (m.binIndex(freq) != (int)index) {
            delete c;
            ::ppl::set_error("BandPassAdapter:  Frequency kill mismatch: index %u does not match frequency %g", index,
                             (double)freq);
            return PPL_E_ARG;
        }
        const float a = (float)c->bp.current.binStart(index), b = (float)c->bp.current.binEnd(index);
        c->killed.emplace_back(std::min(a, b), std::max(a, b));
        // zeroing the file's map: a range ending exactly at nChan is written
        // out of bounds by the reference (BandPass.cpp:118-123)
        int lo = m.binIndex(std::min(a, b)), hi = m.binIndex(std::max(a, b));
        if (hi < lo) std::swap(lo, hi);
        if (lo < 0 || hi == (int)nChan) {
            delete c;
            ::ppl::set_error("BandPassAdapter: kill index %u reaches outside the %u-channel map", index, nChan);
            return PPL_E_ARG;
        }
    }
    c->cfg.kill_index = nullptr;
    c->cfg.n_kill = 0;
    c->startFrequency = (float)c->bp.current.lower;
    c->endFrequency = (float)c->bp.current.endValue();
    *out = c;
    return PPL_OK;
}

extern "C" int ppl_clipper_destroy(ppl_clipper* c)
{
    if (!c) return PPL_OK;
    free_state(c);
    delete c;
    return PPL_OK;
}

// The two apply passes: statistics of the flattened spectra, then the
// elementwise replacement / flatten / normalise.
template <int NT>
static int launch_apply(const BatchIO& io, int nBeams, int maxBins, int maxNT, hipStream_t st)
{
    hipLaunchKernelGGL(clip_stats_kernel<NT>, dim3(cdiv(maxNT, kTileRows), 1, nBeams), dim3(kTileThreads), 0, st,
                       io);
    PPL_LAUNCHED();
    hipLaunchKernelGGL(clip_write_kernel<NT>, dim3(cdiv((unsigned)maxBins / 4, 256), cdiv(maxNT, kWriteRows), nBeams),
                       dim3(256), 0, st, io);
    PPL_LAUNCHED();
    return PPL_OK;
}

// Brings one clipper to the run's shape: the map on its first run, the
// per-spectrum records' capacity, the polarisation count; uploads its
// device descriptor when any of these changed.
static int prepare(ppl_clipper* c, const ppl_clip_job& j)
{
    PPL_CHECK_ARG(c && j.d_I, "null argument");
    const unsigned nBins = j.nSub * j.nCh;
    PPL_CHECK_ARG(nBins % 8 == 0 && nBins > 0, "RFI_Clipper: nBins (%u) must be a positive multiple of 8", nBins);
    PPL_CHECK_ARG(j.nPol >= 1, "no polarisations");
    PPL_CHECK_ARG(nBins <= (unsigned)kMaxBinsWalk, "RFI_Clipper: %u bins exceed the supported %d", nBins,
                  kMaxBinsWalk);
    bool dirty = false;
    if (c->nBins == 0) {
        int rc = setup_map(c, j.nSub, j.nPol, j.nCh);
        if (rc != PPL_OK) return rc;
        dirty = true;
    }
    PPL_CHECK_ARG((unsigned)c->nBins == nBins && (unsigned)c->K.nCh == j.nCh,
                  "RFI_Clipper: spectrum shape changed (%u bins, was %d)", nBins, c->nBins);
    if (c->K.nPol != (int)j.nPol) {
        c->K.nPol = (int)j.nPol;
        dirty = true;
    }
    if ((int)j.nT > c->recCap) {
        (void)hipFree(c->d_rec);
        (void)hipFree(c->d_pre);
        (void)hipFree(c->d_learn);
        (void)hipFree(c->d_stats);
        c->d_rec = nullptr;
        c->d_pre = nullptr;
        c->d_learn = nullptr;
        c->d_stats = nullptr;
        PPL_HIP(hipMalloc(&c->d_rec, sizeof(SpecRec) * j.nT));
        PPL_HIP(hipMalloc(&c->d_pre, sizeof(PreRec) * j.nT));     // a split run precomputes the whole run
        PPL_HIP(hipMalloc(&c->d_learn, sizeof(LearnRec) * j.nT));
        PPL_HIP(hipMalloc(&c->d_stats, sizeof(float2) * j.nT));
        c->recCap = (int)j.nT;
        dirty = true;
    }
    if (dirty || !c->d_bs) {
        BeamState h{};
        h.K = c->K;
        h.S = c->d_S;
        h.meanBuf = c->d_meanBuf;
        h.rmsBuf = c->d_rmsBuf;
        h.lastGood = c->d_lastGood;
        h.lastGoodVer = c->d_lastGoodVer;
        h.bpLo = c->d_bpLo;
        h.bpHi = c->d_bpHi;
        h.meanPrefix = c->d_meanPrefix;
        h.rmsPrefix = c->d_rmsPrefix;
        h.terms = c->d_terms;
        h.kill = c->d_kill;
        h.box = c->d_box;
        h.pre = c->d_pre;
        h.learn = c->d_learn;
        h.rec = c->d_rec;
        h.stats = c->d_stats;
        if (!c->d_bs) PPL_HIP(hipMalloc(&c->d_bs, sizeof(BeamState)));
        PPL_HIP(hipMemcpy(c->d_bs, &h, sizeof h, hipMemcpyHostToDevice));
    }
    return PPL_OK;
}

// Validates and prepares a batch of clippers (beams) for one run, with
// sub-chunks of at most subCap spectra.
static int setup_batch(ppl_clipper* const* clips, const ppl_clip_job* jobs, int n, int subCap, RunSetup& R)
{
    PPL_CHECK_ARG(clips && jobs && n >= 0, "null argument");
    PPL_CHECK_ARG(n <= kMaxBeams, "at most %d beams per clipper batch (got %d)", kMaxBeams, n);
    R = RunSetup{};
    BatchIO& io = R.io;
    int maxIter = 0;
    auto aligned = [](const void* p) { return p == nullptr || (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    for (int k = 0; k < n; ++k) {
        for (int q = 0; q < k; ++q)
            PPL_CHECK_ARG(clips[q] != clips[k], "clipper %d appears twice in one batch", k);
        int rc = prepare(clips[k], jobs[k]);
        if (rc != PPL_OK) return rc;
        const ppl_clip_job& j = jobs[k];
        if (j.nT == 0) continue;
        ppl_clipper* c = clips[k];
        const int nb = R.nb;
        io.st[nb] = c->d_bs;
        io.I[nb] = j.d_I;
        io.W[nb] = j.d_W;
        io.nT[nb] = (int)j.nT;
        io.subLen[nb] = std::min((int)j.nT, subCap);
        // access mode of the row-tiled passes (see load_quad)
        io.ldMode[nb] = !(aligned(j.d_I) && aligned(j.d_W)) ? 2 : j.nPol == 1 ? 0 : j.nCh % 4 == 0 ? 1 : 2;
        R.maxBins = std::max(R.maxBins, c->nBins);
        R.maxNT = std::max(R.maxNT, (int)j.nT);
        maxIter = std::max(maxIter, (int)cdiv(j.nT, (unsigned)io.subLen[nb]));
        R.nTerms = R.nTerms == -2 ? c->K.nTerms : (R.nTerms == c->K.nTerms ? R.nTerms : -1);
        R.used[R.nb++] = c;
    }
    // Sub-chunks: box around the state reached, precompute, walk.  The walk
    // advances S->tNext by up to subLen spectra, or less when the state
    // leaves the box; the last launch walks whatever remains.  Launches past
    // a beam's end return at once.
    R.nIter = maxIter + kExtraIter;
    R.maxSub = std::min(R.maxNT, subCap);
    return PPL_OK;
}

// The run's start: the walk counters' position and first box, and the first
// sub-chunk's precompute.
static int launch_front(const RunSetup& R, hipStream_t st)
{
    const BatchIO& io = R.io;
    hipLaunchKernelGGL(clip_begin_kernel, dim3(cdiv(R.maxBins, 256), 1, R.nb), dim3(256), 0, st, io);
    PPL_LAUNCHED();
    // the first sub-chunk's box; each walk sets up the next one's as it ends
    hipLaunchKernelGGL(clip_bounds_kernel, dim3(cdiv(R.maxBins, 256), 1, R.nb), dim3(256), 0, st, io);
    PPL_LAUNCHED();
    hipLaunchKernelGGL(clip_prep_kernel, dim3(cdiv(R.maxSub, kPreRows), 1, R.nb), dim3(kTileThreads), 0, st, io);
    PPL_LAUNCHED();
    return PPL_OK;
}

// The walks: the first sub-chunk's, then precompute and walk for the others.
static int launch_walks(const RunSetup& R, hipStream_t st)
{
    const BatchIO& io = R.io;
    for (int it = 0; it < R.nIter; ++it) {
        if (it > 0) {
            hipLaunchKernelGGL(clip_prep_kernel, dim3(cdiv(R.maxSub, kPreRows), 1, R.nb), dim3(kTileThreads), 0,
                               st, io);
            PPL_LAUNCHED();
        }
        hipLaunchKernelGGL(clip_walk_kernel, dim3(1, 1, R.nb), dim3(kWalkThreads), 0, st, io,
                           it == R.nIter - 1 ? 1 : 0);
        PPL_LAUNCHED();
    }
    return PPL_OK;
}

// The statistics and write passes over the walked spectra.
static int launch_back(const RunSetup& R, hipStream_t st)
{
    int rc = PPL_OK;
    switch (R.nTerms) {
        case 0: rc = launch_apply<0>(R.io, R.nb, R.maxBins, R.maxNT, st); break;
        case 1: rc = launch_apply<1>(R.io, R.nb, R.maxBins, R.maxNT, st); break;
        case 2: rc = launch_apply<2>(R.io, R.nb, R.maxBins, R.maxNT, st); break;
        case 3: rc = launch_apply<3>(R.io, R.nb, R.maxBins, R.maxNT, st); break;
        case 4: rc = launch_apply<4>(R.io, R.nb, R.maxBins, R.maxNT, st); break;
        default: rc = launch_apply<-1>(R.io, R.nb, R.maxBins, R.maxNT, st);
    }
    if (rc != PPL_OK) return rc;
    for (int k = 0; k < R.nb; ++k) R.used[k]->lastNT = R.io.nT[k];
    return PPL_OK;
}

// One run over a batch of clippers (beams): each stage is one launch for
// all beams (blockIdx.z), the walk one workgroup per beam.  Beams keep
// their own shapes and spectrum counts; shorter beams' blocks return early.
static int run_batch(ppl_clipper* const* clips, const ppl_clip_job* jobs, int n, hipStream_t st)
{
    RunSetup R;
    int rc = setup_batch(clips, jobs, n, kSubLen, R);
    if (rc != PPL_OK || R.nb == 0) return rc;
    if ((rc = launch_front(R, st)) != PPL_OK) return rc;
    if ((rc = launch_walks(R, st)) != PPL_OK) return rc;
    return launch_back(R, st);
}

static int clipper_run(ppl_clipper* c, float* d_I, float* d_W, unsigned nT, unsigned nSub, unsigned nPol,
                       unsigned nCh, hipStream_t st)
{
    ppl_clip_job j{d_I, d_W, nT, nSub, nPol, nCh};
    return run_batch(&c, &j, 1, st);
}

extern "C" int ppl_clipper_run(ppl_clipper* c, float* d_I, float* d_W, unsigned nT, unsigned nSub, unsigned nPol,
                               unsigned nCh, void* stream)
{
    return clipper_run(c, d_I, d_W, nT, nSub, nPol, nCh, stream_of(stream));
}

extern "C" int ppl_clipper_run_batch(ppl_clipper* const* clips, const ppl_clip_job* jobs, int nBeams, void* stream)
{
    return run_batch(clips, jobs, nBeams, stream_of(stream));
}

extern "C" int ppl_clipper_run_front(ppl_clipper* c, float* d_I, float* d_W, unsigned nT, unsigned nSub,
                                     unsigned nPol, unsigned nCh, void* stream)
{
    PPL_CHECK_ARG(c, "null argument");
    PPL_CHECK_ARG(c->phase == 0, "RFI_Clipper: a split run is in progress (front called twice)");
    ppl_clip_job j{d_I, d_W, nT, nSub, nPol, nCh};
    ppl_clipper* one = c;
    const int rc = setup_batch(&one, &j, 1, std::max((int)nT, 1), c->split);
    if (rc != PPL_OK) return rc;
    if (c->split.nb == 0) {
        c->phase = 3;  // nothing to walk: walk and back are no-ops
        return PPL_OK;
    }
    const int r2 = launch_front(c->split, stream_of(stream));
    if (r2 == PPL_OK) c->phase = 1;
    return r2;
}

extern "C" int ppl_clipper_run_walk(ppl_clipper* c, void* stream)
{
    PPL_CHECK_ARG(c, "null argument");
    PPL_CHECK_ARG(c->phase == 1 || c->phase == 3, "RFI_Clipper: walk without a front");
    if (c->phase == 3) {
        c->phase = 4;
        return PPL_OK;
    }
    const int rc = launch_walks(c->split, stream_of(stream));
    if (rc == PPL_OK) c->phase = 2;
    return rc;
}

extern "C" int ppl_clipper_run_back(ppl_clipper* c, void* stream)
{
    PPL_CHECK_ARG(c, "null argument");
    PPL_CHECK_ARG(c->phase == 2 || c->phase == 4, "RFI_Clipper: back without a walk");
    if (c->phase == 4) {
        c->phase = 0;
        return PPL_OK;
    }
    const int rc = launch_back(c->split, stream_of(stream));
    c->phase = 0;
    return rc;
}

extern "C" int ppl_clipper_stats(ppl_clipper* c, float* mean, float* rms)
{
    PPL_CHECK_ARG(c && mean && rms, "null argument");
    if (!c->d_S) {
        *mean = 0.0f;
        *rms = 0.0f;
        return PPL_OK;
    }
    ClipScalars S;
    PPL_HIP(hipMemcpy(&S, c->d_S, sizeof S, hipMemcpyDeviceToHost));
    *mean = S.meanRunAve;
    *rms = S.rmsRunAve;
    return PPL_OK;
}

// The live bandpass model: RFI_Clipper::bandPass() after the spectra run so
// far (setMean / setRMS at the end of every spectrum, RFI_Clipper.cpp:562-563;
// BandPass.cpp:44-74): the constant coefficient, the current map's mean and
// RMS, and optionally its values on the run map (scale * evaluate(v), killed
// bins 0.0), as the device evaluates them.
extern "C" int ppl_clipper_model(ppl_clipper* c, float* params0, float* mean, float* rms, float* values,
                                 unsigned nValues)
{
    PPL_CHECK_ARG(c && params0 && mean && rms, "null argument");
    if (!c->d_S) {
        PPL_CHECK_ARG(!values, "the run map is fixed by the first run; no values before it");
        *params0 = c->bp.params[0];
        *mean = c->bp.curMean;
        *rms = c->bp.curRms;
        return PPL_OK;
    }
    ClipScalars S;
    PPL_HIP(hipMemcpy(&S, c->d_S, sizeof S, hipMemcpyDeviceToHost));
    *params0 = S.params0;
    *mean = S.bpMean;
    *rms = S.bpRms;
    if (values) {
        PPL_CHECK_ARG(nValues == (unsigned)c->nBins, "values: %u entries for a %d-bin map", nValues, c->nBins);
        const int nb = c->nBins, nt = c->K.nTerms;
        for (int i = 0; i < nb; ++i) {
            if (!c->h_kill.empty() && ((c->h_kill[(size_t)i >> 5] >> (i & 31)) & 1u)) {
                values[i] = 0.0f;
                continue;
            }
            float tot = 0.0f;
            tot = (float)((double)tot + (double)S.params0 * 1.0);
            for (int j = 0; j < nt; ++j) tot = (float)((double)tot + c->h_terms[(size_t)j * nb + i]);
            values[i] = c->K.scale * tot;
        }
    }
    return PPL_OK;
}

extern "C" int ppl_clipper_counters(ppl_clipper* c, long long* out, int n)
{
    PPL_CHECK_ARG(c && out && n >= 0, "null argument");
    const int m = std::min(n, (int)PPL_CLIP_NCOUNTERS);
    for (int i = 0; i < m; ++i) out[i] = 0;
    if (!c->d_S) return PPL_OK;
    ClipScalars S;
    PPL_HIP(hipMemcpy(&S, c->d_S, sizeof S, hipMemcpyDeviceToHost));
    for (int i = 0; i < m; ++i) out[i] = S.count[i];
    return PPL_OK;
}

extern "C" int ppl_clipper_flags(ppl_clipper* c, uint8_t* h_good, unsigned nT)
{
    PPL_CHECK_ARG(c && h_good, "null argument");
    PPL_CHECK_ARG((int)nT <= c->lastNT, "asked for %u flags, last run had %d spectra", nT, c->lastNT);
    std::vector<SpecRec> r(nT);
    if (nT) PPL_HIP(hipMemcpy(r.data(), c->d_rec, sizeof(SpecRec) * nT, hipMemcpyDeviceToHost));
    for (unsigned t = 0; t < nT; ++t) h_good[t] = (uint8_t)(r[t].flags & 1);
    return PPL_OK;
}

// Fused variant: clip, then transpose the clipped Stokes-I spectra into a
// dedispersion buffer (ppl_buffer.hip).  The buffer's channel count is
// checked before the clipper runs, so a mismatch leaves the clipper state
// untouched.
unsigned ppl_ddbuf_nchan(const ppl_ddbuf* buf);
int ppl_ddbuf_add_spectra(ppl_ddbuf* buf, const float* d_spectra, unsigned nT, unsigned first, unsigned* consumed,
                          int* full, hipStream_t st);

extern "C" int ppl_clipper_run_into(ppl_clipper* c, float* d_I, unsigned nT, unsigned nSub, unsigned nCh,
                                    ppl_ddbuf* buf, unsigned first, unsigned* consumed, int* full, void* stream)
{
    PPL_CHECK_ARG(c && buf && consumed && full, "null argument");
    PPL_CHECK_ARG(first <= nT, "first spectrum %u beyond the %u clipped", first, nT);
    PPL_CHECK_ARG(nSub * nCh == ppl_ddbuf_nchan(buf),
                  "DedispersionBuffer: input data sample size(%u) does not match buffer sample size (%u)", nSub * nCh,
                  ppl_ddbuf_nchan(buf));
    int rc = clipper_run(c, d_I, nullptr, nT, nSub, 1, nCh, stream_of(stream));
    if (rc != PPL_OK) return rc;
    return ppl_ddbuf_add_spectra(buf, d_I, nT, first, consumed, full, stream_of(stream));
}
