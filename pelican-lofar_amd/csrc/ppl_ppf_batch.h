// ppl_ppf_batch.h -- job table of the batched channeliser launches (shared by
// the FP32 and the split-precision kernels; not installed).
#pragma once

#include "ppl_internal.h"

#include <cstdint>

namespace ppl {

constexpr int kMaxPPFJobs = PPL_PPF_MAX_JOBS;
// Launch sizes that keep every CU busy: waves of the FP32 method (32 per CU),
// workgroups of the split-precision method (8 per CU).
constexpr long kFillWaves = 8192;
constexpr long kFillGroups = 2048;

// One plan's share of a batched launch.  The workgroups [blockStart, next
// job's blockStart) run this job.  64 bytes, so the table of 48 jobs stays a
// kernel argument.
struct PPFJob {
    const uint32_t* in;    // [nSub][2][stride] complex int16 samples
    const uint32_t* hist;  // [nSub][2][TAPS-1][N], the blocks before the chunk
    float* out;            // [nBlk/integ][nSub][1][N]
    const float* coeffs;   // [TAPS][N]
    long stride;           // samples per (sub, pol) stream of `in`
    int nBlk, nSub, runLen, nRuns, integ, blockStart;
};
static_assert(sizeof(PPFJob) == 64, "PPFJob layout");

struct PPFBatch {
    PPFJob job[kMaxPPFJobs];
    const float* tw;  // W_N^k, k < N, of the launch's channel count
    int n;
};

// History update of every job of a batch: the last H blocks of each stream.
struct HistJob {
    const uint32_t* in;
    const uint32_t* oldh;
    uint32_t* newh;
    long stride, nBlk;
    int nStreams, H, N, blockStart;
};
struct HistBatch {
    HistJob job[kMaxPPFJobs];
    int n;
};

// The job a workgroup belongs to (jobs in blockStart order) and the
// workgroup's index inside it.
template <typename B>
__device__ __forceinline__ const auto& find_batch_job(const B& b, int& local)
{
    int jb = 0;
    for (int k = 1; k < b.n; ++k)
        if ((int)blockIdx.x >= b.job[k].blockStart) jb = k;
    local = (int)blockIdx.x - b.job[jb].blockStart;
    return b.job[jb];
}

// Blocks per run of one wave (FP32 method) or workgroup (split-precision
// method): 256, halved while the launch would have fewer than `fill` waves or
// workgroups and the run stays at least `minLen` blocks; always a multiple of
// `mult`.  A run starts by re-reading the TAPS-1 blocks before it, so long
// runs read less; short runs fill the GPU when the chunks are small.  The
// spectra do not depend on the run length.
int ppf_run_length(const long* nBlk, const int* nSub, int n, int unitsPerRun, long fill, int mult, int minLen);

}  // namespace ppl
