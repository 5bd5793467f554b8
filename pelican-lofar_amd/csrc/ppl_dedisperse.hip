// ppl_dedisperse.hip -- brute-force incoherent dedispersion for gfx950.
//
// out[dm][t] = sum_{c=0}^{nChan-1} in[c][t + delay[dm][c]],  t < nOut,
// summed in channel order (the FP32 result is bit-identical to the
// reference's per-thread channel loop, src/lib/DedispersionKernel.cu:32-43).
//
// Kernel shape.  One 256-thread workgroup owns a tile of TD = 4*ND DM trials
// by TT = 64*NT output samples.  Wave w of the workgroup accumulates ND DM
// trials; lane l accumulates samples t0 + l + 64*i (i < NT), so every LDS
// read of a wave is 64 consecutive words (bank-conflict free) and every
// store is a coalesced 256-byte row segment.  Channels are streamed in
// blocks of CB: for each channel the window [t0 + delay[dm0][c],
// t0 + delay[dm0+TD-1][c] + TT) of the input row is staged in LDS, and each
// wave reads its ND delayed 64-wide slices from there.  The FP32 form is
// bound by LDS bandwidth (one 4-byte read per add); FP16 and u8 inputs halve
// and quarter the bytes per add.
#include "ppl_internal.h"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

struct ppl_dedisp_plan {
    ppl_dedisp_config cfg;
    int maxshift = 0, rem = 0, nOut = 0;
    float mstartdm = 0, mdmstep = 0;
    std::vector<float> shift;
    std::vector<int32_t> delays;  // host [tdms][nChan]
    int32_t* d_delayT = nullptr;  // device [nChan][tdmsPad], rows padded with the last trial
    int tdmsPad = 0;              // tdms rounded up to 64, the widest DM tile
    int nd = 8, nt = 8;           // per-thread DM x sample register tile
    int cb = 0;                   // channels per staged block (0: as many as fit)
    int rw = 4;                   // LDS read width per lane, bytes (4, or 16 for FP32 input)
    int spanTD[4] = {0, 0, 0, 0}; // max delay spread inside a DM tile of 8, 16, 32, 64 trials
    long long fusedDiffs = 0;     // delays that differ when the trial value is one fused multiply-add
    // ring tables of the 16-byte kernel, built on first use, for DM tiles
    // of 16 and 32 trials (index 0, 1): per tile and channel the LDS byte
    // offsets of each trial's reads, then per tile and channel the first
    // trial's delay
    int32_t* d_ring[3] = {nullptr, nullptr, nullptr};  // ring tables for DM tiles of 16, 32, 64
    // u8 series, built on first use: per trial dm, the number of channels
    // whose delay is k, for k = 0 .. maxdelay(dm): lenTab[lenOff[dm] + k];
    // lenOff[tdms] is the table's length
    int32_t* d_lenOff = nullptr;
    uint16_t* d_lenTab = nullptr;
};

namespace {

using namespace ppl;

template <typename T> struct Acc;
template <> struct Acc<float> { using type = float; };
template <> struct Acc<_Float16> { using type = float; };
template <> struct Acc<uint8_t> { using type = int32_t; };

// Channel rows are staged as dword streams straight into LDS
// (global_load_lds_dword, no register round trip), in two buffers so that
// the loads of channel block k+1 are in flight while block k is summed.  A
// row starts at the dword holding element t0 + delay[dm0][c]; the element
// misalignment is folded into that channel's per-DM LDS offsets, which the
// staging step also writes (offs[c][j] = delay[dm0+j][c] - delay[dm0][c] +
// misalignment + row base).  Every LDS read of a channel for one DM trial is
// issued before any of them is consumed.
// Workgroups are dispatched to the eight XCDs round robin (workgroup i to
// XCD i % 8), each XCD with its own L2.  The grid is XCD-major: XCD x gets
// whole time tiles (x, x+8, ...), and the DM tiles of one time tile follow
// one another on it, so their overlapping delayed input rows are fetched
// from HBM about once per time tile instead of once per XCD.
constexpr int kXcds = 8;
__device__ __forceinline__ void tile_of(int bid, int nDMTiles, int& tTile, int& dmTile)
{
    const int xcd = bid % kXcds, k = bid / kXcds;
    tTile = (k / nDMTiles) * kXcds + xcd;
    dmTile = k % nDMTiles;
}

// One beam's part of a launch.  A launch covers several beams (jobs): job k
// owns workgroups [blockStart, blockStart + nDMTiles*nTTilesPad), a multiple
// of 8, so a job's own XCD-major tile order is kept.  The table travels in
// the kernel arguments; a workgroup finds its job with a uniform scan.
constexpr int kMaxDDJobs = PPL_DEDISP_MAX_JOBS;
struct DDJob {
    const void* in;
    void* out;
    const int32_t* delayT;
    int nChan, nsamp, tdms, tdmsPad, nOut;
    int rowUnits;   // LDS row length: elements (f32) or dwords per copy (f16, u8)
    int CB;         // channels per staged block
    int nDMTiles;
    int blockStart;
    int pad_;
};
struct DDBatch {
    DDJob job[kMaxDDJobs];
    int n;
    float one;  // 1.0f, kept opaque to the compiler (f16 kernel)
};

__device__ __forceinline__ const DDJob& find_job(const DDBatch& B, int& local)
{
    int jb = 0;
    for (int k = 1; k < B.n; ++k)
        if ((int)blockIdx.x >= B.job[k].blockStart) jb = k;
    local = (int)blockIdx.x - B.job[jb].blockStart;
    return B.job[jb];
}

template <typename Tin, int ND, int NT>
__global__ __launch_bounds__(256) void dedisperse_kernel(const DDBatch B)
{
    using TA = typename Acc<Tin>::type;
    int local;
    const DDJob& J = find_job(B, local);
    const Tin* __restrict__ in = static_cast<const Tin*>(J.in);
    TA* __restrict__ out = static_cast<TA*>(J.out);
    const int32_t* __restrict__ delayT = J.delayT;
    const int nChan = J.nChan, nsamp = J.nsamp, tdms = J.tdms, tdmsPad = J.tdmsPad, nOut = J.nOut;
    const int W = J.rowUnits, CB = J.CB, nDMTiles = J.nDMTiles;
    constexpr int TT = 64 * NT;
    constexpr int TD = 4 * ND;
    constexpr int E = (int)sizeof(Tin);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // layout: rows[2][CB][W] elements (W a multiple of 4 bytes), offs[2][CB][TD]
    const int rowBytes = W * E;
    Tin* rows0 = reinterpret_cast<Tin*>(smem_raw);
    int* offs0 = reinterpret_cast<int*>(smem_raw + 2 * CB * rowBytes);

    int tTile, dmTile;
    tile_of(local, nDMTiles, tTile, dmTile);
    const int t0 = tTile * TT;
    if (t0 >= nOut) return;  // padding of the XCD-major grid
    const int dm0 = dmTile * TD;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const size_t totalBytes = (size_t)nChan * nsamp * E;
    const unsigned char* inBytes = reinterpret_cast<const unsigned char*>(in);

    TA acc[ND][NT];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int i = 0; i < NT; ++i) acc[j][i] = TA(0);

    // stage channels [cb0, cb0+nc) into buffer `buf`: wave w takes channels
    // w, w+4, ... (CB <= 16: at most four).  Loads complete in issue order,
    // so the delays of all of a wave's channels are loaded and used before
    // its first DMA: a delay load issued after a DMA would hold its user
    // behind that DMA.  Rows are whole 64-dword DMA rows, so only a row that
    // reaches the end of the input needs per-lane guards.
    auto stage = [&](int buf, int cb0, int nc) {
        Tin* rows = rows0 + (size_t)buf * CB * W;
        int* offs = offs0 + buf * CB * TD;
        const int rowDwords = rowBytes / 4;
        int dmin[4], dj[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = min(cb0 + min(wave + 4 * i, CB - 1), nChan - 1);
            const int32_t* dl = delayT + (size_t)c * tdmsPad + dm0;
            dmin[i] = __builtin_amdgcn_readfirstlane(dl[0]);
            dj[i] = dl[min(lane, TD - 1)];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int cc = wave + 4 * i;
            if (cc >= nc) break;
            const size_t byteStart = ((size_t)(cb0 + cc) * nsamp + (size_t)t0 + (size_t)dmin[i]) * E;
            const int mis = (int)(byteStart & 3) / E;
            if (lane < TD) offs[cc * TD + lane] = dj[i] - dmin[i] + mis + cc * W;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int cc = wave + 4 * i;
            if (cc >= nc) break;
            const size_t byteStart = ((size_t)(cb0 + cc) * nsamp + (size_t)t0 + (size_t)dmin[i]) * E;
            const size_t byte0 = byteStart & ~(size_t)3;
            unsigned char* dstRow = reinterpret_cast<unsigned char*>(rows + cc * W);
            if (byte0 + 4 * (size_t)rowDwords <= totalBytes) {
                const unsigned char* g = inBytes + byte0 + 4 * lane;
                for (int k0 = 0; k0 < rowDwords; k0 += 64)
                    __builtin_amdgcn_global_load_lds(g + 4 * k0, dstRow + 4 * k0, 4, 0, 0);
            } else {
                for (int k0 = 0; k0 < rowDwords; k0 += 64) {
                    const size_t g = byte0 + 4 * (size_t)(k0 + lane);
                    if (k0 + lane < rowDwords && g + 4 <= totalBytes)
                        __builtin_amdgcn_global_load_lds(inBytes + g, dstRow + 4 * k0, 4, 0, 0);
                }
            }
        }
    };

    const int nBlocks = (nChan + CB - 1) / CB;
    stage(0, 0, min(CB, nChan));
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int blk = 0; blk < nBlocks; ++blk) {
        const int buf = blk & 1;
        const int cb0 = blk * CB;
        const int nc = min(CB, nChan - cb0);
        if (blk + 1 < nBlocks) stage(buf ^ 1, cb0 + CB, min(CB, nChan - cb0 - CB));
        const Tin* rows = rows0 + (size_t)buf * CB * W;
        const int* offs = offs0 + buf * CB * TD + wave * ND;
        for (int cc = 0; cc < nc; ++cc) {
            int o[ND];
#pragma unroll
            for (int j = 0; j < ND; ++j) o[j] = offs[cc * TD + j];
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                const Tin* p = rows + o[j] + lane;
                Tin v[NT];
#pragma unroll
                for (int i = 0; i < NT; ++i) v[i] = p[64 * i];
#pragma unroll
                for (int i = 0; i < NT; ++i) acc[j][i] = acc[j][i] + (TA)v[i];
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    const int dmw = dm0 + wave * ND;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        const int dm = dmw + j;
        if (dm < tdms) {
#pragma unroll
            for (int i = 0; i < NT; ++i) {
                const int t = t0 + lane + 64 * i;
                if (t < nOut) out[(size_t)dm * nOut + t] = acc[j][i];
            }
        }
    }
}

// FP32 input, 16-byte LDS reads from a ring of staged rows.
// gfx950 serves ds_read_b128 at about twice the bytes per clock of
// ds_read_b32 (228 against 114 B/clk/CU measured), so lane l accumulates
// four consecutive samples t0 + 4l + 256q + {0,1,2,3} (q < NQ) from one
// read.  A read must be 16-byte aligned while the delays are arbitrary, so
// each staged channel row is held in four copies shifted by 0..3 elements:
// the read for a delay offset r (from the tile's first trial) is the
// aligned float4 (r >> 2) + lane of copy (r & 3).  The sums are still taken
// in channel order, one FP32 add per sample, so the result is bit-identical
// to the one-sample kernel.
//
// One channel per ring stage, DEPTH stages.  Wave s loads copy s & 3 of the
// row (from element s & 3 on: the same cache lines for all four; with eight
// waves, wave s takes half s >> 2 of it) by LDS DMA, always the same number
// of dwords (twelve or eight DMA instructions).  So every wave has a fixed
// count of DMA instructions per channel in flight, and at the top of
// channel c it waits (explicit vmcnt) only for its DMA into c's stage; the
// workgroup then meets at a bare barrier (every wave's part of the stage
// has landed, every wave is done with channel c-1), and the DMA of channel
// c + DEPTH - 1 goes into the stage channel c-1 used.  The first-trial delay
// that places a row and the wave's ND read offsets of the channel (from a
// per-plan ring table, already in LDS bytes) come by scalar loads
// (lgkmcnt), a channel ahead, so LDS serves only the rows.  No other vector
// memory operation is issued in the channel loop.  The
// compiler's own fence before __syncthreads (a wait for every DMA in
// flight) is kept out with an s_barrier in inline assembly, which also
// keeps the compiler from moving LDS accesses across it.  Row addresses
// are scalar (the DMA's scalar base plus a fixed lane offset) and the loop
// is unrolled by DEPTH, so stage addresses are constants: the vector
// instructions per channel and wave are the 64 adds and 8 read-address adds
// alone.  LDS is the bound: the 16-byte reads and the DMA writes of the four
// copies (a timing variant staging one copy ran 17 % faster), which is why
// the eight-wave form shares each staged row between twice the trials.
// Four-wave workgroups (tiles of 16 and 32 trials) hold 192 float4 per copy
// of a staged row; eight-wave workgroups (64 trials, whose delay spread is
// twice as wide) hold 256, two waves to a copy.
constexpr int ring_quads(int waves) { return waves == 8 ? 256 : 192; }
constexpr int ring_quads_of_tile(int TD) { return ring_quads(TD == 64 ? 8 : 4); }
// Three stages (two channels of DMA lead): 37 KB of LDS a four-wave
// workgroup, so four of them (16 waves) share a CU; four stages (49 KB,
// three workgroups) measured 10 % slower on the LOFAR beam.  An eight-wave
// workgroup takes 49 KB, two to a CU.
#ifndef PPL_RING_DEPTH
#define PPL_RING_DEPTH 3
#endif

// s_waitcnt bounding only the vector memory counter (gfx9 encoding:
// vmcnt[3:0] in bits 3:0, vmcnt[5:4] in bits 15:14; expcnt and lgkmcnt at
// their maxima, i.e. not waited for).
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt out of range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

__device__ __forceinline__ void bare_barrier() { asm volatile("s_barrier" ::: "memory"); }

// Twelve LDS-DMA dword loads of consecutive 64-dword rows: row k from
// global address g + 256k to LDS address lds + 256k (+ 4 * lane each).  The
// address is the uniform row start g in scalar registers plus the lane's
// byte offset (4 * lane, a loop-invariant register), so placing a row costs
// no vector instruction.  The instruction offset applies to both addresses,
// so one address and one M0 serve all of them.  M0 is restored before the
// block ends.  The compiler's hazard checks do not look into inline
// assembly, so the wait state between the M0 write and the first DMA is
// written out.
__device__ __forceinline__ void dma_rows12(const float* g, uint32_t laneBytes, uint32_t lds)
{
    uint32_t saved;
    asm volatile("s_mov_b32 %0, m0\n\t"
                 "s_mov_b32 m0, %3\n\t"
                 "s_nop 0\n\t"  // a scalar write of M0 needs one wait state before an LDS DMA reads it
                 "global_load_lds_dword %1, %2\n\t"
                 "global_load_lds_dword %1, %2 offset:256\n\t"
                 "global_load_lds_dword %1, %2 offset:512\n\t"
                 "global_load_lds_dword %1, %2 offset:768\n\t"
                 "global_load_lds_dword %1, %2 offset:1024\n\t"
                 "global_load_lds_dword %1, %2 offset:1280\n\t"
                 "global_load_lds_dword %1, %2 offset:1536\n\t"
                 "global_load_lds_dword %1, %2 offset:1792\n\t"
                 "global_load_lds_dword %1, %2 offset:2048\n\t"
                 "global_load_lds_dword %1, %2 offset:2304\n\t"
                 "global_load_lds_dword %1, %2 offset:2560\n\t"
                 "global_load_lds_dword %1, %2 offset:2816\n\t"
                 "s_nop 0\n\t"
                 "s_mov_b32 m0, %0"
                 : "=&s"(saved)
                 : "v"(laneBytes), "s"(g), "s"(lds)
                 : "memory");
}

// The same for eight rows (a half copy of an eight-wave stage).
__device__ __forceinline__ void dma_rows8(const float* g, uint32_t laneBytes, uint32_t lds)
{
    uint32_t saved;
    asm volatile("s_mov_b32 %0, m0\n\t"
                 "s_mov_b32 m0, %3\n\t"
                 "s_nop 0\n\t"  // a scalar write of M0 needs one wait state before an LDS DMA reads it
                 "global_load_lds_dword %1, %2\n\t"
                 "global_load_lds_dword %1, %2 offset:256\n\t"
                 "global_load_lds_dword %1, %2 offset:512\n\t"
                 "global_load_lds_dword %1, %2 offset:768\n\t"
                 "global_load_lds_dword %1, %2 offset:1024\n\t"
                 "global_load_lds_dword %1, %2 offset:1280\n\t"
                 "global_load_lds_dword %1, %2 offset:1536\n\t"
                 "global_load_lds_dword %1, %2 offset:1792\n\t"
                 "s_nop 0\n\t"
                 "s_mov_b32 m0, %0"
                 : "=&s"(saved)
                 : "v"(laneBytes), "s"(g), "s"(lds)
                 : "memory");
}

// LDS address (byte offset in the workgroup's LDS) of a __shared__ pointer.
__device__ __forceinline__ uint32_t lds_addr(const void* p)
{
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// Reads of the staged rows in inline assembly, each block waiting for its
// own reads (lgkmcnt(0)) before it ends, so no read is in flight when the
// compiler takes its registers.  The compiler's wait tracking cannot
// follow the four stages' DMA through the loop and would wait for every
// DMA in flight before the first read of a loop iteration.
__device__ __forceinline__ void lds_read_4x1(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, float4 (&v)[8])
{
    asm volatile("ds_read_b128 %0, %4\n\t"
                 "ds_read_b128 %1, %5\n\t"
                 "ds_read_b128 %2, %6\n\t"
                 "ds_read_b128 %3, %7\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3])
                 : "v"(a0), "v"(a1), "v"(a2), "v"(a3)
                 : "memory");
}
__device__ __forceinline__ void lds_read_4x2(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, float4 (&v)[8])
{
    asm volatile("ds_read_b128 %0, %8\n\t"
                 "ds_read_b128 %1, %8 offset:1024\n\t"
                 "ds_read_b128 %2, %9\n\t"
                 "ds_read_b128 %3, %9 offset:1024\n\t"
                 "ds_read_b128 %4, %10\n\t"
                 "ds_read_b128 %5, %10 offset:1024\n\t"
                 "ds_read_b128 %6, %11\n\t"
                 "ds_read_b128 %7, %11 offset:1024\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]), "=&v"(v[6]),
                   "=&v"(v[7])
                 : "v"(a0), "v"(a1), "v"(a2), "v"(a3)
                 : "memory");
}

// WAVES = 4 or 8 waves a workgroup, each summing ND trials: a tile of
// TD = WAVES * ND trials shares every staged row, so eight waves stage a
// row once for twice the trials.
template <int ND, int NQ, int DEPTH, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void dedisperse_ring_kernel(const DDBatch B)
{
    static_assert(WAVES == 4 || WAVES == 8, "four or eight waves");
    constexpr int QUADS = ring_quads(WAVES);
    constexpr int PARTS = WAVES / 4;                // waves per copy
    constexpr int PART_DW = 4 * QUADS / PARTS;      // dwords of a copy each of them stages
    constexpr int DMA = PART_DW / 64;               // DMA instructions per wave and channel
    static_assert(DMA == 12 || DMA == 8, "twelve or eight rows");
    static_assert(ND <= 16, "a wave's offsets fill 64 bytes");
    static_assert(DEPTH >= 3 && (DEPTH - 2) * DMA < 64, "ring depth");
    // The stages are read only by inline assembly and written only by DMA,
    // so one array holds them all.
    __shared__ __align__(16) unsigned char ring[DEPTH][4 * QUADS * 16];
    int local;
    const DDJob& J = find_job(B, local);
    const float* __restrict__ in = static_cast<const float*>(J.in);
    float* __restrict__ out = static_cast<float*>(J.out);
    const int32_t* __restrict__ tab = J.delayT;  // the plan's ring table for this DM tile width
    const int nChan = J.nChan, nsamp = J.nsamp, tdms = J.tdms, nOut = J.nOut;
    const int nDMTiles = J.nDMTiles;
    constexpr int TT = 256 * NQ;
    constexpr int TD = WAVES * ND;

    int tTile, dmTile;
    tile_of(local, nDMTiles, tTile, dmTile);
    const int t0 = tTile * TT;
    if (t0 >= nOut) return;  // padding of the XCD-major grid
    const int dm0 = dmTile * TD;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int copy = wave & 3, part = wave >> 2;  // the copy and the part of it this wave stages
    const int lastElem = nChan * nsamp - 1;  // make_job checks that it fits an int
    // ring table: offs[tile][c][TD] (LDS byte offset of each trial's reads in
    // a stage), then dmin[tile][c] (the tile's first-trial delay).  Both are
    // read through the constant address space: scalar loads, counted by
    // lgkmcnt, not by the DMA's vmcnt, issued a channel ahead.  A wave's ND
    // offsets of a channel are one load; consecutive channels' first-trial
    // delays share a scalar cache line.
    using ConstI32 = const __attribute__((address_space(4))) int32_t;
    ConstI32* offsTab = (ConstI32*)(tab + ((size_t)dmTile * nChan) * TD + wave * ND);
    ConstI32* dminTab = (ConstI32*)(tab + (size_t)nDMTiles * nChan * TD + (size_t)dmTile * nChan);

    float acc[ND][NQ][4];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[j][q][k] = 0.0f;

    // channel c into a stage: part `part` of copy `copy` of the row (DMA
    // instructions, each lane's address clamped to the input rather than
    // masked, so the count is fixed).  Past the last channel: a dummy stage
    // (the last channel again), so that every channel of the loop issues one.
    const uint32_t laneBytes = 4 * lane;
    auto stage = [&](unsigned char* buf, int c, int dmin) {
        c = min(c, nChan - 1);
        const int row0 = c * nsamp + t0 + dmin + copy + part * PART_DW;  // uniform
        float* dst = reinterpret_cast<float*>(buf + copy * QUADS * 16) + part * PART_DW;
        if (row0 + PART_DW - 1 <= lastElem) {
            if constexpr (DMA == 12)
                dma_rows12(in + row0, laneBytes, lds_addr(dst));
            else
                dma_rows8(in + row0, laneBytes, lds_addr(dst));
        } else {
#pragma unroll
            for (int k = 0; k < DMA; ++k) {
                const int rowk = __builtin_amdgcn_readfirstlane(row0 + 64 * k);  // uniform
                __builtin_amdgcn_global_load_lds(in + min(rowk + lane, lastElem), dst + 64 * k, 4, 0, 0);
            }
        }
    };
    auto compute = [&](const unsigned char* buf, const int (&o)[ND]) {
        static_assert((NQ == 1 || NQ == 2) && ND % 4 == 0, "reads in blocks of four trials x NQ float4");
        const uint32_t lb = lds_addr(buf) + 16 * lane;
#pragma unroll
        for (int j0 = 0; j0 < ND; j0 += 4) {
            float4 v[8];
            if constexpr (NQ == 2)
                lds_read_4x2(lb + o[j0], lb + o[j0 + 1], lb + o[j0 + 2], lb + o[j0 + 3], v);
            else
                lds_read_4x1(lb + o[j0], lb + o[j0 + 1], lb + o[j0 + 2], lb + o[j0 + 3], v);
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const float4 x = v[NQ * jj + q];
                    acc[j0 + jj][q][0] = acc[j0 + jj][q][0] + x.x;
                    acc[j0 + jj][q][1] = acc[j0 + jj][q][1] + x.y;
                    acc[j0 + jj][q][2] = acc[j0 + jj][q][2] + x.z;
                    acc[j0 + jj][q][3] = acc[j0 + jj][q][3] + x.w;
                }
        }
    };
    auto dmin_of = [&](int c) { return dminTab[min(c, nChan - 1)]; };
    auto offs_of = [&](int c, int (&o)[ND]) {
        const int cc = min(c, nChan - 1);
#pragma unroll
        for (int j = 0; j < ND; ++j) o[j] = offsTab[cc * TD + j];
    };

    // prologue: channels 0 .. DEPTH-2 into their stages
#pragma unroll
    for (int k = 0; k < DEPTH - 1; ++k) stage(ring[k], k, dmin_of(k));
    // dmin of the channel staged at the next top, loaded a channel ahead
    int nextDmin = dmin_of(DEPTH - 1);
    int oCur[ND];
    offs_of(0, oCur);
    // channel c uses stage c % DEPTH: the loop is unrolled by DEPTH so that
    // every stage address is a constant
    for (int c0 = 0; c0 < nChan; c0 += DEPTH) {
#pragma unroll
        for (int slot = 0; slot < DEPTH; ++slot) {
            const int c = c0 + slot;
            if (c < nChan) {  // uniform
                // this wave's DMA into channel c's stage has landed (the DMA of
                // the next DEPTH-2 stages may be in flight); then the barrier:
                // every wave's part has landed and every wave is done with
                // channel c-1
                wait_vmcnt<(DEPTH - 2) * DMA>();
                bare_barrier();
                stage(ring[(slot + DEPTH - 1) % DEPTH], c + DEPTH - 1, nextDmin);  // channel c-1's stage
                compute(ring[slot], oCur);
                // the next channel's scalars, landing while the wave waits at
                // the next barrier
                nextDmin = dmin_of(c + DEPTH);
                offs_of(c + 1, oCur);
            }
        }
    }
    // no LDS DMA may land after the workgroup has ended
    wait_vmcnt<0>();

    const int dmw = dm0 + wave * ND;
    const bool vec = (nOut & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        const int dm = dmw + j;
        if (dm < tdms) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int t = t0 + 4 * lane + 256 * q;
                float* o = out + (size_t)dm * nOut + t;
                if (vec && t + 3 < nOut) {
                    *reinterpret_cast<float4*>(o) = make_float4(acc[j][q][0], acc[j][q][1], acc[j][q][2], acc[j][q][3]);
                } else {
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (t + k < nOut) o[k] = acc[j][q][k];
                }
            }
        }
    }
}

// FP16 input, packed pairs.  Lane l accumulates the sample pairs
// (t0 + 2l + 128i, t0 + 2l + 128i + 1): one dword LDS read feeds two adds,
// each a v_fma_mix_f32 of a half straight from the dword (fma(h, one, acc)
// with one == 1.0 rounds once, exactly like h + acc).  A pair starting at an
// odd element is read from a second copy of the row shifted by one element,
// built in LDS with alignbyte from the staged dwords.
template <int ND, int NP>
__global__ __launch_bounds__(256) void dedisperse_f16_kernel(const DDBatch B)
{
    int local;
    const DDJob& J = find_job(B, local);
    const _Float16* __restrict__ in = static_cast<const _Float16*>(J.in);
    float* __restrict__ out = static_cast<float*>(J.out);
    const int32_t* __restrict__ delayT = J.delayT;
    const int nChan = J.nChan, nsamp = J.nsamp, tdms = J.tdms, tdmsPad = J.tdmsPad, nOut = J.nOut;
    const int rowDw = J.rowUnits, CB = J.CB, nDMTiles = J.nDMTiles;
    const float one = B.one;
    constexpr int TT = 128 * NP;
    constexpr int TD = 4 * ND;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // per buffer: rows[CB][2][rowDw] dwords (copy A, copy B), then offs[CB][TD]
    uint32_t* lds32 = reinterpret_cast<uint32_t*>(smem_raw);
    const int bufDw = CB * (2 * rowDw + TD);

    int tTile, dmTile;
    tile_of(local, nDMTiles, tTile, dmTile);
    const int t0 = tTile * TT;
    if (t0 >= nOut) return;  // padding of the XCD-major grid
    const int dm0 = dmTile * TD;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const size_t totalBytes = (size_t)nChan * nsamp * 2;
    const unsigned char* inBytes = reinterpret_cast<const unsigned char*>(in);

    float acc[ND][2 * NP];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int i = 0; i < 2 * NP; ++i) acc[j][i] = 0.0f;

    auto rowsOf = [&](int buf) { return lds32 + (size_t)buf * bufDw; };
    auto offsOf = [&](int buf) { return reinterpret_cast<int*>(lds32 + (size_t)buf * bufDw + CB * 2 * rowDw); };
    // wave w stages channels w, w+4, ... (CB <= 16: at most four), their
    // delays loaded and used before the first DMA (see the FP32 kernel)
    auto stage = [&](int buf, int cb0, int nc) {
        uint32_t* rows = rowsOf(buf);
        int* offs = offsOf(buf);
        int dmin[4], dj[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = min(cb0 + min(wave + 4 * i, CB - 1), nChan - 1);
            const int32_t* dl = delayT + (size_t)c * tdmsPad + dm0;
            dmin[i] = __builtin_amdgcn_readfirstlane(dl[0]);
            dj[i] = dl[min(lane, TD - 1)];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int cc = wave + 4 * i;
            if (cc >= nc) break;
            const size_t byteStart = ((size_t)(cb0 + cc) * nsamp + (size_t)t0 + (size_t)dmin[i]) * 2;
            const int mis = (int)(byteStart & 3) / 2;
            if (lane < TD) {
                const int o = dj[i] - dmin[i] + mis;  // element offset in the row
                offs[cc * TD + lane] = cc * 2 * rowDw + ((o & 1) ? rowDw : 0) + (o >> 1);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int cc = wave + 4 * i;
            if (cc >= nc) break;
            const size_t byteStart = ((size_t)(cb0 + cc) * nsamp + (size_t)t0 + (size_t)dmin[i]) * 2;
            const size_t byte0 = byteStart & ~(size_t)3;
            unsigned char* dst = reinterpret_cast<unsigned char*>(rows + (size_t)cc * 2 * rowDw);
            if (byte0 + 4 * (size_t)rowDw <= totalBytes) {  // whole 64-dword rows
                const unsigned char* g = inBytes + byte0 + 4 * lane;
                for (int k0 = 0; k0 < rowDw; k0 += 64)
                    __builtin_amdgcn_global_load_lds(g + 4 * k0, dst + 4 * k0, 4, 0, 0);
            } else {
                for (int k0 = 0; k0 < rowDw; k0 += 64) {
                    const size_t g = byte0 + 4 * (size_t)(k0 + lane);
                    if (k0 + lane < rowDw && g + 4 <= totalBytes) {
                        __builtin_amdgcn_global_load_lds(inBytes + g, dst + 4 * k0, 4, 0, 0);
                    } else if (k0 + lane < rowDw && g < totalBytes) {
                        // the input's last, partial dword (one half)
                        uint32_t v = 0u;
                        for (size_t b = 0; g + b < totalBytes; ++b) v |= (uint32_t)inBytes[g + b] << (8 * b);
                        reinterpret_cast<uint32_t*>(dst)[k0 + lane] = v;
                    }
                }
            }
        }
    };
    // copy B: dword m holds elements 2m+1, 2m+2 of the row
    auto shift_copy = [&](int buf, int nc) {
        uint32_t* rows = rowsOf(buf);
        for (int cc = wave; cc < nc; cc += 4) {
            uint32_t* A = rows + (size_t)cc * 2 * rowDw;
            for (int m = lane; m < rowDw - 1; m += 64) A[rowDw + m] = __builtin_amdgcn_alignbyte(A[m + 1], A[m], 2);
        }
    };

    const int nBlocks = (nChan + CB - 1) / CB;
    stage(0, 0, min(CB, nChan));
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    shift_copy(0, min(CB, nChan));
    __syncthreads();
    for (int blk = 0; blk < nBlocks; ++blk) {
        const int buf = blk & 1;
        const int cb0 = blk * CB;
        const int nc = min(CB, nChan - cb0);
        const bool more = blk + 1 < nBlocks;
        if (more) stage(buf ^ 1, cb0 + CB, min(CB, nChan - cb0 - CB));
        const uint32_t* rows = rowsOf(buf);
        const int* offs = offsOf(buf) + wave * ND;
        for (int cc = 0; cc < nc; ++cc) {
            int o[ND];
#pragma unroll
            for (int j = 0; j < ND; ++j) o[j] = offs[cc * TD + j];
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                const uint32_t* p = rows + o[j] + lane;
                uint32_t v[NP];
#pragma unroll
                for (int i = 0; i < NP; ++i) v[i] = p[64 * i];
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    const _Float16 lo = __builtin_bit_cast(_Float16, (unsigned short)(v[i] & 0xffffu));
                    const _Float16 hi = __builtin_bit_cast(_Float16, (unsigned short)(v[i] >> 16));
                    acc[j][2 * i] = __builtin_fmaf((float)lo, one, acc[j][2 * i]);
                    acc[j][2 * i + 1] = __builtin_fmaf((float)hi, one, acc[j][2 * i + 1]);
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (more) {
            shift_copy(buf ^ 1, min(CB, nChan - cb0 - CB));
            __syncthreads();
        }
    }
    const int dmw = dm0 + wave * ND;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        const int dm = dmw + j;
        if (dm < tdms) {
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const int t = t0 + 2 * lane + 128 * i;
                float* o = out + (size_t)dm * nOut + t;
                if (t + 1 < nOut) {
                    o[0] = acc[j][2 * i];
                    o[1] = acc[j][2 * i + 1];
                } else if (t < nOut) {
                    o[0] = acc[j][2 * i];
                }
            }
        }
    }
}

// u8 input, int32 sums, four samples per dword.  Lane l accumulates samples
// t0 + 4l + 256i + {0,1,2,3}.  The row is held in LDS in four copies shifted
// by 0..3 bytes (built with alignbyte), so every (channel, DM) read is one
// aligned dword.  Bytes 0/2 and 1/3 of the dword are split into the 16-bit
// halves of two 32-bit accumulators (and / perm) and two channels are added
// at once (add3); the halves are flushed into int32 sums every 256 channels,
// before they can overflow.  Integer sums are exact in any order.
template <int ND, int NQ>
__global__ __launch_bounds__(256) void dedisperse_u8_kernel(const DDBatch B)
{
    int local;
    const DDJob& J = find_job(B, local);
    const uint8_t* __restrict__ in = static_cast<const uint8_t*>(J.in);
    int32_t* __restrict__ out = static_cast<int32_t*>(J.out);
    const int32_t* __restrict__ delayT = J.delayT;
    const int nChan = J.nChan, nsamp = J.nsamp, tdms = J.tdms, tdmsPad = J.tdmsPad, nOut = J.nOut;
    const int rowDw = J.rowUnits, CB = J.CB, nDMTiles = J.nDMTiles;
    constexpr int TT = 256 * NQ;
    constexpr int TD = 4 * ND;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // per buffer: rows[CB][4][rowDw] dwords, then offs[CB][TD]
    uint32_t* lds32 = reinterpret_cast<uint32_t*>(smem_raw);
    const int bufDw = CB * (4 * rowDw + TD);

    int tTile, dmTile;
    tile_of(local, nDMTiles, tTile, dmTile);
    const int t0 = tTile * TT;
    if (t0 >= nOut) return;  // padding of the XCD-major grid
    const int dm0 = dmTile * TD;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const size_t totalBytes = (size_t)nChan * nsamp;

    int32_t acc[ND][NQ][4];
    uint32_t h02[ND][NQ], h13[ND][NQ];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            h02[j][i] = h13[j][i] = 0u;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[j][i][q] = 0;
        }
    auto flush = [&]() {
#pragma unroll
        for (int j = 0; j < ND; ++j)
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                acc[j][i][0] += (int32_t)(h02[j][i] & 0xffffu);
                acc[j][i][2] += (int32_t)(h02[j][i] >> 16);
                acc[j][i][1] += (int32_t)(h13[j][i] & 0xffffu);
                acc[j][i][3] += (int32_t)(h13[j][i] >> 16);
                h02[j][i] = h13[j][i] = 0u;
            }
    };

    auto rowsOf = [&](int buf) { return lds32 + (size_t)buf * bufDw; };
    auto offsOf = [&](int buf) { return reinterpret_cast<int*>(lds32 + (size_t)buf * bufDw + CB * 4 * rowDw); };
    // wave w stages channels w, w+4, ... (CB <= 16: at most four), their
    // delays loaded and used before the first DMA (see the FP32 kernel)
    auto stage = [&](int buf, int cb0, int nc) {
        uint32_t* rows = rowsOf(buf);
        int* offs = offsOf(buf);
        int dmin[4], dj[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = min(cb0 + min(wave + 4 * i, CB - 1), nChan - 1);
            const int32_t* dl = delayT + (size_t)c * tdmsPad + dm0;
            dmin[i] = __builtin_amdgcn_readfirstlane(dl[0]);
            dj[i] = dl[min(lane, TD - 1)];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int cc = wave + 4 * i;
            if (cc >= nc) break;
            const size_t byteStart = (size_t)(cb0 + cc) * nsamp + (size_t)t0 + (size_t)dmin[i];
            const int mis = (int)(byteStart & 3);
            if (lane < TD) {
                const int o = dj[i] - dmin[i] + mis;  // byte offset in the row
                offs[cc * TD + lane] = cc * 4 * rowDw + (o & 3) * rowDw + (o >> 2);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int cc = wave + 4 * i;
            if (cc >= nc) break;
            const size_t byteStart = (size_t)(cb0 + cc) * nsamp + (size_t)t0 + (size_t)dmin[i];
            const size_t byte0 = byteStart & ~(size_t)3;
            unsigned char* dst = reinterpret_cast<unsigned char*>(rows + (size_t)cc * 4 * rowDw);
            if (byte0 + 4 * (size_t)rowDw <= totalBytes) {  // whole 64-dword rows
                const uint8_t* g = in + byte0 + 4 * lane;
                for (int k0 = 0; k0 < rowDw; k0 += 64)
                    __builtin_amdgcn_global_load_lds(g + 4 * k0, dst + 4 * k0, 4, 0, 0);
            } else {
                for (int k0 = 0; k0 < rowDw; k0 += 64) {
                    const size_t g = byte0 + 4 * (size_t)(k0 + lane);
                    if (k0 + lane < rowDw && g + 4 <= totalBytes) {
                        __builtin_amdgcn_global_load_lds(in + g, dst + 4 * k0, 4, 0, 0);
                    } else if (k0 + lane < rowDw && g < totalBytes) {
                        // the input's last, partial dword
                        uint32_t v = 0u;
                        for (size_t b = 0; g + b < totalBytes; ++b) v |= (uint32_t)in[g + b] << (8 * b);
                        reinterpret_cast<uint32_t*>(dst)[k0 + lane] = v;
                    }
                }
            }
        }
    };
    auto shift_copies = [&](int buf, int nc) {
        uint32_t* rows = rowsOf(buf);
        for (int cc = wave; cc < nc; cc += 4) {
            uint32_t* A = rows + (size_t)cc * 4 * rowDw;
            for (int m = lane; m < rowDw - 1; m += 64) {
                const uint32_t a0 = A[m], a1 = A[m + 1];
                A[rowDw + m] = __builtin_amdgcn_alignbyte(a1, a0, 1);
                A[2 * rowDw + m] = __builtin_amdgcn_alignbyte(a1, a0, 2);
                A[3 * rowDw + m] = __builtin_amdgcn_alignbyte(a1, a0, 3);
            }
        }
    };

    const int nBlocks = (nChan + CB - 1) / CB;
    stage(0, 0, min(CB, nChan));
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    shift_copies(0, min(CB, nChan));
    __syncthreads();
    int sinceFlush = 0;
    for (int blk = 0; blk < nBlocks; ++blk) {
        const int buf = blk & 1;
        const int cb0 = blk * CB;
        const int nc = min(CB, nChan - cb0);
        const bool more = blk + 1 < nBlocks;
        if (more) stage(buf ^ 1, cb0 + CB, min(CB, nChan - cb0 - CB));
        const uint32_t* rows = rowsOf(buf);
        const int* offs = offsOf(buf) + wave * ND;
        int cc = 0;
        for (; cc + 1 < nc; cc += 2) {
            if (sinceFlush >= 254) {
                flush();
                sinceFlush = 0;
            }
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                const uint32_t* p0 = rows + offs[cc * TD + j] + lane;
                const uint32_t* p1 = rows + offs[(cc + 1) * TD + j] + lane;
#pragma unroll
                for (int i = 0; i < NQ; ++i) {
                    const uint32_t v0 = p0[64 * i], v1 = p1[64 * i];
                    h02[j][i] = h02[j][i] + (v0 & 0x00ff00ffu) + (v1 & 0x00ff00ffu);
                    h13[j][i] = h13[j][i] + __builtin_amdgcn_perm(0u, v0, 0x0c030c01u) +
                                __builtin_amdgcn_perm(0u, v1, 0x0c030c01u);
                }
            }
            sinceFlush += 2;
        }
        if (cc < nc) {
            if (sinceFlush >= 255) {
                flush();
                sinceFlush = 0;
            }
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                const uint32_t* p0 = rows + offs[cc * TD + j] + lane;
#pragma unroll
                for (int i = 0; i < NQ; ++i) {
                    const uint32_t v0 = p0[64 * i];
                    h02[j][i] += (v0 & 0x00ff00ffu);
                    h13[j][i] += __builtin_amdgcn_perm(0u, v0, 0x0c030c01u);
                }
            }
            sinceFlush += 1;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (more) {
            shift_copies(buf ^ 1, min(CB, nChan - cb0 - CB));
            __syncthreads();
        }
    }
    flush();
    const int dmw = dm0 + wave * ND;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        const int dm = dmw + j;
        if (dm < tdms) {
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                const int t = t0 + 4 * lane + 256 * i;
                int32_t* o = out + (size_t)dm * nOut + t;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (t + q < nOut) o[q] = acc[j][i][q];
            }
        }
    }
}

// u8 input through a ring of staged rows, the u8 counterpart of the FP32
// ring kernel.  Lane l sums the dwords l + 64q (q < NQ) of the tile's row
// window, four samples t0 + 4(l + 64q) + {0,1,2,3} each, with the SWAR scheme
// of dedisperse_u8_kernel (bytes 0/2 and 1/3 into the 16-bit halves of two
// accumulators, two channels per v_add3_u32, widened to int32 every 254
// channels).  A stage holds a channel pair; each channel's row is one copy
// staged by LDS DMA from the dword at or below its first byte (so the DMA
// stays dword-aligned), then three copies shifted by 1..3 bytes built in LDS
// with alignbyte, so every read is one aligned dword.  The read offsets
// (copy and dword, with the row's misalignment folded in) and the first-trial
// delays come by scalar loads from a per-plan table, a pair ahead.  Per pair:
// wait for the DMA of pair p+1, a barrier (the shifted copies of pair p are
// complete, every wave is done with pair p-1), the DMA of pair p+2 into pair
// p-1's stage, the shifted copies of pair p+1, then the sums of pair p.
// Integer sums are exact in any order, so the output equals the other u8
// kernel's.
constexpr int kU8CopyDw = 256;                       // dwords per copy of a staged row
constexpr int kU8StageBytes = 2 * 4 * kU8CopyDw * 4;  // a channel pair, four copies each

// One LDS-DMA dword load from a per-lane address (a vector register pair)
// to lds + 4 * lane, M0 saved and restored as in dma_rows12.
__device__ __forceinline__ void dma_dword_lane(const void* g, uint32_t lds)
{
    uint32_t saved;
    asm volatile("s_mov_b32 %0, m0\n\t"
                 "s_mov_b32 m0, %2\n\t"
                 "s_nop 0\n\t"
                 "global_load_lds_dword %1, off\n\t"
                 "s_nop 0\n\t"
                 "s_mov_b32 m0, %0"
                 : "=&s"(saved)
                 : "v"(g), "s"(lds)
                 : "memory");
}

// One LDS-DMA dword load, scalar base plus the lane's byte offset.
__device__ __forceinline__ void dma_dword_s(const void* g, uint32_t laneBytes, uint32_t lds)
{
    uint32_t saved;
    asm volatile("s_mov_b32 %0, m0\n\t"
                 "s_mov_b32 m0, %3\n\t"
                 "s_nop 0\n\t"
                 "global_load_lds_dword %1, %2\n\t"
                 "s_nop 0\n\t"
                 "s_mov_b32 m0, %0"
                 : "=&s"(saved)
                 : "v"(laneBytes), "s"(g), "s"(lds)
                 : "memory");
}

// A barrier after this wave's LDS writes have landed.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Eight dword reads, two from each of four addresses (the second 256 bytes
// on: the q = 1 dword), waiting for them before the block ends.
__device__ __forceinline__ void lds_read4x2(const uint32_t (&a)[4], uint32_t (&v)[8])
{
    asm volatile("ds_read_b32 %0, %8\n\t"
                 "ds_read_b32 %1, %9\n\t"
                 "ds_read_b32 %2, %8 offset:256\n\t"
                 "ds_read_b32 %3, %9 offset:256\n\t"
                 "ds_read_b32 %4, %10\n\t"
                 "ds_read_b32 %5, %11\n\t"
                 "ds_read_b32 %6, %10 offset:256\n\t"
                 "ds_read_b32 %7, %11 offset:256\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]), "=&v"(v[6]),
                   "=&v"(v[7])
                 : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3])
                 : "memory");
}

// Eight dword reads from eight addresses, waiting for them before the block
// ends.
__device__ __forceinline__ void lds_read8(const uint32_t (&a)[8], uint32_t (&v)[8])
{
    asm volatile("ds_read_b32 %0, %8\n\t"
                 "ds_read_b32 %1, %9\n\t"
                 "ds_read_b32 %2, %10\n\t"
                 "ds_read_b32 %3, %11\n\t"
                 "ds_read_b32 %4, %12\n\t"
                 "ds_read_b32 %5, %13\n\t"
                 "ds_read_b32 %6, %14\n\t"
                 "ds_read_b32 %7, %15\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]), "=&v"(v[6]),
                   "=&v"(v[7])
                 : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]), "v"(a[7])
                 : "memory");
}

template <int ND, int NQ>
__global__ __launch_bounds__(256) void dedisperse_u8_ring_kernel(const DDBatch B)
{
    static_assert((NQ == 1 || NQ == 2) && ND % 4 == 0 && ND <= 16, "u8 ring tile");
    constexpr int DEPTH = 3;
    __shared__ __align__(16) unsigned char ring[DEPTH][kU8StageBytes];
    int local;
    const DDJob& J = find_job(B, local);
    const uint8_t* __restrict__ in = static_cast<const uint8_t*>(J.in);
    int32_t* __restrict__ out = static_cast<int32_t*>(J.out);
    const int32_t* __restrict__ tab = J.delayT;
    const int nChan = J.nChan, nsamp = J.nsamp, tdms = J.tdms, nOut = J.nOut;
    const int nDMTiles = J.nDMTiles;
    constexpr int TT = 256 * NQ;
    constexpr int TD = 4 * ND;

    int tTile, dmTile;
    tile_of(local, nDMTiles, tTile, dmTile);
    const int t0 = tTile * TT;
    if (t0 >= nOut) return;  // padding of the XCD-major grid
    const int dm0 = dmTile * TD;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lastDw = (nChan * nsamp - 1) >> 2;  // the input's last dword (make_job bounds the size)
    using ConstI32 = const __attribute__((address_space(4))) int32_t;
    ConstI32* offsTab = (ConstI32*)(tab + ((size_t)dmTile * nChan) * TD + wave * ND);
    ConstI32* dminTab = (ConstI32*)(tab + (size_t)nDMTiles * nChan * TD + (size_t)dmTile * nChan);
    const uint32_t* in32 = reinterpret_cast<const uint32_t*>(in);

    int32_t acc[ND][NQ][4];
    uint32_t h02[ND][NQ], h13[ND][NQ];
#pragma unroll
    for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            h02[j][q] = h13[j][q] = 0u;
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[j][q][k] = 0;
        }
    auto flush = [&]() {
#pragma unroll
        for (int j = 0; j < ND; ++j)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                acc[j][q][0] += (int32_t)(h02[j][q] & 0xffffu);
                acc[j][q][2] += (int32_t)(h02[j][q] >> 16);
                acc[j][q][1] += (int32_t)(h13[j][q] & 0xffffu);
                acc[j][q][3] += (int32_t)(h13[j][q] >> 16);
                h02[j][q] = h13[j][q] = 0u;
            }
    };

    // copy 0 of channel c (past the last channel: the last channel again, so
    // that every pair issues the same DMA) into half h of a stage: wave w
    // loads dwords 64w .. 64w+63 of the row, from the dword holding its first
    // byte (the misalignment is in the read offsets)
    const uint32_t laneBytes = 4 * lane;
    auto dmin_of = [&](int c) { return dminTab[min(c, nChan - 1)]; };
    auto stage_row = [&](unsigned char* buf, int h, int c, int dmin) {
        c = min(c, nChan - 1);
        const int first = c * nsamp + t0 + dmin;  // uniform byte index
        const int dw0 = (first >> 2) + 64 * wave;
        unsigned char* dst = buf + h * (4 * kU8CopyDw * 4) + 256 * wave;
        if (dw0 + 63 <= lastDw)
            dma_dword_s(in32 + dw0, laneBytes, lds_addr(dst));
        else
            dma_dword_lane(in32 + min(dw0 + lane, lastDw), lds_addr(dst));
    };
    auto stage_pair = [&](unsigned char* buf, int p, int dmin0, int dmin1) {
        stage_row(buf, 0, 2 * p, dmin0);
        stage_row(buf, 1, 2 * p + 1, dmin1);
    };
    // copies 1..3 of both rows of a stage from copy 0: dword m of copy s is
    // bytes 4m + s .. 4m + s + 3 of the row (thread t builds dword t of each
    // row's three copies; the last one reads one dword past copy 0, into
    // copy 1, which no read ever reaches).  A pair whose second channel is
    // past the last one gets zeros there (after its DMA has landed), so the
    // sums run the same for every pair.
    auto shift_pair = [&](unsigned char* buf, int p) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint32_t* row = reinterpret_cast<uint32_t*>(buf + h * (4 * kU8CopyDw * 4));
            if (h == 1 && 2 * p + 1 >= nChan) {  // uniform
#pragma unroll
                for (int k = 0; k < 4; ++k) row[k * kU8CopyDw + tid] = 0u;
            } else {
                const uint32_t a = row[tid], b = row[tid + 1];
                row[kU8CopyDw + tid] = __builtin_amdgcn_alignbyte(b, a, 1);
                row[2 * kU8CopyDw + tid] = __builtin_amdgcn_alignbyte(b, a, 2);
                row[3 * kU8CopyDw + tid] = __builtin_amdgcn_alignbyte(b, a, 3);
            }
        }
    };
    auto offs_of = [&](int c, int (&o)[ND]) {
        const int cc = min(c, nChan - 1);
#pragma unroll
        for (int j = 0; j < ND; ++j) o[j] = offsTab[cc * TD + j];
    };
    // the sums of a pair.  v[2 (jj NQ + q) + h] is trial j0 + jj, dword q,
    // row h.
    auto compute = [&](const unsigned char* buf, const int (&o0)[ND], const int (&o1)[ND]) {
        const uint32_t lb0 = lds_addr(buf) + 4 * lane;
        const uint32_t lb1 = lb0 + 4 * kU8CopyDw * 4;
        constexpr int JB = 4 / NQ;  // trials per block of eight reads
#pragma unroll
        for (int j0 = 0; j0 < ND; j0 += JB) {
            uint32_t v[8];
            if constexpr (NQ == 2) {
                const uint32_t a[4] = {lb0 + o0[j0], lb1 + o1[j0], lb0 + o0[j0 + 1], lb1 + o1[j0 + 1]};
                lds_read4x2(a, v);
            } else {
                const uint32_t a[8] = {lb0 + o0[j0],     lb1 + o1[j0],     lb0 + o0[j0 + 1], lb1 + o1[j0 + 1],
                                       lb0 + o0[j0 + 2], lb1 + o1[j0 + 2], lb0 + o0[j0 + 3], lb1 + o1[j0 + 3]};
                lds_read8(a, v);
            }
#pragma unroll
            for (int jj = 0; jj < JB; ++jj)
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const uint32_t x0 = v[2 * (jj * NQ + q)];
                    const uint32_t x1 = v[2 * (jj * NQ + q) + 1];
                    h02[j0 + jj][q] = h02[j0 + jj][q] + (x0 & 0x00ff00ffu) + (x1 & 0x00ff00ffu);
                    h13[j0 + jj][q] = h13[j0 + jj][q] + __builtin_amdgcn_perm(0u, x0, 0x0c030c01u) +
                                      __builtin_amdgcn_perm(0u, x1, 0x0c030c01u);
                }
        }
    };

    const int nPairs = (nChan + 1) / 2;
    // prologue: pairs 0 and 1 staged, pair 0's copies built
    stage_pair(ring[0], 0, dmin_of(0), dmin_of(1));
    stage_pair(ring[1], 1, dmin_of(2), dmin_of(3));
    wait_vmcnt<2>();  // pair 0 landed
    lds_barrier();
    shift_pair(ring[0], 0);
    int o0[ND], o1[ND];
    offs_of(0, o0);
    offs_of(1, o1);
    // first-trial delays of the pair staged next, loaded a pair ahead
    int nd0 = dmin_of(4), nd1 = dmin_of(5);
    int sinceFlush = 0;
    for (int p0 = 0; p0 < nPairs; p0 += DEPTH) {
#pragma unroll
        for (int slot = 0; slot < DEPTH; ++slot) {
            const int p = p0 + slot;
            if (p < nPairs) {  // uniform
                wait_vmcnt<0>();  // pair p+1 landed
                lds_barrier();    // pair p's copies built, every wave done with pair p-1
                stage_pair(ring[(slot + 2) % DEPTH], p + 2, nd0, nd1);
                // pair p+1's shifted copies (its copy 0 landed before the barrier)
                shift_pair(ring[(slot + 1) % DEPTH], p + 1);
                if (sinceFlush >= 254) {
                    flush();
                    sinceFlush = 0;
                }
                compute(ring[slot], o0, o1);
                sinceFlush += 2;
                // the next pair's scalars
                offs_of(2 * p + 2, o0);
                offs_of(2 * p + 3, o1);
                nd0 = dmin_of(2 * p + 6);
                nd1 = dmin_of(2 * p + 7);
            }
        }
    }
    wait_vmcnt<0>();  // no LDS DMA may land after the workgroup has ended
    flush();
    const int dmw = dm0 + wave * ND;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        const int dm = dmw + j;
        if (dm < tdms) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int t = t0 + 4 * lane + 256 * q;
                int32_t* o = out + (size_t)dm * nOut + t;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (t + k < nOut) o[k] = acc[j][q][k];
            }
        }
    }
}

// delayT[c][dm] = trunc(shift[c] * (mstartdm + dm*mdmstep)) for dm < tdms,
// rows padded with the last trial (compiled without contraction, so the
// product and sum round separately, as in the host table).
__global__ void build_delays_kernel(const float* __restrict__ shift, int nChan, int tdms,
                                    int tdmsPad, float mstartdm, float mdmstep,
                                    int32_t* __restrict__ delayT)
{
    const int dm = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (dm >= tdmsPad || c >= nChan) return;
    const int d = dm < tdms ? dm : tdms - 1;
    const float prod = (float)d * mdmstep;
    const float dmval = mstartdm + prod;
    const float p = shift[c] * dmval;
    delayT[(size_t)c * tdmsPad + dm] = (int32_t)p;
}

// The FP32 ring kernel's table for DM tiles of TD trials, from a device delay
// table [nChan][tdmsPad] (tdmsPad a multiple of TD, rows padded with the last
// trial): entry (tile k, channel c, trial j) is the LDS byte offset of the
// trial's reads, r = delay - first trial's delay, as copy (r & 3) of the
// stage at float4 r >> 2; then per (k, c) the first trial's delay.  The same
// layout the host builds from a plan's delays.
__global__ void ring_table_kernel(const int32_t* __restrict__ delayT, int nChan, int tdmsPad, int TD,
                                  int quads, int32_t* __restrict__ tab)
{
    const size_t n = (size_t)(tdmsPad / TD) * nChan * TD;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = (int)(i % TD);
    const size_t kc = i / TD;  // k * nChan + c
    const int c = (int)(kc % nChan), k = (int)(kc / nChan);
    const int32_t* row = delayT + (size_t)c * tdmsPad + (size_t)k * TD;
    const int32_t d0 = row[0];
    const int r = row[j] - d0;
    tab[i] = (r & 3) * quads * 16 + (r >> 2) * 16;
    if (j == 0) tab[n + kc] = d0;
}

struct Shape {
    int nd, nt;  // DM trials per wave, 64-sample groups per lane
    int cb;      // channels per staged block (0: as many as fit)
    int rw = 4;  // LDS read width per lane: 4 bytes, or 16 (FP32 input, four samples)
};

constexpr int kMaxLDS = 64 * 1024;

// LDS bytes per staged channel (both buffers) and the row length for a shape.
template <typename Tin>
void lds_geometry(const Shape& s, int span, int& rowUnits, int& perChannel)
{
    const int TT = 64 * s.nt, TD = 4 * s.nd;
    if constexpr (std::is_same<Tin, _Float16>::value) {
        // rowUnits = dwords per copy: (TT + span + misalignment) halves plus
        // the dword the shifted copy reads past the end
        rowUnits = ((TT + span + 1 + 1) / 2 + 1 + 63) & ~63;  // whole 64-dword DMA rows
        perChannel = 2 * 4 * (2 * rowUnits + TD);
    } else if constexpr (std::is_same<Tin, uint8_t>::value) {
        // four shifted copies of (TT + span + misalignment) bytes, in dwords
        rowUnits = ((TT + span + 3 + 3) / 4 + 1 + 63) & ~63;  // whole 64-dword DMA rows
        perChannel = 2 * 4 * (4 * rowUnits + TD);
    } else {
        // rowUnits = elements, a whole number of 64-dword DMA rows when an
        // element is a dword
        int W = TT + span + 4;
        W = sizeof(Tin) == 4 ? (W + 63) & ~63 : (W + 15) & ~15;
        rowUnits = W;
        perChannel = 2 * (W * (int)sizeof(Tin) + TD * 4);
    }
}

int span_of(const ppl_dedisp_plan* p, int TD)
{
    return p->spanTD[TD <= 8 ? 0 : TD <= 16 ? 1 : TD <= 32 ? 2 : 3];
}

int elem_bytes(int dtype) { return dtype == PPL_F32 ? 4 : dtype == PPL_F16 ? 2 : 1; }

// The ring kernel's table for DM tiles of TD trials (built and uploaded on
// first use; the plan keeps it): for tile k, channel c and trial j of the
// tile, the LDS byte offset of the trial's first read in a ring stage,
// (r & 3) * (one copy) + (r >> 2) * 16 with r = delay[k*TD + j][c] -
// delay[k*TD][c]; then for tile k and channel c, delay[k*TD][c].  Trials
// past tdms take the last trial's delays, as the padded delay table does.
int ring_table(const ppl_dedisp_plan* cp, int TD, const int32_t** out)
{
    ppl_dedisp_plan* p = const_cast<ppl_dedisp_plan*>(cp);
    const int slot = TD == 16 ? 0 : TD == 32 ? 1 : 2;
    const int quads = ring_quads_of_tile(TD);
    const bool u8 = p->cfg.dtype == PPL_U8;  // a plan has one input type, so its tables have one format
    if (!p->d_ring[slot]) {
        if (p->delays.empty()) {
            set_error("dedisperse: no host delay table for the 16-byte kernel");
            return PPL_E_ARG;
        }
        const int tdms = (int)p->cfg.tdms, nChan = (int)p->cfg.nChan;
        const int nTiles = (tdms + TD - 1) / TD;
        std::vector<int32_t> t((size_t)nTiles * nChan * (TD + 1));
        int32_t* dminPart = t.data() + (size_t)nTiles * nChan * TD;
        for (int k = 0; k < nTiles; ++k)
            for (int c = 0; c < nChan; ++c) {
                const int32_t d0 = p->delays[(size_t)(k * TD) * nChan + c];
                dminPart[(size_t)k * nChan + c] = d0;
                // u8: the row's first byte sits mis bytes into its first staged
                // dword (tiles start at multiples of 256 samples)
                const int mis = (int)(((long)c * p->cfg.nsampBuf + d0) & 3);
                for (int j = 0; j < TD; ++j) {
                    const int dm = std::min(k * TD + j, tdms - 1);
                    const int r = p->delays[(size_t)dm * nChan + c] - d0;
                    t[((size_t)k * nChan + c) * TD + j] =
                        u8 ? ((mis + r) & 3) * kU8CopyDw * 4 + ((mis + r) >> 2) * 4 : (r & 3) * quads * 16 + (r >> 2) * 16;
                }
            }
        int32_t* d = nullptr;
        PPL_HIP(hipMalloc(&d, t.size() * sizeof(int32_t)));
        const hipError_t e = hipMemcpy(d, t.data(), t.size() * sizeof(int32_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            set_error("dedisperse: ring table upload: %s", hipGetErrorString(e));
            return PPL_E_HIP;
        }
        p->d_ring[slot] = d;
    }
    *out = p->d_ring[slot];
    return PPL_OK;
}

// Fills a job for plan p under shape s; *ldsBytes receives the job's LDS need.
int make_job(const ppl_dedisp_plan* p, const Shape& s, const void* in, void* out, DDJob& J,
             size_t* ldsBytes)
{
    const int TT = 64 * s.nt, TD = 4 * s.nd;
    const int span = span_of(p, TD);
    int rowUnits = 0, perChannel = 0;
    int budget = kMaxLDS;
    const int32_t* ringTab = nullptr;
    if (s.rw == 16 && p->cfg.dtype == PPL_U8) {
        // the u8 ring kernel: tiles of 16 or 32 trials and 256 or 512 samples
        if ((TD != 16 && TD != 32) || (TT != 256 && TT != 512)) {
            set_error("dedisperse: the u8 ring takes DM tiles of 16 or 32 trials and 256 or 512 samples");
            return PPL_E_ARG;
        }
        if ((long)p->cfg.nChan * p->cfg.nsampBuf >= (1L << 31) || ((long)p->cfg.nChan * p->cfg.nsampBuf) % 4 != 0) {
            set_error("dedisperse: the u8 ring needs nChan*nsampBuf below 2^31 and a multiple of 4");
            return PPL_E_CONFIG;
        }
        // the widest read reaches dword (3 + span) / 4 + TT / 4 - 1 of a copy;
        // the last dword of a shifted copy is not valid
        rowUnits = kU8CopyDw;
        perChannel = (span + 3) / 4 + TT / 4 <= kU8CopyDw - 1 ? 0 : budget + 1;
        if (perChannel == 0) {
            const int rc = ring_table(p, TD, &ringTab);
            if (rc != PPL_OK) return rc;
        }
    } else if (s.rw == 16) {
        if (p->cfg.dtype != PPL_F32 || s.nt % 4 != 0) {
            set_error("dedisperse: 16-byte reads need FP32 or u8 input and a tile of 256k samples");
            return PPL_E_ARG;
        }
        // float4 per copy: the widest read reaches float4 span/4 + TT/4 - 1
        // of a copy, which holds ring_quads_of_tile(TD)
        rowUnits = ring_quads_of_tile(TD);
        perChannel = span / 4 + TT / 4 <= rowUnits ? 0 : budget + 1;
        if ((long)p->cfg.nChan * p->cfg.nsampBuf >= (1L << 31) ||
            (long)p->cfg.nChan * p->tdmsPad >= (1L << 31)) {
            set_error("dedisperse: 16-byte reads need nChan*nsampBuf below 2^31");
            return PPL_E_CONFIG;
        }
        if (TD != 16 && TD != 32 && TD != 64) {
            set_error("dedisperse: 16-byte reads take DM tiles of 16, 32 or 64 trials");
            return PPL_E_ARG;
        }
        if (perChannel == 0) {
            const int rc = ring_table(p, TD, &ringTab);
            if (rc != PPL_OK) return rc;
        }
    } else {
        switch (p->cfg.dtype) {
            case PPL_F32: lds_geometry<float>(s, span, rowUnits, perChannel); break;
            case PPL_F16: lds_geometry<_Float16>(s, span, rowUnits, perChannel); break;
            default: lds_geometry<uint8_t>(s, span, rowUnits, perChannel); break;
        }
    }
    if (perChannel > budget) {
        set_error("dedisperse: delay spread %d too large for tile %dx%d", span, TD, TT);
        return PPL_E_CONFIG;
    }
    int CB = 1;  // the ring kernel stages one channel per stage
    if (s.rw != 16) {
        CB = std::min(16, budget / perChannel);
        if (s.cb > 0) CB = std::min(CB, s.cb);
        CB = std::max(CB, 1);
    }
    J.in = in;
    J.out = out;
    J.delayT = s.rw == 16 ? ringTab : p->d_delayT;
    J.nChan = (int)p->cfg.nChan;
    J.nsamp = (int)p->cfg.nsampBuf;
    J.tdms = (int)p->cfg.tdms;
    J.tdmsPad = p->tdmsPad;
    J.nOut = p->nOut;
    J.rowUnits = rowUnits;
    J.CB = CB;
    J.nDMTiles = (J.tdms + TD - 1) / TD;
    J.blockStart = 0;
    J.pad_ = 0;
    *ldsBytes = s.rw == 16 ? 0 : (size_t)CB * perChannel;  // the 16-byte kernel's LDS is static
    return PPL_OK;
}

int job_blocks(const DDJob& J, const Shape& s)
{
    const int TT = 64 * s.nt;
    const int nTTiles = (J.nOut + TT - 1) / TT;
    return J.nDMTiles * ((nTTiles + kXcds - 1) / kXcds * kXcds);
}

template <typename Tin>
int launch_batch(const Shape& s, const DDBatch& B, unsigned blocks, size_t lds, hipStream_t stream)
{
    dim3 grid(blocks), block(256);
    const dim3 block8(512);  // the eight-wave ring kernel
    if constexpr (std::is_same<Tin, _Float16>::value) {
#define PPL_DD16_CASE(ND_, NT_)                                                                  \
    if (s.nd == ND_ && s.nt == NT_) {                                                            \
        hipLaunchKernelGGL((dedisperse_f16_kernel<ND_, NT_ / 2>), grid, block, lds, stream, B); \
        PPL_LAUNCHED();                                                                          \
        return PPL_OK;                                                                           \
    }
        PPL_DD16_CASE(8, 8)
        PPL_DD16_CASE(8, 4)
        PPL_DD16_CASE(4, 8)
        PPL_DD16_CASE(4, 4)
        PPL_DD16_CASE(2, 8)
        PPL_DD16_CASE(16, 4)
        PPL_DD16_CASE(16, 2)
#undef PPL_DD16_CASE
    } else if constexpr (std::is_same<Tin, uint8_t>::value) {
        if (s.rw == 16) {
#define PPL_DD8R_CASE(ND_, NT_)                                                                   \
    if (s.nd == ND_ && s.nt == NT_) {                                                             \
        hipLaunchKernelGGL((dedisperse_u8_ring_kernel<ND_, NT_ / 4>), grid, block, 0, stream, B); \
        PPL_LAUNCHED();                                                                           \
        return PPL_OK;                                                                            \
    }
            PPL_DD8R_CASE(4, 8)
            PPL_DD8R_CASE(8, 8)
            PPL_DD8R_CASE(4, 4)
            PPL_DD8R_CASE(8, 4)
#undef PPL_DD8R_CASE
            set_error("dedisperse: unsupported u8 ring tile %dx%d", 4 * s.nd, 64 * s.nt);
            return PPL_E_ARG;
        }
#define PPL_DD8_CASE(ND_, NT_)                                                                  \
    if (s.nd == ND_ && s.nt == NT_) {                                                           \
        hipLaunchKernelGGL((dedisperse_u8_kernel<ND_, NT_ / 4>), grid, block, lds, stream, B); \
        PPL_LAUNCHED();                                                                         \
        return PPL_OK;                                                                          \
    }
        PPL_DD8_CASE(8, 8)
        PPL_DD8_CASE(8, 4)
        PPL_DD8_CASE(4, 8)
        PPL_DD8_CASE(4, 4)
        PPL_DD8_CASE(2, 8)
        PPL_DD8_CASE(16, 4)
#undef PPL_DD8_CASE
    } else {
        if (s.rw == 16) {
#define PPL_DDX4_CASE(ND_, NT_)                                                                 \
    if (s.nd == ND_ && s.nt == NT_) {                                                           \
        hipLaunchKernelGGL((dedisperse_ring_kernel<ND_, NT_ / 4, PPL_RING_DEPTH, 4>), grid, block, 0, stream, B); \
        PPL_LAUNCHED();                                                                         \
        return PPL_OK;                                                                          \
    }
            PPL_DDX4_CASE(4, 8)
            PPL_DDX4_CASE(8, 8)
            PPL_DDX4_CASE(4, 4)
            PPL_DDX4_CASE(8, 4)
#undef PPL_DDX4_CASE
            // 64 trials: eight waves of 8 trials (the shape's nd counts trials / 4)
            if (s.nd == 16 && s.nt == 8) {
                hipLaunchKernelGGL((dedisperse_ring_kernel<8, 2, PPL_RING_DEPTH, 8>), grid, block8, 0, stream, B);
                PPL_LAUNCHED();
                return PPL_OK;
            }
            if (s.nd == 16 && s.nt == 4) {
                hipLaunchKernelGGL((dedisperse_ring_kernel<8, 1, PPL_RING_DEPTH, 8>), grid, block8, 0, stream, B);
                PPL_LAUNCHED();
                return PPL_OK;
            }
            set_error("dedisperse: unsupported 16-byte tile %dx%d", 4 * s.nd, 64 * s.nt);
            return PPL_E_ARG;
        }
#define PPL_DD_CASE(ND_, NT_)                                                                       \
    if (s.nd == ND_ && s.nt == NT_) {                                                               \
        hipLaunchKernelGGL((dedisperse_kernel<Tin, ND_, NT_>), grid, block, lds, stream, B);       \
        PPL_LAUNCHED();                                                                             \
        return PPL_OK;                                                                              \
    }
        PPL_DD_CASE(8, 8)
        PPL_DD_CASE(8, 4)
        PPL_DD_CASE(4, 8)
        PPL_DD_CASE(4, 4)
        PPL_DD_CASE(2, 8)
        PPL_DD_CASE(16, 4)
        PPL_DD_CASE(16, 2)
#undef PPL_DD_CASE
    }
    set_error("dedisperse: unsupported tile %dx%d", s.nd, s.nt);
    return PPL_E_ARG;
}

// One launch over n (plan, in, out) jobs of one dtype under shape s.  Jobs
// are ordered by channel count, largest first, so the longest workgroups are
// dispatched first and the short ones fill the tail.
int run_jobs(const ppl_dedisp_job* jobs, int n, const Shape& s, hipStream_t st)
{
    PPL_CHECK_ARG(n >= 1 && n <= kMaxDDJobs, "%d dedispersion jobs (1..%d)", n, kMaxDDJobs);
    const int dtype = jobs[0].plan->cfg.dtype;
    int order[kMaxDDJobs];
    for (int k = 0; k < n; ++k) {
        PPL_CHECK_ARG(jobs[k].plan && jobs[k].d_in && jobs[k].d_out, "job %d: null argument", k);
        PPL_CHECK_ARG(jobs[k].plan->cfg.dtype == dtype, "job %d: dtype differs from job 0", k);
        order[k] = k;
    }
    std::stable_sort(order, order + n, [&](int a, int b) {
        return jobs[a].plan->cfg.nChan > jobs[b].plan->cfg.nChan;
    });
    DDBatch B{};
    B.n = n;
    B.one = 1.0f;
    size_t lds = 0;
    long blocks = 0;
    for (int k = 0; k < n; ++k) {
        const ppl_dedisp_job& jb = jobs[order[k]];
        size_t l = 0;
        const int rc = make_job(jb.plan, s, jb.d_in, jb.d_out, B.job[k], &l);
        if (rc != PPL_OK) return rc;
        B.job[k].blockStart = (int)blocks;
        blocks += job_blocks(B.job[k], s);
        lds = std::max(lds, l);
    }
    PPL_CHECK_ARG(blocks < (1L << 31), "grid too large");
    switch (dtype) {
        case PPL_F32: return launch_batch<float>(s, B, (unsigned)blocks, lds, st);
        case PPL_F16: return launch_batch<_Float16>(s, B, (unsigned)blocks, lds, st);
        case PPL_U8: return launch_batch<uint8_t>(s, B, (unsigned)blocks, lds, st);
    }
    set_error("unknown dtype");
    return PPL_E_ARG;
}

const Shape kShapes[] = {{8, 8, 0}, {8, 4, 0}, {4, 8, 0}, {4, 4, 0}, {2, 8, 0}, {16, 4, 0}, {16, 2, 0},
                         {4, 8, 0, 16}, {8, 8, 0, 16}, {4, 4, 0, 16}, {8, 4, 0, 16}, {16, 8, 0, 16},
                         {16, 4, 0, 16}};

bool shape_known(int nd, int nt, int rw)
{
    for (const Shape& s : kShapes)
        if (s.nd == nd && s.nt == nt && s.rw == rw) return true;
    return false;
}

// The shapes the autotuner tries for a dtype: the ring kernels (read width 16)
// for FP32 and, up to 32 trials, for u8.
bool shape_for(const Shape& s, int dtype)
{
    return s.rw == 4 || dtype == PPL_F32 || (dtype == PPL_U8 && s.nd <= 8);
}

// Largest delay spread inside any DM tile of TD trials.
int tile_span(const std::vector<int32_t>& d, int tdms, int nChan, int TD)
{
    int span = 0;
    for (int dm0 = 0; dm0 < tdms; dm0 += TD) {
        const int dm1 = std::min(dm0 + TD - 1, tdms - 1);
        const int32_t* a = &d[(size_t)dm0 * nChan];
        const int32_t* b = &d[(size_t)dm1 * nChan];
        for (int c = 0; c < nChan; ++c) span = std::max(span, b[c] - a[c]);
    }
    return span;
}

int upload_delays(ppl_dedisp_plan* p)
{
    const int tdms = (int)p->cfg.tdms, nChan = (int)p->cfg.nChan;
    const int pad = ((tdms + 63) / 64) * 64;
    std::vector<int32_t> t((size_t)nChan * pad);
    for (int c = 0; c < nChan; ++c)
        for (int dm = 0; dm < pad; ++dm)
            t[(size_t)c * pad + dm] = p->delays[(size_t)std::min(dm, tdms - 1) * nChan + c];
    if (p->d_delayT) PPL_HIP(hipFree(p->d_delayT));
    p->d_delayT = nullptr;
    PPL_HIP(hipMalloc(&p->d_delayT, t.size() * sizeof(int32_t)));
    PPL_HIP(hipMemcpy(p->d_delayT, t.data(), t.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    p->tdmsPad = pad;
    for (int k = 0; k < 4; ++k) p->spanTD[k] = tile_span(p->delays, tdms, nChan, 8 << k);
    return PPL_OK;
}

}  // namespace

extern "C" int ppl_dedisp_plan_create(ppl_dedisp_plan** plan, const ppl_dedisp_config* cfg)
{
    PPL_CHECK_ARG(plan && cfg, "null argument");
    PPL_CHECK_ARG(cfg->nChan > 0 && cfg->nsampBuf > 0 && cfg->tdms > 0, "empty dedispersion shape");
    PPL_CHECK_ARG(cfg->dtype == PPL_F32 || cfg->dtype == PPL_F16 || cfg->dtype == PPL_U8,
                  "unknown dtype %d", cfg->dtype);
    PPL_CHECK_ARG(cfg->fch1 != 0.0, "frequencyChannel1 must be a positive number");
    PPL_CHECK_ARG(cfg->foff < 0.0,
                  "channelBandwidth must be negative (descending channel order)");
    // the kernels stage each DM tile's rows from its first trial's delay on,
    // so the delays must not decrease with the trial index
    if (!(cfg->dmStep >= 0.0f)) {
        set_error("dedispersion: dmStep %g must not be negative", (double)cfg->dmStep);
        return PPL_E_CONFIG;
    }
    ppl_dedisp_plan* p = new ppl_dedisp_plan;
    p->cfg = *cfg;
    if (p->cfg.dmLow < 0.0f) p->cfg.dmLow = 0.0f;  // DedispersionModule.cpp:57
    const unsigned N = cfg->nChan;
    // DedispersionModule::resize :152-157 (invert = 0)
    p->shift.resize(N);
    for (int c = 0; c < (int)N; ++c) {
        const double f = cfg->fch1 + (cfg->foff * c);
        float val = 4148.741601 * ((1.0 / f / f) - (1.0 / cfg->fch1 / cfg->fch1));
        p->shift[c] = val;
    }
    // :159 (float products, then the double division, truncated to int)
    const float top = p->cfg.dmLow + p->cfg.dmStep * (float)(cfg->tdms - 1);
    p->maxshift = (int)((double)(top * p->shift[N - 1]) / cfg->tsamp);
    if ((int)cfg->nsampBuf <= p->maxshift) {
        set_error("DedispersionModule: maxshift requirements (%d) are bigger than the number "
                  "of samples (%u)", p->maxshift, cfg->nsampBuf);
        delete p;
        return PPL_E_CONFIG;
    }
    p->rem = (int)((cfg->nsampBuf - (unsigned)p->maxshift) % (14u * 16u));  // :162
    p->nOut = (int)cfg->nsampBuf - p->maxshift - p->rem;                      // :274
    // DedispersionKernel members are float (DedispersionModule.h:50-52), :363-364
    const float tsamp_f = (float)cfg->tsamp;
    p->mstartdm = p->cfg.dmLow / tsamp_f;
    p->mdmstep = p->cfg.dmStep / tsamp_f;
    p->delays.resize((size_t)cfg->tdms * N);
    int maxdelay = 0;
    p->fusedDiffs = 0;
    for (unsigned dm = 0; dm < cfg->tdms; ++dm) {
        const float prod = (float)dm * p->mdmstep;
        const float dmval = p->mstartdm + prod;
        // the same trial value with mstartdm + dm*mdmstep as one fused
        // multiply-add, as a compiler contracting DedispersionKernel.cu:29
        // would build it (identical when dmLow == 0: the sum is then exact)
        const float dmvalFused = std::fma((float)dm, p->mdmstep, p->mstartdm);
        for (unsigned c = 0; c < N; ++c) {
            const float v = p->shift[c] * dmval;
            const int32_t d = (int32_t)v;
            p->delays[(size_t)dm * N + c] = d;
            maxdelay = std::max(maxdelay, d);
            if (dmvalFused != dmval && (int32_t)(p->shift[c] * dmvalFused) != d) ++p->fusedDiffs;
        }
    }
    if (p->nOut <= 0 || p->nOut - 1 + maxdelay >= (int)cfg->nsampBuf) {
        set_error("dedispersion: delayed reads (max delay %d, nOut %d) leave the %u-sample buffer",
                  maxdelay, p->nOut, cfg->nsampBuf);
        delete p;
        return PPL_E_CONFIG;
    }
    // default tile: 32 DMs x 512 samples (f32/f16), narrower DM tiles when the
    // delay spread makes the LDS window too wide
    const int elem = cfg->dtype == PPL_F32 ? 4 : cfg->dtype == PPL_F16 ? 2 : 1;
    const Shape prefs[] = {{8, 8}, {4, 8}, {2, 8}};
    for (const Shape& s : prefs) {
        p->nd = s.nd;
        p->nt = s.nt;
        const int TD = 4 * s.nd;
        const int span = tile_span(p->delays, (int)cfg->tdms, (int)N, TD);
        if (4 * 2 * ((64 * s.nt + span + 20) * elem + 4 * TD) <= kMaxLDS) break;
    }
    // FP32: the ring kernel (16-byte reads) at 32 trials x 512 samples, the
    // autotuner's usual choice, where its staged row holds a tile's delay
    // spread and the sizes fit its 32-bit offsets
    if (cfg->dtype == PPL_F32) {
        const long pad = ((long)cfg->tdms + 63) / 64 * 64;
        const int span32 = tile_span(p->delays, (int)cfg->tdms, (int)N, 32);
        if (span32 / 4 + 512 / 4 <= ring_quads_of_tile(32) && (long)N * cfg->nsampBuf < (1L << 31) &&
            (long)N * pad < (1L << 31)) {
            p->nd = 8;
            p->nt = 8;
            p->rw = 16;
        }
    }
    int rc = upload_delays(p);
    if (rc != PPL_OK) {
        delete p;
        return rc;
    }
    *plan = p;
    return PPL_OK;
}

extern "C" int ppl_dedisp_plan_destroy(ppl_dedisp_plan* plan)
{
    if (!plan) return PPL_OK;
    if (plan->d_delayT) (void)hipFree(plan->d_delayT);
    for (int32_t* t : plan->d_ring)
        if (t) (void)hipFree(t);
    if (plan->d_lenOff) (void)hipFree(plan->d_lenOff);
    if (plan->d_lenTab) (void)hipFree(plan->d_lenTab);
    delete plan;
    return PPL_OK;
}

extern "C" int ppl_dedisp_plan_info(const ppl_dedisp_plan* p, ppl_dedisp_info* info)
{
    PPL_CHECK_ARG(p && info, "null argument");
    info->maxshift = p->maxshift;
    info->remainingSamples = p->rem;
    info->nOut = p->nOut;
    info->mstartdm = p->mstartdm;
    info->mdmstep = p->mdmstep;
    info->tileDM = 4 * p->nd;
    info->tileT = 64 * p->nt;
    info->chanBlock = p->cb;
    info->readBytes = p->rw;
    info->fusedDelayDiffs = p->fusedDiffs;
    return PPL_OK;
}

extern "C" int ppl_dedisp_set_tile(ppl_dedisp_plan* p, int tileDM, int tileT)
{
    return ppl_dedisp_set_kernel(p, tileDM, tileT, 0, 4);
}

extern "C" int ppl_dedisp_set_shape(ppl_dedisp_plan* p, int tileDM, int tileT, int chanBlock)
{
    return ppl_dedisp_set_kernel(p, tileDM, tileT, chanBlock, 4);
}

namespace {
// Checks an explicit (tileDM, tileT, chanBlock, readBytes) for a dtype.
int parse_shape(int tileDM, int tileT, int chanBlock, int readBytes, int dtype, Shape& s)
{
    PPL_CHECK_ARG(tileDM % 4 == 0 && tileT % 64 == 0, "tile %dx%d not a multiple of 4x64", tileDM, tileT);
    PPL_CHECK_ARG(chanBlock >= 0 && chanBlock <= 16, "channel block %d outside [0, 16]", chanBlock);
    PPL_CHECK_ARG(readBytes == 4 || readBytes == 16, "read width %d bytes (4 or 16)", readBytes);
    PPL_CHECK_ARG(readBytes == 4 || dtype == PPL_F32 || (dtype == PPL_U8 && tileDM <= 32),
                  "16-byte reads (the ring kernels) need FP32 input, or u8 input with tiles of up to 32 trials");
    PPL_CHECK_ARG(shape_known(tileDM / 4, tileT / 64, readBytes), "tile %dx%d with %d-byte reads not instantiated",
                  tileDM, tileT, readBytes);
    s = Shape{tileDM / 4, tileT / 64, chanBlock, readBytes};
    return PPL_OK;
}
}  // namespace

extern "C" int ppl_dedisp_set_kernel(ppl_dedisp_plan* p, int tileDM, int tileT, int chanBlock, int readBytes)
{
    PPL_CHECK_ARG(p, "null plan");
    if (tileDM == 0 && tileT == 0) return PPL_OK;
    Shape s;
    const int rc = parse_shape(tileDM, tileT, chanBlock, readBytes, p->cfg.dtype, s);
    if (rc != PPL_OK) return rc;
    p->nd = s.nd;
    p->nt = s.nt;
    p->cb = s.cb;
    p->rw = s.rw;
    return PPL_OK;
}

extern "C" int ppl_dedisp_autotune(ppl_dedisp_plan* p, const void* d_in, void* d_out, int reps, void* stream,
                                   float* best_ms)
{
    PPL_CHECK_ARG(p && d_in && d_out, "null argument");
    if (reps < 1) reps = 1;
    hipStream_t st = stream_of(stream);
    hipEvent_t e0, e1;
    PPL_HIP(hipEventCreate(&e0));
    PPL_HIP(hipEventCreate(&e1));
    const int cbs[] = {1, 2, 4, 8, 16};
    float best = 1e30f;
    Shape bestShape{p->nd, p->nt, p->cb, p->rw};
    for (const Shape& base : kShapes) {
        if (!shape_for(base, p->cfg.dtype)) continue;
        for (int cb : cbs) {
            const Shape s{base.nd, base.nt, cb, base.rw};
            DDJob J{};
            size_t lds = 0;
            if (make_job(p, s, d_in, d_out, J, &lds) != PPL_OK || J.CB < cb) continue;
            const ppl_dedisp_job job{p, d_in, d_out};
            if (run_jobs(&job, 1, s, st) != PPL_OK) continue;  // warm
            PPL_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < reps; ++r)
                if (run_jobs(&job, 1, s, st) != PPL_OK) break;
            PPL_HIP(hipEventRecord(e1, st));
            PPL_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            PPL_HIP(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            if (ms < best) {
                best = ms;
                bestShape = s;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (best_ms) *best_ms = best;
    return ppl_dedisp_set_kernel(p, 4 * bestShape.nd, 64 * bestShape.nt, bestShape.cb, bestShape.rw);
}

extern "C" int ppl_dedisp_delays(const ppl_dedisp_plan* p, int32_t* host_out)
{
    PPL_CHECK_ARG(p && host_out, "null argument");
    std::copy(p->delays.begin(), p->delays.end(), host_out);
    return PPL_OK;
}

extern "C" int ppl_dedisperse(const ppl_dedisp_plan* p, const void* d_in, void* d_out, void* stream)
{
    PPL_CHECK_ARG(p && d_in && d_out, "null argument");
    const ppl_dedisp_job job{p, d_in, d_out};
    return run_jobs(&job, 1, Shape{p->nd, p->nt, p->cb, p->rw}, stream_of(stream));
}

extern "C" int ppl_dedisperse_batch(const ppl_dedisp_job* jobs, int n, int tileDM, int tileT, int chanBlock,
                                    int readBytes, void* stream)
{
    PPL_CHECK_ARG(jobs && n >= 1, "no jobs");
    PPL_CHECK_ARG(jobs[0].plan, "job 0: null plan");
    Shape s{jobs[0].plan->nd, jobs[0].plan->nt, jobs[0].plan->cb, jobs[0].plan->rw};
    if (tileDM != 0 || tileT != 0) {
        const int rc = parse_shape(tileDM, tileT, chanBlock, readBytes, jobs[0].plan->cfg.dtype, s);
        if (rc != PPL_OK) return rc;
        return run_jobs(jobs, n, s, stream_of(stream));
    }
    // job 0's own tile; when that is a ring tile some other job's delay
    // spread does not fit (a refusal before any launch), the same tile with
    // one-sample reads
    const int rc = run_jobs(jobs, n, s, stream_of(stream));
    if (rc != PPL_E_CONFIG || s.rw != 16) return rc;
    s.rw = 4;
    return run_jobs(jobs, n, s, stream_of(stream));
}

extern "C" int ppl_dedisp_autotune_batch(const ppl_dedisp_job* jobs, int n, int reps, void* stream,
                                         int* tileDM, int* tileT, int* chanBlock, int* readBytes, float* best_ms)
{
    PPL_CHECK_ARG(jobs && n >= 1 && tileDM && tileT && chanBlock && readBytes, "null argument");
    PPL_CHECK_ARG(jobs[0].plan, "job 0: null plan");
    if (reps < 1) reps = 1;
    hipStream_t st = stream_of(stream);
    hipEvent_t e0, e1;
    PPL_HIP(hipEventCreate(&e0));
    PPL_HIP(hipEventCreate(&e1));
    const int cbs[] = {1, 2, 4, 8, 16};
    float best = 1e30f;
    Shape bestShape{0, 0, 0};
    for (const Shape& base : kShapes) {
        if (!shape_for(base, jobs[0].plan->cfg.dtype)) continue;
        for (int cb : cbs) {
            const Shape s{base.nd, base.nt, cb, base.rw};
            bool fits = true;
            for (int k = 0; k < n && fits; ++k) {
                DDJob J{};
                size_t lds = 0;
                fits = jobs[k].plan && make_job(jobs[k].plan, s, jobs[k].d_in, jobs[k].d_out, J, &lds) == PPL_OK &&
                       J.CB >= cb;
            }
            if (!fits || run_jobs(jobs, n, s, st) != PPL_OK) continue;  // warm
            PPL_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < reps; ++r)
                if (run_jobs(jobs, n, s, st) != PPL_OK) break;
            PPL_HIP(hipEventRecord(e1, st));
            PPL_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            PPL_HIP(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            if (ms < best) {
                best = ms;
                bestShape = s;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (bestShape.nd == 0) {
        set_error("autotune: no tile shape fits every job");
        return PPL_E_CONFIG;
    }
    *tileDM = 4 * bestShape.nd;
    *tileT = 64 * bestShape.nt;
    *chanBlock = bestShape.cb;
    *readBytes = bestShape.rw;
    if (best_ms) *best_ms = best;
    return PPL_OK;
}

extern "C" int ppl_cache_dedisperse_loop(float* outbuff, long outbufSize, float* buff, float mstartdm,
                                         float mdmstep, int tdms, int numSamples,
                                         const float* dmShift, int maxshift, int i_nchans,
                                         void* stream)
{
    PPL_CHECK_ARG(outbuff && buff && dmShift, "null argument");
    PPL_CHECK_ARG(tdms > 0 && i_nchans > 0 && numSamples > maxshift && maxshift >= 0,
                  "bad shape (tdms %d, nchans %d, numSamples %d, maxshift %d)", tdms, i_nchans,
                  numSamples, maxshift);
    // delays must not decrease with the trial index (the kernels stage each
    // DM tile's rows from its first trial's delay on)
    if (!(mstartdm >= 0.0f && mdmstep >= 0.0f)) {
        set_error("cacheDedisperseLoop: mstartdm %g and mdmstep %g must not be negative", (double)mstartdm,
                  (double)mdmstep);
        return PPL_E_CONFIG;
    }
    const int nOut = numSamples - maxshift;
    PPL_CHECK_ARG((long)tdms * nOut * (long)sizeof(float) <= outbufSize,
                  "output buffer (%ld bytes) smaller than tdms*nOut floats", outbufSize);
    hipStream_t st = stream_of(stream);
    const Shape s{8, 8, 0};
    const int TD = 4 * s.nd;
    const int pad = ((tdms + TD - 1) / TD) * TD;
    // delay table on the device, then the spread of the first/last trial
    // per tile is bounded by the spread of the whole DM range
    int32_t* d_delayT = nullptr;
    PPL_HIP(hipMallocAsync((void**)&d_delayT, (size_t)i_nchans * pad * sizeof(int32_t), st));
    dim3 g((unsigned)cdiv(pad, 256), (unsigned)i_nchans);
    hipLaunchKernelGGL(build_delays_kernel, g, dim3(256), 0, st, dmShift, i_nchans, tdms, pad,
                       mstartdm, mdmstep, d_delayT);
    PPL_LAUNCHED();
    std::vector<float> hshift(i_nchans);
    PPL_HIP(hipMemcpyAsync(hshift.data(), dmShift, i_nchans * sizeof(float), hipMemcpyDeviceToHost, st));
    PPL_HIP(hipStreamSynchronize(st));
    int span = 0;
    for (int c = 0; c < i_nchans; ++c) {
        if (!(hshift[c] >= 0.0f)) {
            (void)hipFreeAsync(d_delayT, st);
            set_error("cacheDedisperseLoop: dmShift[%d] = %g is negative (ascending channel frequency)", c,
                      (double)hshift[c]);
            return PPL_E_CONFIG;
        }
        for (int dm0 = 0; dm0 < tdms; dm0 += TD) {
            const int dm1 = std::min(dm0 + TD - 1, tdms - 1);
            const int a = (int)(hshift[c] * (mstartdm + (float)dm0 * mdmstep));
            const int b = (int)(hshift[c] * (mstartdm + (float)dm1 * mdmstep));
            span = std::max(span, b - a);
            if (a < 0 || (long)(nOut - 1) + b >= numSamples) {
                (void)hipFreeAsync(d_delayT, st);
                set_error("cacheDedisperseLoop: delayed read outside the buffer (channel %d)", c);
                return PPL_E_CONFIG;
            }
        }
    }
    const long used = (long)tdms * nOut * (long)sizeof(float);
    if (outbufSize > used)
        PPL_HIP(hipMemsetAsync(reinterpret_cast<char*>(outbuff) + used, 0, outbufSize - used, st));
    ppl_dedisp_plan tmp;
    tmp.cfg.nChan = (unsigned)i_nchans;
    tmp.cfg.nsampBuf = (unsigned)numSamples;
    tmp.cfg.tdms = (unsigned)tdms;
    tmp.cfg.dtype = PPL_F32;
    tmp.nOut = nOut;
    tmp.d_delayT = d_delayT;
    tmp.tdmsPad = pad;
    for (int k = 0; k < 4; ++k) tmp.spanTD[k] = span;
    // The ring kernel (16-byte reads, 32 trials x 512 samples) where a DM
    // tile's delay spread fits its staged row and the sizes fit its 32-bit
    // offsets, with its table built here from the device delays; the
    // one-sample kernel otherwise.  Both give the same sums.
    const Shape ringShape{8, 8, 0, 16};
    const int quads = ring_quads_of_tile(TD);
    const bool ring = span / 4 + 64 * ringShape.nt / 4 <= quads && (long)i_nchans * numSamples < (1L << 31) &&
                      (long)i_nchans * pad < (1L << 31);
    int32_t* d_tab = nullptr;
    if (ring) {
        const size_t n = (size_t)(pad / TD) * i_nchans * TD;
        const hipError_t e = hipMallocAsync((void**)&d_tab, (n + (size_t)(pad / TD) * i_nchans) * sizeof(int32_t), st);
        if (e != hipSuccess) {
            (void)hipFreeAsync(d_delayT, st);
            set_error("cacheDedisperseLoop: ring table: %s", hipGetErrorString(e));
            return PPL_E_HIP;
        }
        hipLaunchKernelGGL(ring_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_delayT,
                           i_nchans, pad, TD, quads, d_tab);
        tmp.d_ring[1] = d_tab;  // the 32-trial slot, so make_job takes this table
    }
    const hipError_t le = hipGetLastError();
    const ppl_dedisp_job job{&tmp, buff, outbuff};
    int rc = PPL_E_HIP;
    if (le != hipSuccess)
        set_error("cacheDedisperseLoop: ring table launch: %s", hipGetErrorString(le));
    else
        rc = run_jobs(&job, 1, ring ? ringShape : s, st);
    if (d_tab) (void)hipFreeAsync(d_tab, st);
    (void)hipFreeAsync(d_delayT, st);
    return rc;
}

// ------------------------------------------------------- u8 series
namespace {

// Rows (spectra) of a u8 buffer with a nonzero offset, in increasing order
// (one workgroup: each thread a contiguous run of rows).
__global__ __launch_bounds__(1024) void exceptional_rows_kernel(const float* __restrict__ rowOff, int nsamp,
                                                                int* __restrict__ rows, int* __restrict__ nRows)
{
    __shared__ int waveTot[16];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int per = (nsamp + 1023) / 1024;
    const int a = min(t * per, nsamp), b = min(a + per, nsamp);
    int n = 0;
    for (int i = a; i < b; ++i) n += rowOff[i] != 0.0f;
    int incl = n;
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
    }
    if (lane == 63) waveTot[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        const int w = lane < 16 ? waveTot[lane] : 0;
        int wi = w;
        for (int d = 1; d < 16; d <<= 1) {
            const int y = __shfl_up(wi, d, 64);
            if (lane >= d) wi += y;
        }
        if (lane < 16) waveTot[lane] = wi - w;
        if (lane == 15) *nRows = wi;
    }
    __syncthreads();
    int o = waveTot[wave] + incl - n;
    for (int i = a; i < b; ++i)
        if (rowOff[i] != 0.0f) rows[o++] = i;
}

// out[dm][t] = float(S - offset) * scale, plus, for every exceptional row e
// with 0 <= e - t <= maxdelay(dm) in increasing order, rowOff[e] times the
// number of channels whose delay is e - t (each product rounded, then
// added).  One workgroup per (256 samples, trial).
__global__ __launch_bounds__(256) void u8_series_kernel(const int32_t* __restrict__ sums,
                                                        const float* __restrict__ rowOff,
                                                        const int* __restrict__ rows, const int* __restrict__ nRowsP,
                                                        const uint16_t* __restrict__ lenTab,
                                                        const int32_t* __restrict__ lenOff, int nOut, int offset,
                                                        float scale, float* __restrict__ out)
{
    const int dm = blockIdx.y;
    const int t0 = blockIdx.x * 256;
    const int t = t0 + threadIdx.x;
    const int base = lenOff[dm];
    const int dmax = lenOff[dm + 1] - base - 1;
    const int nRows = *nRowsP;
    // first exceptional row at or after t0 (uniform)
    int lo = 0, hi = nRows;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rows[mid] < t0) lo = mid + 1;
        else hi = mid;
    }
    float corr = 0.0f;
    for (int i = lo; i < nRows; ++i) {
        const int e = rows[i];
        if (e > t0 + 255 + dmax) break;
        const int k = e - t;
        if (k >= 0 && k <= dmax) {
            const int n = lenTab[base + k];
            if (n != 0) corr = corr + rowOff[e] * (float)n;
        }
    }
    if (t < nOut) {
        const size_t i = (size_t)dm * nOut + t;
        out[i] = (float)(sums[i] - offset) * scale + corr;
    }
}

// The channel-count table of a plan (see ppl_dedisp_plan::d_lenTab).
int len_table(const ppl_dedisp_plan* cp, const int32_t** lenOff, const uint16_t** lenTab)
{
    ppl_dedisp_plan* p = const_cast<ppl_dedisp_plan*>(cp);
    if (!p->d_lenTab) {
        const int tdms = (int)p->cfg.tdms, nChan = (int)p->cfg.nChan;
        std::vector<int32_t> off(tdms + 1);
        long total = 0;
        for (int dm = 0; dm < tdms; ++dm) {
            off[dm] = (int32_t)total;
            int dmax = 0;
            for (int c = 0; c < nChan; ++c) dmax = std::max(dmax, p->delays[(size_t)dm * nChan + c]);
            total += dmax + 1;
            PPL_CHECK_ARG(total < (1L << 31), "u8 series: channel-count table too large");
        }
        off[tdms] = (int32_t)total;
        std::vector<uint16_t> tab((size_t)total, 0);
        PPL_CHECK_ARG(nChan < 65536, "u8 series: %d channels (at most 65535)", nChan);
        for (int dm = 0; dm < tdms; ++dm)
            for (int c = 0; c < nChan; ++c) ++tab[(size_t)off[dm] + p->delays[(size_t)dm * nChan + c]];
        int32_t* dOff = nullptr;
        uint16_t* dTab = nullptr;
        PPL_HIP(hipMalloc(&dOff, off.size() * sizeof(int32_t)));
        if (hipMalloc(&dTab, tab.size() * sizeof(uint16_t)) != hipSuccess) {
            (void)hipFree(dOff);
            set_error("u8 series: allocation of the channel-count table failed");
            return PPL_E_NOMEM;
        }
        if (hipMemcpy(dOff, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dTab, tab.data(), tab.size() * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(dOff);
            (void)hipFree(dTab);
            set_error("u8 series: upload of the channel-count table failed");
            return PPL_E_HIP;
        }
        p->d_lenOff = dOff;
        p->d_lenTab = dTab;
    }
    *lenOff = p->d_lenOff;
    *lenTab = p->d_lenTab;
    return PPL_OK;
}

}  // namespace

extern "C" int ppl_dedisp_u8_series(const ppl_dedisp_plan* p, const int32_t* d_sums, const float* d_rowOffsets,
                                    float scale, float* d_out, void* stream)
{
    PPL_CHECK_ARG(p && d_sums && d_rowOffsets && d_out, "null argument");
    PPL_CHECK_ARG(p->cfg.dtype == PPL_U8, "u8 series of a plan of another dtype");
    const int32_t* lenOff = nullptr;
    const uint16_t* lenTab = nullptr;
    int rc = len_table(p, &lenOff, &lenTab);
    if (rc != PPL_OK) return rc;
    hipStream_t st = stream_of(stream);
    const int nsamp = (int)p->cfg.nsampBuf;
    int* rows = nullptr;
    PPL_HIP(hipMallocAsync((void**)&rows, sizeof(int) * ((size_t)nsamp + 1), st));
    hipLaunchKernelGGL(exceptional_rows_kernel, dim3(1), dim3(1024), 0, st, d_rowOffsets, nsamp, rows + 1, rows);
    PPL_LAUNCHED();
    dim3 grid(cdiv((unsigned)p->nOut, 256), p->cfg.tdms);
    hipLaunchKernelGGL(u8_series_kernel, grid, dim3(256), 0, st, d_sums, d_rowOffsets, rows + 1, rows, lenTab,
                       lenOff, p->nOut, (int)(128 * p->cfg.nChan), scale, d_out);
    PPL_LAUNCHED();
    PPL_HIP(hipFreeAsync(rows, st));
    return PPL_OK;
}

// ---------------------------------------------------------------- converters
namespace {

__global__ void to_f16_kernel(const float* __restrict__ in, _Float16* __restrict__ out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (_Float16)in[i];  // v_cvt_f16_f32: round to nearest even
}

__global__ void quantise_kernel(const float* __restrict__ in, uint8_t* __restrict__ out, size_t n,
                                float inv)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float y = rintf(in[i] * inv);
        if (!(y >= -128.0f)) y = -128.0f;
        if (y > 127.0f) y = 127.0f;
        out[i] = (uint8_t)((int)y + 128);
    }
}

}  // namespace

extern "C" int ppl_convert_f16(const float* d_in, uint16_t* d_out, size_t n, void* stream)
{
    PPL_CHECK_ARG(d_in && d_out, "null argument");
    if (!n) return PPL_OK;
    hipLaunchKernelGGL(to_f16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       stream_of(stream), d_in, reinterpret_cast<_Float16*>(d_out), n);
    PPL_LAUNCHED();
    return PPL_OK;
}

extern "C" int ppl_quantise_u8(const float* d_in, uint8_t* d_out, size_t n, float inv_scale,
                               void* stream)
{
    PPL_CHECK_ARG(d_in && d_out, "null argument");
    if (!n) return PPL_OK;
    hipLaunchKernelGGL(quantise_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       stream_of(stream), d_in, d_out, n, inv_scale);
    PPL_LAUNCHED();
    return PPL_OK;
}
