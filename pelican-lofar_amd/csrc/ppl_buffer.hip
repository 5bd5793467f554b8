// ppl_buffer.hip -- device-resident dedispersion buffer
// (DedispersionBuffer, src/lib/src/DedispersionBuffer.cpp).
//
// Two [nChan][nsampBuf] buffers of the dedispersion dtype.  Spectra
// [t][nChan] (Stokes-I blobs, invert = 0) are transposed in 64x64 LDS tiles
// into the current buffer behind its fill mark, converted to FP16 or to
// offset-binary u8 on the way.  When the buffer is full and dedispersed,
// rotate() copies its last `overlap` = maxshift + remainingSamples samples
// into the other buffer and makes that one current
// (DedispersionBuffer::copy, :65-96; DedispersionModule.cpp:238).
#include "ppl_internal.h"

struct ppl_ddbuf {
    unsigned nChan = 0, nsamp = 0, overlap = 0, fill = 0;
    int dtype = PPL_F32;
    float inv_scale = 1.0f;
    void* d_buf[2] = {nullptr, nullptr};
    // u8 only: per buffer row (spectrum), the offset it was quantised
    // relative to (0 for an ordinary row), beside each of the two buffers
    float* d_rowOff[2] = {nullptr, nullptr};
    int cur = 0;
};

namespace {

using namespace ppl;

template <typename T> __device__ __forceinline__ T convert(float v, float inv);
template <> __device__ __forceinline__ float convert<float>(float v, float) { return v; }
template <> __device__ __forceinline__ _Float16 convert<_Float16>(float v, float) { return (_Float16)v; }
template <> __device__ __forceinline__ uint8_t convert<uint8_t>(float v, float inv)
{
    float y = rintf(v * inv);
    if (!(y >= -128.0f)) y = -128.0f;
    if (y > 127.0f) y = 127.0f;
    return (uint8_t)((int)y + 128);
}

// u8 mode, exceptional rows.  A spectrum more than half of whose samples
// fall outside the 8-bit range (rint(x * inv) outside [-128, 127], NaN
// included) -- in this chain the RFI clipper's rejected spectra, which it
// leaves at -0.2 - bandpass - dataModel, thousands of units below the
// normalised data -- is quantised relative to its middle channel's value
// (channel nChan/2), which is kept per row; the dedispersed sums then get
// that offset back once for every channel that lands on the row
// (ppl_dedisp_u8_series).  Every other row has offset 0 and is quantised as
// it stands.  One workgroup per spectrum.
__global__ __launch_bounds__(256) void row_check_kernel(const float* __restrict__ spectra, long t0, int nChan, int nCh,
                                                        int nPol, float inv, float* __restrict__ rowOff)
{
    __shared__ int part[4];
    const long rowStride = (long)nChan * nPol;
    const float* row = spectra + (t0 + blockIdx.x) * rowStride;
    auto at = [&](int c) { return nPol == 1 ? row[c] : row[(long)(c / nCh) * nCh * nPol + c % nCh]; };
    int n = 0;
    for (int c = threadIdx.x; c < nChan; c += 256) {
        const float y = rintf(at(c) * inv);
        n += (!(y >= -128.0f) || y > 127.0f) ? 1 : 0;
    }
    for (int d = 32; d >= 1; d >>= 1) n += __shfl_down(n, d, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int total = part[0] + part[1] + part[2] + part[3];
        rowOff[blockIdx.x] = 2 * total > nChan ? at(nChan / 2) : 0.0f;
    }
}

// Polarisation 0 of spectra [t][sub][pol][ch] (the SpectrumDataSet layout),
// rows [t0, t0+nT), channel c = sub*nCh + ch -> buf[c][col0 + (t - t0)],
// each value less its row's offset (u8 mode; rowOff = nullptr otherwise).
// With nPol = 1 the rows are plain [t][nChan].
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ spectra, long t0, int nT,
                                                        int nChan, int nCh, int nPol, T* __restrict__ buf,
                                                        int nsamp, int col0, float inv,
                                                        const float* __restrict__ rowOff)
{
    __shared__ float tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    const int cBase = blockIdx.x * 64, tBase = blockIdx.y * 64;
    const long rowStride = (long)nChan * nPol;
    const int c = cBase + tx;
    const long cOff = nPol == 1 ? c : (long)(c / nCh) * nCh * nPol + c % nCh;
    for (int r = ty; r < 64; r += 4) {
        const int t = tBase + r;
        tile[r][tx] = (t < nT && c < nChan) ? spectra[(t0 + t) * rowStride + cOff] : 0.0f;
    }
    __syncthreads();
    const float off = (rowOff != nullptr && tBase + tx < nT) ? rowOff[col0 + tBase + tx] : 0.0f;
    for (int r = ty; r < 64; r += 4) {
        const int c = cBase + r, t = tBase + tx;
        if (c < nChan && t < nT) {
            const float v = tile[tx][r];
            buf[(long)c * nsamp + col0 + t] = convert<T>(rowOff != nullptr ? v - off : v, inv);
        }
    }
}

template <typename T>
__global__ void carry_kernel(const T* __restrict__ src, T* __restrict__ dst, int nChan, int nsamp, int overlap)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nChan * overlap) return;
    const long c = i / overlap, k = i % overlap;
    dst[c * nsamp + k] = src[c * nsamp + (nsamp - overlap) + k];
}

// The row offsets of the carried samples go with them.
__global__ void carry_rows_kernel(const float* __restrict__ src, float* __restrict__ dst, int nsamp, int overlap)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < overlap) dst[k] = src[(nsamp - overlap) + k];
}

size_t elem_size(int dtype) { return dtype == PPL_F32 ? 4 : dtype == PPL_F16 ? 2 : 1; }

}  // namespace

extern "C" int ppl_ddbuf_create(ppl_ddbuf** out, unsigned nChan, unsigned nsampBuf, unsigned overlap, int dtype,
                                float inv_scale)
{
    PPL_CHECK_ARG(out, "null argument");
    PPL_CHECK_ARG(nChan > 0 && nsampBuf > 0, "empty buffer");
    PPL_CHECK_ARG(overlap < nsampBuf, "overlap %u must be smaller than the buffer (%u)", overlap, nsampBuf);
    PPL_CHECK_ARG(dtype == PPL_F32 || dtype == PPL_F16 || dtype == PPL_U8, "unknown dtype %d", dtype);
    ppl_ddbuf* b = new ppl_ddbuf;
    b->nChan = nChan;
    b->nsamp = nsampBuf;
    b->overlap = overlap;
    b->dtype = dtype;
    b->inv_scale = inv_scale;
    const size_t bytes = (size_t)nChan * nsampBuf * elem_size(dtype);
    for (int i = 0; i < 2; ++i) {
        if (hipMalloc(&b->d_buf[i], bytes) != hipSuccess) {
            (void)hipFree(b->d_buf[0]);
            delete b;
            set_error("dedispersion buffer: allocation of %zu bytes failed", bytes);
            return PPL_E_NOMEM;
        }
    }
    if (dtype == PPL_U8) {
        for (int i = 0; i < 2; ++i) {
            if (hipMalloc(&b->d_rowOff[i], sizeof(float) * nsampBuf) != hipSuccess ||
                hipMemset(b->d_rowOff[i], 0, sizeof(float) * nsampBuf) != hipSuccess) {
                (void)hipFree(b->d_buf[0]);
                (void)hipFree(b->d_buf[1]);
                (void)hipFree(b->d_rowOff[0]);
                (void)hipFree(b->d_rowOff[1]);
                delete b;
                set_error("dedispersion buffer: row offsets not allocated");
                return PPL_E_NOMEM;
            }
        }
    }
    *out = b;
    return PPL_OK;
}

extern "C" int ppl_ddbuf_destroy(ppl_ddbuf* b)
{
    if (!b) return PPL_OK;
    (void)hipFree(b->d_buf[0]);
    (void)hipFree(b->d_buf[1]);
    (void)hipFree(b->d_rowOff[0]);
    (void)hipFree(b->d_rowOff[1]);
    delete b;
    return PPL_OK;
}

static int add_blob(ppl_ddbuf* b, const float* d_spectra, unsigned nT, unsigned first, unsigned nCh, unsigned nPol,
                    unsigned* consumed, int* full, hipStream_t st)
{
    PPL_CHECK_ARG(b && d_spectra && consumed && full, "null argument");
    PPL_CHECK_ARG(first <= nT, "first sample %u beyond %u", first, nT);
    PPL_CHECK_ARG(nPol >= 1 && nCh >= 1 && b->nChan % nCh == 0,
                  "DedispersionBuffer: %u channels per subband do not divide the buffer's %u", nCh, b->nChan);
    const unsigned room = b->nsamp - b->fill;
    const unsigned n = std::min(room, nT - first);
    if (n > 0) {
        dim3 grid(cdiv(b->nChan, 64), cdiv(n, 64));
        void* dst = b->d_buf[b->cur];
        switch (b->dtype) {
            case PPL_F32:
                hipLaunchKernelGGL(transpose_kernel<float>, grid, dim3(256), 0, st, d_spectra, (long)first, (int)n,
                                   (int)b->nChan, (int)nCh, (int)nPol, static_cast<float*>(dst), (int)b->nsamp, (int)b->fill,
                                   b->inv_scale, nullptr);
                break;
            case PPL_F16:
                hipLaunchKernelGGL(transpose_kernel<_Float16>, grid, dim3(256), 0, st, d_spectra, (long)first,
                                   (int)n, (int)b->nChan, (int)nCh, (int)nPol, static_cast<_Float16*>(dst), (int)b->nsamp,
                                   (int)b->fill, b->inv_scale, nullptr);
                break;
            default: {
                float* off = b->d_rowOff[b->cur];
                hipLaunchKernelGGL(row_check_kernel, dim3(n), dim3(256), 0, st, d_spectra, (long)first, (int)b->nChan,
                                   (int)nCh, (int)nPol, b->inv_scale, off + b->fill);
                PPL_LAUNCHED();
                hipLaunchKernelGGL(transpose_kernel<uint8_t>, grid, dim3(256), 0, st, d_spectra, (long)first,
                                   (int)n, (int)b->nChan, (int)nCh, (int)nPol, static_cast<uint8_t*>(dst), (int)b->nsamp,
                                   (int)b->fill, b->inv_scale, static_cast<const float*>(off));
            }
        }
        PPL_LAUNCHED();
    }
    b->fill += n;
    *consumed = n;
    *full = b->fill == b->nsamp;
    return PPL_OK;
}

unsigned ppl_ddbuf_nchan(const ppl_ddbuf* b) { return b ? b->nChan : 0u; }

int ppl_ddbuf_add_spectra(ppl_ddbuf* b, const float* d_spectra, unsigned nT, unsigned first, unsigned* consumed,
                          int* full, hipStream_t st)
{
    return add_blob(b, d_spectra, nT, first, b ? b->nChan : 1u, 1u, consumed, full, st);
}

extern "C" int ppl_ddbuf_add(ppl_ddbuf* b, const float* d_spectra, unsigned nT, unsigned first, unsigned* consumed,
                             int* full, void* stream)
{
    return ppl_ddbuf_add_spectra(b, d_spectra, nT, first, consumed, full, stream_of(stream));
}

extern "C" int ppl_ddbuf_add_blob(ppl_ddbuf* b, const float* d_spectra, unsigned nT, unsigned first, unsigned nSub,
                                  unsigned nPol, unsigned nCh, unsigned* consumed, int* full, void* stream)
{
    PPL_CHECK_ARG(b, "null argument");
    PPL_CHECK_ARG(nSub * nCh == b->nChan,
                  "DedispersionBuffer: input data sample size(%u) does not match buffer sample size (%u)", nSub * nCh,
                  b->nChan);
    return add_blob(b, d_spectra, nT, first, nCh, nPol, consumed, full, stream_of(stream));
}

extern "C" int ppl_ddbuf_rotate(ppl_ddbuf* b, void* stream)
{
    PPL_CHECK_ARG(b, "null argument");
    const int next = b->cur ^ 1;
    const long n = (long)b->nChan * b->overlap;
    if (n > 0) {
        hipStream_t st = stream_of(stream);
        const unsigned blocks = (unsigned)((n + 255) / 256);
        switch (b->dtype) {
            case PPL_F32:
                hipLaunchKernelGGL(carry_kernel<float>, dim3(blocks), dim3(256), 0, st,
                                   static_cast<const float*>(b->d_buf[b->cur]), static_cast<float*>(b->d_buf[next]),
                                   (int)b->nChan, (int)b->nsamp, (int)b->overlap);
                break;
            case PPL_F16:
                hipLaunchKernelGGL(carry_kernel<_Float16>, dim3(blocks), dim3(256), 0, st,
                                   static_cast<const _Float16*>(b->d_buf[b->cur]),
                                   static_cast<_Float16*>(b->d_buf[next]), (int)b->nChan, (int)b->nsamp,
                                   (int)b->overlap);
                break;
            default:
                hipLaunchKernelGGL(carry_kernel<uint8_t>, dim3(blocks), dim3(256), 0, st,
                                   static_cast<const uint8_t*>(b->d_buf[b->cur]),
                                   static_cast<uint8_t*>(b->d_buf[next]), (int)b->nChan, (int)b->nsamp,
                                   (int)b->overlap);
                PPL_LAUNCHED();
                hipLaunchKernelGGL(carry_rows_kernel, dim3(cdiv(b->overlap, 256)), dim3(256), 0, st,
                                   static_cast<const float*>(b->d_rowOff[b->cur]), b->d_rowOff[next], (int)b->nsamp,
                                   (int)b->overlap);
        }
        PPL_LAUNCHED();
    }
    b->cur = next;
    b->fill = b->overlap;
    return PPL_OK;
}

extern "C" void* ppl_ddbuf_data(ppl_ddbuf* b) { return b ? b->d_buf[b->cur] : nullptr; }

extern "C" const float* ppl_ddbuf_row_offsets(const ppl_ddbuf* b)
{
    return b && b->dtype == PPL_U8 ? b->d_rowOff[b->cur] : nullptr;
}

extern "C" unsigned ppl_ddbuf_fill(const ppl_ddbuf* b) { return b ? b->fill : 0u; }
