"""CPU restatement of BandPassRecorder (src/lib/src/BandPassRecorder.cpp) and
BandPassOutput (src/lib/src/BandPassOutput.cpp), in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports this module.

Arithmetic follows the reference's C++11 expression types: float running
sums, std::pow(float, int) evaluated in double, a float total updated with a
double term narrowed back to float.  Three places of the reference cannot run
as written, and the restatement takes the evident intent instead:
  - LAPACK sgels is called with (const char*)'N' as its first argument (a
    pointer made from a character, :18), a design matrix stored by rows where
    LAPACK reads columns (:60), and a right-hand side sized for the
    coefficients, not the bins (:53, filled at :154-156); the least-squares
    solution is what it was asked for, here numpy.linalg.lstsq (SVD) in
    float64 on the float design matrix;
  - the trim loop never resets its counter (:106-129), so it would never end
    once a point is trimmed; here each pass counts its own trims;
  - a trimmed point is set to fit * totalSamples (:126), a value of the
    un-averaged sums, in a spectrum already averaged (:101-103); here it is set
    to the fit's value, which the comment above that line says it should be.
"""
import numpy as np

F32 = np.float32


def accumulate(blobs, sums=None):
    """Integrated spectrum (:85-97): sum[s*nCh + c] += I[t][s][0][c] over t
    in order, float32.  blobs: iterable of float32 [t][s][pol][c] arrays."""
    for blob in blobs:
        blob = np.asarray(blob, F32)
        n_t, n_sub, _, n_ch = blob.shape
        if sums is None:
            sums = np.zeros(n_sub * n_ch, F32)
        for t in range(n_t):          # elementwise float32 adds, time order
            sums += blob[t, :, 0, :].reshape(-1)
    return sums


def bin_freqs(n_bins, start, end):
    """BinMap(nBins), setStart(start), setEnd(end) (BinMap.cpp): the bin
    centres start + width*i in double, width (end - start)/nBins, as the
    float frequencies the recorder keeps (:56-62)."""
    width = (np.float64(F32(end)) - np.float64(F32(start))) / n_bins
    return (np.float64(F32(start)) + width * np.arange(n_bins)).astype(F32), width


def evaluate(params, v):
    """_theFit / BandPass::_evaluate: tot(float) += p[i] * pow(double(v), i)."""
    v = np.asarray(v, F32).astype(np.float64)
    tot = np.zeros(v.shape, F32)
    for i, p in enumerate(params):
        tot = (tot.astype(np.float64) + np.float64(F32(p)) * v ** i).astype(F32)
    return tot


def poly_fit(freq, y, n_coeffs):
    """Least-squares coefficients of sum_p c[p] freq^p through y, the design
    matrix entries float(pow(freq, p)) (:56-62); returned as float32."""
    A = np.stack([(np.asarray(freq, F32).astype(np.float64) ** p).astype(F32) for p in range(n_coeffs)],
                 axis=1).astype(np.float64)
    c, *_ = np.linalg.lstsq(A, np.asarray(y, F32).astype(np.float64), rcond=None)
    return c.astype(F32)


def _seq_sum(x):
    """float32 sum in index order."""
    return np.cumsum(np.asarray(x, F32), dtype=F32)[-1] if len(x) else F32(0)


def _seq_sum_sq(x):
    """float acc = float(double(acc) + pow(double(x), 2)) in index order."""
    acc = F32(0)
    for v in np.asarray(x, F32).astype(np.float64):
        acc = F32(np.float64(acc) + v * v)
    return acc


def _spread(sq, s, n):
    """sqrt(sq/n - pow(s/n, 2)): float quotients, the rest in double, to float."""
    return F32(np.sqrt(np.float64(F32(sq / F32(n))) - np.float64(F32(s / F32(n))) ** 2))


def fit_spectrum(spectrum, start, end, n_coeffs):
    """The fitting stage (:101-137) on an averaged spectrum.  Returns
    (params, rms, cleaned spectrum, fits made)."""
    y = np.array(spectrum, F32)
    n = len(y)
    freq, _ = bin_freqs(n, start, end)
    fits = 0
    while True:
        fit = poly_fit(freq, y, n_coeffs)
        fits += 1
        model = evaluate(fit, freq)
        residual = (y - model).astype(F32)
        margin = F32(2.0) * _spread(_seq_sum_sq(residual), _seq_sum(residual), n)
        out = residual > margin
        y[out] = model[out]
        if not out.any() or fits > n:
            break
    rms = _spread(_seq_sum_sq(y), _seq_sum(y), n)
    return fit, rms, y, fits


def average(sums, total):
    """sum /= totalSamples (:101-103): float division by float(total)."""
    return (np.asarray(sums, F32) / F32(total)).astype(F32)


def bandpass_stats(params, n_bins, start, end):
    """BandPass::setData (BandPass.cpp:26-42) on the recorder's map: the model
    at every bin, its float mean in bin order and element n/2 of the sorted
    values as the median."""
    freq, _ = bin_freqs(n_bins, start, end)
    data = evaluate(params, freq)
    mean = F32(_seq_sum(data) / F32(n_bins))
    median = np.partition(data, n_bins // 2)[n_bins // 2]
    return data, mean, F32(median)


def bp_file_text(n_bins, start, end, rms, median, params):
    """BandPassOutput::sendStream (BandPassOutput.cpp:51-69): QTextStream's
    default 6 significant digits, start/end as floats, width as a double."""
    _, width = bin_freqs(n_bins, start, end)
    start_v = np.float64(F32(start))
    end_v = F32(start_v + width * n_bins)
    lines = ["# Generated by BandPassOutput", "%d" % n_bins, "%.6g" % np.float64(F32(start_v)),
             "%.6g" % np.float64(end_v), "%.6g" % width, "%.6g" % np.float64(F32(rms)),
             "%.6g" % np.float64(F32(median))] + ["%.6g" % np.float64(F32(p)) for p in params]
    return "\n".join(lines) + "\n"
