"""Product FIR coefficients (host FP64) against the oracle restatement."""
import numpy as np
import pytest

from conftest import load_ppl


@pytest.mark.parametrize("window", ["kaiser", "gaussian", "blackman", "hamming", "hann"])
@pytest.mark.parametrize("n_taps,n_ch", [(8, 16), (8, 64), (4, 256), (16, 32)])
def test_coefficients_match_oracle(orc, window, n_taps, n_ch):
    ppl = load_ppl()
    a = ppl.ppf_coefficients(n_taps, n_ch, window)
    b = orc.ppf_coefficients(n_taps, n_ch, window)
    assert np.allclose(a, b, rtol=1e-10, atol=1e-15)


def test_unknown_window_rejected():
    ppl = load_ppl()
    with pytest.raises(KeyError):
        ppl.ppf_coefficients(8, 16, "rectangular")
