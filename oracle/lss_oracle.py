"""CPU FP64 restatement of the reference's large-scale-structure functions on top of the solver
output: the FFTLog correlation function, Savitzky-Golay smoothing of P(k), the Kaiser anisotropic
spectrum and its multipoles (src/discoeb/perturbations.py:1409-1569, util.py:291-328).
TEST INFRASTRUCTURE: only tests/ use it, as the checker of the HIP path; the product package
never imports it.

Pinning (tests/test_lss_oracle.py): scipy.signal.savgol_coeffs for the filter coefficients (to
its own rounding) and exact reproduction of cubics by the filter; the Legendre projection of the Kaiser spectrum,
integrated by Gauss-Legendre quadrature, against the multipole formulas; the FFTLog correlation
function against direct quadrature of k^2 P(k) j_l(k r) / (2 pi^2) for a smooth spectrum.
"""
import numpy as np

from oracle.cmb_oracle import lngamma_complex


def savgol_coeffs(window_length, polyorder):
    """Filter taps of util.py:291-328: the minimum-norm solution c of A c = e_0 with
    A[i, j] = x_j^i, x the window positions measured from its centre (half-integers for even
    windows) in decreasing order, so that np.convolve(y, c, 'same') evaluates the local
    least-squares polynomial at the centre.  Computed exactly, in rational arithmetic, as
    c = A^T (A A^T)^-1 e_0, then rounded: the reference's floating-point lstsq on the unscaled
    Vandermonde matrix (like scipy.signal.savgol_coeffs) carries its own rounding, up to 8e-10
    relative at window 101, order 4."""
    from fractions import Fraction
    half, rem = divmod(int(window_length), 2)
    pos = Fraction(2 * half - 1, 2) if rem == 0 else Fraction(half)
    x = [pos - t for t in range(int(window_length))]
    m = polyorder + 1
    M = [[sum(v ** (i + j) for v in x) for j in range(m)] + [Fraction(int(i == 0))] for i in range(m)]
    for c in range(m):
        piv = next(r for r in range(c, m) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        for r in range(m):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [a - f * b for a, b in zip(M[r], M[c])]
    z = [M[i][m] / M[i][i] for i in range(m)]
    return np.array([float(sum(z[i] * v ** i for i in range(m))) for v in x])


def savgol_filter(y, window_length, polyorder):
    """util.py:291-328: np.convolve with the taps, 'same' length (zero beyond the ends)."""
    return np.convolve(np.asarray(y, dtype=np.float64), savgol_coeffs(window_length, polyorder), mode="same")


def get_power(k, y, idx, param):
    """perturbations.py:1446-1468."""
    k = np.asarray(k, dtype=np.float64)
    return 2 * np.pi ** 2 * param["A_s"] * (k / param["k_p"]) ** (param["n_s"] - 1) * k ** (-3.0) * np.asarray(y)[..., idx] ** 2


def smoothing_window(k, dlogk):
    """The odd window length of get_power_smoothed (perturbations.py:1496-1497): dlogk over the first
    log-k spacing, rounded half to even, made odd."""
    w = int(np.round(dlogk / (np.log(k[1]) - np.log(k[0]))))
    return w + (w + 1) % 2


def get_power_smoothed(k, y, dlogk, idx, param):
    """perturbations.py:1473-1505: exp of the filtered log P (cubic, the window of smoothing_window);
    the first w // 2 and the last (w + 1) // 2 points keep the unsmoothed P (the reference's
    slices [:w // 2] and [-w // 2:], the latter flooring -w / 2)."""
    w = smoothing_window(k, dlogk)
    pm = get_power(k, y, idx, param)
    pms = np.exp(savgol_filter(np.log(pm), w, 3))
    pms[:w // 2] = pm[:w // 2]
    pms[-w // 2:] = pm[-w // 2:]
    return pms


def power_kaiser(y, kmodes, bias, nmu, param, mu_sampling=True, smooth_dlogk=None):
    """perturbations.py:1507-1543: P(k, mu) = (b delta_m - mu^2 theta_m)^2 with delta_m, theta_m the
    transfer functions scaled by the primordial amplitude (theta_m = theta / aH already carries
    the growth rate), or, with smoothing, sqrt of the smoothed spectra (theta_m negative).
    mu: nmu points uniform in mu on [-1, 1], or cos of nmu points uniform on [0, pi]."""
    kmodes = np.asarray(kmodes, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    mu = np.linspace(-1.0, 1.0, nmu) if mu_sampling else np.cos(np.linspace(0.0, np.pi, nmu))
    if smooth_dlogk is None:
        amp = np.sqrt(2 * np.pi ** 2 * param["A_s"] * (kmodes / param["k_p"]) ** (param["n_s"] - 1) * kmodes ** (-3.0))
        dm, tm = amp * y[:, 4], amp * y[:, 5]
    else:
        dm = np.sqrt(get_power_smoothed(kmodes, y, smooth_dlogk, 4, param))
        tm = -np.sqrt(get_power_smoothed(kmodes, y, smooth_dlogk, 5, param))
    return (bias * dm[:, None] - mu[None, :] ** 2 * tm[:, None]) ** 2, mu


def power_multipoles(y, kmodes, b, param):
    """perturbations.py:1547-1569: monopole, quadrupole, hexadecapole of the Kaiser spectrum."""
    kmodes = np.asarray(kmodes, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    amp = np.sqrt(2 * np.pi ** 2 * param["A_s"] * (kmodes / param["k_p"]) ** (param["n_s"] - 1) * kmodes ** (-3.0))
    dm, tm = amp * y[:, 4], amp * y[:, 5]
    p0 = b ** 2 * dm ** 2 - 2 * b / 3 * dm * tm + tm ** 2 / 5
    p2 = -4 * b / 3 * dm * tm + 4 / 7 * tm ** 2
    p4 = 8 / 35 * tm ** 2
    return p0, p2, p4


def get_xi_from_P(k, Pk, ell=0):
    """perturbations.py:1409-1443 (Talman 1978, Hamilton 2000): rfft of P k^1.5 on the log-uniform
    grid, phases exp(2 i (Im lnGamma((1.5 + l) / 2 + i kappa_m) - kappa_m ln pi)) with
    kappa_m = pi m / ln(k_max / k_min), irfft, times Re(i^l) / (2 pi r)^1.5 at r = 2 pi / k.
    N = len(k) must be even (for odd N the reference's irfft returns N - 1 points, which do not
    broadcast against its N radii).  Returns (xi, r), both in increasing r."""
    k = np.asarray(k, dtype=np.float64)
    Pk = np.asarray(Pk, dtype=np.float64)
    N = k.size
    assert N % 2 == 0
    L = np.log(k[-1] / k[0])
    f = np.fft.rfft(Pk * k ** 1.5)
    ki = np.pi * np.arange(N // 2 + 1) / L
    theta = np.imag(lngamma_complex((1.5 + ell) / 2 + 1j * ki))
    f = f * np.exp(2j * (theta - np.log(np.pi) * ki))
    r = 2 * np.pi / k
    xi = np.real(1j ** ell * np.fft.irfft(f, N) / (2 * np.pi * r) ** 1.5)
    return xi[::-1], r[::-1]
