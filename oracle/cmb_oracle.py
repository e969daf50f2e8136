"""CPU FP64 restatement of the CMB source-function and line-of-sight chain of the reference
(src/discoeb/cmb.py, util.py:13-45, spline_interpolation.py).  TEST INFRASTRUCTURE: only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker of the
HIP path; the product package never imports it.

Every function works on plain numpy float64 arrays (the background tables are passed in as
arrays), so the same inputs can be handed to the HIP kernels and to this restatement.

Pinning (tests/test_cmb_oracle.py): the reference's visibility-function fixture (background
only, exact inputs), its D_ell fixture (closed form), scipy.special.loggamma for the complex
log-gamma, brute-force spherical-Bessel quadrature for the FFTLog line of sight, the RHS's own
eta' for the metric recovered from the Einstein equations, a finite-difference derivative of
alpha for alpha', and scipy's natural cubic spline for the spline helpers.
"""
import numpy as np


# ---------------------------------------------------------------------------------------------
# natural cubic splines (spline_interpolation.py:6-276)
# ---------------------------------------------------------------------------------------------
def _fill_forward(v):
    """NaNs replaced by the last non-NaN value along the last axis (spline_interpolation.py:32-40)."""
    v = np.array(v, dtype=np.float64, copy=True)
    if not np.isnan(v).any():
        return v
    flat = v.reshape(-1, v.shape[-1])
    for row in flat:
        last = row[0]
        for i in range(row.size):
            if np.isnan(row[i]):
                row[i] = last
            last = row[i]
    return flat.reshape(v.shape)


def spline_second_derivatives(x, y):
    """S (S[0] = S[-1] = 0) of the natural cubic spline through y[..., n] on knots x[n], by the
    Thomas algorithm on the interior system (spline_interpolation.py:41-82)."""
    x = np.asarray(x, dtype=np.float64)
    y = _fill_forward(y)
    n = x.size
    S = np.zeros(y.shape)
    m = n - 2
    if m <= 0:
        return S
    h = np.diff(x)
    a, b, c = h[:-1], 2.0 * (h[:-1] + h[1:]), h[1:]
    d = 6.0 * ((y[..., 2:] - y[..., 1:-1]) / h[1:] - (y[..., 1:-1] - y[..., :-2]) / h[:-1])
    cp = np.zeros(m)
    dp = np.zeros(y.shape[:-1] + (m,))
    cp[0] = c[0] / b[0]
    dp[..., 0] = d[..., 0] / b[0]
    for i in range(1, m):
        den = b[i] - a[i] * cp[i - 1]
        cp[i] = c[i] / den if i < m - 1 else 0.0
        dp[..., i] = (d[..., i] - a[i] * dp[..., i - 1]) / den
    Si = np.zeros_like(dp)
    Si[..., m - 1] = dp[..., m - 1]
    for i in range(m - 2, -1, -1):
        Si[..., i] = dp[..., i] - cp[i] * Si[..., i + 1]
    S[..., 1:-1] = Si
    return S


def spline_index(x, xq):
    """count(x < xq) - 1 clipped to [0, n - 2] (spline_interpolation.py:131-145)."""
    return np.clip(np.searchsorted(x, xq, side="left") - 1, 0, x.size - 2)


def spline_eval(x, y, S, xq):
    """Value of the natural cubic spline (spline_interpolation.py:147-176)."""
    x = np.asarray(x, dtype=np.float64)
    xq = np.asarray(xq, dtype=np.float64)
    i = spline_index(x, xq)
    h = x[i + 1] - x[i]
    t = (xq - x[i]) / h
    A, B = 1.0 - t, t
    return A * y[..., i] + B * y[..., i + 1] + ((A ** 3 - A) * S[..., i] + (B ** 3 - B) * S[..., i + 1]) * (h ** 2) / 6.0


def spline_deriv(x, y, S, xq):
    """First derivative of the natural cubic spline (spline_interpolation.py:219-239)."""
    x = np.asarray(x, dtype=np.float64)
    xq = np.asarray(xq, dtype=np.float64)
    i = spline_index(x, xq)
    h = x[i + 1] - x[i]
    d = xq - x[i]
    b = (y[..., i + 1] - y[..., i]) / h - h * (S[..., i + 1] + 2.0 * S[..., i]) / 6.0
    c = S[..., i] / 2.0
    dd = (S[..., i + 1] - S[..., i]) / (6.0 * h)
    return b + 2.0 * c * d + 3.0 * dd * d ** 2


def spline(x, y, xq, kind=0):
    x = np.asarray(x, dtype=np.float64)
    y = _fill_forward(np.asarray(y, dtype=np.float64))
    S = spline_second_derivatives(x, y)
    return spline_eval(x, y, S, xq) if kind == 0 else spline_deriv(x, y, S, xq)


# ---------------------------------------------------------------------------------------------
# background pieces used by the sources
# ---------------------------------------------------------------------------------------------
def aprimeoa(a, sc, rhonu):
    """Conformal Hubble rate (background.py:101-123); sc: dict of background scalars, rhonu =
    rho_nu/rho_nu0 at a (from the logrhonu_of_loga spline)."""
    rho_Q = a ** (-3 * (1 + sc["w_DE_0"] + sc["w_DE_a"])) * np.exp(3 * (a - 1) * sc["w_DE_a"])
    grho = (sc["grhom"] * sc["Omegam"] / a + (sc["grhog"] + sc["grhor"] * (sc["Neff"] + sc["Nmnu"] * rhonu)) / a ** 2 +
            sc["grhom"] * sc["OmegaDE"] * rho_Q * a ** 2 + sc["grhom"] * sc["Omegak"])
    return np.sqrt(grho / 3.0)


def background_quantities(a, sc, rhonu):
    """rho_nu, rho_Q, w_Q at a (background.py:348-406)."""
    rho_Q = a ** (-3 * (1 + sc["w_DE_0"] + sc["w_DE_a"])) * np.exp(3 * (a - 1) * sc["w_DE_a"])
    return {"rhonu": rhonu, "rho_Q": rho_Q, "w_Q": sc["w_DE_0"] + sc["w_DE_a"] * (1.0 - a)}


def visibility_functions(tau, tau_tab, opac, gvis, gvisprime, gvispprime, optical_depth):
    """Opacity, its derivative, g, g', g'' and the optical depth at conformal times tau from
    splines in log tau of the thermal-history tables (cmb.py:92-166)."""
    lt = np.log(tau_tab)
    lq = np.log(np.asarray(tau, dtype=np.float64))
    out = {"opac": spline(lt, opac, lq), "opacprime": spline(lt, opac, lq, kind=1) / tau}
    out["gvis"] = spline(lt, gvis, lq)
    out["gvisprime"] = spline(lt, gvisprime, lq)
    out["gvispprime"] = spline(lt, gvispprime, lq)
    out["optical_depth"] = np.exp(spline(lt, np.log(np.maximum(optical_depth, 1e-10)), lq))
    return out


# ---------------------------------------------------------------------------------------------
# source function (cmb.py:17-89, 169-633)
# ---------------------------------------------------------------------------------------------
def extract_perturbations(yout, youtprime, lmaxg, lmaxgp, lmaxr):
    """Named views of the state and its time derivative (cmb.py:17-89)."""
    idxb, idxg = 5, 7
    idxgp = 7 + (lmaxg + 1)
    idxr = 9 + lmaxg + lmaxgp
    y, yp = yout, youtprime
    return {"idxb": idxb, "idxg": idxg, "idxgp": idxgp, "idxr": idxr, "iq0": 10 + lmaxg + lmaxgp + lmaxr,
            "eta": y[..., 2], "deltac": y[..., 3], "thetac": y[..., 4], "deltab": y[..., idxb], "thetab": y[..., idxb + 1],
            "deltabprime": yp[..., idxb], "thetabprime": yp[..., idxb + 1],
            "deltag": y[..., idxg], "thetag": y[..., idxg + 1], "shearg": y[..., idxg + 2] / 2.0,
            "deltagprime": yp[..., idxg], "thetagprime": yp[..., idxg + 1], "sheargprime": yp[..., idxg + 2] / 2.0,
            "deltar": y[..., idxr], "thetar": y[..., idxr + 1], "shearr": y[..., idxr + 2] / 2.0,
            "deltarprime": yp[..., idxr], "thetarprime": yp[..., idxr + 1], "shearrprime": yp[..., idxr + 2] / 2.0,
            "deltaq": y[..., -2], "thetaq": y[..., -1]}


def neutrino_perturbations(yout, yprime, a, amnu, q, w, iq0, aH):
    """delta rho, delta p, flux, shear of the massive neutrinos and the time derivatives of delta
    rho and shear, by quadrature over the momentum bins (perturbations.py:29-85,
    cmb.py:169-240).  a, aH: (n_tau,); yout, yprime: (n_k, n_tau, nvar)."""
    nq = q.size
    p0 = yout[..., iq0:iq0 + nq]
    p1 = yout[..., iq0 + nq:iq0 + 2 * nq]
    p2 = yout[..., iq0 + 2 * nq:iq0 + 3 * nq]
    p0p = yprime[..., iq0:iq0 + nq]
    p2p = yprime[..., iq0 + 2 * nq:iq0 + 3 * nq]
    aq = a[:, None] * amnu / q[None, :]
    v = 1.0 / np.sqrt(1.0 + aq ** 2)
    vprime = -0.5 * aq * aH[:, None] * v ** 3
    return {"drhonu": np.sum(w * p0 / v, axis=-1), "dpnu": np.sum(w * p0 * v, axis=-1) / 3,
            "fnu": np.sum(w * p1, axis=-1), "shearnu": np.sum(w * p2 * v, axis=-1) * 2 / 3,
            "rho_nu_prime": np.sum(w * (p0p / v - p0 / v ** 2 * vprime), axis=-1),
            "shear_nu_prime": np.sum(w * (p2p * v + p2 * vprime), axis=-1) * 2 / 3, "aprimeoa": aH}


def metric_perturbations(pt, nu, bq, sc, kmodes, a):
    """h', eta', alpha, alpha', alpha'' from the Einstein equations (cmb.py:243-345).  As in the
    reference, the time derivative of the total density uses the perturbed neutrino density's
    derivative rho_nu_prime."""
    k = np.asarray(kmodes, dtype=np.float64)[:, None]
    Omegac = sc["Omegam"] - sc["Omegab"]
    rhonu, rho_Q, w_Q = bq["rhonu"], bq["rho_Q"], bq["w_Q"]
    aH = nu["aprimeoa"]
    rpptq = (1 + w_Q) * rho_Q * sc["grhom"] * sc["OmegaDE"] * pt["thetaq"] * a ** 2
    grhoprime = (-sc["grhom"] * sc["Omegam"] / a - 2 * (sc["grhog"] + sc["grhor"] * (sc["Neff"] + sc["Nmnu"] * rhonu)) / a ** 2 +
                 sc["grhom"] * sc["OmegaDE"] * (-3 * (1 + sc["w_DE_0"] + (1 - a) * sc["w_DE_a"]) * rho_Q * a ** 2 +
                                                2 * rho_Q * a ** 2)) * aH + sc["grhor"] * sc["Nmnu"] * nu["rho_nu_prime"] / a ** 2
    aHprime = grhoprime / (6 * aH)
    dgrho = (sc["grhom"] * (Omegac * pt["deltac"] + sc["Omegab"] * pt["deltab"]) / a +
             (sc["grhog"] * pt["deltag"] + sc["grhor"] * (sc["Neff"] * pt["deltar"] + sc["Nmnu"] * nu["drhonu"])) / a ** 2 +
             sc["grhom"] * sc["OmegaDE"] * pt["deltaq"] * rho_Q * a ** 2)
    dgtheta = (sc["grhom"] * (Omegac * pt["thetac"] + sc["Omegab"] * pt["thetab"]) / a +
               4.0 / 3.0 * (sc["grhog"] * pt["thetag"] + sc["Neff"] * sc["grhor"] * pt["thetar"]) / a ** 2 +
               sc["Nmnu"] * sc["grhor"] * k * nu["fnu"] / a ** 2 + rpptq)
    dgshear = (4.0 / 3.0 * (sc["grhog"] * pt["shearg"] + sc["Neff"] * sc["grhor"] * pt["shearr"]) / a ** 2 +
               sc["Nmnu"] * sc["grhor"] * nu["shearnu"] / a ** 2)
    dgshearprime = (4.0 / 3.0 * (sc["grhog"] * pt["sheargprime"] + sc["Neff"] * sc["grhor"] * pt["shearrprime"]) / a ** 2 +
                    sc["Nmnu"] * sc["grhor"] * nu["shear_nu_prime"] / a ** 2 - 2 * aH * dgshear)
    hprime = (2.0 * k ** 2 * pt["eta"] + dgrho) / aH
    etaprime = 0.5 * dgtheta / k ** 2
    alpha = (hprime + 6.0 * etaprime) / 2.0 / k ** 2
    alphaprime = -3 * dgshear / (2 * k ** 2) + pt["eta"] - 2 * aH * alpha
    alphapprime = -3 * dgshearprime / (2 * k ** 2) + etaprime - 2 * (aHprime * alpha + aH * alphaprime)
    return {"hprime": hprime, "etaprime": etaprime, "alpha": alpha, "alphaprime": alphaprime, "alphapprime": alphapprime,
            "aprimeoa": aH, "aprimeoaprime": aHprime}


def polarization_terms(pt, metric, vis, yout, yprime, kmodes):
    """Pi = F2 + G0 + G2 and its first two derivatives (cmb.py:349-395)."""
    k = np.asarray(kmodes, dtype=np.float64)[:, None]
    idxg, idxgp = pt["idxg"], pt["idxgp"]
    couplprime = (8 * (pt["thetagprime"] + k ** 2 * metric["alphaprime"]) / 15 -
                  k * 0.6 * (yprime[:, :, idxg + 3] + yprime[:, :, idxgp + 1] + yprime[:, :, idxgp + 3]))
    P = yout[:, :, idxg + 2] + yout[:, :, idxgp] + yout[:, :, idxgp + 2]
    Pp = yprime[:, :, idxg + 2] + yprime[:, :, idxgp] + yprime[:, :, idxgp + 2]
    Ppp = couplprime - 0.3 * (vis["opacprime"] * P + vis["opac"] * Pp)
    return {"polarization_term": P, "polarization_termprime": Pp, "polarization_termpprime": Ppp}


def source_function(pt, metric, vis, pol, kmodes, tau, isw="spline"):
    """S1 (ISW), S2 (Sachs-Wolfe), S3 (Doppler), S4 (polarisation) and their sum
    (cmb.py:398-633).  isw="spline": S1 = e^-tau_c d/dtau(eta + alpha') from a natural spline in
    tau (the reference's current compute_source_term_isw, cmb.py:437-449); "algebraic": S1 =
    e^-tau_c (eta' + alpha''), the form its docstring states (cmb.py:406-408)."""
    k = np.asarray(kmodes, dtype=np.float64)[:, None]
    tau = np.asarray(tau, dtype=np.float64)
    if isw == "algebraic":
        s1 = metric["etaprime"] + metric["alphapprime"]
    else:
        pot = pt["eta"] + metric["alphaprime"]
        s1 = spline(tau, pot, tau, kind=1)
    S1 = np.exp(-vis["optical_depth"]) * s1
    S2 = vis["gvis"] * (0.25 * pt["deltag"] + 2 * metric["alphaprime"] + pol["polarization_term"] / 16 +
                        (pt["thetabprime"] + 3 / 16 * pol["polarization_termpprime"]) / k ** 2)
    S3 = vis["gvisprime"] * (metric["alpha"] + (pt["thetab"] + 3 / 8 * pol["polarization_termprime"]) / k ** 2)
    S4 = vis["gvispprime"] * 3 / 16 * pol["polarization_term"] / k ** 2
    return {"S1": S1, "S2": S2, "S3": S3, "S4": S4, "S": S1 + S2 + S3 + S4}


def sources_from_states(yout, yprime, kmodes, aexp, tau, bgd, lmaxg, lmaxgp, lmaxr, nqmax, isw="spline"):
    """The whole chain of compute_Cell_spectrum_from_cosmo_params steps 6-10
    (cmb.py:1050-1073) for states yout / yprime (n_k, n_tau, nvar).  bgd: dict with the
    background scalars (grhom, grhog, grhor, Omegam, Omegab, Neff, Nmnu, OmegaDE, Omegak,
    w_DE_0, w_DE_a, amnu), rhonu (at aexp), the thermal tables tau_tab, opac, gvis,
    gvisprime, gvispprime, optical_depth, and the momentum bins q, w."""
    a = np.asarray(aexp, dtype=np.float64)
    pt = extract_perturbations(yout, yprime, lmaxg, lmaxgp, lmaxr)
    bq = background_quantities(a, bgd, bgd["rhonu"])
    aH = aprimeoa(a, bgd, bgd["rhonu"])
    vis = visibility_functions(tau, bgd["tau_tab"], bgd["opac"], bgd["gvis"], bgd["gvisprime"], bgd["gvispprime"],
                               bgd["optical_depth"])
    nu = neutrino_perturbations(yout, yprime, a, bgd["amnu"], bgd["q"], bgd["w"], pt["iq0"], aH)
    metric = metric_perturbations(pt, nu, bq, bgd, kmodes, a)
    pol = polarization_terms(pt, metric, vis, yout, yprime, kmodes)
    out = source_function(pt, metric, vis, pol, kmodes, tau, isw=isw)
    out.update(metric=metric, vis=vis, pol=pol, nu=nu)
    return out


# ---------------------------------------------------------------------------------------------
# line of sight (util.py:13-45, cmb.py:636-927)
# ---------------------------------------------------------------------------------------------
_LANCZOS_7 = np.array([0.99999999999980993227684700473478, 676.520368121885098567009190444019,
                       -1259.13921672240287047156078755283, 771.3234287776530788486528258894,
                       -176.61502916214059906584551354, 12.507343278686904814458936853,
                       -0.13857109526572011689554707, 9.984369578019570859563e-6, 1.50563273514931155834e-7])


def lngamma_complex(z):
    """log Gamma(z) for complex z by the Lanczos approximation (g = 7, 9 terms) with the
    reflection formula for Re z <= 1/2 (util.py:13-45)."""
    z = np.asarray(z, dtype=np.complex128)

    def lanczos(zz):
        zz = zz - 1.0
        j = np.arange(1, 9)
        t = zz[..., None] + j
        Ag = _LANCZOS_7[0] + np.sum(_LANCZOS_7[1:] / np.abs(t) ** 2 * np.conj(t), axis=-1)
        return (zz + 0.5) * np.log(zz + 7.5) - (zz + 7.5) + 0.9189385332046727418 + np.log(Ag)

    lnpi = 1.14472988584940017414342735135
    refl = np.real(z) <= 0.5
    out = np.empty_like(z)
    if refl.any():
        zr = z[refl]
        with np.errstate(over="ignore"):  # |Im z| > ~226: sin overflows and Gamma -> 0, as in the reference
            out[refl] = lnpi - np.log(np.sin(np.pi * zr)) - lanczos(1.0 - zr)
    if (~refl).any():
        out[~refl] = lanczos(z[~refl])
    return out


def bessel_integral_coefficients(ell, p):
    """M_l(p) = 2^(p-1) sqrt(pi) Gamma((1+l+p)/2) / Gamma((2+l-p)/2), the coefficient of
    int_0^inf chi^p j_l(k chi) d chi = M_l(p) k^-(p+1) (cmb.py:636-677); (n_ell, n_p)."""
    ell = np.atleast_1d(np.asarray(ell, dtype=np.float64))[:, None]
    p = np.atleast_1d(np.asarray(p, dtype=np.complex128))[None, :]
    a1 = (1.0 + ell + p) / 2.0
    a2 = (2.0 + ell - p) / 2.0
    return np.exp((p - 1.0) * np.log(2.0) + 0.5 * np.log(np.pi) + lngamma_complex(a1) - lngamma_complex(a2))


def densify_k(kmodes, S, n_k_dense):
    """S interpolated along log k (natural cubic spline per time slice) onto n_k_dense
    log-spaced wavenumbers (cmb.py:740-751)."""
    lk = np.log(kmodes)
    lkd = np.linspace(lk[0], lk[-1], n_k_dense)
    Sd = spline(lk, np.asarray(S).T, lkd).T
    return np.exp(lkd), Sd


def fftlog_grid(tau, tau0, n_fftlog):
    """Log-spaced comoving-distance grid of the FFTLog expansion (cmb.py:753-775)."""
    chi = (tau0 - np.asarray(tau, dtype=np.float64))[::-1]
    lmin = np.log(max(chi[0], 1.0)) - 2.0
    lmax = np.log(chi[-1]) + 2.0
    dl = (lmax - lmin) / n_fftlog
    return chi, lmin, dl, np.exp(lmin + np.arange(n_fftlog) * dl)


def fftlog_coefficients(kmodes, tau, S, tau0, n_fftlog):
    """c_n k^-(p_n + 1) of the biased FFTLog expansion of every row of S (cmb.py:753-803):
    (n_k, n_fftlog/2 + 1) complex, and p_n."""
    chi, lmin, dl, chi_f = fftlog_grid(tau, tau0, n_fftlog)
    Sc = np.asarray(S, dtype=np.float64)[:, ::-1]
    lco = np.log(np.maximum(chi, 1e-10))
    Sf = spline(lco, Sc, np.log(chi_f))
    bias = -1.1
    nf = n_fftlog // 2 + 1
    eta = 2.0 * np.pi * np.arange(nf) / (n_fftlog * dl)
    p = bias + 1j * eta
    c = np.fft.rfft(Sf * chi_f[None, :] ** (-bias), axis=-1) * np.exp(-1j * eta * lmin) / n_fftlog
    ck = c * np.exp(-(p[None, :] + 1.0) * np.log(np.asarray(kmodes, dtype=np.float64))[:, None])
    return ck, p


def source_E(gvis, pol, kmodes, tau, tau0):
    """E-mode source 3 g Pi / (16 (k chi)^2), chi = tau0 - tau, zero for chi < 1 Mpc (no counterpart
    in the reference; Zaldarriaga & Seljak 1997 with Pi = F_2 + G_0 + G_2)."""
    chi = tau0 - np.asarray(tau, dtype=np.float64)
    k = np.asarray(kmodes, dtype=np.float64)[:, None]
    safe = np.maximum(chi, 1.0)[None, :]
    return np.where(chi[None, :] >= 1.0, 3.0 / 16.0 * np.asarray(gvis)[None, :] * pol / (k ** 2 * safe ** 2), 0.0)


def spin2_factor(ellmax):
    """sqrt((l + 2)! / (l - 2)!) for l = 0..ellmax, zero below l = 2."""
    ell = np.arange(ellmax + 1, dtype=np.float64)
    return np.sqrt(np.where(ell >= 2, (ell + 2) * (ell + 1) * ell * (ell - 1), 0.0))


def compute_theta_ell(ellmax, kmodes, tau, S, tau0, n_fftlog=16384, n_k_dense=0, field="T"):
    """Theta_l(k) = int S(k, tau) j_l(k (tau0 - tau)) d tau by FFTLog (cmb.py:680-835); field "E"
    multiplies by spin2_factor (the E-mode projection, no counterpart in the reference)."""
    kmodes = np.asarray(kmodes, dtype=np.float64)
    if n_k_dense > 0:
        kmodes, S = densify_k(kmodes, S, n_k_dense)
    ck, p = fftlog_coefficients(kmodes, tau, S, tau0, n_fftlog)
    M = bessel_integral_coefficients(np.arange(ellmax + 1, dtype=np.float64), p)
    th = np.outer(ck[:, 0].real, M[:, 0].real)
    if n_fftlog % 2 == 0:
        th = th + 2.0 * (ck[:, 1:-1].real @ M[:, 1:-1].real.T - ck[:, 1:-1].imag @ M[:, 1:-1].imag.T)
        th = th + np.real(np.outer(ck[:, -1], M[:, -1]))
    else:
        th = th + 2.0 * (ck[:, 1:].real @ M[:, 1:].real.T - ck[:, 1:].imag @ M[:, 1:].imag.T)
    if field == "E":
        th = th * spin2_factor(ellmax)[None, :]
    return th, kmodes


def compute_Cell(theta_ell, kmodes, n_s, k_p, theta_ell2=None):
    """C_l = trapezoid over ln k of (k/k_p)^(n_s-1) Theta_l(k)^2 (cmb.py:838-873); with theta_ell2
    the cross spectrum."""
    kmodes = np.asarray(kmodes, dtype=np.float64)
    t2 = np.asarray(theta_ell if theta_ell2 is None else theta_ell2)
    f = (kmodes[:, None] / k_p) ** (n_s - 1) * np.asarray(theta_ell) * t2
    lk = np.log(kmodes)
    return np.sum(np.diff(lk)[:, None] * (f[1:] + f[:-1]) / 2.0, axis=0)


def compute_Dell(Cell, A_s, Tcmb, ellmax=None):
    """D_l = l (l + 1) C_l A_s 2 T_cmb^2 1e12 for l >= 2 (cmb.py:876-927)."""
    Cell = np.asarray(Cell, dtype=np.float64)
    if ellmax is None:
        ellmax = Cell.shape[0] - 1
    ell = np.arange(ellmax + 1)
    return ell[2:], ell[2:] * (ell[2:] + 1) * Cell[2:ellmax + 1] * A_s * 2 * Tcmb ** 2 * 1e12


def cmb_aexp_out(n_early=32, n_pre_recomb=64, n_recomb=128, n_post=96):
    """Output scale factors of compute_Cell_spectrum_from_cosmo_params: log-spaced pieces
    split at z = 3000, 1400, 600 (cmb.py:999-1014)."""
    from oracle.oracle import geomspace
    a1, a2, a3 = 1.0 / 3001.0, 1.0 / 1401.0, 1.0 / 601.0
    return np.concatenate([geomspace(1e-4, a1, n_early, endpoint=False), geomspace(a1, a2, n_pre_recomb, endpoint=False),
                           geomspace(a2, a3, n_recomb, endpoint=False), geomspace(a3, 1.0, n_post)])
