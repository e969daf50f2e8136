"""ctypes front-end of the CPU oracle (oracle/deb_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker of the HIP path.  The product package never imports it.

The parameter dictionary follows the reference's ``param`` keys
(src/discoeb/background.py:141-189, src/discoeb/approximations.py:9-20).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DEB_ORACLE_LIB names another build of the same source (bench.py's -O3 -march=native CPU baseline)
_LIB_PATH = os.environ.get("DEB_ORACLE_LIB") or os.path.join(_HERE, "_build", "libdeb_oracle.so")

PARAM_KEYS = [
    "Omegam", "Omegab", "H0", "Tcmb", "YHe", "Neff", "Nmnu", "mnu", "w_DE_0", "w_DE_a", "cs2_DE", "Omegak",
    "A_s", "n_s", "k_p",
    "use_tca", "tca_tau_c_over_tau_h_trigger", "tca_tau_c_over_tau_k_trigger",
    "use_rsa", "rsa_tau_c_over_tau_trigger", "rsa_tau_over_tau_k_trigger",
    "use_ur_fluid", "ur_fluid_tau_over_tau_k_trigger",
]
# approximations.py:9-20 defaults
_DEFAULTS = {
    "Omegak": 0.0, "k_p": 0.05, "cs2_DE": 1.0,
    "use_tca": False, "tca_tau_c_over_tau_h_trigger": 0.015, "tca_tau_c_over_tau_k_trigger": 0.010,
    "use_rsa": True, "rsa_tau_c_over_tau_trigger": 10.0, "rsa_tau_over_tau_k_trigger": 80.0,
    "use_ur_fluid": True, "ur_fluid_tau_over_tau_k_trigger": 120.0,
}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_bg_new.restype = ctypes.c_void_p
    lib.orc_bg_new.argtypes = [_dp, ctypes.c_int]
    lib.orc_bg_free.argtypes = [ctypes.c_void_p]
    lib.orc_bg_from_tables.restype = ctypes.c_void_p
    lib.orc_bg_from_tables.argtypes = [_dp, ctypes.c_int, ctypes.POINTER(ctypes.c_long)]
    lib.orc_bg_scalar.restype = ctypes.c_double
    lib.orc_bg_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    lib.orc_bg_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp]
    lib.orc_bg_spline.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _dp, ctypes.c_int, _dp]
    lib.orc_spline_eval.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp]
    lib.orc_nu_bins.argtypes = [ctypes.c_int, _dp, _dp]
    lib.orc_romb_test.restype = ctypes.c_double
    lib.orc_romb_test.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.orc_layout.argtypes = [ctypes.c_int] * 5
    lib.orc_rhs.argtypes = [ctypes.c_void_p, _ip, ctypes.c_double, _dp, ctypes.c_double, _dp]
    lib.orc_jacobian.argtypes = [ctypes.c_void_p, _ip, ctypes.c_double, _dp, ctypes.c_double, _dp, _dp]
    lib.orc_ics.argtypes = [ctypes.c_void_p, _ip, ctypes.c_double, ctypes.c_double, _dp]
    lib.orc_start_time.restype = ctypes.c_double
    lib.orc_start_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.orc_convert.argtypes = [ctypes.c_void_p, _ip, ctypes.c_double, _dp, _dp]
    lib.orc_rodas5_step.argtypes = [ctypes.c_void_p, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
    lib.orc_evolve.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, _dp, _ip, _ip, _ip]
    lib.orc_evolve_replay.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, _ip,
                                      ctypes.c_int, _dp, _dp]
    lib.orc_start_tau.restype = ctypes.c_double
    lib.orc_start_tau.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, ctypes.c_int]
    lib.orc_evolve_mesh.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, _dp, _ip]
    lib.orc_replay_norm.argtypes = [ctypes.c_void_p, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _ip, ctypes.c_int,
                                    ctypes.c_int, _dp]
    lib.orc_adjoint.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), _dp, _ip, _dp, ctypes.c_int, _dp, _dp,
                                ctypes.c_int, _dp, _ip, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
    _lib = lib
    return lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def param_vector(param):
    vals = []
    for k in PARAM_KEYS:
        v = param.get(k, _DEFAULTS.get(k))
        if v is None:
            raise KeyError(f"missing cosmological parameter {k!r}")
        vals.append(float(v))
    return np.array(vals, dtype=np.float64)


def geomspace(start, stop, num, endpoint=True):
    """jax.numpy.geomspace arithmetic: 10**(log10(start)*(1-s) + log10(stop)*s)."""
    l0, l1 = np.log10(start), np.log10(stop)
    div = (num - 1) if endpoint else num
    s = np.arange(num, dtype=np.float64) / div
    lin = l0 * (1.0 - s) + l1 * s
    if endpoint:
        lin[-1] = l1
    return 10.0 ** lin


class Background:
    """Oracle counterpart of evolve_background(thermo_module='RECFAST') (background.py:191-345)."""

    def __init__(self, param, num_thermo=256):
        self._lib = _load()
        self.param = dict(param)
        self.num_thermo = int(num_thermo)
        pv, pp = _d(param_vector(param))
        self._h = self._lib.orc_bg_new(pp, self.num_thermo)

    @classmethod
    def from_tables(cls, block, num_thermo, offsets, param=None):
        """Background taken verbatim from a HIP table block (numpy float64 array) whose array
        offsets (in doubles) are given for: nu_loga, nu_logrho, nu_logrho_S, nu_logp, nu_logp_S,
        nu_logpp, nu_logpp_S, aexp, loga, tau, xe, cs2a, tempba, S_tau_of_a, S_a_of_tau,
        S_xe_of_tau, S_cs2a_of_tau, S_tempba_of_tau, S_xe_of_loga, S_cs2a_of_loga."""
        self = cls.__new__(cls)
        self._lib = _load()
        self.param = dict(param or {})
        self.num_thermo = int(num_thermo)
        self._block = np.ascontiguousarray(block, dtype=np.float64)
        offs = (ctypes.c_long * len(offsets))(*[int(o) for o in offsets])
        self._h = self._lib.orc_bg_from_tables(self._block.ctypes.data_as(_dp), self.num_thermo, offs)
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.orc_bg_free(h)
            self._h = None

    def scalar(self, name):
        return self._lib.orc_bg_scalar(self._h, name.encode())

    def array(self, name):
        n = self._lib.orc_bg_array(self._h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.empty(n)
        self._lib.orc_bg_array(self._h, name.encode(), out.ctypes.data_as(_dp))
        return out

    def spline(self, name, x, kind=0):
        x, xp = _d(np.atleast_1d(x))
        out = np.empty_like(x)
        if self._lib.orc_bg_spline(self._h, name.encode(), int(kind), xp, x.size, out.ctypes.data_as(_dp)) != 0:
            raise KeyError(name)
        return out

    # --- per-mode pieces ------------------------------------------------------------------
    def rhs(self, lmax, tau, y, k):
        lm, lp = _i(lmax)
        y, yp = _d(y)
        f = np.empty_like(y)
        self._lib.orc_rhs(self._h, lp, float(tau), yp, float(k), f.ctypes.data_as(_dp))
        return f

    def jacobian(self, lmax, tau, y, k):
        lm, lp = _i(lmax)
        y, yp = _d(y)
        n = y.size
        J = np.empty((n, n))
        dT = np.empty(n)
        self._lib.orc_jacobian(self._h, lp, float(tau), yp, float(k), J.ctypes.data_as(_dp), dT.ctypes.data_as(_dp))
        return J, dT

    def ics(self, lmax, tau, k):
        lm, lp = _i(lmax)
        n = self._lib.orc_layout(*[int(v) for v in lmax])
        y = np.empty(n)
        self._lib.orc_ics(self._h, lp, float(tau), float(k), y.ctypes.data_as(_dp))
        return y

    def start_time(self, k):
        return self._lib.orc_start_time(self._h, float(k))

    def start_tau(self, k, aexp_out):
        """Initial conformal time of one mode: 0.99 min(min tau_out, start_time(k)) (perturbations.py:897)."""
        tau, tp = _d(self.spline("tau_of_a", np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))))
        return self._lib.orc_start_tau(self._h, float(k), tp, tau.size)

    def convert(self, lmax, k, y):
        lm, lp = _i(lmax)
        y, yp = _d(y)
        out = np.empty(20)
        self._lib.orc_convert(self._h, lp, float(k), yp, out.ctypes.data_as(_dp))
        return out

    def rodas5_step(self, lmax, k, t0, dt, y0):
        lm, lp = _i(lmax)
        y0, yp = _d(y0)
        y1 = np.empty_like(y0)
        err = np.empty_like(y0)
        self._lib.orc_rodas5_step(self._h, lp, float(k), float(t0), float(dt), yp, y1.ctypes.data_as(_dp),
                                  err.ctypes.data_as(_dp))
        return y1, err

    def evolve(self, kmodes, aexp_out, lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3, rtol=1e-4, atol=1e-4,
               pcoeff=0.25, icoeff=0.80, dcoeff=0.0, factormax=20.0, factormin=0.3, max_steps=2048,
               return_full=False, batch_size=0, nthreads=None):
        """evolve_perturbations (batch_size=0) / evolve_perturbations_batched on given k-modes.

        Returns (y[num_k, nout, 20 or nvar], tau_out, status, n_accepted, n_rejected)."""
        kmodes, kp = _d(kmodes)
        aexp_out = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
        tau_out, tp = _d(self.spline("tau_of_a", aexp_out))
        lmax = [lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax]
        lm, lp = _i(lmax)
        n = self._lib.orc_layout(*lmax)
        w = n if return_full else 20
        sv, sp = _d([rtol, atol, pcoeff, icoeff, dcoeff, factormax, factormin])
        nk = kmodes.size
        y = np.empty((nk, tau_out.size, w))
        st = np.zeros(nk, dtype=np.int32)
        na = np.zeros(nk, dtype=np.int32)
        nr = np.zeros(nk, dtype=np.int32)
        if nthreads is None:
            nthreads = os.cpu_count() or 1
        rc = self._lib.orc_evolve(self._h, kp, nk, tp, tau_out.size, lp, sp, int(max_steps), int(bool(return_full)),
                                  int(batch_size), int(nthreads), y.ctypes.data_as(_dp), st.ctypes.data_as(_ip),
                                  na.ctypes.data_as(_ip), nr.ctypes.data_as(_ip))
        if rc != 0:
            raise ValueError("num_k must be a multiple of batch_size")
        return y, tau_out, st, na, nr


def _replay(self, k, aexp_out, mesh, keep, lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3, rtol=1e-4, atol=1e-4,
            pcoeff=0.25, icoeff=0.80, dcoeff=0.0, factormax=20.0, factormin=0.3, return_full=False):
    """Take exactly the steps of `mesh` (nsteps x 2: t_prev, t_next; `keep` accepted flags) for one
    mode, from its initial conditions at mesh[0, 0]; returns (y[nout, 20 or nvar], own error norm
    per step)."""
    lib = _load()
    aexp_out = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    tau_out, tp = _d(self.spline("tau_of_a", aexp_out))
    lmax = [lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax]
    lm, lp = _i(lmax)
    n = lib.orc_layout(*lmax)
    w = n if return_full else 20
    mesh, mp = _d(mesh)
    keep, kp = _i(keep)
    sv, sp = _d([rtol, atol, pcoeff, icoeff, dcoeff, factormax, factormin])
    y = np.full((tau_out.size, w), np.nan)
    en = np.empty(keep.size)
    lib.orc_evolve_replay(self._h, float(k), tp, tau_out.size, lp, sp, int(bool(return_full)), mp, kp, keep.size,
                          y.ctypes.data_as(_dp), en.ctypes.data_as(_dp))
    return y, en


Background.replay = _replay


def _mesh(self, k, tau_out, lmax, rtol, atol, max_steps=2048, pcoeff=0.25, icoeff=0.80, dcoeff=0.0, factormax=20.0,
          factormin=0.3):
    """Adaptive solve of one mode at the given output times; returns (outputs (nout, 20), accepted
    step ends as fractions s = (t - t0) / (t1 - t0) of the mode's range, the last exactly 1)."""
    lib = _load()
    tau_out, tp = _d(tau_out)
    lm, lp = _i(lmax)
    sv, sp = _d([rtol, atol, pcoeff, icoeff, dcoeff, factormax, factormin])
    y = np.empty((tau_out.size, 20))
    acc = np.empty(max_steps)
    na = np.zeros(1, dtype=np.int32)
    st = lib.orc_evolve_mesh(self._h, float(k), tp, tau_out.size, lp, sp, int(max_steps), y.ctypes.data_as(_dp),
                             acc.ctypes.data_as(_dp), na.ctypes.data_as(_ip))
    if st != 0:
        raise RuntimeError(f"oracle solve failed with status {st}")
    t0 = lib.orc_start_tau(self._h, float(k), tp, tau_out.size)
    t1 = float(tau_out.max())
    s = (acc[:na[0]] - t0) / (t1 - t0)
    s[-1] = 1.0
    return y, s


def _replay_norm(self, kk, tau_out, mesh, mlen, lmax, nthreads=None):
    """Frozen-mesh replay of modes kk on normalized meshes (rows of mesh, lengths mlen): (nk, nout, 20)."""
    lib = _load()
    kk, kp = _d(np.atleast_1d(kk))
    tau_out, tp = _d(tau_out)
    mesh, mp = _d(np.atleast_2d(mesh))
    mlen, mlp = _i(np.atleast_1d(mlen))
    lm, lp = _i(lmax)
    y = np.full((kk.size, tau_out.size, 20), np.nan)
    bad = lib.orc_replay_norm(self._h, lp, kp, kk.size, tp, tau_out.size, mp, mlp, mesh.shape[1],
                              int(nthreads or os.cpu_count() or 1), y.ctypes.data_as(_dp))
    if bad:
        raise RuntimeError("replay did not reach every output time")
    return y


def _adjoint(self, shifted, dth, kk, tau_out, tau_out_sh, mesh, mlen, gbar, lmax, nthreads=None):
    """Discrete adjoint of the frozen-mesh replay (oracle/deb_oracle.cpp adjoint_replay).

    shifted: 12 Backgrounds (parameter j at +d_j: 2 j, at -d_j: 2 j + 1); gbar: (nk, nout, 20)
    output cotangents.  Returns (grad (nk, 6), parts (nk, 4, 6): direct ODE, outputs at fixed
    state, initial state, mesh and output times; lam0 (nk, n); (T0bar, T1bar) (nk, 2))."""
    lib = _load()
    kk, kp = _d(np.atleast_1d(kk))
    nk = kk.size
    tau_out, tp = _d(tau_out)
    tsh, tshp = _d(tau_out_sh)
    mesh, mp = _d(np.atleast_2d(mesh))
    mlen, mlp = _i(np.atleast_1d(mlen))
    lm, lp = _i(lmax)
    dth, dp = _d(dth)
    gbar, gp = _d(gbar)
    hs = (ctypes.c_void_p * 12)(*[b._h for b in shifted])
    n = lib.orc_layout(*[int(v) for v in lmax])
    grad = np.empty((nk, 6))
    parts = np.empty((nk, 4, 6))
    lam0 = np.empty((nk, n))
    tb = np.empty((nk, 2))
    bad = lib.orc_adjoint(self._h, hs, dp, lp, kp, nk, tp, tshp, tau_out.size, mp, mlp, mesh.shape[1], gp,
                          int(nthreads or os.cpu_count() or 1), grad.ctypes.data_as(_dp), parts.ctypes.data_as(_dp),
                          lam0.ctypes.data_as(_dp), tb.ctypes.data_as(_dp))
    if bad:
        raise RuntimeError("adjoint: an output time lies outside the mesh")
    return grad, parts, lam0, tb


Background.mesh = _mesh
Background.replay_norm = _replay_norm
Background.adjoint = _adjoint


def nvar(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax):
    return _load().orc_layout(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax)


def nu_bins(nq):
    q = np.empty(nq)
    w = np.empty(nq)
    _load().orc_nu_bins(nq, q.ctypes.data_as(_dp), w.ctypes.data_as(_dp))
    return q, w


def spline_eval(x, y, xq, kind=0, uniform=False):
    x, xp = _d(x)
    y, yp = _d(y)
    xq, xqp = _d(np.atleast_1d(xq))
    out = np.empty_like(xq)
    _load().orc_spline_eval(xp, yp, x.size, int(uniform), int(kind), xqp, xq.size, out.ctypes.data_as(_dp))
    return out


def romb_test(which, a, b):
    return _load().orc_romb_test(which, a, b)


def get_power(k, y, idx, param):
    """P(k) = 2 pi^2 A_s (k/k_p)^(n_s-1) k^-3 y[..., idx]^2 (perturbations.py:1446-1468)."""
    return 2 * np.pi ** 2 * param["A_s"] * (k / param.get("k_p", 0.05)) ** (param["n_s"] - 1) * k ** (-3.0) * y[..., idx] ** 2
