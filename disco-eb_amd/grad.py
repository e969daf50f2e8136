"""P(k) and its derivatives with respect to nine sampling parameters.

Parameters theta = (omega_c, omega_b, h, n_s, ln(1e10 A_s), m_nu, w0, wa, tau_reion), mapped
to the model's Omega_m = (omega_c + omega_b) / h^2, Omega_b = omega_b / h^2, H0 = 100 h,
A_s = exp(theta_4) / 1e10.  The reference differentiates its whole pipeline with jax.jacfwd
(notebooks; DirectAdjoint in perturbations.py:919,1079).  Here the derivative is that of the
frozen-mesh map: the adaptive base solve fixes each mode's accepted step ends (as fractions of
its integration range).  The six parameters that enter the background and perturbations
(omega_c, omega_b, h, m_nu, w0, wa) are differentiated
  * method="adjoint" (default): by the discrete adjoint kernel (deb_evolve_adjoint): one reverse
    sweep per mode over checkpointed states gives all six derivatives; the background's
    parameter dependence enters through tables differenced over twelve shifted cosmologies
    (one background launch), as do the initial state and the integration range;
  * method="fd": by central differences of the frozen-mesh map, all twelve shifted cosmologies
    replayed in one launch.
n_s and A_s enter P(k) only through the primordial spectrum (analytic columns), and tau_reion
does not enter the reference's model (zero column).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .background import background_tables
from .batch import power_units, tau_out_all, units_for
from .perturbations import _check_mesh, _raise_on_status, evolve_units, nvar

PARAM9 = ("omega_c", "omega_b", "h", "n_s", "ln10As", "m_nu", "w0", "wa", "tau_reion")
ODE_PARAMS = (0, 1, 2, 5, 6, 7)
# Relative central-difference step.  The frozen-mesh map still holds discrete decisions (the
# approximation switches, the thermal-history solver's own adaptive steps); where one flips
# inside +-step the difference quotient carries the jump.  Steps of 1e-5 and 3e-5 agree to
# about 4e-5 relative; larger steps cross more such decisions.
REL_STEP = 3e-5


def theta_to_param(theta, base):
    """Model parameter dict for theta (9 values), other keys from base."""
    oc, ob, h, ns, lnA, mnu, w0, wa, _tr = [float(v) for v in theta]
    p = dict(base)
    p.update(Omegam=(oc + ob) / h ** 2, Omegab=ob / h ** 2, H0=100.0 * h, n_s=ns, A_s=math.exp(lnA) / 1e10, mnu=mnu, w_DE_0=w0,
             w_DE_a=wa)
    return p


def param_to_theta(param, tau_reion=0.054):
    h = param["H0"] / 100.0
    ob = param["Omegab"] * h ** 2
    oc = param["Omegam"] * h ** 2 - ob
    return np.array([oc, ob, h, param["n_s"], math.log(1e10 * param["A_s"]), param["mnu"], param["w_DE_0"], param["w_DE_a"],
                     tau_reion])


def _accepted_mesh(log, nacc, t1):
    """Accepted step ends of every unit as fractions s = (t - t0) / (t1 - t0), the last exactly 1.

    log: (n_units, max_steps, 3) step log (t_prev, t_next, error norm; NaN past the last
    attempt); nacc: accepted steps per unit; t1: (n_units,) end of each unit's range.  Returns
    (mesh (n_units, max nacc) with 1 past each unit's end, mesh_len) device tensors."""
    t1 = torch.as_tensor(t1, dtype=torch.float64, device=log.device)
    en = log[:, :, 2]
    ok = torch.isfinite(en) & (en < 1.0)
    t0 = log[:, 0, 0]
    s = (log[:, :, 1] - t0[:, None]) / (t1 - t0)[:, None]
    pos = torch.cumsum(ok.to(torch.int64), dim=1) - 1
    L = max(1, int(nacc.max()))  # at least one column: units that never stepped get a one-step mesh
    mesh = torch.ones((log.shape[0], L), dtype=torch.float64, device=log.device)
    rows = torch.arange(log.shape[0], device=log.device)[:, None].expand_as(ok)
    mesh[rows[ok], pos[ok]] = s[ok]
    mlen = nacc.to(torch.int32)
    mesh[torch.arange(log.shape[0], device=log.device), (mlen.to(torch.int64) - 1).clamp(min=0)] = 1.0
    return mesh, mlen


# Launch of the replayed shifted cosmologies (twelve times the k-modes): largest k first, two
# wavefronts per unit, so that two units share a compute unit (for 12 x 256 units on one
# MI355X: 274 ms per gradient at 128 threads, 415 ms at 256, 376 ms at 64).
REPLAY_SCHEDULE = {"order": "k_desc", "threads": 128}


def _base_schedule(n_units):
    """Launch of the adaptive base solves: largest k first; up to 512 units (one launch's worth of
    slots on one MI355X) at the library's default width, where the slowest unit's latency
    decides, and beyond that at 128 threads, where units in flight decide (two per compute unit;
    12.9k k-solves/s against 8.1k at 256 threads for 15,518 units).  Results do not depend on
    the width."""
    return {"order": "k_desc"} if n_units <= 512 else {"order": "k_desc", "threads": 128}


# Per-cosmology status of the gradient functions: 0 ok; 4 the cosmology's w(a) crosses -1 (no
# P(k), no derivatives: NaN); 6 a finite-difference partner of one column crosses -1 while the
# cosmology does not (that column is NaN).
GRAD_OK, GRAD_PHANTOM, GRAD_COLUMN_UNDEFINED = 0, 4, 6


def _is_lambda(theta):
    return theta[6] == -1.0 and theta[7] == 0.0


def _shift_pair(theta, j, d):
    """The two cosmologies whose central difference gives column j.  At an exact cosmological
    constant (w0 = -1, wa = 0) every wa shift leaves w(1) = -1 (the fluid equations' 1 / (1 + w)
    singularity at the last output), so the wa column is taken at w0 = -1 + 2 d, where both
    partners keep w(a) > -1 on (0, 1] (a joint shift, error O(d) in the mixed derivative)."""
    hi, lo = theta.copy(), theta.copy()
    hi[j] += d
    lo[j] -= d
    if j == 7 and _is_lambda(theta):
        hi[6] = lo[6] = -1.0 + 2.0 * d
    return hi, lo


def _cosmo_status(unit_status, C, nk):
    """Per-cosmology flag from the per-unit status of C x nk units (4: phantom divide)."""
    return (unit_status.reshape(C, nk) == 4).any(dim=1)


def evolve_units_adjoint(tables, num_thermo, shifted, dtheta, unit_cosmo, unit_k, tau_out, tau_out_shifted, mesh, mesh_len, gbar, *,
                         lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8, max_ckpt_bytes=8 << 30, return_parts=False):
    """Low-level launch of the discrete adjoint (deb_evolve_adjoint) over n_units units.

    tables (C, size) and shifted (C * 12, size) background tables, shifted block 2 p (+) and 2 p + 1
    (-) of cosmology c at theta_p +- dtheta[c, p]; tau_out (C, nout), tau_out_shifted (C * 12, nout);
    mesh / mesh_len as evolve_units(replay=...); gbar (n_units, nout, 20) output-field cotangents.
    Units are launched largest k first, in chunks whose checkpoints fit max_ckpt_bytes.
    Returns (grad (n_units, 6), status (n_units,)), plus (lam0 (n_units, nvar), tbar (n_units, 2))
    when return_parts is set."""
    L = _lib.lib()
    dev = tables.device
    C = tables.shape[0]
    U = int(unit_k.numel())
    nout = int(tau_out.shape[-1])
    n = nvar(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax)
    prob = _lib.DebProblem(int(lmaxg), int(lmaxgp), int(lmaxr), int(lmaxnu), int(nqmax), 0)
    mesh = mesh.to(device=dev, dtype=torch.float64).contiguous()
    mlen = mesh_len.to(device=dev, dtype=torch.int32).contiguous()
    _check_mesh(mesh, mlen)
    Lm = int(mlen.max()) if U else 1
    gbar = gbar.to(device=dev, dtype=torch.float64).contiguous()
    if gbar.shape != (U, nout, 20):
        raise ValueError("gbar must be (n_units, nout, 20)")
    if shifted.shape[0] != 12 * C or tau_out_shifted.shape != (12 * C, nout) or tuple(dtheta.shape) != (C, 6):
        raise ValueError("twelve shifted cosmologies per cosmology are required")
    uc = unit_cosmo.to(device=dev, dtype=torch.int32).contiguous()
    uk = unit_k.to(device=dev, dtype=torch.float64).contiguous()
    dth = torch.as_tensor(dtheta, dtype=torch.float64, device=dev).contiguous()
    grad = torch.empty((U, 6), dtype=torch.float64, device=dev)
    status = torch.empty(U, dtype=torch.int32, device=dev)
    lam0 = torch.empty((U, n), dtype=torch.float64, device=dev) if return_parts else None
    tbar = torch.empty((U, 2), dtype=torch.float64, device=dev) if return_parts else None
    chunk = max(1, min(U, int(max_ckpt_bytes // ((Lm + 1) * n * 8))))
    to = tau_out.to(device=dev, dtype=torch.float64).contiguous()
    tsh = tau_out_shifted.to(device=dev, dtype=torch.float64).contiguous()
    for s0 in range(0, U, chunk):
        s1 = min(U, s0 + chunk)
        cnt = s1 - s0
        work = torch.empty(L.deb_adjoint_workspace_doubles(cnt, C, int(num_thermo), Lm, ctypes.byref(prob)), dtype=torch.float64, device=dev)
        order = torch.argsort(uk[s0:s1], descending=True, stable=True).to(torch.int32)
        sched = _lib.DebSchedule(order.data_ptr(), 256)
        _lib.check(L.deb_evolve_adjoint(_lib.ptr(tables), int(num_thermo), C, _lib.ptr(shifted), _lib.ptr(dth), _lib.ptr(uc[s0:s1]),
                                        _lib.ptr(uk[s0:s1]), cnt, _lib.ptr(to), _lib.ptr(tsh), nout, ctypes.byref(prob),
                                        _lib.ptr(mesh[s0:s1]), mesh.shape[1], Lm, _lib.ptr(mlen[s0:s1]), _lib.ptr(gbar[s0:s1]),
                                        _lib.ptr(grad[s0:s1]), _lib.ptr(status[s0:s1]), _lib.ptr(work),
                                        _lib.ptr(lam0[s0:s1]) if lam0 is not None else None,
                                        _lib.ptr(tbar[s0:s1]) if tbar is not None else None, ctypes.byref(sched), _lib.stream()))
        del order, work  # stream-ordered reuse only after the launch has run
    if return_parts:
        return grad, status, lam0, tbar
    return grad, status


def _setup(thetas, base, kmodes, aexp, lm, solver, num_thermo, rel_step):
    """Base solves (with step logs) of C cosmologies on one k-grid, their accepted meshes, and the
    twelve shifted cosmologies per cosmology (tables and output times)."""
    C = thetas.shape[0]
    p0s = [theta_to_param(th, base) for th in thetas]
    dev = _lib.device()
    knp = np.asarray(kmodes, dtype=np.float64)
    nk = knp.size
    t0 = background_tables(p0s, num_thermo)
    tau0 = tau_out_all(t0, num_thermo, aexp)
    uc0, uk0, _ = units_for([knp] * C, dev)
    y, st, na, nr, log = evolve_units(t0, num_thermo, uc0, uk0, tau0, **lm, **solver, step_log=True, schedule=_base_schedule(C * nk))
    _raise_on_status(st, allow_phantom=True)
    bad = _cosmo_status(st, C, nk).cpu().numpy()
    t1 = tau0.amax(dim=1)[uc0.to(torch.int64)]
    mesh, mlen = _accepted_mesh(log, na, t1)
    # phantom units never stepped: give them a one-step mesh (nothing is replayed for them)
    mlen = torch.where(mlen < 1, torch.ones_like(mlen), mlen)
    mesh[:, 0] = torch.where(na < 1, torch.ones_like(mesh[:, 0]), mesh[:, 0])
    del log
    steps = np.zeros((C, len(ODE_PARAMS)))
    params = []
    for c in range(C):
        for i, j in enumerate(ODE_PARAMS):
            d = rel_step * max(abs(thetas[c, j]), 1e-2)
            steps[c, i] = d
            # a cosmology on the phantom divide replays nothing: its partners are itself (skipped too)
            pair = (thetas[c], thetas[c]) if bad[c] else _shift_pair(thetas[c], j, d)
            for th in pair:
                params.append(theta_to_param(th, base))
    tab = background_tables(params, num_thermo)
    tau = tau_out_all(tab, num_thermo, aexp)
    return dict(p0s=p0s, t0=t0, tau0=tau0, uc0=uc0, uk0=uk0, y=y, bad=bad, mesh=mesh, mlen=mlen, steps=steps, tab=tab, tau=tau, nk=nk,
                kk=torch.as_tensor(knp, device=dev))


def _power_gbar(y, pk, idx, cot):
    """Output-field cotangents of sum cot * P (P = prefactor y[idx]^2): 2 cot P / y[idx] at field idx."""
    g = torch.zeros((y.shape[0], y.shape[1], 20), dtype=torch.float64, device=y.device)
    g[:, :, idx] = 2.0 * cot * pk / y[:, :, idx]
    return g


def pk_jacobian_many(thetas, base, kmodes, aexp_out=(1.0,), idx=4, rel_step=REL_STEP, num_thermo=256, schedule=None, lmaxg=12,
                     lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8, return_status=False, method="adjoint", **solver):
    """P(k) at aexp_out and its Jacobian with respect to theta for C cosmologies at once.

    thetas: (C, 9); every cosmology uses the same k-modes.  The adaptive base solves of all
    C x n_k units run in one launch (with their step logs).  method="adjoint": one adjoint launch
    per output time gives every unit's six ODE-parameter derivatives (the rows of the Jacobian);
    method="fd": one replay launch of the 12 C x n_k shifted units.  The result of cosmology c
    does not depend on the others (units are independent).  A cosmology whose w(a) crosses -1
    does not stop the batch: its P(k) and Jacobian are NaN (status 4); with method="fd" a column
    whose shifted partner crosses -1 is NaN (status 6).  At an exact cosmological constant the w0
    and wa columns are taken by differences (method "fd"): the fluid's perturbations vanish there,
    while their response to w is first order.
    Returns (pk (C, n_k, nout), J (C, n_k, nout, 9)) device tensors, and the per-cosmology
    status (C,) int32 numpy array when return_status is set."""
    if method not in ("adjoint", "fd"):
        raise ValueError("method must be 'adjoint' or 'fd'")
    thetas = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
    if thetas.ndim != 2 or thetas.shape[1] != 9:
        raise ValueError("thetas must be (C, 9)")
    C = thetas.shape[0]
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    aexp = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    S = _setup(thetas, base, kmodes, aexp, lm, solver, num_thermo, rel_step)
    nk, kk, bad = S["nk"], S["kk"], S["bad"]
    dev = kk.device
    pku = power_units(S["t0"], num_thermo, S["uc0"], S["uk0"], S["y"], idx)
    pk = pku.reshape(C, nk, -1)
    nout = pk.shape[2]
    J = torch.zeros((C, nk, nout, 9), dtype=torch.float64, device=dev)
    status = np.zeros(C, dtype=np.int32)
    nan = float("nan")
    reps = 2 * len(ODE_PARAMS)
    if method == "fd":
        uc, uk, _ = units_for([kk.cpu().numpy()] * (C * reps), dev)
        Lm = S["mesh"].shape[1]
        rmesh = S["mesh"].reshape(C, 1, nk, Lm).expand(C, reps, nk, Lm).reshape(C * reps * nk, Lm)
        rlen = S["mlen"].reshape(C, 1, nk).expand(C, reps, nk).reshape(-1)
        yr, str_, _, _ = evolve_units(S["tab"], num_thermo, uc, uk, S["tau"], **lm, **solver,
                                      schedule=REPLAY_SCHEDULE if schedule is None else schedule, replay=(rmesh, rlen))
        _raise_on_status(str_, allow_phantom=True)
        colbad = _cosmo_status(str_, C * reps, nk).reshape(C, len(ODE_PARAMS), 2).any(dim=2).cpu().numpy()
        pkr = power_units(S["tab"], num_thermo, uc, uk, yr, idx).reshape(C, reps, nk, -1)
        for c in range(C):
            for i, j in enumerate(ODE_PARAMS):
                J[c, :, :, j] = (pkr[c, 2 * i] - pkr[c, 2 * i + 1]) / (2.0 * S["steps"][c, i])
                if colbad[c, i]:
                    J[c, :, :, j] = nan
                    status[c] = GRAD_COLUMN_UNDEFINED
    else:
        dth = torch.as_tensor(S["steps"], dtype=torch.float64, device=dev)
        for jo in range(nout):
            cot = torch.zeros((C * nk, nout), dtype=torch.float64, device=dev)
            cot[:, jo] = 1.0
            gbar = _power_gbar(S["y"], pku, idx, cot)
            g, st = evolve_units_adjoint(S["t0"], num_thermo, S["tab"], dth, S["uc0"], S["uk0"], S["tau0"], S["tau"], S["mesh"], S["mlen"],
                                         gbar, **lm)
            J[:, :, jo, list(ODE_PARAMS)] = g.reshape(C, nk, len(ODE_PARAMS))
        lam = [c for c in range(C) if _is_lambda(thetas[c]) and not bad[c]]
        if lam:  # w0, wa at a cosmological constant: differences of the frozen-mesh map
            _, Jf, stf = pk_jacobian_many(thetas[lam], base, kmodes, aexp_out=aexp_out, idx=idx, rel_step=rel_step, num_thermo=num_thermo,
                                          schedule=schedule, return_status=True, method="fd", **lm, **solver)
            for i, c in enumerate(lam):
                J[c, :, :, 6:8] = Jf[i, :, :, 6:8]
                status[c] = max(status[c], int(stf[i]))
    for c in range(C):
        J[c, :, :, 3] = pk[c] * torch.log(kk / S["p0s"][c].get("k_p", 0.05))[:, None]
        J[c, :, :, 4] = pk[c]
        if bad[c]:
            pk[c] = nan
            J[c] = nan
            status[c] = GRAD_PHANTOM
    if return_status:
        return pk, J, status
    return pk, J


def pk_and_vjp_many(thetas, base, kmodes, cotangents, aexp_out=(1.0,), idx=4, rel_step=REL_STEP, num_thermo=256, lmaxg=12, lmaxgp=12,
                    lmaxr=12, lmaxnu=17, nqmax=8, **solver):
    """P(k) and the 9-parameter VJP of sum_k,z cotangent * P for C cosmologies: one base launch and
    one adjoint launch for all of them.  cotangents: (C, n_k, nout).  Returns (pk (C, n_k, nout),
    vjp (C, 9), status (C,)).  At an exact cosmological constant the w0 and wa entries come from
    the difference Jacobian (see pk_jacobian_many)."""
    thetas = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
    C = thetas.shape[0]
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    aexp = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    S = _setup(thetas, base, kmodes, aexp, lm, solver, num_thermo, rel_step)
    nk, kk, bad = S["nk"], S["kk"], S["bad"]
    dev = kk.device
    pku = power_units(S["t0"], num_thermo, S["uc0"], S["uk0"], S["y"], idx)
    pk = pku.reshape(C, nk, -1)
    cot = torch.as_tensor(np.asarray(cotangents, dtype=np.float64), device=dev).reshape(C * nk, -1)
    gbar = _power_gbar(S["y"], pku, idx, cot)
    dth = torch.as_tensor(S["steps"], dtype=torch.float64, device=dev)
    g, st = evolve_units_adjoint(S["t0"], num_thermo, S["tab"], dth, S["uc0"], S["uk0"], S["tau0"], S["tau"], S["mesh"], S["mlen"], gbar,
                                 **lm)
    v = torch.zeros((C, 9), dtype=torch.float64, device=dev)
    v[:, list(ODE_PARAMS)] = g.reshape(C, nk, len(ODE_PARAMS)).sum(dim=1)
    cr = cot.reshape(C, nk, -1)
    logk = torch.log(kk)[None, :, None] - torch.log(torch.tensor([p.get("k_p", 0.05) for p in S["p0s"]], dtype=torch.float64,
                                                                  device=dev))[:, None, None]
    v[:, 3] = (cr * pk * logk).sum(dim=(1, 2))
    v[:, 4] = (cr * pk).sum(dim=(1, 2))
    status = np.where(bad, GRAD_PHANTOM, GRAD_OK).astype(np.int32)
    lam = [c for c in range(C) if _is_lambda(thetas[c]) and not bad[c]]
    if lam:
        _, Jf, stf = pk_jacobian_many(thetas[lam], base, kmodes, aexp_out=aexp_out, idx=idx, rel_step=rel_step, num_thermo=num_thermo,
                                      return_status=True, method="fd", **lm, **solver)
        for i, c in enumerate(lam):
            v[c, 6:8] = torch.einsum("kon,ko->n", Jf[i, :, :, 6:8], cr[c])
            status[c] = max(status[c], int(stf[i]))
    v[torch.as_tensor(bad, device=dev)] = float("nan")
    return pk, v, status


def pk_jacobian(theta, base, kmodes, aexp_out=(1.0,), idx=4, rel_step=REL_STEP, num_thermo=256, schedule=None, lmaxg=12, lmaxgp=12, lmaxr=12,
                lmaxnu=17, nqmax=8, method="adjoint", **solver):
    """P(k) at aexp_out and its Jacobian with respect to theta (one cosmology).

    Returns (pk (n_k, nout), J (n_k, nout, 9)) device tensors."""
    pk, J, st = pk_jacobian_many(np.asarray(theta, dtype=np.float64)[None, :], base, kmodes, aexp_out=aexp_out, idx=idx,
                                 rel_step=rel_step, num_thermo=num_thermo, schedule=schedule, lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr,
                                 lmaxnu=lmaxnu, nqmax=nqmax, return_status=True, method=method, **solver)
    _raise_on_grad_status(st[0])
    return pk[0], J[0]


def _raise_on_grad_status(st):
    if st == GRAD_PHANTOM:
        raise ValueError("w(a) = w_DE_0 + w_DE_a (1 - a) crosses -1 in (0, 1]: the dark-energy fluid equations are singular "
                         "at the phantom divide, so P(k) and its gradient are undefined")
    if st == GRAD_COLUMN_UNDEFINED:
        raise ValueError("a finite-difference partner of a dark-energy column crosses w = -1 while the cosmology does not: "
                         "that derivative is undefined at this step size (move the cosmology off the divide)")


class PkFunction(torch.autograd.Function):
    """P(k, a_out) as a differentiable function of theta (9,).

    forward runs the adaptive base solve and keeps its accepted step mesh, the background tables
    and the twelve shifted cosmologies' tables; backward launches the discrete adjoint kernel with
    the incoming cotangent (one reverse sweep per mode) and adds the closed-form n_s and A_s
    entries.  opts: keyword arguments of pk_jacobian (aexp_out, idx, truncations, solver)."""

    @staticmethod
    def forward(ctx, theta, kmodes, base, opts):
        th = np.asarray(theta.detach().cpu().numpy(), dtype=np.float64)[None, :]
        o = dict(opts)
        idx = o.pop("idx", 4)
        rel_step = o.pop("rel_step", REL_STEP)
        num_thermo = o.pop("num_thermo", 256)
        aexp = np.atleast_1d(np.asarray(o.pop("aexp_out", (1.0,)), dtype=np.float64))
        lm = {k: o.pop(k) for k in ("lmaxg", "lmaxgp", "lmaxr", "lmaxnu", "nqmax") if k in o}
        lm = dict(dict(lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8), **lm)
        o.pop("method", None)
        o.pop("schedule", None)
        S = _setup(th, base, kmodes, aexp, lm, o, num_thermo, rel_step)
        if S["bad"][0]:
            _raise_on_grad_status(GRAD_PHANTOM)
        pku = power_units(S["t0"], num_thermo, S["uc0"], S["uk0"], S["y"], idx)
        ctx.S, ctx.pku, ctx.idx, ctx.lm, ctx.num_thermo, ctx.th = S, pku, idx, lm, num_thermo, th
        ctx.fd = dict(kmodes=kmodes, base=base, aexp_out=aexp, idx=idx, rel_step=rel_step, num_thermo=num_thermo, solver=o)
        return pku.to(theta.device)

    @staticmethod
    def backward(ctx, g):
        S, pku = ctx.S, ctx.pku
        dev = pku.device
        cot = g.to(device=dev, dtype=torch.float64).reshape(pku.shape)
        gbar = _power_gbar(S["y"], pku, ctx.idx, cot)
        dth = torch.as_tensor(S["steps"], dtype=torch.float64, device=dev)
        gr, _ = evolve_units_adjoint(S["t0"], ctx.num_thermo, S["tab"], dth, S["uc0"], S["uk0"], S["tau0"], S["tau"], S["mesh"], S["mlen"],
                                     gbar, **ctx.lm)
        v = torch.zeros(9, dtype=torch.float64, device=dev)
        v[list(ODE_PARAMS)] = gr.sum(dim=0)
        logk = torch.log(S["kk"] / S["p0s"][0].get("k_p", 0.05))[:, None]
        v[3] = (cot * pku * logk).sum()
        v[4] = (cot * pku).sum()
        if _is_lambda(ctx.th[0]):
            f = ctx.fd
            _, Jf = pk_jacobian(ctx.th[0], f["base"], f["kmodes"], aexp_out=f["aexp_out"], idx=f["idx"], rel_step=f["rel_step"],
                                num_thermo=f["num_thermo"], method="fd", **ctx.lm, **f["solver"])
            v[6:8] = torch.einsum("kon,ko->n", Jf[:, :, 6:8], cot)
        return v.to(g.device), None, None, None


def pk_and_vjp(theta, base, kmodes, cotangent, method="adjoint", **opts):
    """(P(k), d<cotangent, P>/d theta): P at aexp_out (default a = 1) and the 9-parameter VJP.

    method="adjoint": one base launch and one adjoint launch; method="fd": contraction of the
    difference Jacobian."""
    if method == "fd":
        pk, J = pk_jacobian(theta, base, kmodes, method="fd", **opts)
        c = torch.as_tensor(np.asarray(cotangent, dtype=np.float64), device=J.device).reshape(pk.shape)
        return pk, torch.einsum("kon,ko->n", J, c)
    opts.pop("schedule", None)
    pk, v, st = pk_and_vjp_many(np.asarray(theta, dtype=np.float64)[None, :], base, kmodes, np.asarray(cotangent)[None], **opts)
    _raise_on_grad_status(int(st[0]))
    return pk[0], v[0]


__all__ = ["PARAM9", "theta_to_param", "param_to_theta", "pk_jacobian", "pk_jacobian_many", "pk_and_vjp_many", "PkFunction", "pk_and_vjp",
           "evolve_units_adjoint"]
