"""P(k) and its derivatives with respect to nine sampling parameters.

Parameters theta = (omega_c, omega_b, h, n_s, ln(1e10 A_s), m_nu, w0, wa, tau_reion), mapped
to the model's Omega_m = (omega_c + omega_b) / h^2, Omega_b = omega_b / h^2, H0 = 100 h,
A_s = exp(theta_4) / 1e10.  The reference differentiates its whole pipeline with jax.jacfwd
(notebooks; DirectAdjoint in perturbations.py:919,1079).  Here the derivative is that of the
frozen-mesh map: the adaptive base solve fixes each mode's accepted step ends (as fractions of
its integration range), and the six parameters that enter the background and perturbations
(omega_c, omega_b, h, m_nu, w0, wa) are differentiated by central differences of that map --
all twelve shifted cosmologies in one replay launch.  n_s and A_s enter P(k) only through the
primordial spectrum (analytic columns), and tau_reion does not enter the reference's model
(zero column).
"""
import math

import numpy as np
import torch

from . import _lib
from .background import background_tables
from .batch import power_units, tau_out_all, units_for
from .perturbations import _raise_on_status, evolve_units

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
    L = int(nacc.max())
    mesh = torch.ones((log.shape[0], L), dtype=torch.float64, device=log.device)
    rows = torch.arange(log.shape[0], device=log.device)[:, None].expand_as(ok)
    mesh[rows[ok], pos[ok]] = s[ok]
    mlen = nacc.to(torch.int32)
    mesh[torch.arange(log.shape[0], device=log.device), mlen.to(torch.int64) - 1] = 1.0
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


def pk_jacobian_many(thetas, base, kmodes, aexp_out=(1.0,), idx=4, rel_step=REL_STEP, num_thermo=256, schedule=None, lmaxg=12,
                     lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8, return_status=False, **solver):
    """P(k) at aexp_out and its Jacobian with respect to theta for C cosmologies at once.

    thetas: (C, 9); every cosmology uses the same k-modes.  Two perturbation launches serve all
    of them: the adaptive base solves of the C x n_k units (with their step logs), then the
    12 C x n_k replays of the shifted cosmologies, each on its base unit's accepted mesh.  The
    result of cosmology c is bitwise that of pk_jacobian(thetas[c], ...): units are independent.
    A cosmology whose w(a) crosses -1 does not stop the batch: its P(k) and Jacobian are NaN
    (status 4); a column whose shifted partner crosses -1 is NaN (status 6).
    Returns (pk (C, n_k, nout), J (C, n_k, nout, 9)) device tensors, and the per-cosmology
    status (C,) int32 numpy array when return_status is set."""
    thetas = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
    if thetas.ndim != 2 or thetas.shape[1] != 9:
        raise ValueError("thetas must be (C, 9)")
    C = thetas.shape[0]
    p0s = [theta_to_param(th, base) for th in thetas]
    dev = _lib.device()
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    knp = np.asarray(kmodes, dtype=np.float64)
    kk = torch.as_tensor(knp, device=dev)
    nk = kk.numel()
    aexp = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    # base solves with their step logs
    t0 = background_tables(p0s, num_thermo)
    tau0 = tau_out_all(t0, num_thermo, aexp)
    uc0, uk0, _ = units_for([knp] * C, dev)
    y, st, na, nr, log = evolve_units(t0, num_thermo, uc0, uk0, tau0, **lm, **solver, step_log=True, schedule=_base_schedule(C * nk))
    _raise_on_status(st, allow_phantom=True)
    bad = _cosmo_status(st, C, nk).cpu().numpy()
    pk = power_units(t0, num_thermo, uc0, uk0, y, idx).reshape(C, nk, -1)
    t1 = tau0.amax(dim=1)[uc0.to(torch.int64)]
    mesh, mlen = _accepted_mesh(log, na, t1)
    # phantom units never stepped: give them a one-step mesh (their replays are skipped as well)
    mlen = torch.where(mlen < 1, torch.ones_like(mlen), mlen)
    mesh[:, 0] = torch.where(na < 1, torch.ones_like(mesh[:, 0]), mesh[:, 0])
    del log
    # twelve shifted cosmologies per base cosmology, replayed on its mesh, all in one launch
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
    reps = 2 * len(ODE_PARAMS)
    tab = background_tables(params, num_thermo)
    tau = tau_out_all(tab, num_thermo, aexp)
    uc, uk, _ = units_for([knp] * len(params), dev)
    L = mesh.shape[1]
    rmesh = mesh.reshape(C, 1, nk, L).expand(C, reps, nk, L).reshape(C * reps * nk, L)
    rlen = mlen.reshape(C, 1, nk).expand(C, reps, nk).reshape(-1)
    yr, str_, _, _ = evolve_units(tab, num_thermo, uc, uk, tau, **lm, **solver, schedule=REPLAY_SCHEDULE if schedule is None else schedule,
                                  replay=(rmesh, rlen))
    _raise_on_status(str_, allow_phantom=True)
    colbad = _cosmo_status(str_, C * reps, nk).reshape(C, len(ODE_PARAMS), 2).any(dim=2).cpu().numpy()
    pkr = power_units(tab, num_thermo, uc, uk, yr, idx).reshape(C, reps, nk, -1)
    J = torch.zeros((C, nk, pk.shape[2], 9), dtype=torch.float64, device=dev)
    status = np.zeros(C, dtype=np.int32)
    nan = float("nan")
    for c in range(C):
        for i, j in enumerate(ODE_PARAMS):
            J[c, :, :, j] = (pkr[c, 2 * i] - pkr[c, 2 * i + 1]) / (2.0 * steps[c, i])
            if colbad[c, i]:
                J[c, :, :, j] = nan
                status[c] = GRAD_COLUMN_UNDEFINED
        J[c, :, :, 3] = pk[c] * torch.log(kk / p0s[c].get("k_p", 0.05))[:, None]
        J[c, :, :, 4] = pk[c]
        if bad[c]:
            pk[c] = nan
            J[c] = nan
            status[c] = GRAD_PHANTOM
    if return_status:
        return pk, J, status
    return pk, J


def pk_jacobian(theta, base, kmodes, aexp_out=(1.0,), idx=4, rel_step=REL_STEP, num_thermo=256, schedule=None, lmaxg=12, lmaxgp=12, lmaxr=12,
                lmaxnu=17, nqmax=8, **solver):
    """P(k) at aexp_out and its Jacobian with respect to theta (one cosmology).

    Returns (pk (n_k, nout), J (n_k, nout, 9)) device tensors."""
    pk, J, st = pk_jacobian_many(np.asarray(theta, dtype=np.float64)[None, :], base, kmodes, aexp_out=aexp_out, idx=idx,
                                 rel_step=rel_step, num_thermo=num_thermo, schedule=schedule, lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr,
                                 lmaxnu=lmaxnu, nqmax=nqmax, return_status=True, **solver)
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
    """P(k, a_out) as a differentiable function of theta (9,): backward returns J^T g."""

    @staticmethod
    def forward(ctx, theta, kmodes, base, opts):
        pk, J = pk_jacobian(theta.detach().cpu().numpy(), base, kmodes, **opts)
        ctx.save_for_backward(J)
        return pk.to(theta.device)

    @staticmethod
    def backward(ctx, g):
        (J,) = ctx.saved_tensors
        gt = torch.einsum("kon,ko->n", J, g.to(J.device))
        return gt.to(g.device), None, None, None


def pk_and_vjp(theta, base, kmodes, cotangent, **opts):
    """(P(k), d<cotangent, P>/d theta): P at aexp_out (default a = 1) and the 9-parameter VJP."""
    pk, J = pk_jacobian(theta, base, kmodes, **opts)
    c = torch.as_tensor(np.asarray(cotangent, dtype=np.float64), device=J.device).reshape(pk.shape)
    return pk, torch.einsum("kon,ko->n", J, c)


__all__ = ["PARAM9", "theta_to_param", "param_to_theta", "pk_jacobian", "pk_jacobian_many", "PkFunction", "pk_and_vjp"]
