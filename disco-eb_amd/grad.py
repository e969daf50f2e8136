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
n_s and A_s enter P(k) only through the primordial spectrum (analytic columns).  tau_reion does
not enter the reference's model: its column is zero unless the base parameters switch on
reionization (an extension: base["tau_reion"] > 0, see evolve_background), in which case theta's
tau_reion sets the reionization optical depth and its column is a frozen-mesh central difference.
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


def _reionized(base):
    return float(base.get("tau_reion", 0.0)) > 0.0


def theta_to_param(theta, base):
    """Model parameter dict for theta (9 values), other keys from base.  theta's tau_reion becomes
    the reionization optical depth only when base has reionization switched on (tau_reion > 0)."""
    oc, ob, h, ns, lnA, mnu, w0, wa, tr = [float(v) for v in theta]
    p = dict(base)
    p.update(Omegam=(oc + ob) / h ** 2, Omegab=ob / h ** 2, H0=100.0 * h, n_s=ns, A_s=math.exp(lnA) / 1e10, mnu=mnu, w_DE_0=w0,
             w_DE_a=wa)
    if _reionized(base):
        p["tau_reion"] = tr
    return p


def _reion_column(S, thetas, base, aexp, lm, solver, num_thermo, rel_step, idx):
    """dP/d tau_reion of every unit by a central difference of the frozen-mesh map: each cosmology's
    two tau_reion-shifted copies replayed on its base mesh.  (U, nout) tensor."""
    C = thetas.shape[0]
    offs, uc0, uk0 = S["offs"], S["uc0"], S["uk0"]
    dev = uk0.device
    nks = np.diff(offs)
    steps = np.array([rel_step * max(abs(thetas[c, 8]), 1e-2) for c in range(C)])
    params = []
    for c in range(C):
        for sgn in (1.0, -1.0):
            th = thetas[c].copy()
            th[8] += sgn * steps[c]
            params.append(theta_to_param(th, base))
    tab = background_tables(params, num_thermo)
    tau = tau_out_all(tab, num_thermo, aexp)
    U = int(uk0.numel())
    bi = torch.as_tensor(np.concatenate([np.tile(np.arange(offs[c], offs[c + 1]), 2) for c in range(C)]).astype(np.int64), device=dev)
    uc = torch.as_tensor(np.concatenate([np.repeat(2 * c + np.arange(2), nks[c]) for c in range(C)]).astype(np.int32), device=dev)
    yr, st, _, _ = evolve_units(tab, num_thermo, uc, uk0[bi].contiguous(), tau, **lm, **solver, schedule=REPLAY_SCHEDULE,
                                replay=(S["mesh"][bi].contiguous(), S["mlen"][bi].contiguous()))
    _raise_on_status(st, allow_phantom=True)
    pkr = power_units(tab, num_thermo, uc, uk0[bi].contiguous(), yr, idx)
    start = np.concatenate([[0], np.cumsum(2 * nks)])[:-1]
    ucn = np.repeat(np.arange(C), nks)
    pos = np.arange(U) - np.repeat(offs[:-1], nks)
    plus = torch.as_tensor(start[ucn] + pos, device=dev)
    minus = torch.as_tensor(start[ucn] + nks[ucn] + pos, device=dev)
    d = torch.as_tensor(steps[ucn], dtype=torch.float64, device=dev)
    return (pkr[plus] - pkr[minus]) / (2.0 * d[:, None])


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


def evolve_units_adjoint(tables, num_thermo, shifted, dtheta, unit_cosmo, unit_k, tau_out, tau_out_shifted, mesh, mesh_len, gbar, *,
                         lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8, max_ckpt_bytes=32 << 30, return_parts=False):
    """Low-level launch of the discrete adjoint (deb_evolve_adjoint) over n_units units.

    tables (C, size) and shifted (C * 12, size) background tables, shifted block 2 p (+) and 2 p + 1
    (-) of cosmology c at theta_p +- dtheta[c, p]; tau_out (C, nout), tau_out_shifted (C * 12, nout);
    mesh / mesh_len as evolve_units(replay=...); gbar (n_units, nout, 20) output-field cotangents.
    Units are launched largest k first, in consecutive chunks of that order whose checkpoints (per
    unit the states at the chunk's longest mesh's step starts and one record of stages and
    factorisation per step, deb_adjoint_workspace_doubles) fit max_ckpt_bytes.
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
    to = tau_out.to(device=dev, dtype=torch.float64).contiguous()
    tsh = tau_out_shifted.to(device=dev, dtype=torch.float64).contiguous()
    order = torch.argsort(uk, descending=True, stable=True)
    ml = mlen[order].cpu().numpy().astype(np.int64) if U else np.zeros(0, dtype=np.int64)
    base_ws = L.deb_adjoint_workspace_doubles(0, C, int(num_thermo), 1, ctypes.byref(prob))
    per_unit = lambda m: (L.deb_adjoint_workspace_doubles(1, C, int(num_thermo), m, ctypes.byref(prob)) - base_ws) * 8
    s0 = 0
    while s0 < U:  # greedy chunks over the k-descending order
        m, s1 = int(ml[s0]), s0 + 1
        while s1 < U and (s1 - s0 + 1) * per_unit(max(m, int(ml[s1]))) <= max_ckpt_bytes:
            m = max(m, int(ml[s1]))
            s1 += 1
        idx = order[s0:s1]
        cnt = s1 - s0
        cuc, cuk = uc[idx].contiguous(), uk[idx].contiguous()
        cmesh, cmlen, cgbar = mesh[idx, :m].contiguous(), mlen[idx].contiguous(), gbar[idx].contiguous()
        cgrad = torch.empty((cnt, 6), dtype=torch.float64, device=dev)
        cst = torch.empty(cnt, dtype=torch.int32, device=dev)
        clam = torch.empty((cnt, n), dtype=torch.float64, device=dev) if return_parts else None
        ctb = torch.empty((cnt, 2), dtype=torch.float64, device=dev) if return_parts else None
        work = torch.empty(L.deb_adjoint_workspace_doubles(cnt, C, int(num_thermo), m, ctypes.byref(prob)), dtype=torch.float64, device=dev)
        sched = _lib.DebSchedule(None, 256)
        _lib.check(L.deb_evolve_adjoint(_lib.ptr(tables), int(num_thermo), C, _lib.ptr(shifted), _lib.ptr(dth), _lib.ptr(cuc), _lib.ptr(cuk),
                                        cnt, _lib.ptr(to), _lib.ptr(tsh), nout, ctypes.byref(prob), _lib.ptr(cmesh), m, m,
                                        _lib.ptr(cmlen), _lib.ptr(cgbar), _lib.ptr(cgrad), _lib.ptr(cst), _lib.ptr(work),
                                        _lib.ptr(clam) if clam is not None else None, _lib.ptr(ctb) if ctb is not None else None,
                                        ctypes.byref(sched), _lib.stream()))
        grad[idx], status[idx] = cgrad, cst
        if return_parts:
            lam0[idx], tbar[idx] = clam, ctb
        del work, cmesh  # stream-ordered reuse only after the launch has run
        s0 = s1
    if return_parts:
        return grad, status, lam0, tbar
    return grad, status


def _kgrids(kmodes, C):
    """Per-cosmology k-grids: a list of C one-dimensional grids (ragged), or one grid shared by all.
    Returns (list of C float64 arrays, ragged flag)."""
    if isinstance(kmodes, (list, tuple)) and len(kmodes) == C and C > 0 and np.ndim(kmodes[0]) == 1:
        return [np.asarray(k, dtype=np.float64).reshape(-1) for k in kmodes], True
    knp = np.asarray(kmodes, dtype=np.float64).reshape(-1)
    return [knp] * C, False


def _setup(thetas, base, kmodes, aexp, lm, solver, num_thermo, rel_step):
    """Base solves (with step logs) of C cosmologies, each on its own k-grid or on one shared grid,
    their accepted meshes, and the twelve shifted cosmologies per cosmology (tables and output
    times).  Units are cosmology-major: cosmology c owns units offs[c]..offs[c + 1] - 1."""
    C = thetas.shape[0]
    p0s = [theta_to_param(th, base) for th in thetas]
    dev = _lib.device()
    kgrids, ragged = _kgrids(kmodes, C)

    def shifted_params(bad):
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
        return steps, params

    # the base and the shifted cosmologies' tables in one launch (one workgroup each): the
    # phantom-divide test the kernels make, (1 + w0)(1 + w0 + wa) <= 0 unless w0 = -1 and wa = 0,
    # decides the shifts before the base solve reports it
    bad_pre = np.array([(not _is_lambda(th)) and (1.0 + th[6]) * (1.0 + th[6] + th[7]) <= 0.0 for th in thetas])
    steps, params = shifted_params(bad_pre)
    t_all = background_tables(p0s + params, num_thermo)
    t0, tab = t_all[:C], t_all[C:]
    tau0 = tau_out_all(t0, num_thermo, aexp)
    uc0, uk0, offs = units_for(kgrids, dev)
    U = int(uk0.numel())
    y, st, na, nr, log = evolve_units(t0, num_thermo, uc0, uk0, tau0, **lm, **solver, step_log=True, schedule=_base_schedule(U))
    _raise_on_status(st, allow_phantom=True)
    bad = (torch.zeros(C, dtype=torch.float64, device=dev).index_add_(0, uc0.to(torch.int64), (st == 4).to(torch.float64)) > 0).cpu().numpy()
    t1 = tau0.amax(dim=1)[uc0.to(torch.int64)]
    mesh, mlen = _accepted_mesh(log, na, t1)
    # phantom units never stepped: give them a one-step mesh (nothing is replayed for them)
    mlen = torch.where(mlen < 1, torch.ones_like(mlen), mlen)
    mesh[:, 0] = torch.where(na < 1, torch.ones_like(mesh[:, 0]), mesh[:, 0])
    del log
    if not np.array_equal(bad, bad_pre):  # the solve disagrees with the parameters' test: take its word
        steps, params = shifted_params(bad)
        tab = background_tables(params, num_thermo)
    tau = tau_out_all(tab, num_thermo, aexp)
    return dict(p0s=p0s, t0=t0, tau0=tau0, uc0=uc0, uk0=uk0, y=y, bad=bad, mesh=mesh, mlen=mlen, steps=steps, tab=tab, tau=tau,
                offs=offs, kgrids=kgrids, ragged=ragged)


def _power_gbar(y, pk, idx, cot):
    """Output-field cotangents of sum cot * P (P = prefactor y[idx]^2): 2 cot P / y[idx] at field idx."""
    g = torch.zeros((y.shape[0], y.shape[1], 20), dtype=torch.float64, device=y.device)
    g[:, :, idx] = 2.0 * cot * pk / y[:, :, idx]
    return g


def _per_cosmology(S, x):
    """Unit-major tensor x (n_units, ...) as (C, n_k, ...) on a shared grid, else a list of C tensors."""
    C = len(S["p0s"])
    if not S["ragged"]:
        return x.reshape(C, -1, *x.shape[1:])
    o = S["offs"]
    return [x[o[c]:o[c + 1]] for c in range(C)]


def pk_jacobian_many(thetas, base, kmodes, aexp_out=(1.0,), idx=4, rel_step=REL_STEP, num_thermo=256, schedule=None, lmaxg=12,
                     lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8, return_status=False, method="adjoint", **solver):
    """P(k) at aexp_out and its Jacobian with respect to theta for C cosmologies at once.

    thetas: (C, 9).  kmodes: one k-grid shared by every cosmology, or a list of C grids (each
    cosmology its own, of any length).  The adaptive base solves of all units run in one launch
    (with their step logs).  method="adjoint": one adjoint launch per output time gives every
    unit's six ODE-parameter derivatives (the rows of the Jacobian); method="fd": one replay
    launch of the twelve shifted copies of every unit.  The result of cosmology c does not depend
    on the others (units are independent).  A cosmology whose w(a) crosses -1 does not stop the
    batch: its P(k) and Jacobian are NaN (status 4); with method="fd" a column whose shifted
    partner crosses -1 is NaN (status 6).  At an exact cosmological constant the w0 and wa
    columns are taken by differences (method "fd"): the fluid's perturbations vanish there,
    while their response to w is first order.
    Returns (pk (C, n_k, nout), J (C, n_k, nout, 9)) device tensors on a shared grid, lists of C
    tensors (n_k_c, nout) and (n_k_c, nout, 9) on ragged grids, and the per-cosmology status (C,)
    int32 numpy array when return_status is set."""
    if method not in ("adjoint", "fd"):
        raise ValueError("method must be 'adjoint' or 'fd'")
    thetas = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
    if thetas.ndim != 2 or thetas.shape[1] != 9:
        raise ValueError("thetas must be (C, 9)")
    C = thetas.shape[0]
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    aexp = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    S = _setup(thetas, base, kmodes, aexp, lm, solver, num_thermo, rel_step)
    bad, offs, uc0, uk0 = S["bad"], S["offs"], S["uc0"], S["uk0"]
    dev = uk0.device
    U = int(uk0.numel())
    pku = power_units(S["t0"], num_thermo, uc0, uk0, S["y"], idx)
    nout = pku.shape[1]
    J = torch.zeros((U, nout, 9), dtype=torch.float64, device=dev)
    status = np.zeros(C, dtype=np.int32)
    nan = float("nan")
    reps = 2 * len(ODE_PARAMS)
    nks = np.diff(offs)
    if method == "fd":
        # replay units: cosmology-major, then shifted copy r, then the cosmology's k-modes
        base_idx = np.concatenate([np.tile(np.arange(offs[c], offs[c + 1]), reps) for c in range(C)]).astype(np.int64)
        rep_uc = np.concatenate([np.repeat(c * reps + np.arange(reps), nks[c]) for c in range(C)]).astype(np.int32)
        start = np.concatenate([[0], np.cumsum(reps * nks)])[:-1]
        pos = np.arange(U) - np.repeat(offs[:-1], nks)
        ucn = np.repeat(np.arange(C), nks)
        bi = torch.as_tensor(base_idx, device=dev)
        uc = torch.as_tensor(rep_uc, device=dev)
        yr, str_, _, _ = evolve_units(S["tab"], num_thermo, uc, uk0[bi].contiguous(), S["tau"], **lm, **solver,
                                      schedule=REPLAY_SCHEDULE if schedule is None else schedule,
                                      replay=(S["mesh"][bi].contiguous(), S["mlen"][bi].contiguous()))
        _raise_on_status(str_, allow_phantom=True)
        rbad = (torch.zeros(C * reps, dtype=torch.float64, device=dev).index_add_(0, uc.to(torch.int64), (str_ == 4).to(torch.float64))
                > 0).reshape(C, len(ODE_PARAMS), 2).any(dim=2).cpu().numpy()
        pkr = power_units(S["tab"], num_thermo, uc, uk0[bi].contiguous(), yr, idx)
        stp = torch.as_tensor(S["steps"], dtype=torch.float64, device=dev)[torch.as_tensor(ucn, device=dev)]
        for i, j in enumerate(ODE_PARAMS):
            plus = torch.as_tensor(start[ucn] + (2 * i) * nks[ucn] + pos, device=dev)
            minus = torch.as_tensor(start[ucn] + (2 * i + 1) * nks[ucn] + pos, device=dev)
            J[:, :, j] = (pkr[plus] - pkr[minus]) / (2.0 * stp[:, i:i + 1])
        for c in range(C):
            for i, j in enumerate(ODE_PARAMS):
                if rbad[c, i]:
                    J[offs[c]:offs[c + 1], :, j] = nan
                    status[c] = GRAD_COLUMN_UNDEFINED
    else:
        dth = torch.as_tensor(S["steps"], dtype=torch.float64, device=dev)
        for jo in range(nout):
            cot = torch.zeros((U, nout), dtype=torch.float64, device=dev)
            cot[:, jo] = 1.0
            gbar = _power_gbar(S["y"], pku, idx, cot)
            g, st = evolve_units_adjoint(S["t0"], num_thermo, S["tab"], dth, uc0, uk0, S["tau0"], S["tau"], S["mesh"], S["mlen"], gbar,
                                         **lm)
            J[:, jo, list(ODE_PARAMS)] = g
        lam = [c for c in range(C) if _is_lambda(thetas[c]) and not bad[c]]
        if lam:  # w0, wa at a cosmological constant: differences of the frozen-mesh map
            sub = [S["kgrids"][c] for c in lam] if S["ragged"] else S["kgrids"][0]
            _, Jf, stf = pk_jacobian_many(thetas[lam], base, sub, aexp_out=aexp_out, idx=idx, rel_step=rel_step, num_thermo=num_thermo,
                                          schedule=schedule, return_status=True, method="fd", **lm, **solver)
            for i, c in enumerate(lam):
                J[offs[c]:offs[c + 1], :, 6:8] = Jf[i][:, :, 6:8]
                status[c] = max(status[c], int(stf[i]))
    kp = torch.tensor([p.get("k_p", 0.05) for p in S["p0s"]], dtype=torch.float64, device=dev)[uc0.to(torch.int64)]
    J[:, :, 3] = pku * torch.log(uk0 / kp)[:, None]
    J[:, :, 4] = pku
    if _reionized(base):
        J[:, :, 8] = _reion_column(S, thetas, base, aexp, lm, solver, num_thermo, rel_step, idx)
    for c in range(C):
        if bad[c]:
            pku[offs[c]:offs[c + 1]] = nan
            J[offs[c]:offs[c + 1]] = nan
            status[c] = GRAD_PHANTOM
    pk, J = _per_cosmology(S, pku), _per_cosmology(S, J)
    if return_status:
        return pk, J, status
    return pk, J


def pk_and_vjp_many(thetas, base, kmodes, cotangents, aexp_out=(1.0,), idx=4, rel_step=REL_STEP, num_thermo=256, lmaxg=12, lmaxgp=12,
                    lmaxr=12, lmaxnu=17, nqmax=8, **solver):
    """P(k) and the 9-parameter VJP of sum_k,z cotangent * P for C cosmologies: one base launch and
    one adjoint launch for all of them.  kmodes: one shared grid, with cotangents (C, n_k, nout);
    or a list of C grids, with a list of C cotangent arrays (n_k_c, nout).  Returns (pk, vjp (C, 9),
    status (C,)), pk shaped as the cotangents.  At an exact cosmological constant the w0 and wa
    entries come from the difference Jacobian (see pk_jacobian_many)."""
    thetas = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
    C = thetas.shape[0]
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    aexp = np.atleast_1d(np.asarray(aexp_out, dtype=np.float64))
    S = _setup(thetas, base, kmodes, aexp, lm, solver, num_thermo, rel_step)
    bad, offs, uc0, uk0 = S["bad"], S["offs"], S["uc0"], S["uk0"]
    dev = uk0.device
    U = int(uk0.numel())
    pku = power_units(S["t0"], num_thermo, uc0, uk0, S["y"], idx)
    nout = pku.shape[1]
    if S["ragged"]:
        cot = torch.cat([torch.as_tensor(np.asarray(c, dtype=np.float64), device=dev).reshape(-1, nout) for c in cotangents])
    else:
        cot = torch.as_tensor(np.asarray(cotangents, dtype=np.float64), device=dev).reshape(-1, nout)
    if cot.shape[0] != U:
        raise ValueError("cotangents must match the k-grids")
    gbar = _power_gbar(S["y"], pku, idx, cot)
    dth = torch.as_tensor(S["steps"], dtype=torch.float64, device=dev)
    g, st = evolve_units_adjoint(S["t0"], num_thermo, S["tab"], dth, uc0, uk0, S["tau0"], S["tau"], S["mesh"], S["mlen"], gbar, **lm)
    ucl = uc0.to(torch.int64)
    v = torch.zeros((C, 9), dtype=torch.float64, device=dev)
    v[:, list(ODE_PARAMS)] = torch.zeros((C, len(ODE_PARAMS)), dtype=torch.float64, device=dev).index_add_(0, ucl, g)
    kp = torch.tensor([p.get("k_p", 0.05) for p in S["p0s"]], dtype=torch.float64, device=dev)[ucl]
    cp = (cot * pku).sum(dim=1)
    v[:, 3] = torch.zeros(C, dtype=torch.float64, device=dev).index_add_(0, ucl, cp * torch.log(uk0 / kp))
    v[:, 4] = torch.zeros(C, dtype=torch.float64, device=dev).index_add_(0, ucl, cp)
    if _reionized(base):
        dr = (cot * _reion_column(S, thetas, base, aexp, lm, solver, num_thermo, rel_step, idx)).sum(dim=1)
        v[:, 8] = torch.zeros(C, dtype=torch.float64, device=dev).index_add_(0, ucl, dr)
    status = np.where(bad, GRAD_PHANTOM, GRAD_OK).astype(np.int32)
    lam = [c for c in range(C) if _is_lambda(thetas[c]) and not bad[c]]
    if lam:
        sub = [S["kgrids"][c] for c in lam] if S["ragged"] else S["kgrids"][0]
        _, Jf, stf = pk_jacobian_many(thetas[lam], base, sub, aexp_out=aexp_out, idx=idx, rel_step=rel_step, num_thermo=num_thermo,
                                      return_status=True, method="fd", **lm, **solver)
        for i, c in enumerate(lam):
            v[c, 6:8] = torch.einsum("kon,ko->n", Jf[i][:, :, 6:8], cot[offs[c]:offs[c + 1]])
            status[c] = max(status[c], int(stf[i]))
    v[torch.as_tensor(bad, device=dev)] = float("nan")
    for c in range(C):
        if bad[c]:
            pku[offs[c]:offs[c + 1]] = float("nan")
    return _per_cosmology(S, pku), v, status


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
        logk = torch.log(S["uk0"] / S["p0s"][0].get("k_p", 0.05))[:, None]
        v[3] = (cot * pku * logk).sum()
        v[4] = (cot * pku).sum()
        f = ctx.fd
        if _reionized(f["base"]):
            v[8] = (cot * _reion_column(S, ctx.th, f["base"], f["aexp_out"], ctx.lm, f["solver"], ctx.num_thermo, f["rel_step"],
                                        ctx.idx)).sum()
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


def _cell_weights(weights, ellmax, dev):
    """Weights per C_l index 0..ellmax: given for l = 2..ellmax (ellmax - 1 values) or for every l."""
    w = torch.as_tensor(np.asarray(weights, dtype=np.float64), device=dev).reshape(-1)
    if w.numel() == ellmax - 1:
        w = torch.cat([torch.zeros(2, dtype=torch.float64, device=dev), w])
    if w.numel() != ellmax + 1:
        raise ValueError(f"weights must hold ellmax - 1 = {ellmax - 1} values (l = 2..ellmax) or ellmax + 1 (every l)")
    return w


def cell_vjp(theta, base, kmodes, weights, *, ellmax=2500, aexp_out=None, n_fftlog=16384, rel_step=REL_STEP, num_thermo=1024, lmaxg=24,
             lmaxgp=24, lmaxr=12, lmaxnu=8, nqmax=3, rtol=1e-4, atol=1e-4, isw="spline", return_cell=False):
    """L = sum_l weights_l C_l^TT and its gradient with respect to theta (9,).

    The chain is compute_Cell_spectrum_from_cosmo_params' (cmb.py:930-1090): background
    (num_thermo knots), per-mode adaptive perturbations at kmodes with the source function formed
    in the solver's output epilogue (compute_sources_fused), line of sight to ellmax at n_fftlog,
    C_l = trapezoid over ln k of (k / k_p)^(n_s - 1) Theta_l^2.  The derivative is that of the
    frozen-mesh map, as for P(k): the base solve's accepted steps are replayed for the twelve
    shifted cosmologies of the six parameters that enter the background and the perturbations
    (omega_c, omega_b, h, m_nu, w0, wa), each taken by a central difference of the whole chain
    (background, sources on the frozen mesh, line of sight, C_l); n_s by the exact cross
    spectrum of Theta_l and ln(k / k_p) Theta_l; ln(1e10 A_s) does not enter C_l as the reference
    defines it (A_s enters D_l); tau_reion enters only with reionization switched on in base
    (base["tau_reion"] > 0; then by a central difference like the six).  weights: for l =
    2..ellmax or for every l.
    Returns (L, grad (9,)) as a float and a device tensor, plus C_l when return_cell."""
    from .background import attach_background, evolve_background
    from .cmb import compute_Cell, compute_sources_fused, compute_theta_ell
    theta = np.asarray(theta, dtype=np.float64).reshape(-1)
    from .cmb import _driver_aexp
    aexp = _driver_aexp(base, aexp_out)
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    dev = _lib.device()
    kk = torch.as_tensor(np.asarray(kmodes, dtype=np.float64), device=dev).reshape(-1).contiguous()
    w = _cell_weights(weights, int(ellmax), dev)

    def chain(th, replay=None, p=None):
        if p is None:
            p = evolve_background(param=theta_to_param(th, base), num_thermo=num_thermo)
        src = compute_sources_fused(p, kk, aexp, rtol=rtol, atol=atol, isw=isw, step_log=replay is None, replay=replay, **lm)
        tau0 = float(p["tau_of_a_spline"].evaluate(np.array([1.0])).cpu()[0])
        th_l, kout = compute_theta_ell(ellmax, kk, src["tau"], src["S"], tau0, n_fftlog=n_fftlog)
        cl = compute_Cell(th_l, kout, p["n_s"], p.get("k_p", 0.05))
        return p, src, th_l, kout, cl

    p0, src0, th0, k0, cl0 = chain(theta)
    L0 = float((w * cl0).sum())
    t1 = src0["tau"].max().expand(kk.numel())
    mesh, mlen = _accepted_mesh(src0["step_log"], src0["n_accept"], t1)
    del src0
    g = torch.zeros(9, dtype=torch.float64, device=dev)
    # the shifted cosmologies' backgrounds in one launch (one workgroup each), then their chains
    cols = ODE_PARAMS + ((8,) if _reionized(base) else ())
    steps = [rel_step * max(abs(theta[j]), 1e-2) for j in cols]
    shifted = [th for j, d in zip(cols, steps) for th in _shift_pair(theta, j, d)]
    tabs = background_tables([theta_to_param(th, base) for th in shifted], num_thermo)
    for c, (j, d) in enumerate(zip(cols, steps)):
        Ls = []
        for h in range(2):
            i = 2 * c + h
            p = attach_background(theta_to_param(shifted[i], base), tabs, num_thermo, i)
            _, _, _, _, cl = chain(shifted[i], replay=(mesh, mlen), p=p)
            Ls.append((w * cl).sum())
        g[j] = (Ls[0] - Ls[1]) / (2.0 * d)
    # n_s: d C_l / d n_s = trapezoid over ln k of ln(k / k_p) (k / k_p)^(n_s - 1) Theta_l^2, the cross
    # spectrum of Theta_l with ln(k / k_p) Theta_l
    kp = p0.get("k_p", 0.05)
    dcl = compute_Cell(th0, k0, p0["n_s"], kp, theta_ell2=th0 * torch.log(k0 / kp)[:, None])
    g[3] = (w * dcl).sum()
    if return_cell:
        return L0, g, cl0
    return L0, g


__all__ = ["PARAM9", "theta_to_param", "param_to_theta", "pk_jacobian", "pk_jacobian_many", "pk_and_vjp_many", "PkFunction", "pk_and_vjp",
           "evolve_units_adjoint", "cell_vjp"]
