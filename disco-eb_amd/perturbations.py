"""evolve_perturbations, get_power and compute_time_derivatives on the GPU.

Mirrors src/discoeb/perturbations.py:1124-1226 (evolve_perturbations), :1446-1468 (get_power)
and :1339-1406 (compute_time_derivatives): same keyword arguments, same output layout
(num_k, nout, 20 fields or nvar), same param-dict annotations and error behaviour.  Each k-mode
is integrated independently (the reference's evolve_one_mode) by one wavefront of the HIP
kernel behind deb_evolve.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _geomspace(start, stop, num, endpoint=True):
    # jax.numpy.geomspace arithmetic: 10**(log10(start)*(1-s) + log10(stop)*s)
    l0, l1 = np.log10(start), np.log10(stop)
    div = (num - 1) if endpoint else num
    if num == 1:
        return np.array([start], dtype=np.float64)
    s = np.arange(num, dtype=np.float64) / div
    lin = l0 * (1.0 - s) + l1 * s
    if endpoint:
        lin[-1] = l1
    return 10.0 ** lin


def _linspace(start, stop, num, endpoint=True):
    div = (num - 1) if endpoint else num
    if num == 1:
        return np.array([start], dtype=np.float64)
    s = np.arange(num, dtype=np.float64) / div
    out = start * (1.0 - s) + stop * s
    if endpoint:
        out[-1] = stop
    return out


def k_grid(param, kmin, kmax, num_k, k_sampling_method="camb"):
    """k-sampling of perturbations.py:1177-1199 ('log', 'linear' or the CAMB-like hybrid)."""
    if k_sampling_method == "log":
        return _geomspace(kmin, kmax, num_k)
    if k_sampling_method == "linear":
        return _linspace(kmin, kmax, num_k)
    if k_sampling_method == "camb":
        if "tau_maxvis" not in param:
            raise ValueError("param dictionary must contain 'tau_maxvis' for 'camb' k-sampling.")
        taurst = float(param["tau_maxvis"])
        q1, q2 = 8.0 / taurst, 30.0 / taurst
        n1, n2 = int(num_k * 0.2), int(num_k * 0.4)
        n3 = num_k - n1 - n2
        return np.concatenate([_geomspace(kmin, q1, n1, endpoint=False), _linspace(q1, q2, n2, endpoint=False),
                               _geomspace(q2, kmax, n3)])
    raise ValueError(f"Unknown k_sampling_method: {k_sampling_method}")


def nvar(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax):
    return 7 + (lmaxg + 1) + (lmaxgp + 1) + (lmaxr + 1) + nqmax * (lmaxnu + 1) + 2


def evolve_units(tables, num_thermo, unit_cosmo, unit_k, tau_out, *, lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3,
                 rtol=1e-4, atol=1e-4, pcoeff=0.25, icoeff=0.80, dcoeff=0.0, factormax=20.0, factormin=0.3, max_steps=2048,
                 return_full=False, out=None, step_log=False, schedule=None, replay=None):
    """Low-level launch: independent adaptive solves of (cosmology, k) units.

    tables: (n_cosmo, size) device tensor from background_tables; unit_cosmo (int32) and unit_k
    (float64) device tensors of length n_units; tau_out (n_cosmo, nout) device tensor.
    schedule: None (units launched in index order, one workgroup each) or a dict with keys
    order: None, "k_desc" (largest k first) or an int32 device tensor permuting the units, and
    threads: 64, 128 or 256 threads per unit (absent or 0: the library default).
    replay: None, or (mesh (n_units, L) float64, mesh_len (n_units,) int32) device tensors: each
    unit then takes exactly its accepted steps, ending at t0 + s (t1 - t0) for the listed s
    (the frozen-mesh map, see deb_evolve_replay).
    Returns (y, status, n_accept, n_reject) device tensors, plus the step mesh
    (n_units, max_steps, 3: t_prev, t_next, error norm) when step_log is set; asynchronous on
    the current stream."""
    L = _lib.lib()
    dev = tables.device
    n_units = int(unit_k.numel())
    nout = int(tau_out.shape[-1])
    nv = nvar(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax)
    w = nv if return_full else 20
    y = out if out is not None else torch.empty((n_units, nout, w), dtype=torch.float64, device=dev)
    status = torch.empty(n_units, dtype=torch.int32, device=dev)
    nacc = torch.empty(n_units, dtype=torch.int32, device=dev)
    nrej = torch.empty(n_units, dtype=torch.int32, device=dev)
    work = torch.empty(L.deb_evolve_workspace_doubles(), dtype=torch.float64, device=dev)
    prob = _lib.DebProblem(int(lmaxg), int(lmaxgp), int(lmaxr), int(lmaxnu), int(nqmax), int(bool(return_full)))
    cfg = _lib.DebSolverCfg(float(rtol), float(atol), float(pcoeff), float(icoeff), float(dcoeff), float(factormax),
                            float(factormin), int(max_steps))
    uc = unit_cosmo.to(device=dev, dtype=torch.int32).contiguous()
    uk = unit_k.to(device=dev, dtype=torch.float64).contiguous()
    to = tau_out.to(device=dev, dtype=torch.float64).contiguous()
    log = torch.full((n_units, int(max_steps), 3), float("nan"), dtype=torch.float64, device=dev) if step_log else None
    sched, order = _schedule(schedule, uk)
    sref = ctypes.byref(sched) if sched is not None else None
    if replay is not None:
        if step_log:
            raise ValueError("step_log and replay are exclusive")
        mesh = replay[0].to(device=dev, dtype=torch.float64).contiguous()
        mlen = replay[1].to(device=dev, dtype=torch.int32).contiguous()
        if mesh.ndim != 2 or mesh.shape[0] != n_units or mlen.numel() != n_units:
            raise ValueError("replay mesh must be (n_units, L) with one length per unit")
        _check_mesh(mesh, mlen)
        _lib.check(L.deb_evolve_replay(_lib.ptr(tables), int(num_thermo), _lib.ptr(uc), _lib.ptr(uk), n_units, _lib.ptr(to), nout,
                                       ctypes.byref(prob), ctypes.byref(cfg), _lib.ptr(mesh), mesh.shape[1], _lib.ptr(mlen), _lib.ptr(y),
                                       _lib.ptr(status), _lib.ptr(nacc), _lib.ptr(nrej), _lib.ptr(work), sref, _lib.stream()))
    else:
        _lib.check(L.deb_evolve(_lib.ptr(tables), int(num_thermo), _lib.ptr(uc), _lib.ptr(uk), n_units, _lib.ptr(to), nout,
                                ctypes.byref(prob), ctypes.byref(cfg), _lib.ptr(y), _lib.ptr(status), _lib.ptr(nacc),
                                _lib.ptr(nrej), _lib.ptr(work), _lib.ptr(log) if log is not None else None, sref, _lib.stream()))
    del order  # stream-ordered: the allocator reuses it only after the launch has run
    if step_log:
        return y, status, nacc, nrej, log
    return y, status, nacc, nrej


def evolve_units_batched(tables, num_thermo, unit_cosmo, unit_k, tau_out, *, batch_size=16, lmaxg=11, lmaxgp=11, lmaxr=11,
                         lmaxnu=8, nqmax=3, rtol=1e-4, atol=1e-4, pcoeff=0.25, icoeff=0.80, dcoeff=0.0, factormax=20.0,
                         factormin=0.3, max_steps=2048, return_full=False, step_log=False, threads=0, static_scratch=False):
    """Low-level launch of the batched shared-step solve (deb_evolve_batched): the units
    g * batch_size .. (g + 1) * batch_size - 1, which must share a cosmology, step together.
    Arguments as evolve_units; threads: 64, 128 or 256 threads per batch (0: library default).
    static_scratch: call deb_evolve_batched_noalloc, whose workspace is the library's static
    per-device array (no workspace tensor; at most deb_noalloc_scratch_doubles() doubles, i.e.
    n_units x 2 nvar + 32; calls that overlap on one device race).  Returns (y, status,
    n_accept, n_reject) device tensors, plus the per-batch step mesh (n_units / batch_size,
    max_steps, 3) when step_log is set."""
    L = _lib.lib()
    dev = tables.device
    n_units = int(unit_k.numel())
    if batch_size < 1 or n_units % batch_size != 0:
        raise ValueError(f"num_k = {n_units} must be a multiple of batch_size = {batch_size}")
    nout = int(tau_out.shape[-1])
    nv = nvar(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax)
    y = torch.empty((n_units, nout, nv if return_full else 20), dtype=torch.float64, device=dev)
    status, nacc, nrej = (torch.empty(n_units, dtype=torch.int32, device=dev) for _ in range(3))
    prob = _lib.DebProblem(int(lmaxg), int(lmaxgp), int(lmaxr), int(lmaxnu), int(nqmax), int(bool(return_full)))
    cfg = _lib.DebSolverCfg(float(rtol), float(atol), float(pcoeff), float(icoeff), float(dcoeff), float(factormax),
                            float(factormin), int(max_steps))
    work = (None if static_scratch else
            torch.empty(L.deb_evolve_batched_workspace_doubles(n_units, ctypes.byref(prob)), dtype=torch.float64, device=dev))
    uc = unit_cosmo.to(device=dev, dtype=torch.int32).contiguous()
    uk = unit_k.to(device=dev, dtype=torch.float64).contiguous()
    to = tau_out.to(device=dev, dtype=torch.float64).contiguous()
    log = (torch.full((n_units // batch_size, int(max_steps), 3), float("nan"), dtype=torch.float64, device=dev)
           if step_log else None)
    sched, _ = _schedule({"threads": threads}, uk)
    lg, sr = _lib.ptr(log) if log is not None else None, ctypes.byref(sched) if sched is not None else None
    if static_scratch:
        _lib.check(L.deb_evolve_batched_noalloc(_lib.ptr(tables), int(num_thermo), _lib.ptr(uc), _lib.ptr(uk), n_units, int(batch_size),
                                                _lib.ptr(to), nout, ctypes.byref(prob), ctypes.byref(cfg), _lib.ptr(y), _lib.ptr(status),
                                                _lib.ptr(nacc), _lib.ptr(nrej), lg, sr, _lib.stream()))
    else:
        _lib.check(L.deb_evolve_batched(_lib.ptr(tables), int(num_thermo), _lib.ptr(uc), _lib.ptr(uk), n_units, int(batch_size),
                                        _lib.ptr(to), nout, ctypes.byref(prob), ctypes.byref(cfg), _lib.ptr(y), _lib.ptr(status),
                                        _lib.ptr(nacc), _lib.ptr(nrej), _lib.ptr(work), lg, sr, _lib.stream()))
    if step_log:
        return y, status, nacc, nrej, log
    return y, status, nacc, nrej


def _schedule(schedule, uk):
    """DebSchedule for a schedule dict (see evolve_units) and the order tensor it points to."""
    if schedule is None:
        return None, None
    threads = int(schedule.get("threads", 0) or 0)
    if threads not in (0, 64, 128, 256):
        raise ValueError(f"threads per unit must be 64, 128 or 256 (got {threads})")
    order = schedule.get("order")
    if isinstance(order, str):
        if order != "k_desc":
            raise ValueError(f"unknown unit order {order!r} (expected 'k_desc', None or a tensor)")
        order = torch.argsort(uk, descending=True, stable=True).to(torch.int32)
    elif order is not None:
        order = torch.as_tensor(order, dtype=torch.int32, device=uk.device).contiguous()
        if order.numel() != uk.numel():
            raise ValueError("order must list every unit once")
        if not torch.cuda.is_current_stream_capturing():  # indices reach the kernel unchecked
            srt = torch.sort(order.long()).values.cpu()
            if not torch.equal(srt, torch.arange(uk.numel())):
                raise ValueError("order must be a permutation of 0..n_units-1")
    if order is None:
        return (_lib.DebSchedule(None, threads) if threads else None), None
    return _lib.DebSchedule(order.data_ptr(), threads), order


def _check_mesh(mesh, mlen):
    """Host check of a replay mesh (skipped under graph capture): the kernel reads mesh[u, i] for
    every i < mlen[u], so each length must lie in 1..L; each row must increase from above 0 and
    end at 1."""
    if torch.cuda.is_current_stream_capturing():
        return
    L = mesh.shape[1]
    lens = mlen.to(torch.int64)
    if lens.numel() and (int(lens.min()) < 1 or int(lens.max()) > L):
        raise ValueError(f"replay mesh lengths must lie in 1..{L}")
    cols = torch.arange(L, device=mesh.device)[None, :]
    inside = cols < lens[:, None]
    prev = torch.cat([torch.zeros_like(mesh[:, :1]), mesh[:, :-1]], dim=1)
    if bool((inside & ~(mesh > prev)).any()):
        raise ValueError("replay mesh rows must increase from above 0")
    last = mesh.gather(1, (lens - 1)[:, None]).flatten()
    if bool((last != 1.0).any()):
        raise ValueError("replay mesh rows must end at 1 (the last output time)")


def _raise_on_status(status, allow_phantom=False):
    st = status.cpu()
    if allow_phantom:
        st = torch.where(st == 4, torch.zeros_like(st), st)
    if (st == 1).any():
        raise RuntimeError("The maximum number of solver steps was reached. Try increasing `max_steps`.")
    if (st == 2).any():
        raise RuntimeError("Non-finite error estimate in the perturbation solve.")
    if (st == 3).any():
        raise RuntimeError("internal error: LDS carve-up exceeds the launch size")
    if (st == 5).any():
        raise ValueError("the units of one batch must belong to one cosmology")
    if (st == 4).any():
        raise ValueError("w(a) = w_DE_0 + w_DE_a (1 - a) crosses -1 in (0, 1]: the dark-energy fluid equations are singular "
                         "at the phantom divide (use w_DE_0 = -1, w_DE_a = 0 for a cosmological constant)")


# Launch order of evolve_perturbations: largest k (most steps) first, so the longest solves
# start first when there are more units than resident workgroups.  Results do not depend on it.
DEFAULT_SCHEDULE = {"order": "k_desc"}


def evolve_perturbations(*, param, aexp_out, kmin: float, kmax: float, num_k: int, lmaxg: int = 11, lmaxgp: int = 11,
                         lmaxr: int = 11, lmaxnu: int = 8, nqmax: int = 3, rtol: float = 1e-4, atol: float = 1e-4,
                         pcoeff: float = 0.25, icoeff: float = 0.80, dcoeff: float = 0.0, factormax: float = 20.0,
                         factormin: float = 0.3, max_steps: int = 2048, return_full: bool = False,
                         k_sampling_method: str = "camb"):
    """Synchronous-gauge perturbations on a k-grid (perturbations.py:1124-1226).

    Returns (y[num_k, nout, 20 or nvar], kmodes, param) as float64 device tensors; annotates
    param with lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax, nout and tau_out as the reference does."""
    kmodes = k_grid(param, kmin, kmax, num_k, k_sampling_method)
    tables = param["_deb_tables"]
    dev = tables.device
    aexp = torch.as_tensor(np.atleast_1d(np.asarray(aexp_out, dtype=np.float64)) if not torch.is_tensor(aexp_out) else aexp_out,
                           dtype=torch.float64, device=dev).reshape(-1)
    tau_out = param["tau_of_a_spline"].evaluate(aexp)
    cosmo = int(param.get("_deb_cosmo", 0))
    n_cosmo = tables.shape[0] if tables.ndim == 2 else 1
    tau_all = torch.zeros((n_cosmo, tau_out.numel()), dtype=torch.float64, device=dev)
    tau_all[cosmo] = tau_out
    kt = torch.as_tensor(kmodes, dtype=torch.float64, device=dev)
    uc = torch.full((kt.numel(),), cosmo, dtype=torch.int32, device=dev)
    y, status, nacc, nrej = evolve_units(tables, param["_deb_num_thermo"], uc, kt, tau_all, lmaxg=lmaxg, lmaxgp=lmaxgp,
                                         lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax, rtol=rtol, atol=atol, pcoeff=pcoeff,
                                         icoeff=icoeff, dcoeff=dcoeff, factormax=factormax, factormin=factormin,
                                         max_steps=max_steps, return_full=return_full, schedule=DEFAULT_SCHEDULE)
    _raise_on_status(status)
    param["lmaxg"], param["lmaxgp"], param["lmaxr"], param["lmaxnu"], param["nqmax"] = lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax
    param["nout"] = int(aexp.numel())
    param["tau_out"] = tau_out
    param["_deb_steps"] = (nacc, nrej)
    return y, kt, param


def evolve_perturbations_batched(*, param, aexp_out, kmin: float, kmax: float, num_k: int, lmaxg: int = 11, lmaxgp: int = 11,
                                 lmaxr: int = 11, lmaxnu: int = 8, nqmax: int = 3, rtol: float = 1e-4, atol: float = 1e-4,
                                 pcoeff: float = 0.25, icoeff: float = 0.80, dcoeff: float = 0.0, factormax: float = 20.0,
                                 factormin: float = 0.3, max_steps: int = 2048, batch_size: int = 16, return_full: bool = False,
                                 k_sampling_method: str = "camb"):
    """Perturbations with batch_size consecutive k-modes on one shared step sequence
    (perturbations.py:1228-1337); num_k must be a multiple of batch_size.  Same returns and
    param annotations as evolve_perturbations."""
    kmodes = k_grid(param, kmin, kmax, num_k, k_sampling_method)
    if batch_size < 1 or num_k % batch_size != 0:
        raise ValueError(f"num_k = {num_k} must be a multiple of batch_size = {batch_size}")
    tables = param["_deb_tables"]
    dev = tables.device
    aexp = torch.as_tensor(np.atleast_1d(np.asarray(aexp_out, dtype=np.float64)) if not torch.is_tensor(aexp_out) else aexp_out,
                           dtype=torch.float64, device=dev).reshape(-1)
    tau_out = param["tau_of_a_spline"].evaluate(aexp)
    cosmo = int(param.get("_deb_cosmo", 0))
    n_cosmo = tables.shape[0] if tables.ndim == 2 else 1
    tau_all = torch.zeros((n_cosmo, tau_out.numel()), dtype=torch.float64, device=dev)
    tau_all[cosmo] = tau_out
    kt = torch.as_tensor(kmodes, dtype=torch.float64, device=dev)
    uc = torch.full((kt.numel(),), cosmo, dtype=torch.int32, device=dev)
    y, status, nacc, nrej = evolve_units_batched(tables, param["_deb_num_thermo"], uc, kt, tau_all, batch_size=batch_size, lmaxg=lmaxg,
                                                 lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax, rtol=rtol, atol=atol,
                                                 pcoeff=pcoeff, icoeff=icoeff, dcoeff=dcoeff, factormax=factormax,
                                                 factormin=factormin, max_steps=max_steps, return_full=return_full)
    _raise_on_status(status)
    param["lmaxg"], param["lmaxgp"], param["lmaxr"], param["lmaxnu"], param["nqmax"] = lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax
    param["nout"] = int(aexp.numel())
    param["tau_out"] = tau_out
    param["_deb_steps"] = (nacc, nrej)
    return y, kt, param


def get_power(*, k, y, idx: int, param):
    """P(k) = 2 pi^2 A_s (k/k_p)^(n_s-1) k^-3 y[..., idx]^2 (perturbations.py:1446-1468), on the GPU."""
    yt = torch.as_tensor(y, dtype=torch.float64).contiguous()
    dev = yt.device
    kt = torch.as_tensor(k, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    nf = yt.shape[-1]
    rows = yt.numel() // nf
    krep = max(rows // max(kt.numel(), 1), 1)
    out = torch.empty(rows, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().deb_power(_lib.ptr(kt), krep, _lib.ptr(yt), rows, nf, int(idx), float(param["A_s"]),
                                    float(param["n_s"]), float(param.get("k_p", 0.05)), _lib.ptr(out), _lib.stream()))
    return out.reshape(yt.shape[:-1])


def compute_time_derivatives(yout, tau, kmodes, param, lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax):
    """dy/dtau re-evaluated at saved states (perturbations.py:1339-1406)."""
    L = _lib.lib()
    tables = param["_deb_tables"]
    dev = tables.device
    y = torch.as_tensor(yout, dtype=torch.float64, device=dev).contiguous()
    tt = torch.as_tensor(tau, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    kt = torch.as_tensor(kmodes, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    if y.shape != (kt.numel(), tt.numel(), nvar(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax)):
        raise ValueError("yout must have shape (n_kmodes, n_times, nvar)")
    dy = torch.empty_like(y)
    work = torch.empty(L.deb_evolve_workspace_doubles(), dtype=torch.float64, device=dev)
    prob = _lib.DebProblem(int(lmaxg), int(lmaxgp), int(lmaxr), int(lmaxnu), int(nqmax), 1)
    _lib.check(L.deb_time_derivatives(_lib.ptr(tables), int(param["_deb_num_thermo"]), int(param.get("_deb_cosmo", 0)),
                                      _lib.ptr(y), _lib.ptr(tt), tt.numel(), _lib.ptr(kt), kt.numel(), ctypes.byref(prob),
                                      _lib.ptr(dy), _lib.ptr(work), _lib.stream()))
    return dy
