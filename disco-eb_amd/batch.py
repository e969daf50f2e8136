"""Many cosmologies at once: one background launch for all of them, one launch for every
(cosmology, k) unit in cost order, P(k) for all units; the launch-order autotuner; graph
capture of the whole P(k) pipeline.

The reference runs one cosmology per call (evolve_background -> evolve_perturbations ->
get_power, src/discoeb/background.py:191-345, perturbations.py:1124-1226, 1446-1468) and
vmaps over k; here the cosmology index is part of the unit, so a Fisher-style batch of
cosmologies with ragged k-grids is one launch.
"""
import json
import os
import time

import numpy as np
import torch

from . import _lib
from .background import background_tables
from .perturbations import _raise_on_status, evolve_units, nvar


def tau_out_all(tables, num_thermo, aexp_out):
    """Conformal times tau(a_out) of every cosmology: (n_cosmo, nout) device tensor."""
    aexp = torch.as_tensor(aexp_out, dtype=torch.float64, device=tables.device).reshape(-1).contiguous()
    nc = tables.shape[0]
    tau = torch.empty((nc, aexp.numel()), dtype=torch.float64, device=tables.device)
    _lib.check(_lib.lib().deb_tau_out(_lib.ptr(tables), int(num_thermo), nc, _lib.ptr(aexp), aexp.numel(), _lib.ptr(tau),
                                      _lib.stream()))
    return tau


def units_for(kgrids, device):
    """(unit_cosmo int32, unit_k float64, offsets) for a list of per-cosmology k-grids."""
    counts = [len(k) for k in kgrids]
    uc = torch.repeat_interleave(torch.arange(len(kgrids), dtype=torch.int32), torch.tensor(counts)).to(device)
    uk = torch.as_tensor(np.concatenate([np.asarray(k, dtype=np.float64) for k in kgrids]), device=device)
    offsets = np.concatenate([[0], np.cumsum(counts)])
    return uc, uk, offsets


def power_units(tables, num_thermo, unit_cosmo, unit_k, y, idx):
    """P(k) of every unit and output time, with each unit's own A_s, n_s, k_p: (n_units, nout)."""
    yt = y.contiguous()
    n_units, nout, w = yt.shape
    pk = torch.empty((n_units, nout), dtype=torch.float64, device=yt.device)
    _lib.check(_lib.lib().deb_power_units(_lib.ptr(tables), int(num_thermo), _lib.ptr(unit_cosmo), _lib.ptr(unit_k), n_units, nout,
                                          _lib.ptr(yt), w, int(idx), _lib.ptr(pk), _lib.stream()))
    return pk


def evolve_many(params, kgrids, aexp_out, *, num_thermo=256, idx=4, schedule="auto", lmaxg=11, lmaxgp=11, lmaxr=11,
                lmaxnu=8, nqmax=3, return_full=False, **solver):
    """Perturbations and P(k) of several cosmologies, each on its own k-grid, in one launch.

    params: list of parameter dicts; kgrids: list of 1-d arrays (one per cosmology);
    schedule: "auto" (autotuned launch order), None (index order) or a schedule dict (see
    evolve_units).  Returns (ys, pks, status): lists of per-cosmology device tensors
    (n_k, nout, 20 or nvar) and (n_k, nout), views into single batch arrays, and the per-unit
    status.  Cosmologies whose w(a) crosses -1 do not stop the batch: their units carry
    status 4 and NaN outputs; any other failure raises."""
    if len(params) != len(kgrids):
        raise ValueError("one k-grid per cosmology is required")
    tables = background_tables(params, num_thermo)
    tau = tau_out_all(tables, num_thermo, aexp_out)
    uc, uk, offs = units_for(kgrids, tables.device)
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    if isinstance(schedule, str):
        if schedule != "auto":
            raise ValueError("schedule must be 'auto', None or a dict")
        schedule = autotune_schedule(tables, num_thermo, uc, uk, tau, **lm, **solver)
    y, st, na, nr = evolve_units(tables, num_thermo, uc, uk, tau, **lm, return_full=return_full, schedule=schedule, **solver)
    _raise_on_status(st, allow_phantom=True)
    pk = power_units(tables, num_thermo, uc, uk, y, idx)
    ys = [y[offs[i]:offs[i + 1]] for i in range(len(params))]
    pks = [pk[offs[i]:offs[i + 1]] for i in range(len(params))]
    return ys, pks, st


# ---------------------------------------------------------------------------------------------
# autotuner
# ---------------------------------------------------------------------------------------------
_TUNED = {}
CACHE_ENV = "DISCOEB_AMD_TUNE_CACHE"  # optional JSON file that keeps tuned schedules across runs


def _cache_key(n_units, nout, lm, solver):
    return json.dumps([n_units, nout, sorted(lm.items()), sorted((k, float(v)) for k, v in solver.items())])


def _load_cache():
    path = os.environ.get(CACHE_ENV)
    if path and os.path.exists(path):
        with open(path) as f:
            _TUNED.update(json.load(f))


def _store_cache():
    path = os.environ.get(CACHE_ENV)
    if path:
        with open(path, "w") as f:
            json.dump(_TUNED, f)


def schedule_candidates(n_units):
    """Schedules the autotuner tries: index order (units grouped by cosmology, k ascending) or
    largest k first across the whole batch, each with 64, 128 or 256 threads per unit (more
    threads shorten every step of a unit; fewer keep more units in flight)."""
    return [{"order": order, "threads": nt} for order in (None, "k_desc") for nt in (64, 128, 256)]


def autotune_schedule(tables, num_thermo, unit_cosmo, unit_k, tau_out, *, sample=1024, repeats=1, **kw):
    """Pick the fastest schedule for this problem shape by timing each candidate with HIP events
    on a strided sample of the units (every unit's cost depends on k only, so the sample keeps
    the cost distribution).  Results are cached per (n_units, nout, truncations, solver)."""
    _load_cache()
    lm = {k: kw[k] for k in ("lmaxg", "lmaxgp", "lmaxr", "lmaxnu", "nqmax") if k in kw}
    solver = {k: v for k, v in kw.items() if k not in lm and k != "return_full"}
    n_units = int(unit_k.numel())
    key = _cache_key(n_units, int(tau_out.shape[-1]), lm, solver)
    if key in _TUNED:
        return _decode(_TUNED[key])
    stride = max(1, n_units // sample)
    sel = torch.arange(0, n_units, stride, device=unit_k.device)
    uc, uk = unit_cosmo[sel].contiguous(), unit_k[sel].contiguous()
    best, best_ms, timings = None, float("inf"), []
    for cand in schedule_candidates(n_units):
        sch = None if cand is None else dict(cand)
        ms = []
        for _ in range(repeats):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            evolve_units(tables, num_thermo, uc, uk, tau_out, schedule=sch, **kw)
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        t = float(np.median(ms))
        timings.append((cand, t))
        if t < best_ms:
            best, best_ms = sch, t
    _TUNED[key] = {"schedule": best, "ms": best_ms, "sample_units": int(sel.numel()), "tuned_at": time.time(),
                   "timings": [[c, t] for c, t in timings]}
    _store_cache()
    return best


def _decode(entry):
    return entry["schedule"]


# ---------------------------------------------------------------------------------------------
# graph capture
# ---------------------------------------------------------------------------------------------
class PkGraph:
    """Background -> tau_out -> perturbations -> P(k) for fixed shapes, captured once into a HIP
    graph and replayed.  New parameter values go into the static parameter tensor
    (set_params), and replay() recomputes everything on the GPU with no host work besides the
    single graph launch.  Shapes (number of cosmologies, k-grids, output times, truncations)
    are fixed at capture."""

    def __init__(self, params, kgrids, aexp_out, *, num_thermo=256, idx=4, schedule=None, lmaxg=11, lmaxgp=11, lmaxr=11,
                 lmaxnu=8, nqmax=3, **solver):
        dev = _lib.device()
        self.num_thermo, self.idx = int(num_thermo), int(idx)
        self.lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
        self.solver = solver
        self.params = torch.tensor([_lib.param_vector(p) for p in params], dtype=torch.float64, device=dev)
        self.aexp = torch.as_tensor(aexp_out, dtype=torch.float64, device=dev).reshape(-1).contiguous()
        self.uc, self.uk, self.offsets = units_for(kgrids, dev)
        nc, nout = len(params), self.aexp.numel()
        L = _lib.lib()
        self.tables = torch.empty((nc, L.deb_tables_size(self.num_thermo)), dtype=torch.float64, device=dev)
        self.tau = torch.empty((nc, nout), dtype=torch.float64, device=dev)
        n_units = self.uk.numel()
        self.y = torch.empty((n_units, nout, 20), dtype=torch.float64, device=dev)
        self.schedule = None
        if schedule is not None:
            self.schedule = dict(schedule)
            if isinstance(self.schedule.get("order"), str):  # sort outside the capture
                self.schedule["order"] = torch.argsort(self.uk, descending=True, stable=True).to(torch.int32)
        self._run()  # warm-up outside the capture (kernel attributes, allocator pools)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()  # default priority
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(self.graph, stream=s):
                self.status, self.pk = self._run()
        torch.cuda.current_stream().wait_stream(s)

    def _run(self):
        L = _lib.lib()
        nc = self.params.shape[0]
        _lib.check(L.deb_background(_lib.ptr(self.params), nc, self.num_thermo, _lib.ptr(self.tables), _lib.stream()))
        _lib.check(L.deb_tau_out(_lib.ptr(self.tables), self.num_thermo, nc, _lib.ptr(self.aexp), self.aexp.numel(),
                                 _lib.ptr(self.tau), _lib.stream()))
        y, st, na, nr = evolve_units(self.tables, self.num_thermo, self.uc, self.uk, self.tau, **self.lm, out=self.y,
                                     schedule=self.schedule, **self.solver)
        pk = power_units(self.tables, self.num_thermo, self.uc, self.uk, y, self.idx)
        return st, pk

    def set_params(self, params):
        """Copy new parameter values (same number of cosmologies) into the captured input."""
        v = torch.tensor([_lib.param_vector(p) for p in params], dtype=torch.float64)
        if v.shape != self.params.shape:
            raise ValueError("the number of cosmologies is fixed at capture")
        self.params.copy_(v)

    def replay(self):
        """Run the captured pipeline; returns (P(k) (n_units, nout), status) device tensors."""
        self.graph.replay()
        return self.pk, self.status

    def split(self, t):
        return [t[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.offsets) - 1)]


__all__ = ["evolve_many", "power_units", "tau_out_all", "units_for", "autotune_schedule", "schedule_candidates", "PkGraph",
           "nvar"]
