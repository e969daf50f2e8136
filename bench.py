"""Throughput of the perturbation path on one MI355X.

Workload (one "step"): background + RECFAST tables, then 512 k-modes log-spaced in
1e-5..5 Mpc^-1 evolved independently by the adaptive Rodas5 kernel (rtol 1e-5, atol 1e-10,
lmax 12 for photons and massless neutrinos, one massive neutrino of 0.06 eV on an 8-point
momentum grid with lmax_nu 17, w0 = -0.95, wa = 0.1), then P(k) at z = 2, 1, 0.

Prints one JSON line: k-mode ODE solves/s, the ODE kernel's roofline position (FP64 FLOP/s
from the kernel's own step counters, timed with HIP events on the launch stream) and the
CPU oracle timed on a bounded sample of the same modes.  After the timed steps (outside
them) it also measures, unless switched off:
  * gradients: P(k, z = 0) and its 9-parameter VJP for 256 k-modes log-spaced in 1e-4..1/Mpc
    (same cosmology and solver settings), evaluations/s;
  * batched gradients: the P(k, z = 0) Jacobian of 16 cosmologies around the workload's in two
    perturbation launches, against one call per cosmology (results bitwise equal);
  * line of sight: Theta_l(k) for 3000 k x 2500 conformal times, l = 0..2500, 16384-point
    FFTLog (synthetic smooth sources), and its FP64 matrix-core contraction on its own.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)


def _load():
    import importlib.util
    pkg_dir = os.path.join(ROOT, "disco-eb_amd")
    spec = importlib.util.spec_from_file_location("discoeb_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["discoeb_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def workload_param():
    h = 0.674
    return dict(Omegam=(0.120 + 0.0224) / h ** 2, Omegab=0.0224 / h ** 2, H0=100 * h, Tcmb=2.7255, YHe=0.248, Neff=2.046,
                Nmnu=1.0, mnu=0.06, w_DE_0=-0.95, w_DE_a=0.1, cs2_DE=1.0, Omegak=0.0, A_s=2.1e-9, n_s=0.965, k_p=0.05)


LMAX = dict(lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8)
SOLVER = dict(rtol=1e-5, atol=1e-10)
AOUT = [1.0 / 3.0, 0.5, 1.0]
NK = 512


def work_per_step(lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax):
    """Executed FP64 operations of one attempted step of the structured Rodas5 solve (the kernel's
    own algorithm, counted item by item; a fused multiply-add counts as 2).

    n state entries; m = 17 + 3 nq dense unknowns; nce = n - 1 - m tail-chain entries; NRED = 19.
      R    one right-hand side: metric (60 + 8 nq), core rows (15 each), tail rows (5 each)
      Jacobian point: the scale-factor column and d f / dtau in one two-direction dual pass (3 R);
           unit-vector columns: metric per column plus the rows it reads (17 core rows, and 3
           rows of its own momentum bin for a neutrino column; one parent row per chain head);
           tail coefficients and the bottom-up chain elimination (8 per chain entry)
      bordered reduction: nq 3x3 inverses (90 each), Q (6 per neutrino unknown), the reduced
           matrix (W_cn Q, I - G_n^T Q) and P = [W_cn; G_n^T] M^-1
      Gauss-Jordan inverse of the 19 x 19 reduced matrix: 2 NRED^3
      8 stages: stage inputs (34 n in all), stage right-hand sides (7 R, the first shares the
           Jacobian point's background), assembly of the dt-weighted and C-weighted terms and the
           row-0 column (82 n in all), and 8 structured solves (chain sweeps 6 per entry,
           M^-1 b_n and P b_n, the 19 x 19 mat-vec, x_n)
      candidate, error norm: 3 n
    Returns (operations of a step, n, m, the Jacobian point's share): an attempt that retries a
    rejected one from the same point skips the dual pass, the unit-vector columns and the tail
    coefficients (4 per chain entry), so it executes the first figure less the last."""
    n = 7 + (lmaxg + 1) + (lmaxgp + 1) + (lmaxr + 1) + nqmax * (lmaxnu + 1) + 2
    nq = nqmax
    m = 17 + 3 * nq
    nch = 3 + nq
    nce = n - 1 - m
    NRED = 19
    metric = 60 + 8 * nq
    R = metric + 15 * m + 5 * nce
    jac_point = 3 * R + (17 * (metric + 15 * 17) + 3 * nq * (metric + 15 * 20) + nch * (metric + 15)) + 4 * nce
    jac = jac_point + 4 * nce
    reduce = 90 * nq + 6 * 3 * nq + (17 * 2 * 3 * nq * 2 + 2 * 2 * 3 * nq * 2) + NRED * 3 * nq * 5
    gj = 2 * NRED ** 3
    solve = 6 * nce + 5 * 3 * nq + 2 * NRED * 3 * nq + 2 * NRED * NRED + 4 * 3 * nq
    stages = 34 * n + 7 * R + 82 * n + 8 * solve
    return jac + reduce + gj + stages + 3 * n, n, m, jac_point


def measured_traffic(key):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE passes, scaled by a device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        entry = json.load(f).get(key)
    return None if entry is None else float(entry["bytes_per_launch"])


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_cores():
    """Cores this process may use: its affinity set, capped by OMP_NUM_THREADS when that is set
    (the share of a shared host given to one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else n


def cpu_baseline(kmodes, sample, sample_1core):
    """The oracle's C++ FP64 Rodas5 (same tableau, controller and interpolation as the HIP path, but
    with a dense LU of the full n x n matrix per step) built -O3 -march=native on this host and
    timed over all available cores (OpenMP over modes) and on one core."""
    import subprocess
    odir = os.path.join(ROOT, "oracle")
    subprocess.check_call(["make", "-s", "-C", odir, "native"])
    os.environ["DEB_ORACLE_LIB"] = os.path.join(odir, "_build", "libdeb_oracle_native.so")
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    P = workload_param()

    def run(nsample, threads):
        idx = np.unique(np.linspace(0, len(kmodes) - 1, nsample).round().astype(int))
        ks = kmodes[idx]
        bg = O.Background(P, 256)
        t = time.perf_counter()
        bg.evolve(ks, AOUT, **LMAX, **SOLVER, nthreads=threads)
        return len(ks), time.perf_counter() - t

    cores = _host_cores()
    n_all, dt_all = run(sample, cores)
    n_one, dt_one = run(sample_1core, 1)
    return {"value": n_all / dt_all, "unit": "k-mode solves/s", "cores": cores, "kind": "port",
            "value_1core": n_one / dt_one, "cpu_model": _cpu_model(),
            "sample": f"{n_all} of the {len(kmodes)} k-modes (evenly strided over the log k grid) on {cores} threads in "
                      f"{dt_all:.1f} s, and {n_one} modes on 1 thread in {dt_one:.1f} s; background included.  The oracle's "
                      f"C++ FP64 Rodas5 built -O3 -march=native (no FP contraction): same tableau, PI controller and output "
                      f"interpolation as the GPU path, dense LU of the full state matrix per step (the GPU uses the "
                      f"structured solve), OpenMP over modes; not the reference's JAX code, which cannot run here"}


def measure_gradient(deb, nk, repeats=2):
    """P(k, z=0) plus its 9-parameter VJP (seeded standard-normal cotangent) per second, by the
    discrete adjoint (one base launch, one adjoint launch) and, beside it, by differences."""
    import torch
    P = workload_param()
    theta = deb.param_to_theta(P)
    k = deb.k_grid({}, 1e-4, 1.0, nk, "log")
    cot = np.random.default_rng(0).standard_normal(nk)
    opts = dict(**LMAX, **SOLVER, aexp_out=(1.0,))
    out = {"k_modes": nk, "params": list(deb.PARAM9)}
    for method in ("adjoint", "fd"):
        deb.pk_and_vjp(theta, P, k, cot, method=method, **opts)  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(repeats):
            pk, g = deb.pk_and_vjp(theta, P, k, cot, method=method, **opts)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / repeats
        out[method] = {"gradients_per_s": 1.0 / dt, "ms_per_gradient": 1e3 * dt, "vjp": [float(v) for v in g.cpu().numpy()]}
    out["gradients_per_s"] = out["adjoint"]["gradients_per_s"]
    out["method"] = ("discrete adjoint: reverse Rodas5 sweep per mode over checkpointed states on the frozen accepted-step mesh; "
                     "background dependence through tables differenced over 12 shifted cosmologies; n_s, A_s closed form")
    return out


def measure_gradient_many(deb, n_cosmo, nk, n_plain=2):
    """P(k, z=0) and its 9-parameter Jacobian for n_cosmo cosmologies (theta within +-5% of the
    workload's): batched by the adjoint (one base launch + one adjoint launch for all), batched by
    differences (one base launch + one replay launch of the 12 n_cosmo shifted cosmologies), and
    the plain form -- one cosmology at a time, one launch per shifted cosmology -- on n_plain of
    them.  Rates are gradients (full Jacobians) per second."""
    import torch
    P = workload_param()
    th0 = deb.param_to_theta(P)
    rng = np.random.default_rng(1)
    scale = np.where(np.isin(np.arange(9), (3, 4, 8)), 0.0, 0.05)  # n_s, A_s, tau_reion held
    thetas = th0 * (1.0 + scale * (2.0 * rng.random((n_cosmo, 9)) - 1.0))
    k = deb.k_grid({}, 1e-4, 1.0, nk, "log")
    opts = dict(**LMAX, **SOLVER, aexp_out=(1.0,))
    res = {"cosmologies": n_cosmo, "k_modes": nk}
    Js = {}
    for method in ("adjoint", "fd"):
        torch.cuda.synchronize()
        t = time.perf_counter()
        pkm, Jm = deb.pk_jacobian_many(thetas, P, k, method=method, **opts)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        Js[method] = Jm
        res[method] = {"s": dt, "gradients_per_s": n_cosmo / dt}
    # plain form: per cosmology, the base solve and each of the 12 shifted replays in its own launch
    t = time.perf_counter()
    for c in range(n_plain):
        _plain_jacobian(deb, thetas[c], P, k, opts)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / n_plain
    res["plain"] = {"s_per_gradient": dt, "gradients_per_s": 1.0 / dt, "sample": n_plain,
                    "form": "one cosmology per call, base solve + 12 shifted-cosmology replays, one launch each"}
    res["speedup_adjoint_vs_plain"] = res["adjoint"]["gradients_per_s"] / res["plain"]["gradients_per_s"]
    cols = [0, 1, 2, 5, 6, 7]
    res["adjoint_vs_fd_median_rel"] = float((Js["adjoint"][..., cols] - Js["fd"][..., cols]).abs().median() /
                                            Js["fd"][..., cols].abs().max())
    return res


def _plain_jacobian(deb, theta, P, k, opts):
    """The unbatched difference gradient: one launch per shifted cosmology."""
    import torch
    aexp = np.atleast_1d(opts["aexp_out"])
    lm = {x: opts[x] for x in ("lmaxg", "lmaxgp", "lmaxr", "lmaxnu", "nqmax")}
    sv = {x: opts[x] for x in ("rtol", "atol")}
    S = deb.grad._setup(np.asarray(theta)[None, :], P, k, aexp, lm, sv, 256, deb.grad.REL_STEP)
    kt = S["uk0"]
    uc = torch.zeros(kt.numel(), dtype=torch.int32, device=kt.device)
    for r in range(12):
        tab = S["tab"][r:r + 1]
        tau = S["tau"][r:r + 1]
        deb.evolve_units(tab, 256, uc, kt, tau, **lm, **sv, replay=(S["mesh"], S["mlen"]), schedule={"order": "k_desc"})


def measure_los(deb, nk=3000, nt=2500, ellmax=2500, nfft=16384):
    """Line-of-sight projection at the CMB shape, and the contraction kernel alone."""
    import torch
    dev = torch.device("cuda", 0)
    P = dict(workload_param(), Neff=3.046, Nmnu=0.0, mnu=0.0, w_DE_0=-1.0, w_DE_a=0.0)
    p = deb.evolve_background(param=P, num_thermo=1024)
    aexp = np.geomspace(1e-4, 1.0, nt)
    tau = p["tau_of_a_spline"].evaluate(aexp)
    k = torch.as_tensor(np.geomspace(1e-5, 0.5, nk), device=dev)
    tv = tau.reshape(1, -1)
    S = torch.exp(-0.5 * ((tv - 280.0) / 20.0) ** 2) * torch.cos(k.reshape(-1, 1) * tv) * 1e-2  # smooth synthetic sources
    tau0 = float(p["tau0"])
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    deb.compute_theta_ell(ellmax, k, tau, S, tau0, n_fftlog=nfft)
    torch.cuda.synchronize()
    e0.record()
    th, kk = deb.compute_theta_ell(ellmax, k, tau, S, tau0, n_fftlog=nfft)
    e1.record()
    torch.cuda.synchronize()
    los_ms = e0.elapsed_time(e1)
    K = ((nfft + 2 + 15) // 16) * 16
    A = torch.randn((nk, K), dtype=torch.float64, device=dev)
    B = torch.randn((ellmax + 1, K), dtype=torch.float64, device=dev)
    C = torch.empty((nk, ellmax + 1), dtype=torch.float64, device=dev)
    L = deb._lib.lib()
    L.deb_contract(deb._lib.ptr(A), deb._lib.ptr(B), nk, ellmax + 1, K, deb._lib.ptr(C), deb._lib.stream())
    torch.cuda.synchronize()
    reps = 3
    e0.record()
    for _ in range(reps):
        deb._lib.check(L.deb_contract(deb._lib.ptr(A), deb._lib.ptr(B), nk, ellmax + 1, K, deb._lib.ptr(C), deb._lib.stream()))
    e1.record()
    torch.cuda.synchronize()
    cms = e0.elapsed_time(e1) / reps
    flop = 2.0 * nk * (ellmax + 1) * K
    tf = flop / (cms * 1e-3) / 1e12
    return {"shape": {"k_modes": nk, "n_tau": nt, "ellmax": ellmax, "n_fftlog": nfft}, "theta_ell_ms": los_ms,
            "contract": {"ms": cms, "flop": flop, "achieved_tflops": tf, "peak_tflops": FP64_PEAK_TFLOPS, "frac": tf / FP64_PEAK_TFLOPS,
                         "traffic": measured_traffic(f"contract_kernel {nk} x {ellmax + 1} x {K}")}}


def measure_cmb_chain(deb, nk=3000, ellmax=2500, nfft=16384, k_chunk=500):
    """The CMB workload end to end at its own sizes: flat LCDM with massless neutrinos, photon
    hierarchies to l = 24 (lmaxr 12), 3000 k on geomspace(1e-5, 0.5), 2500 output times on the
    recombination-weighted split, rtol = atol = 1e-4 (the C_l driver's tolerances), perturbations
    -> time derivatives -> sources, line of sight for l <= 2500 at n_fftlog 16384, C_l^TT.  The
    sources are formed in the solver's output epilogue ("fused": only the source planes are
    stored) or by the staged chain streamed over k in chunks of k_chunk modes (full states and
    derivatives of one chunk at a time); per-mode adaptive steps, and the reference's form:
    modes in batches on shared steps (8 per batch, since 3000 is not a multiple of the
    reference's 16).  Every form gives the same planes bit for bit on the same solver."""
    import torch
    dev = torch.device("cuda", 0)
    P = dict(workload_param(), Neff=3.046, Nmnu=0.0, mnu=0.0, w_DE_0=-1.0, w_DE_a=0.0)
    aexp = deb.cmb_aexp_out(250, 500, 1000, 750)
    lm = dict(lmaxg=24, lmaxgp=24, lmaxr=12, lmaxnu=8, nqmax=3)
    k = torch.as_tensor(np.geomspace(1e-5, 0.5, nk), dtype=torch.float64, device=dev)
    res = {"shape": {"k_modes": nk, "n_tau": int(aexp.size), "ellmax": ellmax, "n_fftlog": nfft, "nvar": 99, "k_chunk": k_chunk,
                     "rtol": 1e-4, "atol": 1e-4}}
    deb.evolve_background(param=dict(P), num_thermo=1024)  # warm-up: first use of the background kernel at this size
    for solver, form in (("per_mode", "fused"), ("per_mode", "streamed"), ("batched", "fused")):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        t = [time.perf_counter()]
        p = deb.evolve_background(param=dict(P), num_thermo=1024)
        torch.cuda.synchronize()
        t.append(time.perf_counter())
        if form == "fused":
            extra = dict(batch_size=8) if solver == "batched" else {}
            src = deb.compute_sources_fused(p, k, aexp, solver=solver, **extra, **lm)
        else:
            extra = dict(batch_size=8, k_chunk=504) if solver == "batched" else dict(k_chunk=k_chunk)
            src = deb.compute_sources_streamed(p, k, aexp, solver=solver, **extra, **lm)
        torch.cuda.synchronize()
        t.append(time.perf_counter())
        th, kk = deb.compute_theta_ell(ellmax, k, src["tau"], src["S"], float(p["tau0"]), n_fftlog=nfft)
        cl = deb.compute_Cell(th, kk, p["n_s"], p["k_p"])
        torch.cuda.synchronize()
        t.append(time.perf_counter())
        total = t[-1] - t[0]
        res[f"{solver}_{form}"] = {"s": total, "background_s": t[1] - t[0], "perturbations_sources_s": t[2] - t[1], "los_cell_s": t[3] - t[2],
                       "k_solves_per_s": nk / (t[2] - t[1]), "accepted_steps_mean": float(src["n_accept"].double().mean()),
                       "peak_device_bytes": int(torch.cuda.max_memory_allocated()), "cell_finite": bool(torch.isfinite(cl[2:]).all()),
                       **extra}
        del src, th, cl
    res["full_states_bytes"] = nk * int(aexp.size) * 99 * 8
    # the gradient of L = sum_{l=2}^{2500} w_l C_l^TT (w_l = default_rng(1).uniform(0, 1, 2499)) with
    # respect to the nine parameters: frozen-mesh differences of the chain (fused sources)
    w = np.random.default_rng(1).uniform(0.0, 1.0, ellmax - 1)
    theta = deb.param_to_theta(P)
    torch.cuda.synchronize()
    t = time.perf_counter()
    L, g = deb.cell_vjp(theta, P, k.cpu().numpy(), w, ellmax=ellmax, aexp_out=aexp, n_fftlog=nfft, **lm)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    res["cell_gradient"] = {"s": dt, "gradients_per_s": 1.0 / dt, "L": L, "grad": g.cpu().tolist(), "params": list(deb.PARAM9),
                            "weights": "default_rng(1).uniform(0, 1, 2499) for l = 2..2500",
                            "method": "frozen-mesh central differences of the whole chain (12 shifted cosmologies replayed with "
                                      "fused sources, line of sight, C_l); n_s exact; A_s and tau_reion do not enter C_l"}
    return res


def fisher_workload(n_cosmo=256):
    """The many-cosmology workload: n_cosmo parameter vectors theta = (omega_c, omega_b, h, n_s,
    ln 10^10 A_s, m_nu, w0, wa, tau_reion) from a Latin hypercube (scipy.stats.qmc, d = 9,
    seed 2) spanning +-10% of the single-cosmology workload's values, and per cosmology a log
    k-grid over 1e-4..2/Mpc of n_k = default_rng(3).integers(128, 769) modes."""
    from scipy.stats import qmc
    th0 = np.array([0.120, 0.0224, 0.674, 0.965, np.log(21.0), 0.06, -0.95, 0.1, 0.054])
    u = qmc.LatinHypercube(d=9, seed=2).random(n_cosmo)
    thetas = th0 + (2.0 * u - 1.0) * 0.1 * np.abs(th0)
    nks = np.random.default_rng(3).integers(128, 769, size=n_cosmo)
    return thetas, [np.geomspace(1e-4, 2.0, int(n)) for n in nks]


def _note(msg):
    """A progress line on stderr (the JSON result stays the one line on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def measure_fisher(deb, n_cosmo=256, n_latency=16, n_plain=3):
    """The many-cosmology workload end to end: P(k, z = 0) and its 9-column Jacobian for every
    cosmology (adaptive base solves of all units in one launch, the discrete adjoint of all units
    in k-ordered launches sized to their checkpoints), with the forward alone as a captured graph
    beside it; per-request latency (one cosmology per call) p50/p99 over a sample; the plain form
    (one cosmology per call, base solve and twelve shifted-cosmology replay launches) on a
    sample, scaled to the batch."""
    import torch
    P = workload_param()
    thetas, grids = fisher_workload(n_cosmo)
    units = int(sum(g.size for g in grids))
    res = {"cosmologies": n_cosmo, "units": units, "n_k_range": [int(min(g.size for g in grids)), int(max(g.size for g in grids))],
           "seeds": {"latin_hypercube": 2, "n_k": 3}, "truncation": "lmax 12, nq 8, lmax_nu 17", "rtol": 1e-5, "atol": 1e-10}
    # forward only, captured once and replayed
    params = [deb.theta_to_param(th, P) for th in thetas]
    g = deb.PkGraph(params, grids, (1.0,), schedule={"order": "k_desc", "threads": 128}, **LMAX, **SOLVER)
    torch.cuda.synchronize()
    t = time.perf_counter()
    pk, st = g.replay()
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t
    phantom = int(sum(bool((x == 4).any()) for x in g.split(st)))
    res["forward_graph"] = {"s": fwd, "k_solves_per_s": units / fwd, "phantom_cosmologies": phantom}
    _note(f"fisher: forward graph {fwd:.2f} s")
    del g
    # forward + Jacobian of every cosmology
    torch.cuda.synchronize()
    t = time.perf_counter()
    pks, Js, status = deb.pk_jacobian_many(thetas, P, grids, aexp_out=(1.0,), return_status=True, method="adjoint", **LMAX, **SOLVER)
    torch.cuda.synchronize()
    tot = time.perf_counter() - t
    valid = [c for c in range(n_cosmo) if status[c] == 0]
    res["jacobian"] = {"s": tot, "gradients_per_s": n_cosmo / tot, "k_modes_per_s": units / tot,
                       "status_counts": {str(k): int(v) for k, v in zip(*np.unique(status, return_counts=True))},
                       "phantom_cosmologies": int((status == 4).sum()), "finite": bool(all(torch.isfinite(Js[c]).all() for c in valid))}
    _note(f"fisher: forward + Jacobian {tot:.2f} s")
    # per-request latency: one cosmology per call
    lat = []
    for c in valid[:n_latency]:
        torch.cuda.synchronize()
        t = time.perf_counter()
        deb.pk_jacobian(thetas[c], P, grids[c], aexp_out=(1.0,), **LMAX, **SOLVER)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
    res["latency_one_per_call"] = {"sample": len(lat), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                                   "n_k": [int(grids[c].size) for c in valid[:n_latency]]}
    _note("fisher: latency sample done")
    # the plain form on a sample, scaled by modes
    opts = dict(aexp_out=(1.0,), **LMAX, **SOLVER)
    tp, kp = 0.0, 0
    for c in valid[:n_plain]:
        torch.cuda.synchronize()
        t = time.perf_counter()
        _plain_jacobian(deb, thetas[c], P, grids[c], opts)
        torch.cuda.synchronize()
        tp += time.perf_counter() - t
        kp += grids[c].size
    est = tp / kp * sum(grids[c].size for c in valid)
    res["plain"] = {"sample": n_plain, "s_sample": tp, "s_estimated_for_batch": est,
                    "form": "one cosmology per call, base solve + 12 shifted-cosmology replays, one launch each (difference Jacobian)"}
    res["speedup_vs_plain"] = est / tot
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=256, help="k-modes timed on the CPU oracle, all cores (0: skip)")
    ap.add_argument("--cpu-sample-1core", type=int, default=24, help="k-modes timed on the CPU oracle on one core")
    ap.add_argument("--grad-modes", type=int, default=256, help="k-modes of the gradient measurement (0: skip)")
    ap.add_argument("--grad-cosmologies", type=int, default=16,
                    help="cosmologies of the batched gradient measurement, at --grad-modes k-modes (0: skip)")
    ap.add_argument("--los", type=int, default=1, help="measure the line-of-sight projection (0: skip)")
    ap.add_argument("--cmb-chain", type=int, default=1, help="measure the CMB chain at 3000 k x 2500 tau, l <= 2500 (0: skip)")
    ap.add_argument("--fisher", type=int, default=0,
                    help="also measure the 256-cosmology forward + Jacobian workload (about two minutes; 0: skip)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (a known-size kernel for scaling HBM counters)")
    args = ap.parse_args()
    rank = int(os.environ.get("RANK", "0"))
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        if rank == 0:
            print("bench.py measures one GPU only: there is no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    deb = _load()
    torch.cuda.set_device(0)
    P = workload_param()
    kmodes = deb.k_grid({}, 1e-5, 5.0, NK, "log")
    dev = torch.device("cuda", 0)
    kt = torch.as_tensor(kmodes, dtype=torch.float64, device=dev)
    uc = torch.zeros(NK, dtype=torch.int32, device=dev)
    aout = torch.tensor(AOUT, dtype=torch.float64, device=dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kernel_ms = []

    def one_step(record):
        tables = deb.background_tables([P], 256)
        p = deb.attach_background(dict(P), tables, 256, 0)
        tau = p["tau_of_a_spline"].evaluate(aout).reshape(1, -1)
        if record:
            ev0.record()
        y, st, na, nr = deb.evolve_units(tables, 256, uc, kt, tau, **LMAX, **SOLVER, schedule=deb.perturbations.DEFAULT_SCHEDULE)
        if record:
            ev1.record()
        pk = deb.get_power(k=kt, y=y, idx=4, param=P)
        return st, na, nr, pk

    for _ in range(args.warmup):
        st, na, nr, pk = one_step(False)
        torch.cuda.synchronize()
        if (st != 0).any():
            raise RuntimeError(f"solver failures in warm-up: status {st.cpu().numpy().tolist()}")
    step_ms = []
    counts = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        st, na, nr, pk = one_step(True)
        torch.cuda.synchronize()
        step_ms.append(1e3 * (time.perf_counter() - ts))
        kernel_ms.append(ev0.elapsed_time(ev1))
        counts.append((na, nr))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # step counts after the timed region (a first reduction would load its kernel inside it)
    steps_attempted = [int((na + nr).sum().item()) for na, nr in counts]
    retries = [int(nr.sum().item()) for na, nr in counts]  # each rejection is retried from its own point
    ok = bool((st == 0).all())
    ms_per_step = 1e3 * wall / args.steps
    value = NK * args.steps / wall
    W, n, m, Wj = work_per_step(**LMAX)
    flop = float(np.mean(steps_attempted)) * W - float(np.mean(retries)) * Wj
    kmean = float(np.mean(kernel_ms))
    achieved = flop / (kmean * 1e-3) / 1e12
    out = {
        "metric": "k-mode ODE solves/s (FP64, rtol 1e-5); P(k)+9-param adjoint gradients/s",
        "value": value, "unit": "k-mode solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "512 k-modes log 1e-5..5/Mpc, LCDM + 1 massive nu (0.06 eV, nq 8, lmax_nu 17) + w0wa "
                               "(-0.95, 0.1), lmax 12, rtol 1e-5 atol 1e-10, P(k) at z=2,1,0",
                   "k_modes": NK, "nvar": n, "dense_block": m, "num_thermo": 256, "solver": "Rodas5, per-mode adaptive",
                   "launch": {"order": "largest k first", "threads_per_unit": deb._lib.lib().deb_default_threads_per_unit()},
                   "all_modes_ok": ok, "gradients_per_s": None, "step_ms": [round(v, 2) for v in step_ms]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic("perturb_kernel<8, 256> 512 k-modes nq 8 lmax_nu 17 rtol 1e-5"),
                     "kernel": "perturb_kernel", "kernel_ms": kmean, "attempted_steps_per_launch": float(np.mean(steps_attempted)),
                     "retries_per_launch": float(np.mean(retries)), "flop_per_step": W, "flop_skipped_per_retry": Wj},
        "cpu_baseline": None,
    }
    if args.copy_calibration_mib > 0:
        src = torch.ones(args.copy_calibration_mib * (1 << 20) // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = src.numel() * 8
    _note(f"k-mode solves: {value:.0f}/s")
    if args.grad_modes > 0:
        out["gradient"] = measure_gradient(deb, args.grad_modes)
        out["config"]["gradients_per_s"] = out["gradient"]["gradients_per_s"]
        out["config"]["gradient_method"] = "adjoint"
        _note("gradient done")
        if args.grad_cosmologies > 0:
            out["gradient_many"] = measure_gradient_many(deb, args.grad_cosmologies, args.grad_modes)
            _note("gradient of many cosmologies done")
    if args.los:
        out["line_of_sight"] = measure_los(deb)
        _note("line of sight done")
    if args.cmb_chain:
        out["cmb_chain"] = measure_cmb_chain(deb)
        _note("CMB chain done")
    if args.fisher:
        out["fisher"] = measure_fisher(deb)
    if args.cpu_sample > 0 and rank == 0:
        out["cpu_baseline"] = cpu_baseline(kmodes, args.cpu_sample, args.cpu_sample_1core)
        _note("CPU baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
