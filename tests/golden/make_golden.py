"""Extract the reference's own golden data files into compact .npz fixtures.

Reads data files held by the reference's test-suite (CSV / JSON numeric tables only) from
/root/reference and writes them, unchanged in value, to tests/golden/*.npz.  The fixtures
are data: inputs and expected outputs.  Run once in the build container:

    python tests/golden/make_golden.py
"""
import json
import os

import numpy as np
import pandas as pd

REF = "/root/reference/tests"
HERE = os.path.dirname(os.path.abspath(__file__))

# cosmology of tests/test_cmb.py:72-90 ("DISCO-Notebook")
DISCO_NOTEBOOK = dict(Omegam=0.3099, Omegab=0.0488911, H0=67.742, n_s=0.96822, A_s=2.1064e-09, mnu=0.06, Tcmb=2.7255,
                      YHe=0.248, Neff=2.046, Nmnu=1.0, w_DE_0=-0.99, w_DE_a=0.0, cs2_DE=1.0, Omegak=0.0, k_p=0.05)


def _cosmo_arrays():
    keys = sorted(DISCO_NOTEBOOK)
    return np.array(keys), np.array([DISCO_NOTEBOOK[k] for k in keys])


def main():
    keys, vals = _cosmo_arrays()

    # tests/test_cmb/test_benchmark_evolve_background.csv (test_cmb.py:244-276, num_thermo=1024);
    # the xe column holds exp(xe_of_loga(log a)).
    df = pd.read_csv(f"{REF}/test_cmb/test_benchmark_evolve_background.csv")
    np.savez_compressed(f"{HERE}/evolve_background.npz", a=np.array([0.001, 0.01, 0.1, 0.5, 1.0]),
                        tau=df["tau_at_test_points"].values, exp_xe=df["xe_at_test_points"].values,
                        grhom=df["grhom"].values[0], num_thermo=1024, cosmo_keys=keys, cosmo_vals=vals)

    # tests/test_cmb/test_benchmark_evolve_perturbations_batched.csv (test_cmb.py:311-369):
    # full state of modes 0, 31, 63 of 64, 320 outputs x 72 variables.
    df = pd.read_csv(f"{REF}/test_cmb/test_benchmark_evolve_perturbations_batched.csv")
    np.savez_compressed(f"{HERE}/evolve_perturbations_batched.npz", kmodes=df["kmodes"].dropna().values,
                        y_k0=df["yout_k0_flat"].values.reshape(320, 72), y_k31=df["yout_k31_flat"].values.reshape(320, 72),
                        y_k63=df["yout_k63_flat"].values.reshape(320, 72), cosmo_keys=keys, cosmo_vals=vals)

    # tests/test_cmb/test_benchmark_compute_background_quantities.csv (rhonu, rho_Q, w_Q at 320 a)
    df = pd.read_csv(f"{REF}/test_cmb/test_benchmark_compute_background_quantities.csv")
    np.savez_compressed(f"{HERE}/background_quantities.npz", rhonu=df["rhonu"].values, rho_Q=df["rho_Q"].values,
                        w_Q=df["w_Q"].values)

    # tests/resources/CLASS_data.json: CLASS P_bc(k) at a = 0.01 (tests/test_perturbations.py)
    d = json.load(open(f"{REF}/resources/CLASS_data.json"))
    np.savez_compressed(f"{HERE}/class_pk.npz", k=np.array(d["k"]), Pkbc=np.array(d["Pkbc"]))

    # tests/resources/RECFAST_DISCO_EB_data.json: x_e(a) of DISCO-EB v0.1.0 (tests/test_background.py)
    d = json.load(open(f"{REF}/resources/RECFAST_DISCO_EB_data.json"))
    np.savez_compressed(f"{HERE}/recfast_v010.npz", a=np.array(d["a"]), xe=np.array(d["xe"]))

    # tests/test_cmb/test_benchmark_compute_Dell.csv: D_ell for C_ell = 1 (closed form)
    df = pd.read_csv(f"{REF}/test_cmb/test_benchmark_compute_Dell.csv")
    np.savez_compressed(f"{HERE}/compute_dell.npz", ell=df["ell"].values, Dell=df["Dell"].values)

    # CMB chain fixtures of tests/test_cmb.py:93-240 (DISCO-Notebook, num_thermo=1024, output
    # a = geomspace(3e-4, 5e-3, 256, endpoint=False) ++ geomspace(5e-3, 1, 64)).  The
    # visibility functions depend on the background only; the others come from the reference's
    # batched perturbation run (128 camb-sampled k, rtol = atol = 1e-4), sampled at k0/k63/k127.
    df = pd.read_csv(f"{REF}/test_cmb/test_benchmark_compute_visibility_functions.csv")
    np.savez_compressed(f"{HERE}/visibility_functions.npz", gvis=df["gvis"].values, optical_depth=df["optical_depth"].values,
                        gvisprime=df["gvisprime"].values)
    out = {}
    for name in ["compute_theta_ell", "compute_Cell", "compute_source_term_sachs_wolfe", "compute_source_term_doppler",
                 "compute_source_term_polarization", "compute_metric_perturbations", "compute_polarization_terms",
                 "compute_neutrino_perturbations", "extract_perturbations"]:
        df = pd.read_csv(f"{REF}/test_cmb/test_benchmark_{name}.csv")
        for col in df.columns[1:]:
            out[col] = df[col].values
    np.savez_compressed(f"{HERE}/cmb_chain.npz", **out)

    # CAMB / CLASS D_ell^TT for DISCO-Notebook (tests/test_cmb.py:1410-1472, 1764-1859) and the
    # reference's own per-ell relative differences to them
    for code in ("camb", "class"):
        df = pd.read_csv(f"{REF}/resources/{code}_benchmarks/DISCO-Notebook.csv")
        rel = pd.read_csv(f"{REF}/test_cmb/test_cmb_vs_{code}_DISCO_Notebook_.csv")["relative_error"].values
        np.savez_compressed(f"{HERE}/{code}_dell.npz", ref_relative_error=rel, **{c: df[c].values for c in df.columns})
    print("written:", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
