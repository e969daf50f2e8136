"""Host-side logic of the gradient path (no GPU): extraction of the accepted-step mesh."""
import numpy as np
import torch

from conftest import load_package


def _mesh_by_loop(log, nacc, t1):
    """Accepted step ends of each unit as fractions of its range, written out step by step."""
    n_units = log.shape[0]
    L = int(nacc.max())
    mesh = np.ones((n_units, L))
    for u in range(n_units):
        ends = [r[1] for r in log[u] if np.isfinite(r[2]) and r[2] < 1.0]
        t0 = log[u, 0, 0]
        s = [(e - t0) / (t1[u] - t0) for e in ends]
        s[-1] = 1.0
        mesh[u, :len(s)] = s
    return mesh


def test_accepted_mesh_keeps_accepted_ends_only():
    deb = load_package()
    rng = np.random.default_rng(3)
    n_units, max_steps = 5, 40
    log = np.full((n_units, max_steps, 3), np.nan)
    nacc = np.zeros(n_units, dtype=np.int32)
    t1 = np.zeros(n_units)
    for u in range(n_units):
        t, n_att = 0.1 * (u + 1), int(rng.integers(3, max_steps + 1))
        for a in range(n_att):
            dt = rng.uniform(0.01, 0.2)
            en = rng.uniform(0.0, 2.0)  # > 1: rejected, the next attempt starts at the same t
            log[u, a] = (t, t + dt, en)
            if en < 1.0:
                t += dt
                nacc[u] += 1
        if nacc[u] == 0:  # every unit ends with an accepted step
            log[u, n_att - 1, 2] = 0.5
            t = log[u, n_att - 1, 1]
            nacc[u] = 1
        t1[u] = t
    mesh, mlen = deb.grad._accepted_mesh(torch.tensor(log), torch.tensor(nacc), t1)
    assert np.array_equal(mlen.numpy(), nacc)
    np.testing.assert_array_equal(mesh.numpy(), _mesh_by_loop(log, nacc, t1))
    for u in range(n_units):
        s = mesh[u, :nacc[u]].numpy()
        assert s[-1] == 1.0 and np.all(np.diff(s) > 0) and s[0] > 0
