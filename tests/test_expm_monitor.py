"""The host stop rule of the adaptive Krylov exponential (`eigensolver.ExpmMonitor`) driven by a
plain numpy Lanczos loop (full reorthogonalisation), against tnkit's recorded runs: the same
Krylov dimension, cap flag, frozen estimate and output vector."""
import json
import os

import numpy as np
import pytest

from tnkit_amd.eigensolver import ExpmMonitor

from conftest import GOLDEN

with open(os.path.join(GOLDEN, "golden_expm.json")) as f:
    CASES = json.load(f)["lanczos_expm"]


def _case(c):
    if c.get("diag"):
        v0 = np.zeros(40, dtype=np.complex128)
        v0[[2, 9, 30]] = [1.0, 2.0, -1.0]
        return np.diag(np.arange(1.0, 41.0)), v0
    rng = np.random.default_rng(c["seed"])
    A = rng.standard_normal((c["n"], c["n"]))
    A = 0.5 * (A + A.T)
    return A, rng.standard_normal(c["n"]) + 1j * rng.standard_normal(c["n"])


def _run(A, v, tau, m, tol, adaptive, info):
    beta0 = float(np.linalg.norm(v))
    mon = ExpmMonitor(tau, beta0, tol, adaptive, info, 2)
    V = [v / beta0]
    alphas, betas, capped = [], [], False
    for k in range(m):
        w = A @ V[-1] - (betas[-1] * V[-2] if betas else 0)
        alphas.append(float(np.vdot(V[-1], w).real))
        w = w - alphas[-1] * V[-1]
        for _ in range(2):
            for u in V:
                w = w - np.vdot(u, w) * u
        beta = float(np.linalg.norm(w))
        stop = mon.check(k, alphas, betas, beta)
        if stop == "breakdown":
            break
        betas.append(beta)
        V.append(w / beta)
        if stop:
            break
    else:
        capped = mon.adaptive
    c = mon.finish(alphas, betas, capped)
    return sum(ci * vi for ci, vi in zip(c, V))


@pytest.mark.parametrize("c", CASES, ids=lambda c: f"dt{c['dt']}-m{c['m']}")
def test_monitor_reproduces_tnkit_runs(c):
    A, v0 = _case(c)
    info = {}
    out = _run(A, v0, -1j * c["dt"], c["m"], c["tol"], c["adaptive"], info)
    assert info["krylov_dim"] == c["krylov_dim"] and info["capped"] == c["capped"]
    assert info["est_ref_j"] == c["est_ref_j"]
    er = float.fromhex(c["est_ref"])
    assert abs(info["est_ref"] - er) <= 1e-8 * abs(er) + 1e-14
    ref = np.array([float.fromhex(x) for x in c["out_re"]]) + \
        1j * np.array([float.fromhex(x) for x in c["out_im"]])
    assert np.abs(out - ref).max() < 1e-10


def test_monitor_breakdown_and_environment_switch(monkeypatch):
    """An eigenvector breaks the space down at j = 1: est_ref is 0 there and the output is the
    exact exp(tau lam) v; TNKIT_EXPM_ADAPTIVE=0 turns the adaptive stop off."""
    A = np.diag([1.0, 2.0, 3.0])
    v = np.array([0.0, 1.0, 0.0], dtype=np.complex128)
    info = {}
    out = _run(A, v, -0.3j, 3, 1e-12, True, info)
    assert info["krylov_dim"] == 1 and info["est_ref"] == 0.0 and info["est_ref_j"] == 1
    assert np.allclose(out, np.exp(-0.6j) * v, atol=1e-14)
    monkeypatch.setenv("TNKIT_EXPM_ADAPTIVE", "0")
    assert ExpmMonitor(-0.1j, 1.0, 1e-12, None, None, 2).adaptive is False
    monkeypatch.setenv("TNKIT_EXPM_ADAPTIVE", "1")
    assert ExpmMonitor(-0.1j, 1.0, 1e-12, None, None, 2).adaptive is True
