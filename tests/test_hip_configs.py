"""The five benchmark configurations at their own sizes on the device.

* Config 2: the chi=1024 bulk matvec of the XXZ chain N=100 against the oracle (float64 and the
  float32 operator), and the 64 -> 512 -> 1024 bond ramp with the mixed solver against the
  float64 one at every stage.
* Config 3: the chi=2048 U(1) block matvec over 31 sectors against the oracle's dense
  contraction of the densified operands, and the local ground energy mixed vs float64.
* Config 4: 32 TFIM quenches N=64 with per-leg caps 64..256 and svd_tol=0, run until the caps
  bind (2 steps of dt=0.05), against the exact free-fermion <Z_j>(t) and against single-leg tdvp2_evolve.
* Config 5: dE/dg_i of the per-site-field TFIM N=64 at chi=256 against -<Z_i> and against
  central differences of the exact free-fermion energy.
The numpy oracle runs with up to 16 BLAS threads here (threadpoolctl)."""
import sys

import numpy as np
import pytest
import torch
from threadpoolctl import threadpool_limits

import bench
from oracle import blocks as OB
from oracle import dense as D
from oracle import exact as X
from tnkit_amd import block_ops as BO
from tnkit_amd import kernels
from tnkit_amd.eigensolver import lanczos_ground_mixed, lanczos_ground_restart

pytestmark = pytest.mark.gpu

CPU_THREADS = 16


def _say(*a):
    # straight to the process's stderr: long tests report progress past pytest's capture
    print(*a, file=sys.__stderr__, flush=True)


def _rel(got, ref):
    return float(np.abs(got - ref).max() / np.abs(ref).max())


# ---------------------------------------------------------------------------------------------
# config 2
# ---------------------------------------------------------------------------------------------


def test_config2_matvec_chi1024_vs_oracle(hip):
    """XXZ N=100, bulk bond chi=1024 (D=5, d=2): float64 to 1e-12, the float32 operator with
    float64 vectors to 2e-5 (max-normalised)."""
    _say("config2 matvec chi=1024 N=100")
    pr = bench.build_local_problem(100, 1024)
    i0 = pr["i0"]
    L, R, th = pr["L"], pr["R"], pr["theta"]
    W1, W2 = pr["mpo"][i0], pr["mpo"][i0 + 1]
    assert tuple(th.shape) == (1024, 2, 2, 1024) and tuple(L.shape) == (1024, 5, 1024)
    with threadpool_limits(CPU_THREADS):
        ref = D.apply_two_site_fused(th.cpu().numpy(),
                                     *D.fused_two_site(L.cpu().numpy(), W1.cpu().numpy(),
                                                       W2.cpu().numpy(), R.cpu().numpy()))
    y64 = kernels.apply_two_site_fused(th, *kernels.fused_two_site(L, W1, W2, R))
    y32 = kernels.apply_two_site_fused(th, *kernels.fused_two_site(L, W1, W2, R,
                                                                   dtype=torch.float32))
    e64, e32 = _rel(y64.cpu().numpy(), ref), _rel(y32.cpu().numpy(), ref)
    _say(f"  rel err f64 {e64:.2e}, f32 operator {e32:.2e}")
    assert e64 < 1e-12 and e32 < 2e-5


RAMP = [(64, 4), (512, 4), (1024, 2)]
_RAMP_FP64 = {}


def _ramp(eig):
    _say(f"config2 ramp N=100 {RAMP} eig={eig}")
    out = bench.config2(100, RAMP, eig, "gram")
    for r in out:
        _say(f"  chi={r['chi']} E={r['E']:.12f} sweeps {[round(s, 2) for s in r['sweep_s']]} "
             f"matvecs {r['matvec_calls']}")
    return out


def test_config2_ramp_fp64(hip):
    """The float64 ramp: bonds reach each stage's chi and the energy does not rise from stage to
    stage (by more than 1e-10: at chi=512 it is already converged to ~1e-12)."""
    out = _ramp("lanczos-restart")
    assert [r["max_bond"] for r in out] == [c for c, _ in RAMP]
    es = [r["E"] for r in out]
    assert es[1] < es[0] and es[2] <= es[1] + 1e-10
    _RAMP_FP64["E"] = es


def test_config2_ramp_mixed_matches_fp64(hip):
    """The mixed solver (float32 operator to 1e-4, then float64) ends every stage within 1e-9 of
    the float64 run's energy."""
    if "E" not in _RAMP_FP64:
        pytest.skip("needs test_config2_ramp_fp64 in the same session")
    out = _ramp("lanczos-mixed")
    assert all(r["mixed_counts"]["lo"] > 0 for r in out)
    for r, e64 in zip(out, _RAMP_FP64["E"]):
        _say(f"  chi={r['chi']} |E_mixed - E_fp64| = {abs(r['E'] - e64):.2e}")
        assert abs(r["E"] - e64) <= 1e-9


# ---------------------------------------------------------------------------------------------
# config 3
# ---------------------------------------------------------------------------------------------


def test_config3_block_matvec_chi2048_vs_dense_oracle(hip):
    """U(1) XXZ, bulk bond chi=2048 over 31 Gaussian sectors: the block matvec (all sector GEMMs
    of each contraction in one grouped launch) equals the oracle's dense contraction of the
    densified operands to 1e-12 (float64) and 5e-6 (float32 operator); the local ground energy
    of the mixed solver is within 1e-9 of the float64 one."""
    _say("config3 block matvec chi=2048")
    pr = bench.build_block_problem(2048)
    L, R, W, th = pr["L"], pr["R"], pr["W"], pr["theta"]
    assert len(pr["sectors"]) == 31 and sum(pr["sectors"].values()) == 2048
    op64 = BO.BlockTwoSite(L, W, W, R)
    op32 = BO.BlockTwoSite(L, W, W, R, dtype=torch.float32)
    y64, y32 = op64.apply(th.data), op32.apply(th.data)
    T = lambda x: tuple((g.charges, g.dims, g.sign) for g in x.legs)  # noqa: E731
    dn = lambda x: OB.canonical_to_dense(T(x), x.data.cpu().numpy())  # noqa: E731
    with threadpool_limits(CPU_THREADS):
        Wd = dn(W)
        yd = D.apply_two_site_fused(dn(th), *D.fused_two_site(dn(L), Wd, Wd, dn(R)))
    ref = OB.dense_to_canonical(yd, T(th))
    e64, e32 = _rel(y64.cpu().numpy(), ref), _rel(y32.cpu().numpy(), ref)
    _say(f"  rel err f64 {e64:.2e}, f32 operator {e32:.2e}")
    assert e64 < 1e-12 and e32 < 5e-6
    e_hi, _, _ = lanczos_ground_restart(op64.apply, th.data.clone(), tol=1e-9, max_iter=8192)
    cnt = {"lo": 0, "hi": 0}
    e_mx, _, _ = lanczos_ground_mixed(op32.apply, op64.apply, th.data.clone(), tol=1e-9,
                                      max_iter=8192, counts=cnt)
    _say(f"  local E f64 {e_hi:.14f} mixed {e_mx:.14f} (lo {cnt['lo']}, hi {cnt['hi']})")
    assert abs(e_mx - e_hi) <= 1e-9 and cnt["lo"] > 0


# ---------------------------------------------------------------------------------------------
# config 4
# ---------------------------------------------------------------------------------------------

# bonds at most double per half-sweep from chi=16, so every cap (<= 256) binds by step 2
C4_LEGS, C4_N, C4_DT, C4_STEPS = 32, 64, 0.05, 2


@pytest.fixture(scope="module")
def quench_states(hip):
    """32 ground states at chi=16 for g0 ~ U[1.2, 2.0] (default_rng(1)), complex128."""
    from tnkit_amd import _dmrg as H
    from tnkit_amd.models import tfim_mpo
    _say("config4 ground states")
    g0 = np.random.default_rng(1).uniform(1.2, 2.0, C4_LEGS)
    states = []
    for j, g in enumerate(g0):
        _, mps, _ = H.dmrg(tfim_mpo(C4_N, float(g)), 16, n_sweeps=4, tol=1e-12, rtol=0.0,
                           seed=j, backend="hip-real")
        states.append([t.to(torch.complex128) for t in mps])
    caps = [int(c) for c in np.linspace(64, 256, C4_LEGS).round()]
    return g0, states, caps


_C4 = {}


def _quench(states, caps, low):
    from tnkit_amd import tdvp as T
    from tnkit_amd.models import tfim_mpo
    from tnkit_amd.observables import local_expectations
    Z = np.diag([1.0, -1.0])
    stats = {}
    _, rec = T.evolve_batch(states, tfim_mpo(C4_N, 1.0), C4_DT, C4_STEPS, chi_max=caps,
                            krylov_m=20, keep_device=True, svd_tol=0.0, stats=stats,
                            measure=lambda ss: [local_expectations(s, Z) for s in ss],
                            meas_every=1, low_precision=low,
                            progress=lambda k, t: _say(f"  step {k} t={t:.2f}"))
    return rec, stats


def test_config4_batched_quench_caps_bind(hip, quench_states):
    """TDVP2 dt=0.05 with svd_tol=0: every leg's widest bond reaches its cap within the run
    (caps 64..256 are uneven), and <Z_j>(t) of every leg follows the exact free-fermion curve
    to 2e-5 at every step."""
    from tnkit_amd import free_fermion as F
    g0, states, caps = quench_states
    _say(f"config4 batch c128 caps {caps[0]}..{caps[-1]}, {C4_STEPS} steps, svd_tol=0")
    rec, stats = _quench(states, caps, False)
    widest = [max(bond[j] for bond in stats["bonds"]) for j in range(C4_LEGS)]
    _say(f"  widest bond per leg {widest}")
    assert widest == caps
    times = [k * C4_DT for k in range(C4_STEPS + 1)]
    err = 0.0
    for j in range(C4_LEGS):
        ff = F.quench_z(C4_N, float(g0[j]), 1.0, times)
        for k, (_, zs) in enumerate(rec):
            err = max(err, float(np.abs(np.asarray(zs[j]) - ff[k]).max()))
    _say(f"  max |<Z> - exact| = {err:.2e}")
    assert err < 2e-5
    _C4["rec"] = rec


def test_config4_legs_equal_sequential_tdvp2(hip, quench_states):
    """Two legs (the smallest and the largest cap) evolved alone by tdvp2_evolve give the batched
    run's <Z_j>(t) to 1e-9."""
    from tnkit_amd import tdvp as T
    from tnkit_amd.models import tfim_mpo
    from tnkit_amd.observables import local_expectations
    if "rec" not in _C4:
        pytest.skip("needs test_config4_batched_quench_caps_bind in the same session")
    _, states, caps = quench_states
    Z = np.diag([1.0, -1.0])
    for j in (0, C4_LEGS - 1):
        _say(f"config4 sequential leg {j} cap {caps[j]}")
        _, r1 = T.tdvp2_evolve(states[j], tfim_mpo(C4_N, 1.0), caps[j], C4_DT, C4_STEPS,
                               krylov_m=20, measure=lambda m: local_expectations(m, Z),
                               meas_every=1, svd_tol=0.0)
        d = max(float(np.abs(np.asarray(_C4["rec"][k][1][j]) - np.asarray(r1[k][1])).max())
                for k in range(len(r1)))
        _say(f"  max |<Z> batch - sequential| = {d:.2e}")
        assert d < 1e-9


def test_config4_complex64_operator(hip, quench_states):
    """The complex64 operator with complex128 Krylov vectors: within 1e-5 of the complex128
    batch at every step."""
    if "rec" not in _C4:
        pytest.skip("needs test_config4_batched_quench_caps_bind in the same session")
    _, states, caps = quench_states
    _say("config4 batch c64 operator")
    rec, _ = _quench(states, caps, True)
    d = max(float(np.abs(np.asarray(rec[k][1][j]) - np.asarray(_C4["rec"][k][1][j])).max())
            for k in range(len(rec)) for j in range(C4_LEGS))
    _say(f"  max |<Z> c64 - c128| = {d:.2e}")
    assert d < 1e-5


# ---------------------------------------------------------------------------------------------
# config 5
# ---------------------------------------------------------------------------------------------


def test_config5_field_gradient_n64_chi256(hip):
    """TFIM N=64, g_i = 1 + 0.1 N(0,1) (default_rng(2)), two-site DMRG chi=256 float64: dE/dg_i
    from the backward kernels equals -<Z_i> and the central differences (h = 1e-5) of the exact
    free-fermion energy at every site, to 1e-7; the energy is the free-fermion one to 1e-9."""
    from tnkit_amd import _dmrg as H
    from tnkit_amd.autodiff import field_gradient
    from tnkit_amd.models import tfim_mpo_fields
    from tnkit_amd.observables import local_expectations
    N, h = 64, 1e-5
    _say("config5 N=64 chi=256")
    g = 1.0 + 0.1 * np.random.default_rng(2).standard_normal(N)
    E, mps, info = H.dmrg(tfim_mpo_fields(g), 256, n_sweeps=6, tol=1e-12, rtol=0.0, seed=0,
                          backend="hip-real", trunc="gram")
    e_exact = X.tfim_free_fermion_energy(N, g)
    _say(f"  E {E:.13f} exact {e_exact:.13f} sweeps {info['sweeps']}")
    assert abs(E - e_exact) < 1e-9
    dg = field_gradient(list(mps), torch.as_tensor(g)).cpu().numpy()
    z = local_expectations(list(mps), np.diag([1.0, -1.0]))
    fd = np.empty(N)
    for i in range(N):
        e = np.zeros(N)
        e[i] = h
        fd[i] = (X.tfim_free_fermion_energy(N, g + e)
                 - X.tfim_free_fermion_energy(N, g - e)) / (2 * h)
    _say(f"  max |dg + <Z>| {np.abs(dg + z).max():.2e}, max |dg - fd| {np.abs(dg - fd).max():.2e}")
    assert np.abs(dg + z).max() < 1e-7 and np.abs(dg - fd).max() < 1e-7
