"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle after a dense cross-check).

CPU tests: the oracle still reproduces every fixture, and the package's own
pulse sampling (emulators_amd.workloads / emu_base.sequence_data_from_samples)
produces the fixture's per-step tables.  GPU tests: the HIP path against the
fixtures through the C ABI.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import emu_sv_oracle as orc

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _c(x) -> torch.Tensor:
    return torch.as_tensor(x).to(torch.complex128)


@pytest.mark.parametrize("tag", ["a", "b"])
def test_oracle_reproduces_hpsi_fixture(tag):
    d = load("hpsi_random")
    ham = orc.OracleHamiltonian(_c(d[f"{tag}_omega"]), _c(d[f"{tag}_delta"]), _c(d[f"{tag}_phi"]),
                                torch.as_tensor(d[f"{tag}_u"]))
    got = ham * torch.as_tensor(d[f"{tag}_psi"])
    assert torch.allclose(got, torch.as_tensor(d[f"{tag}_hpsi"]), atol=1e-11, rtol=0)


def test_oracle_reproduces_krylov_fixture():
    d = load("krylov_step")
    ham = orc.OracleHamiltonian(_c(d["omega"]), _c(d["delta"]), _c(d["phi"]), torch.as_tensor(d["u"]))
    dt, tol = float(d["dt"]), float(d["tol"])
    r = orc.krylov_exp_impl(lambda x: -1j * dt * (ham * x), torch.as_tensor(d["psi"]).clone(),
                            is_hermitian=True, exp_tolerance=tol, norm_tolerance=tol)
    assert r.converged and r.iteration_count == int(d["iterations"])
    assert torch.allclose(r.result, torch.as_tensor(d["result"]), atol=tol, rtol=0)


def _chain10_tables(d):
    n = d["u"].shape[0]
    om = _c(d["omega_k"])[:, None].repeat(1, n)
    de = _c(d["delta_k"])[:, None].repeat(1, n)
    return om, de, torch.zeros_like(om), torch.as_tensor(d["u"]).reshape(1, n, n)


def test_oracle_reproduces_chain10_fixture():
    d = load("chain10")
    om, de, ph, ut = _chain10_tables(d)
    run = orc.run_sequence(om, de, ph, lambda t: ut, list(d["target_times"]))
    k = om.shape[0]
    assert run.iterations == list(d["iterations"])
    assert torch.allclose(run.states[k], torch.as_tensor(d["final_state"]), atol=1e-12, rtol=0)
    assert torch.allclose(run.occupation[k], torch.as_tensor(d["occupation"]), atol=1e-12, rtol=0)
    assert abs(run.energy[k] - float(d["energy"])) < 1e-10


def test_package_sampling_matches_chain10_tables():
    """The package's pulse sampling (PCHIP midpoints, C6/r^6) gives the
    fixture's per-step tables and interaction matrix."""
    from emulators_amd import workloads

    d = load("chain10")
    seq = workloads.chain10()
    assert list(seq.target_times) == list(d["target_times"])
    np.testing.assert_allclose(seq.omega.real.numpy(), np.repeat(d["omega_k"][:, None], 10, 1),
                               atol=1e-12, rtol=0)
    np.testing.assert_allclose(seq.delta.real.numpy(), np.repeat(d["delta_k"][:, None], 10, 1),
                               atol=1e-12, rtol=0)
    assert float(seq.phi.abs().max()) == 0.0
    np.testing.assert_allclose(seq.interaction_matrix(0.0)[0].numpy(), d["u"], rtol=1e-14, atol=0)


# ---------------------------------------------------------------------------
# GPU: the HIP path against the fixtures
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["a", "b"])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_gpu_hpsi_vs_fixture(gpu_device, tag, dtype):
    from emulators_amd.emu_sv import RydbergHamiltonian

    d = load("hpsi_random")
    ham = RydbergHamiltonian(_c(d[f"{tag}_omega"]), _c(d[f"{tag}_delta"]), _c(d[f"{tag}_phi"]),
                             torch.as_tensor(d[f"{tag}_u"]), device=gpu_device)
    psi = torch.as_tensor(d[f"{tag}_psi"]).to(device=gpu_device, dtype=dtype)
    got = ham.apply(psi).cpu().to(torch.complex128)
    ref = torch.as_tensor(d[f"{tag}_hpsi"])
    rel = float((got - ref).norm() / ref.norm())
    assert rel < (1e-13 if dtype == torch.complex128 else 1e-6), rel


@pytest.mark.gpu
def test_gpu_krylov_step_vs_fixture(gpu_device):
    from emulators_amd.emu_sv.time_evolution import EvolveStateVector

    d = load("krylov_step")
    tol = float(d["tol"])
    psi = torch.as_tensor(d["psi"]).to(gpu_device)
    out, _ = EvolveStateVector().apply(float(d["dt"]), _c(d["omega"]), _c(d["delta"]), _c(d["phi"]),
                                       torch.as_tensor(d["u"]), psi, tol, [])
    assert torch.allclose(out.cpu(), torch.as_tensor(d["result"]), atol=tol, rtol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_gpu_chain10_vs_fixture(gpu_device, precision):
    """The N=10 chain: complex128 to fidelity >= 1 - 1e-10 and observables
    within 1e-9; complex64 (FP64 Lanczos scalars) to 1 - F <= 1e-6 and
    observables within 1e-4."""
    from emulators_amd import workloads
    from emulators_amd.emu_sv import Energy, Occupation, StateResult, SVBackend, SVConfig

    d = load("chain10")
    seq = workloads.chain10()
    cfg = SVConfig(observables=[StateResult(evaluation_times=[1.0]),
                                Occupation(evaluation_times=[1.0]), Energy(evaluation_times=[1.0])],
                   log_level=40, precision=precision)
    res = SVBackend._run_from_sequence_data(seq, cfg)
    psi = res.state[-1].data.cpu()
    fid = float(orc.fidelity(psi, torch.as_tensor(d["final_state"])))
    occ = torch.as_tensor(d["occupation"])
    if precision == "complex128":
        assert 1.0 - fid <= 1e-10, 1.0 - fid
        assert torch.allclose(res.occupation[-1], occ, atol=1e-9, rtol=0)
        assert abs(float(res.energy[-1]) - float(d["energy"])) < 1e-9 * abs(float(d["energy"]))
    else:
        assert 1.0 - fid <= 1e-6, 1.0 - fid
        assert torch.allclose(res.occupation[-1], occ, atol=1e-4, rtol=0)
        assert abs(float(res.energy[-1]) - float(d["energy"])) < 1e-4 * abs(float(d["energy"]))


def _sketch(state: torch.Tensor, d: dict) -> torch.Tensor:
    """The fixture's sign sketch applied to a state (make_golden.sign_sketch)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.sign_sketch(state, int(d["sketch_rows"]), int(d["sketch_seed"]))


def test_sign_sketch_estimates_distances():
    """The sketch that pins the square20 state: ||S d|| is within 35% of
    ||d|| for random and for sparse (few-amplitude) differences, and S is
    linear, so a sketch distance bounds the state distance."""
    d = {"sketch_rows": 128, "sketch_seed": 20}
    g = torch.Generator().manual_seed(5)
    for trial in range(4):
        x = torch.randn(2**12, dtype=torch.complex128, generator=g)
        if trial % 2:
            x[torch.randperm(2**12, generator=g)[8:]] = 0.0  # 8 nonzero amplitudes
        ratio = float(_sketch(x, d).norm() / x.norm())
        assert 0.65 < ratio < 1.35, ratio
    a = torch.randn(2**10, dtype=torch.complex128, generator=g)
    b = torch.randn(2**10, dtype=torch.complex128, generator=g)
    assert torch.allclose(_sketch(a + 2j * b, d), _sketch(a, d) + 2j * _sketch(b, d), atol=1e-12)


def test_package_sampling_matches_square20_tables():
    from emulators_amd import workloads

    d = load("square20")
    seq = workloads.square20()
    assert list(seq.target_times) == list(d["target_times"])
    np.testing.assert_allclose(seq.omega.real.numpy(), np.repeat(d["omega_k"][:, None], 20, 1),
                               atol=1e-12, rtol=0)
    np.testing.assert_allclose(seq.delta.real.numpy(), np.repeat(d["delta_k"][:, None], 20, 1),
                               atol=1e-12, rtol=0)
    np.testing.assert_allclose(seq.interaction_matrix(0.0)[0].numpy(), d["u"], rtol=1e-14, atol=0)
    assert len(d["iterations"]) == 200 and abs(float(d["final_norm"]) - 1.0) < 1e-9


def test_package_inputs_match_random16_fixture():
    from emulators_amd import workloads

    d = load("random16_grad")
    pos, om, de = workloads.random16_parameters()
    np.testing.assert_array_equal(pos, d["positions"])
    np.testing.assert_array_equal(om, d["omega_k"])
    np.testing.assert_array_equal(de, d["delta_k"])
    u = workloads.interaction_matrix_from_positions(pos, workloads.C6)
    np.testing.assert_allclose(u, d["u"], rtol=1e-14, atol=0)
    assert d["grad_omega"].shape == (100,) and d["grad_delta"].shape == (100,)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_gpu_square20_vs_fixture(gpu_device, precision):
    """The N=20 lattice through SVBackend (device-driven segments) against the
    oracle's run: complex128 observables within 1e-9 and the state within
    1e-8 through the fixture's sign sketch; complex64 (FP64 Lanczos scalars)
    observables within 1e-4 and the sketch distance within 2e-3 (1 - F of a
    few 1e-6 at most)."""
    from emulators_amd import workloads
    from emulators_amd.emu_sv import Energy, Occupation, StateResult, SVBackend, SVConfig

    d = load("square20")
    cfg = SVConfig(observables=[StateResult(evaluation_times=[1.0]),
                                Occupation(evaluation_times=[1.0]), Energy(evaluation_times=[1.0])],
                   log_level=40, precision=precision)
    res = SVBackend._run_from_sequence_data(workloads.square20(), cfg)
    psi = res.state[-1].data.cpu().to(torch.complex128)
    dist = float((_sketch(psi, d) - torch.as_tensor(d["state_sketch"])).norm())
    occ = torch.as_tensor(d["occupation"])
    en = float(d["energy"])
    if precision == "complex128":
        assert dist <= 1e-8, dist
        assert torch.allclose(res.occupation[-1], occ, atol=1e-9, rtol=0)
        assert abs(float(res.energy[-1]) - en) <= 1e-9 * max(1.0, abs(en))
    else:
        assert dist <= 2e-3, dist
        assert torch.allclose(res.occupation[-1], occ, atol=1e-4, rtol=0)
        assert abs(float(res.energy[-1]) - en) <= 1e-4 * max(1.0, abs(en))


@pytest.mark.gpu
def test_gpu_random16_gradient_vs_fixture(gpu_device):
    """All 200 gradient components of the 16-atom workload (100 steps, global
    Omega_k / delta_k, loss E(T) + sum <n_i>(T)): the hand-written adjoint
    (checkpoints every 10 steps) against the oracle's chain of the
    reference's per-step autograd.  Each component within 1e-6 of its value
    relative, floored at 1e-2 of the largest component (components near zero
    carry the Krylov-tolerance noise of the larger ones)."""
    from emulators_amd.emu_sv.adjoint import EnergyExpectation, evolve_pulse, total_occupation

    d = load("random16_grad")
    u = torch.as_tensor(d["u"])
    n, s = u.shape[0], len(d["omega_k"])
    om = torch.as_tensor(d["omega_k"]).clone().requires_grad_(True)
    de = torch.as_tensor(d["delta_k"]).clone().requires_grad_(True)
    omegas, deltas = om[:, None].expand(s, n), de[:, None].expand(s, n)
    phis = torch.zeros(s, n, dtype=torch.float64)
    psi0 = torch.zeros(2**n, dtype=torch.complex128, device=gpu_device)
    psi0[0] = 1.0
    tol = float(d["tolerance"])
    psi = evolve_pulse(omegas, deltas, phis, u, psi0, torch.full((s,), 0.01, dtype=torch.float64),
                       tol, 10, 100)
    loss = EnergyExpectation.apply(omegas[-1], deltas[-1], phis[-1], u, psi) + total_occupation(psi)
    loss.backward()
    assert abs(float(loss.detach()) - float(d["loss"])) <= 1e-9 * abs(float(d["loss"]))
    for got, ref in ((om.grad, d["grad_omega"]), (de.grad, d["grad_delta"])):
        ref = torch.as_tensor(ref)
        floor = 1e-2 * float(ref.abs().max())
        bound = 1e-6 * torch.clamp(ref.abs(), min=floor)
        bad = torch.nonzero((got - ref).abs() > bound).flatten().tolist()
        assert not bad, [(k, float(got[k]), float(ref[k])) for k in bad[:5]]
