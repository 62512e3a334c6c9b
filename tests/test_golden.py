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
