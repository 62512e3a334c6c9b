"""Host-side logic and the C ABI surface (no GPU needed).

* the shared library loads and exports every symbol include/emu_sv_hip.h
  declares, and the ctypes table binds each of them;
* host numerics of the library: expm, parameter packing, the pass planner;
* the Python mirror of the reference interface: PCHIP, target times,
  evaluation times, SequenceData production, the generic krylov_exp on CPU
  tensors against the oracle, SVConfig and Results plumbing.
"""

import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg
import torch

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "emu_sv_hip.h"


def header_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(emu_sv_[a-z_0-9]+)\s*\(",
                                 text, flags=re.M)))


def test_library_exports_every_header_symbol():
    from emulators_amd import _native as native
    syms = header_symbols()
    assert len(syms) >= 13
    out = subprocess.run(["nm", "-D", "--defined-only", str(native.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    assert set(syms) == set(native.SIGNATURES), "ctypes table and header disagree"
    for s in syms:
        assert getattr(native.LIB, s) is not None


def test_version_and_workspace():
    from emulators_amd import _native as native
    assert "gfx950" in native.version()
    assert native.LIB.emu_sv_workspace_bytes(26, 0) >= (1 << 18) * 64
    assert native.LIB.emu_sv_workspace_bytes(0, 0) == 0


def test_error_codes_map_to_exceptions():
    from emulators_amd import _native as native
    with pytest.raises(ValueError, match="qubits"):
        native.plan_passes(40, torch.complex128)
    with pytest.raises(ValueError):
        native.dtype_code(torch.float32)
    rc = native.LIB.emu_sv_hpsi(10, 0, None, 1.0, 0.0, None, 1.0, None, 1.0, None, None, None, 0,
                                0, 0, None)
    assert rc == native.EMU_SV_EINVAL
    assert "null" in native.LIB.emu_sv_last_error().decode()
    # the workspace is required even without reductions (the first pass's
    # diagonal table lives there); fake device pointers are never touched
    import ctypes
    fake, fake2 = ctypes.c_void_p(4096), ctypes.c_void_p(8192)
    rc = native.LIB.emu_sv_hpsi(10, 0, fake, 1.0, 0.0, fake, 1.0, None, 1.0, fake2, None, None, 0,
                                0, 0, None)
    assert rc == native.EMU_SV_EINVAL
    assert "workspace" in native.LIB.emu_sv_last_error().decode()


@pytest.mark.parametrize("m,scale", [(1, 1.0), (3, 0.01), (7, 1.0), (12, 8.0), (40, 30.0)])
def test_native_expm_vs_scipy(m, scale):
    from emulators_amd import _native as native
    rng = np.random.default_rng(m)
    a = scale * (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / math.sqrt(m)
    ref = scipy.linalg.expm(a)
    got = native.expm(a)
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_native_expm_extended_krylov_matrix():
    """The shape expm is used on: -i dt T (tridiagonal) with the Expokit extension."""
    from emulators_amd import _native as native
    m = 12
    rng = np.random.default_rng(3)
    t = np.zeros((m + 2, m + 2), dtype=np.complex128)
    for j in range(m):
        t[j, j] = -1j * rng.normal()
        if j + 1 < m:
            b = abs(rng.normal()) + 0.1
            t[j + 1, j] = b
            t[j, j + 1] = -1j * b
    t[m, m - 1] = 0.3
    t[m + 1, m] = 1.0
    assert np.allclose(native.expm(t), torch.linalg.matrix_exp(torch.from_numpy(t)).numpy(),
                       atol=1e-13, rtol=0)


def test_pack_hamiltonian_layout():
    from emulators_amd import _native as native
    n = 4
    om = np.array([1.0, 2.0, 3.0, 4.0])
    de = np.array([0.1, 0.2, 0.3, 0.4])
    ph = np.array([0.0, 0.5, 0.0, 0.0])
    u = np.arange(16, dtype=np.float64).reshape(4, 4)
    block, cplx = native.pack_hamiltonian(om, de, ph, u)
    assert cplx and block[1120] == 1.0
    for q in range(n):
        b = n - 1 - q  # qubit q lives in bit n-1-q
        assert block[0 + b] == pytest.approx(0.5 * om[q] * math.cos(ph[q]))
        assert block[32 + b] == pytest.approx(0.5 * om[q] * math.sin(ph[q]))
        assert block[64 + b] == de[q]
    # only the upper triangle U[i][j], i<j, is read, mirrored into bit order
    for i in range(n):
        for j in range(n):
            bi, bj = n - 1 - i, n - 1 - j
            exp = 0.0 if i == j else u[min(i, j), max(i, j)]
            assert block[96 + bi * 32 + bj] == exp


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_planner_covers_every_bit(dtype):
    """The pass count is the minimum for the LDS tile and contiguity limits."""
    from emulators_amd import _native as native
    for n in range(1, 33):
        for tl, cmin in [(8, 1), (10, 3), (12, 2), (13, 3)]:
            p = native.plan_passes(n, dtype, tl, cmin, kernel="lds")
            if n <= tl:
                assert p == 1
            else:
                assert p == 1 + math.ceil((n - tl) / (tl - cmin))


# --- interface mirror -----------------------------------------------------------
def test_pchip_matches_oracle():
    from emulators_amd.emu_base.math.pchip_torch import PCHIP1D
    from oracle.emu_sv_oracle import pchip_eval
    rng = np.random.default_rng(4)
    x = torch.tensor(np.cumsum(rng.uniform(0.2, 2.0, 30)), dtype=torch.float64)
    y = torch.tensor(rng.standard_normal(30), dtype=torch.float64)
    xq = torch.linspace(float(x[0]) - 1, float(x[-1]) + 1, 301, dtype=torch.float64)
    assert torch.allclose(PCHIP1D(x, y)(xq), pchip_eval(x, y, xq), atol=1e-13, rtol=0)


def test_pchip_reference_helper_formulas():
    """test_pchip.py:13-60 cases."""
    from emulators_amd.emu_base.math import pchip_torch as p
    s0, s1 = torch.tensor([2.0, -2.0]), torch.tensor([4.0, -1.0])
    h0, h1 = torch.tensor([1.0, 2.0]), torch.tensor([3.0, 1.0])
    assert torch.allclose(p._endpoint_slope(s0, s1, h0, h1), ((2 * h0 + h1) * s0 - h0 * s1) / (h0 + h1))
    got = p._limit_endpoint(torch.tensor([1.0, -2.0, 3.0]), torch.tensor([-1.0, 2.0, 3.0]),
                            torch.tensor([-1.0, 2.0, 4.0]))
    assert torch.allclose(got, torch.tensor([0.0, 0.0, 3.0]))
    got = p._limit_endpoint(torch.tensor([10.0, -10.0, 2.0]), torch.tensor([2.0, -2.0, 2.0]),
                            torch.tensor([-1.0, 1.0, 3.0]))
    assert torch.allclose(got, torch.tensor([6.0, -6.0, 2.0]))
    d = p._pchip_derivatives(torch.tensor([1.0]), torch.tensor([2.5]))
    assert torch.allclose(d, torch.tensor([2.5, 2.5]))
    with pytest.raises(ValueError):
        p.PCHIP1D(torch.tensor([0.0, 0.0]), torch.tensor([1.0, 2.0]))
    with pytest.raises(TypeError):
        p.PCHIP1D(torch.tensor([0, 1]), torch.tensor([1.0, 2.0]))


class _Obs:
    def __init__(self, times):
        self.evaluation_times = times


class _Cfg:
    def __init__(self, observables, default=(1.0,)):
        self.observables = observables
        self.default_evaluation_times = default


def test_target_times_reference_cases():
    """test_adapter.py:905-1010 cases with duration 20 and dt 5.5 / 5."""
    from emulators_amd.emu_base import get_target_times
    evals = [0, 0.9, 11.9, 18.1]
    tt = get_target_times(20, _Cfg([_Obs([t / 20 for t in evals])]), 5.5)
    assert np.allclose(tt, sorted(set([0, 5.5, 11, 16.5, 20] + evals)))
    tt = get_target_times(20, _Cfg([_Obs(None)], default=[0.5]), 5.5)
    assert np.allclose(tt, [0, 5.5, 10, 11, 16.5, 20])
    tt = get_target_times(20, _Cfg([_Obs([0.55 + 1e-14])]), 5.5)
    assert tt == [0, 5.5, 11, 16.5, 20]
    tt = get_target_times(20, _Cfg([_Obs([0.55 + 1e-9])]), 5.5)
    assert tt == [0, 5.5, 11, (0.55 + 1e-9) * 20, 16.5, 20]
    assert get_target_times(20, _Cfg([_Obs([1.0 - 1e-14])]), 5.0) == [0.0, 5.0, 10.0, 15.0, 20.0]


def test_is_evaluation_time_tolerance():
    """test_adapter.py:1012-1028: +-1e-13 matches, +-1e-9 does not."""
    from emulators_amd.emu_base import is_evaluation_time
    cfg = _Cfg([], default=[0.7])
    for obs, t in [(_Obs([0.3]), 0.3), (_Obs(None), 0.7)]:
        assert is_evaluation_time(cfg, obs, t)
        assert is_evaluation_time(cfg, obs, t + 1e-13)
        assert is_evaluation_time(cfg, obs, t - 1e-13)
        assert not is_evaluation_time(cfg, obs, t + 1e-9)
        assert not is_evaluation_time(cfg, obs, t - 1e-9)


def test_sequence_data_from_samples_midpoints():
    """Midpoint tables against the oracle's PCHIP; last amplitude clamped >= 0;
    interaction cutoff and SLM mask as PulserData.get_sequences."""
    from emulators_amd.emu_base import sequence_data_from_samples
    from emulators_amd.emu_sv import SVConfig
    from oracle.emu_sv_oracle import midpoint_samples
    d = 95
    amp = np.concatenate([np.linspace(0, 10, 40), np.full(30, 10.0), np.linspace(10, -0.5, 25)])
    det = np.linspace(-5, 7, d)
    ph = np.full(d, 0.2)
    samples = {"a": {"amp": amp, "det": det, "phase": ph}, "b": {"amp": amp, "det": det, "phase": ph}}
    u = torch.tensor([[0.0, 1e-12], [1e-12, 0.0]], dtype=torch.float64)
    cfg = SVConfig(observables=[], interaction_cutoff=1e-10, log_level=40)
    sd = sequence_data_from_samples(samples, ["a", "b", "c"][:2], u, cfg, 10.0, slm_targets=[1],
                                    slm_end_time=30.0)
    assert sd.target_times == [0.0, 10.0, 20.0, 30.0, 40.0, 50.0, 60.0, 70.0, 80.0, 90.0, 95.0]
    exp_amp = midpoint_samples(torch.tensor(amp), sd.target_times, True)
    assert torch.allclose(sd.omega[:, 0].real, exp_amp, atol=1e-14)
    assert sd.omega[-1, 0].real >= 0.0
    assert torch.allclose(sd.delta[:, 1].real, midpoint_samples(torch.tensor(det), sd.target_times,
                                                               False), atol=1e-14)
    assert sd.omega.dtype == torch.complex128
    assert torch.all(sd.interaction_matrix(100.0) == 0.0)  # cutoff removed 1e-12


def test_interaction_matrix_c6():
    from emulators_amd.emu_base import interaction_matrix_from_positions
    u = interaction_matrix_from_positions([[0.0, 0.0], [5.0, 0.0], [0.0, 10.0]], 5420158.53)
    assert u[0, 1] == pytest.approx(5420158.53 / 5**6)
    assert u[0, 2] == pytest.approx(5420158.53 / 10**6)
    assert torch.all(torch.diagonal(u) == 0) and torch.allclose(u, u.T)


def test_generic_krylov_on_cpu_matches_oracle():
    """krylov_exp with an arbitrary operator (Arnoldi and Lanczos) on CPU
    tensors: same iteration counts as the oracle on the reference's seeded
    1000x1000 cases (test_krylov_exp.py:134-191)."""
    from emulators_amd.emu_base.math.krylov_exp import krylov_exp_impl
    from oracle.emu_sv_oracle import krylov_exp_impl as oracle_impl

    def uniform(*dims):
        return (torch.rand(*dims, dtype=torch.complex128) + 1j * torch.rand(*dims, dtype=torch.complex128)
                - (0.5 + 0.5j) * torch.ones(*dims))

    torch.random.manual_seed(1234)
    m = 0.001 * uniform(1000, 1000)
    v = uniform(1000, 1)
    v /= v.norm()
    for herm, etol in [(True, 1e-13), (False, 1e-15)]:
        op = lambda x: torch.tensordot(m, x, dims=1)  # noqa: E731
        got = krylov_exp_impl(op, v.clone(), herm, etol, 1e-20, 80)
        ref = oracle_impl(op, v.clone(), herm, etol, 1e-20, 80)
        assert got.iteration_count == ref.iteration_count == 7
        assert torch.allclose(got.result, ref.result, atol=1e-14, rtol=0)


def test_sv_config_defaults_and_validation():
    from emulators_amd.emu_sv import Occupation, SVConfig
    cfg = SVConfig(log_level=40)
    assert (cfg.dt, cfg.max_krylov_dim, cfg.krylov_tolerance, cfg.gpu, cfg.interaction_cutoff) == (
        10.0, 100, 1e-10, None, 0.0)
    assert cfg.precision == "complex128"
    assert "krylov_tolerance" in cfg._expected_kwargs()
    with pytest.raises(ValueError):
        SVConfig(precision="bf16")
    with pytest.raises(ValueError):
        SVConfig(observables=[Occupation(), Occupation()], log_level=40)
    with pytest.raises(ValueError, match="unexpected keyword arguments"):
        SVConfig(not_an_option=1, log_level=40)


def test_results_protocol():
    from emulators_amd.emu_sv import Energy, Results, SVConfig
    cfg = SVConfig(log_level=40)
    e = Energy(evaluation_times=[0.0, 0.5, 1.0])

    class H:
        def expect(self, state):
            return torch.tensor(float(state))

    res = Results(atom_order=("a",), total_duration=10)
    for t, s in [(0.0, 1.0), (0.25, 2.0), (0.5, 3.0), (1.0, 4.0)]:
        e(cfg, t, s, H(), res)
    assert res.get_result_times("energy") == [0.0, 0.5, 1.0]
    assert [float(x) for x in res.energy] == [1.0, 3.0, 4.0]
    assert float(res.get_result("energy", 0.5)) == 3.0
    # the observable itself selects its results too, as in pulser's Results
    # (the reference's tests call results.get_result(state_result, 1.0))
    assert res.get_result_times(e) == [0.0, 0.5, 1.0]
    assert float(res.get_result(e, 1.0)) == 4.0
    with pytest.raises(ValueError, match="another observable"):
        res.get_result(Energy(evaluation_times=[1.0]), 1.0)
    agg = Results.aggregate([res, res])
    assert [float(x) for x in agg.energy] == [1.0, 3.0, 4.0]
    assert float(agg.get_result(e, 0.0)) == 1.0


def test_sampling_helpers():
    from emulators_amd.emu_sv.utils import index_to_bitstring, sample_indices
    assert index_to_bitstring(4, 5) == "0101"
    torch.manual_seed(0)
    p = torch.zeros(64, dtype=torch.float64)
    p[3] = 1.0
    assert set(sample_indices(p, 50).tolist()) == {3}
    assert set(sample_indices(p, 50, max_categories=8).tolist()) == {3}


def test_backend_refuses_cpu():
    from emulators_amd import workloads
    from emulators_amd.emu_sv import SVBackend, SVConfig
    seq = workloads.chain10()
    with pytest.raises((ValueError, RuntimeError)):
        SVBackend._run_from_sequence_data(seq, SVConfig(gpu=False, log_level=40))


@pytest.mark.parametrize("n,with_phase,dt", [(6, False, 0.05), (8, True, 0.02), (9, False, 0.2)])
def test_lanczos_exp_matches_dense_extended_expm(n, with_phase, dt):
    """emu_sv_lanczos_exp on real Lanczos matrices of the oracle: the Expokit
    terms and the combination coefficients equal those of the dense expm of
    the extended matrix (krylov_exp.py:77-90), at every iteration."""
    from emulators_amd import _native as native
    from oracle import emu_sv_oracle as orc
    g = torch.Generator().manual_seed(n)
    om = 10 * torch.randn(n, dtype=torch.float64, generator=g)
    de = 5 * torch.randn(n, dtype=torch.float64, generator=g)
    ph = torch.randn(n, dtype=torch.float64, generator=g) if with_phase else torch.zeros(n, dtype=torch.float64)
    u = torch.triu(30 * torch.rand(n, n, dtype=torch.float64, generator=g), 1)
    u = u + u.T
    h = orc.OracleHamiltonian(om, de, ph, u)
    psi = torch.randn(2**n, dtype=torch.complex128, generator=g)
    sigma = -1j * dt
    r = orc.krylov_exp_impl(lambda x: sigma * (h * x), psi.clone(), True, 1e-10, 1e-10)
    T = r.T.numpy()
    for m in range(1, r.iteration_count + 1):
        text = np.zeros((m + 2, m + 2), dtype=np.complex128)
        text[: m + 1, :m] = T[: m + 1, :m]
        text[m + 1, m] = 1.0
        ref = torch.linalg.matrix_exp(torch.from_numpy(text)).numpy()
        e1, e2, c = native.lanczos_exp(text, m, sigma, T[m, m - 1].real, True)
        assert abs(e1 - abs(ref[m, 0])) < 1e-13 * max(1.0, abs(ref[m, 0]))
        assert abs(e2 - abs(ref[m + 1, 0])) < 1e-13 * max(1.0, abs(ref[m + 1, 0]))
        assert np.max(np.abs(c - ref[: m + 1, 0])) < 1e-12


def test_lanczos_exp_single_and_pure_phase():
    """m = 1 and a diagonal J: exp(T) e_0 = exp(sigma alpha_0); with sigma
    real the phases are trivial; with beta = 0 the extension entries vanish."""
    from emulators_amd import _native as native
    t = np.zeros((3, 3), dtype=np.complex128)
    t[0, 0] = -0.3j * 2.0
    e1, e2, c = native.lanczos_exp(t, 1, -0.3j, 0.0, True)
    assert abs(c[0] - np.exp(-0.3j * 2.0)) < 1e-15 and e1 == 0.0 and e2 == 0.0


def test_forced_schedule_env(monkeypatch):
    from emulators_amd import autotune
    monkeypatch.setenv("EMU_SV_SCHEDULE", "c128=lds:11:2, c64=wave:10:3")
    assert autotune.forced_schedule(torch.complex128) == autotune.Schedule("lds", 11, 2)
    assert autotune.forced_schedule(torch.complex64) == autotune.Schedule("wave", 10, 3)
    monkeypatch.delenv("EMU_SV_SCHEDULE")
    assert autotune.forced_schedule(torch.complex64) is None


def test_pack_hamiltonian_steps_matches_single_blocks():
    import ctypes

    from emulators_amd import _native as native

    rng = np.random.default_rng(3)
    n, s = 5, 4
    om, de, ph = (np.ascontiguousarray(rng.normal(size=(s, n))) for _ in range(3))
    u = rng.uniform(size=(n, n))
    u = np.ascontiguousarray(u + u.T)
    blocks = np.zeros((s, native.EMU_SV_HAM_DOUBLES))
    dp = ctypes.POINTER(ctypes.c_double)
    native.check(native.LIB.emu_sv_pack_hamiltonian_steps(
        n, s, om.ctypes.data_as(dp), de.ctypes.data_as(dp), ph.ctypes.data_as(dp),
        u.ctypes.data_as(dp), blocks.ctypes.data_as(dp)), "pack steps")
    gen = native.EMU_SV_HAM_DOUBLES - 1
    for k in range(s):
        ref, _ = native.pack_hamiltonian(om[k], de[k], ph[k], u)
        assert np.array_equal(blocks[k][:gen], ref[:gen])
    # every packed block gets its own generation (the stamp of the kept diagonal table)
    gens = [blocks[k][gen] for k in range(s)]
    assert all(g > 0 and g == int(g) for g in gens) and len(set(gens)) == s


def test_pack_generation_is_unique_and_positive():
    """Slot 1135 of a packed block is a process-unique positive integer, so a
    diagonal table kept in a workspace is never taken for another block, also
    when the new block sits at the old one's device address."""
    from emulators_amd import _native as native

    om, de, ph = np.ones(4), np.zeros(4), np.zeros(4)
    u = np.zeros((4, 4))
    gens = [native.pack_hamiltonian(om, de, ph, u)[0][native.EMU_SV_HAM_DOUBLES - 1]
            for _ in range(5)]
    assert all(g >= 1 and g == int(g) for g in gens)
    assert len(set(gens)) == 5 and gens == sorted(gens)
    block, _ = native.pack_hamiltonian(om, de, ph, u)
    assert not np.any(block[1130:1135])  # reserved slots stay zero


def test_batch_engine_rejects_cpu_and_bad_registers():
    import ctypes

    from emulators_amd import _native as native
    from emulators_amd.emu_sv.batch import BatchRegister, _Reg, evolve_batch

    reg = BatchRegister(np.ones((2, 3)), np.zeros((2, 3)), np.zeros((2, 3)), np.zeros((3, 3)),
                        np.full(2, 0.01))
    with pytest.raises(ValueError):
        evolve_batch([reg], device="cpu")
    bad = (_Reg * 1)()  # null pointers: the workspace query refuses it
    assert native.LIB.emu_sv_batch_workspace_bytes(1, native.EMU_SV_C64, bad, 0, 0) == 0
    assert native.LIB.emu_sv_batch_destroy(None) == 0


def test_new_entry_points_validate_arguments_on_the_host():
    """Argument checks of the noisy-path, density-matrix and sampling entry
    points fail before any device work (null device pointers never reach a
    kernel)."""
    import ctypes

    from emulators_amd import _native as native
    lib, ok = native.LIB, native.EMU_SV_OK
    dp = ctypes.POINTER(ctypes.c_double)
    eight = (ctypes.c_double * 8)(*([0.0] * 8))
    block = np.zeros(native.EMU_SV_HAM_DOUBLES)
    assert lib.emu_sv_pack_noise(eight, block.ctypes.data_as(dp)) == ok
    assert block[1121] == 1.0
    bad = (ctypes.c_double * 8)(float("nan"), *([0.0] * 7))
    assert lib.emu_sv_pack_noise(bad, block.ctypes.data_as(dp)) == native.EMU_SV_EINVAL
    fake = ctypes.c_void_p(4096)  # never dereferenced: every call below fails validation
    op32 = (ctypes.c_double * 32)()
    assert lib.emu_sv_apply_pair(4, 0, fake, fake, 1, 1, op32, 1, None) == native.EMU_SV_EINVAL
    assert lib.emu_sv_apply_pair(4, 0, fake, fake, 0, 4, op32, 1, None) == native.EMU_SV_EINVAL
    assert lib.emu_sv_apply_1q(4, 0, fake, 4, eight, None) == native.EMU_SV_EINVAL
    assert lib.emu_sv_antihermitian(17, 0, fake, fake, 1.0, 0.0, None) == native.EMU_SV_EINVAL
    vecs = (ctypes.c_void_p * 1)(fake)
    sc = (ctypes.c_double * 1)(1.0)
    assert lib.emu_sv_arnoldi_orthogonalize(4, 0, 0, vecs, sc, fake, fake, fake, None) == \
        native.EMU_SV_EINVAL
    assert lib.emu_sv_arnoldi_orthogonalize(4, 0, 1, vecs, sc, fake, fake, fake, None) == \
        native.EMU_SV_EINVAL  # w aliases the basis vector
    assert lib.emu_sv_sample_workspace_bytes(20) == 256 * 8
    assert lib.emu_sv_sample(4, 0, fake, ctypes.c_uint64(1), 2, fake, 1, fake, fake,
                             ctypes.c_size_t(8), None) == native.EMU_SV_EINVAL
    assert lib.emu_sv_sample(20, 0, fake, ctypes.c_uint64(1), 0, fake, 1, fake, fake,
                             ctypes.c_size_t(8), None) == native.EMU_SV_EINVAL  # workspace too small
    assert lib.emu_sv_sample(4, 0, fake, ctypes.c_uint64(1), 0, fake, 0, fake, fake,
                             ctypes.c_size_t(8), None) == ok  # no draws: nothing launched
    flags = native.EMU_SV_FLAG_NOISE | native.KERNELS["tile"]
    assert lib.emu_sv_lanczos_matvec(20, 0, fake, 0.0, -0.01, fake, 1.0, None, 1.0, 0.0, 0.0,
                                     ctypes.c_void_p(8192), 1.0, 0.0, 0.0, ctypes.c_void_p(12288),
                                     fake, fake, fake, 0, 0, flags, None) == native.EMU_SV_EINVAL
    # device coefficients: the wave kernel (and the default kind) cannot read them
    for kind in ("wave", "default"):
        assert lib.emu_sv_lanczos_matvec_dev(20, 0, fake, 0.0, -0.01, ctypes.c_void_p(8192), None,
                                             ctypes.c_void_p(12288), fake, ctypes.c_void_p(16384),
                                             fake, fake, fake, 0, 0, native.KERNELS[kind],
                                             None) == native.EMU_SV_EINVAL
    assert lib.emu_sv_lanczos_matvec_dev(20, 0, fake, 0.0, -0.01, ctypes.c_void_p(8192), None,
                                         ctypes.c_void_p(12288), fake, ctypes.c_void_p(16384),
                                         fake, fake, fake, 14, 2, native.KERNELS["lds"],
                                         None) == native.EMU_SV_EINVAL  # c128 LDS tiles <= 2^13


def test_device_iteration_wrappers_check_shapes_before_launch():
    """The device-coefficient Lanczos wrappers refuse host tensors and
    undersized records before anything reaches the library."""
    from emulators_amd import _native as native
    n = 4
    ham = torch.zeros(native.EMU_SV_HAM_DOUBLES, dtype=torch.float64)
    v = torch.zeros(2**n, dtype=torch.complex128)
    rec = torch.zeros(12, dtype=torch.float64)
    coef = torch.zeros(6, dtype=torch.float64)
    with pytest.raises(ValueError, match="parameter block"):
        native.lanczos_iterate_dev(n, ham, -0.01j, v, None, v.clone(), coef, v.clone(), rec,
                                   coef.clone(), coef.clone())
    with pytest.raises(ValueError, match="parameter block"):
        native.lanczos_matvec_dev(n, ham, -0.01j, v, None, v.clone(), coef, v.clone(), rec)


@pytest.mark.parametrize("m", [1, 3, 12, 30, 62])
@pytest.mark.parametrize("scale", [0.01, 1.0, 8.0, 40.0, 400.0])
def test_expm_column_vs_scipy(m, scale):
    """expm(A)[:, 0] from the Taylor-substep host routine against scipy's
    Pade expm, on Hessenberg matrices like the Arnoldi ones (nearly
    skew-Hermitian plus a dissipative part); beyond norm 256 it takes the
    Pade path."""
    import scipy.linalg as sl

    from emulators_amd import _native as native
    rng = np.random.default_rng(m * 1000 + int(scale * 10))
    a = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) * scale / np.sqrt(m)
    a = np.triu(a, -1)
    a = 0.5 * (a - a.conj().T) - 0.05 * abs(scale) * np.eye(m)
    ref = sl.expm(a)[:, 0]
    got = native.expm_column(a)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_emu_base_public_names():
    """The names the reference's emu_base exports (emu_base/__init__.py:1-36)
    are all importable from this package's emu_base."""
    import emulators_amd.emu_base as eb
    for name in ("get_max_rss_cpu", "get_max_rss_gpu", "compute_noise_from_lindbladians",
                 "matmul_2x2_with_batched", "PulserData", "SequenceData", "is_evaluation_time",
                 "find_root_brents", "krylov_exp", "HamiltonianType", "DEFAULT_MAX_KRYLOV_DIM",
                 "DEVICE_COUNT", "apply_measurement_errors", "unix_like", "init_logging"):
        assert hasattr(eb, name), name
        assert name in eb.__all__, name


def test_matmul_2x2_with_batched():
    """test/emu_base/math/test_matmul.py:5-9: the batched product left @ right;
    plus the shape check."""
    from emulators_amd.emu_base import matmul_2x2_with_batched
    g = torch.Generator().manual_seed(9)
    left = torch.randn(2, 2, dtype=torch.complex128, generator=g)
    right = torch.randn(10, 2, 5, dtype=torch.complex128, generator=g)
    assert torch.allclose(left @ right, matmul_2x2_with_batched(left, right))
    with pytest.raises(ValueError):
        matmul_2x2_with_batched(torch.eye(3, dtype=torch.complex128), right)
