#!/usr/bin/env python
"""Krylov time-steps/s of the N=26 triangular-lattice workload on one MI355X.

Workload: 26 atoms on a triangular lattice (5.5 um, rows of 7/6/7/6),
Omega = 2 pi rad/us with sin^2 rise/fall over 200 ns, delta ramp -20 -> 15
rad/us over 3000 ns, dt = 10 ns (300 Krylov steps), Lanczos tolerance 1e-10,
initial state |g...g>.  A "step" is one Krylov time step exp(-i dt H) psi.

Timed: ``--steps K`` consecutive steps.  When K covers the whole pulse
(300 steps) one run from t = 0 is timed; otherwise the window is the K steps
centred on the middle of the pulse: both precisions are first evolved
(untimed) from t = 0 to the window's first step, and the window is then timed
``--repeats R`` times (default 5) from that state; ``value`` is from the
median.  complex128 gives ``value``; complex64 with FP64 Lanczos scalars is
reported beside it with its fidelity against the complex128 state at the end
of the window (each precision evolved from t = 0 on its own).

Also measured: the fused Lanczos iteration (all passes, with its reductions)
timed with HIP events on the stream it runs on, against HBM peak and against
a plain copy kernel timed in the same run; and the CPU baseline (the oracle,
a restatement of the reference's torch-CPU algorithm) running one whole
Krylov step of the window on the host cores this process may use.

Output: one JSON line on stdout.  Progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
METRIC = "Krylov steps/s + H·ψ HBM GB/s at N=26 (c64 vs c128 at stated fidelity)"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_steps(seq, psi, k0, count, dtype, tol, stats):
    import torch

    from emulators_amd.emu_sv.time_evolution import EvolveStateVector

    stepper = EvolveStateVector()
    u = seq.interaction_matrix(0.0)[0]
    for k in range(k0, k0 + count):
        dt = (seq.target_times[k + 1] - seq.target_times[k]) * 1e-3
        psi, _ = stepper.apply(dt, seq.omega[k], seq.delta[k], seq.phi[k], u, psi, tol, [])
        if (k + 1) % 50 == 0:
            log(f"{dtype}: step {k + 1}, matvecs so far {stats.matvecs}")
    return psi


def ground_state(n, dtype):
    import torch

    psi = torch.zeros(2**n, dtype=dtype, device="cuda")
    psi[0] = 1.0
    return psi


def timed_window(seq, dtype, k0, steps, repeats, warmup, tol):
    """Evolve untimed from t = 0 to step k0, then time ``steps`` steps from
    there ``repeats`` times (each from the same state).  Returns the state
    after the window, the wall times and the matvec count of one window."""
    import torch

    from emulators_amd.emu_base.math.krylov_exp import KrylovStats

    n = seq.omega.shape[1]
    psi0 = ground_state(n, dtype)
    if warmup:
        run_steps(seq, psi0.clone(), k0, min(warmup, steps), dtype, tol, KrylovStats)
    if k0:
        log(f"{dtype}: evolving (untimed) to the window's first step {k0}")
        psi0 = run_steps(seq, psi0, 0, k0, dtype, tol, KrylovStats)
    torch.cuda.synchronize()
    walls, matvecs, psi = [], None, None
    for rep in range(repeats):
        KrylovStats.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        psi = run_steps(seq, psi0.clone(), k0, steps, dtype, tol, KrylovStats)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
        if matvecs is None:
            matvecs = KrylovStats.matvecs
        elif KrylovStats.matvecs != matvecs:
            log(f"{dtype}: repeat {rep} took {KrylovStats.matvecs} matvecs against {matvecs}")
        log(f"{dtype}: window repeat {rep + 1}/{repeats} in {walls[-1]:.3f} s")
    return psi, psi0, walls, matvecs


def median(xs):
    ys = sorted(xs)
    mid = len(ys) // 2
    return ys[mid] if len(ys) % 2 else 0.5 * (ys[mid - 1] + ys[mid])


def copy_bandwidth(n, reps):
    """A plain device copy of one complex128 vector (read 16*2^n, write
    16*2^n bytes), timed with HIP events: the attainable streaming rate."""
    import torch

    from emulators_amd import _native as native

    src = torch.zeros(2**n, dtype=torch.complex128, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        native.convert(n, src, dst)
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        native.convert(n, src, dst)
    stop.record(stream)
    stop.synchronize()
    ms = start.elapsed_time(stop) / reps
    del src, dst
    return 2 * 16 * (2**n) / (ms * 1e-3) / 1e9, ms


def time_matvec(seq, dtype, reps, step):
    """Average duration of one Lanczos iteration's launch sequence
    (emu_sv_lanczos_matvec: orthogonalise + all H.psi passes + reductions),
    timed with HIP events on the stream it runs on.  As in the Krylov loop,
    the first-pass diagonal table is built once (the first call) and kept
    by the later ones (EMU_SV_FLAG_KEEP_DIAG)."""
    import torch

    from emulators_amd import _native as native
    from emulators_amd import autotune
    from emulators_amd.emu_sv import RydbergHamiltonian

    n = seq.omega.shape[1]
    ham = RydbergHamiltonian(seq.omega[step], seq.delta[step], seq.phi[step],
                             seq.interaction_matrix(0.0)[0], device="cuda")
    params = ham.device_params(torch.device("cuda"))
    sched = autotune.best_schedule(n, dtype, torch.device("cuda"), params, ham.complex)
    g = torch.Generator(device="cuda").manual_seed(0)
    w = torch.randn(2**n, dtype=dtype, device="cuda", generator=g)
    pa = torch.randn(2**n, dtype=dtype, device="cuda", generator=g)
    pb = torch.randn(2**n, dtype=dtype, device="cuda", generator=g)
    r = torch.empty_like(w)

    def call(keep=True):
        # zero coefficients keep w unchanged; the traffic is that of a real iteration
        native.lanczos_matvec(n, params, -0.01j, w, 1.0, pa, 1.0, 0j, pb, 1.0, 0j, r,
                              complex_coef=ham.complex, keep_diag=keep, **sched.kwargs())

    call(keep=False)
    for _ in range(2):
        call()
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        call()
    stop.record(stream)
    stop.synchronize()
    ms = start.elapsed_time(stop) / reps
    bpe = 16 if dtype == torch.complex128 else 8
    alg_bytes = 3 * (2**n) * bpe  # read v_j, read v_{j-1}, write v_{j+1}
    passes = native.plan_passes(n, dtype, sched.tile_log, sched.min_contig, kernel=sched.kernel,
                                partners=sched.partners, first_log=sched.first_log)
    del w, pa, pb, r
    return ms, alg_bytes, passes, sched


def tile_text(sched, btile_log):
    """The tile size of a schedule, for the bench line."""
    if sched.kernel == "btile":
        return f"2^{btile_log}-amplitude tiles in registers, >= {sched.min_contig} contiguous bits in pass B"
    if sched.kernel == "default" or sched.tile_log == 0:
        return "library default tiles"
    return f"tile 2^{sched.tile_log}, >= {sched.min_contig} contiguous bits"


def measured_traffic(tag, sched, n):
    """HBM bytes per Lanczos iteration from the committed rocprofv3 counter
    runs, when one was taken on this schedule and size (else None)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "lanczos_matvec_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("n_qubits") != n:
        return None
    ents = rec.get(tag) or []
    for ent in ents if isinstance(ents, list) else [ents]:
        if {"partners": 0, "first_log": 0, **ent.get("schedule", {})} == sched.kwargs():
            return float(ent["bytes_per_call"])
    return None


def host_cpus():
    """CPUs this process may run on: the affinity mask, capped by a cgroup
    CPU quota when one is set; plus os.cpu_count() and the CPU model."""
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    threads = usable if quota is None else max(1, min(usable, int(quota)))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": threads, "os_cpu_count": os.cpu_count(), "affinity": usable,
            "cgroup_quota_cpus": quota, "model": model}


def cpu_baseline(seq, step, state, tol):
    """One whole Krylov step of the workload on the oracle (torch CPU,
    complex128, the reference's index_add_ H.psi and Lanczos control flow):
    step ``step`` from ``state`` (the GPU's state at that step), on every
    CPU this process may use."""
    import torch

    from oracle import emu_sv_oracle as orc

    cpus = host_cpus()
    torch.set_num_threads(cpus["threads"])
    dt = (seq.target_times[step + 1] - seq.target_times[step]) * 1e-3
    u = seq.interaction_matrix(0.0)[0]
    t0 = time.perf_counter()
    ham = orc.OracleHamiltonian(seq.omega[step], seq.delta[step], seq.phi[step], u)
    t_diag = time.perf_counter() - t0
    log(f"CPU baseline: diagonal built in {t_diag:.1f} s on {cpus['threads']} threads")
    v = state.to("cpu", torch.complex128)
    t1 = time.perf_counter()
    res = orc.krylov_exp_impl(lambda x: -1j * dt * (ham * x), v, is_hermitian=True,
                              exp_tolerance=tol, norm_tolerance=tol)
    t_krylov = time.perf_counter() - t1
    t_step = t_diag + t_krylov
    log(f"CPU baseline: Krylov step {step} in {t_step:.1f} s ({res.iteration_count} iterations)")
    out_state = res.result
    del res, v, ham
    # the N=10 chain (100 steps) run whole on the oracle: the reference CPU path's own workload
    from emulators_amd import workloads
    seq10 = workloads.chain10()
    t0 = time.perf_counter()
    orc.run_sequence(seq10.omega, seq10.delta, seq10.phi, seq10.interaction_matrix, seq10.target_times)
    t_chain10 = time.perf_counter() - t0
    log(f"CPU baseline: N=10 chain (100 steps) on the oracle in {t_chain10:.2f} s")
    return {
        "value": 1.0 / t_step,
        "unit": "Krylov steps/s",
        "cores": cpus["threads"],
        "kind": "port",
        "sample": (f"oracle (torch-CPU complex128 restatement of the reference) running one whole "
                   f"Krylov step of the workload, step {step} from the GPU's state at that step: "
                   f"diagonal build {t_diag:.2f} s + Krylov exponential {t_krylov:.2f} s "
                   f"({out_state.numel()} amplitudes)"),
        "host": cpus,
        "chain10_seconds": t_chain10,
        "_state": out_state,
    }


def adjoint_cost(reps: int = 2) -> dict:
    """The 16-atom gradient workload: 16 atoms at seeded random positions
    (30x30 um box, min spacing 5 um), 100 piecewise-constant 10 ns steps with
    global Omega_k ~ U(0, 4 pi), delta_k ~ U(-10, 10) (200 parameters), loss =
    final energy + sum of occupations.  Times one forward run (100 Krylov
    steps) against one full gradient (forward with checkpoints every 10 steps
    + hand-written adjoint sweep), best of ``reps`` after one warmup each, for
    the three backward modes (recompute, the default: the forward Lanczos
    decompositions kept within a memory budget -- all of them at N=16 --
    and segments beyond it re-run from their checkpoints; reverse: the state
    reverse-propagated beside the adjoint, reset at checkpoints; batched: the
    same two recurrences sharing launches in the batch engine)."""
    import torch

    from emulators_amd import workloads
    from emulators_amd.emu_base.math.krylov_exp import ScaledHamiltonian, krylov_exp
    from emulators_amd.emu_sv import RydbergHamiltonian
    from emulators_amd.emu_sv.adjoint import EnergyExpectation, evolve_pulse, total_occupation

    pos, om_np, de_np = workloads.random16_parameters()
    n, s = pos.shape[0], len(om_np)
    u = torch.as_tensor(workloads.interaction_matrix_from_positions(pos, workloads.C6))
    tol = 1e-10
    zeros = torch.zeros(n, dtype=torch.float64)

    def psi0():
        p = torch.zeros(2**n, dtype=torch.complex128, device="cuda")
        p[0] = 1.0
        return p

    def forward():
        psi = psi0()
        for k in range(s):
            ham = RydbergHamiltonian(torch.full((n,), float(om_np[k])), torch.full((n,), float(de_np[k])),
                                     zeros, u, device="cuda")
            psi = krylov_exp(ScaledHamiltonian(ham, -0.01j), psi, tol, tol)
        torch.cuda.synchronize()

    def gradient(mode):
        om = torch.as_tensor(om_np).clone().requires_grad_(True)
        de = torch.as_tensor(de_np).clone().requires_grad_(True)
        omegas, deltas = om[:, None].expand(s, n), de[:, None].expand(s, n)
        phis = torch.zeros(s, n, dtype=torch.float64)
        psi = evolve_pulse(omegas, deltas, phis, u, psi0(), torch.full((s,), 0.01), tol, 10, 100, mode)
        loss = EnergyExpectation.apply(omegas[-1], deltas[-1], phis[-1], u, psi) + total_occupation(psi)
        loss.backward()
        torch.cuda.synchronize()
        return om.grad, de.grad

    def best(fn, *a):
        fn(*a)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn(*a)
            times.append(time.perf_counter() - t0)
        return min(times)

    t_fwd = best(forward)
    out = {"workload": "random16: N=16, 100 steps x 10 ns, global Omega_k/delta_k (200 parameters), "
                       "loss = E(T) + sum_i <n_i>(T), complex128, tolerance 1e-10, checkpoints every 10",
           "forward_s": t_fwd}
    for mode in ("recompute", "reverse", "batched"):
        t = best(gradient, mode)
        out[mode] = {"gradient_s": t, "ratio_vs_forward": t / t_fwd}
    out["ratio_vs_forward"] = out["recompute"]["ratio_vs_forward"]
    log(f"adjoint: forward {t_fwd:.3f} s, gradient x{out['recompute']['ratio_vs_forward']:.2f} "
        f"(recompute, forward bases kept), x{out['reverse']['ratio_vs_forward']:.2f} (reverse), "
        f"x{out['batched']['ratio_vs_forward']:.2f} (batched)")
    return out


def batch_throughput(sample_every: int = 16) -> dict:
    """The batching workload: 256 seeded registers, N ~ U{8..20} at random
    positions (min spacing 5 um), Omega ~ U(1, 4 pi) rad/us, detuning ramp
    -U(5,20) -> +U(5,20) rad/us, 500..2000 ns at dt = 10 ns, complex64 with
    FP64 Lanczos scalars.  All 256 run in the device-driven batch engine
    (graph-replayed batched iterations); every ``sample_every``-th register
    also runs alone through the single-register engine (complex64), which
    gives the sequential instances/s and a fidelity check."""
    import torch

    from emulators_amd import workloads
    from emulators_amd.emu_base.math.krylov_exp import ScaledHamiltonian, krylov_exp
    from emulators_amd.emu_sv import RydbergHamiltonian
    from emulators_amd.emu_sv.batch import evolve_batch, registers_from_sequences

    inst = workloads.batch_instances()
    regs = registers_from_sequences([workloads.batch_sequence(x) for x in inst])
    log(f"batch: {len(regs)} registers, N {min(r.nqubits for r in regs)}..{max(r.nqubits for r in regs)}, "
        f"{sum(r.steps for r in regs)} steps in total")
    warm = evolve_batch(regs[:8], dtype=torch.complex64)  # compile / warm the graph path
    del warm
    from emulators_amd import autotune
    t0 = time.perf_counter()
    tuned = autotune.with_batch_schedules(regs, torch.complex64, torch.device("cuda"), 1e-10, 60)
    t_tune = time.perf_counter() - t0
    schedules = {int(r.nqubits): [int(r.tile_log), int(r.min_contig)] for r in tuned}
    log(f"batch: schedules tuned per N in {t_tune:.1f} s: {schedules}")
    res = evolve_batch(tuned, dtype=torch.complex64, check_every=64)
    log(f"batch: {len(regs)} registers in {res.seconds:.2f} s ({res.iterations} batched iterations), "
        f"failed {len(res.failed)}")
    sample = list(range(0, len(regs), sample_every))
    t_seq = 0.0
    worst = 0.0
    for i in sample:
        reg = regs[i]
        n = reg.nqubits
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        psi = torch.zeros(2**n, dtype=torch.complex64, device="cuda")
        psi[0] = 1.0
        for k in range(reg.steps):
            ham = RydbergHamiltonian(torch.as_tensor(reg.omegas[k]), torch.as_tensor(reg.deltas[k]),
                                     torch.as_tensor(reg.phis[k]),
                                     torch.as_tensor(reg.interaction_matrix), device="cuda")
            psi = krylov_exp(ScaledHamiltonian(ham, -1j * float(reg.dts[k])), psi, 1e-10, 1e-10)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t0
        a = psi.to(torch.complex128)
        b = res.states[i].to(torch.complex128)
        fid = float(torch.abs(torch.vdot(a, b)) ** 2 / (torch.vdot(a, a).real * torch.vdot(b, b).real))
        worst = max(worst, 1.0 - fid)
    seq_rate = len(sample) / t_seq
    out = {
        "workload": "256 registers, N ~ U{8..20}, random positions (min 5 um), Omega ~ U(1, 4pi), "
                    "delta -U(5,20) -> +U(5,20), 500-2000 ns at dt=10 ns, complex64 + FP64 scalars",
        "instances": len(regs),
        "batched_s": res.seconds,
        "batched_instances_per_s": len(regs) / res.seconds,
        "batched_iterations": res.iterations,
        "sequential_sample": len(sample),
        "sequential_instances_per_s": seq_rate,
        "speedup": (len(regs) / res.seconds) / seq_rate,
        "max_fidelity_deficit_vs_sequential": worst,
        "failed": len(res.failed),
        "schedules_per_n": {str(k): v for k, v in sorted(schedules.items())},
        "schedule_note": "(tile_log, min_contig) of the LDS pass kernel per atom count, picked by "
                         "autotune.tune_batch on this register set cut to its first 4 steps (best "
                         "uniform schedule, then per-N changes that save > 2%); [0, 0] = library "
                         "default (one tile)",
        "tuning_s": t_tune,
    }
    log(f"batch: {out['batched_instances_per_s']:.1f} instances/s batched vs "
        f"{seq_rate:.2f} sequential (x{out['speedup']:.1f}), 1-F <= {worst:.1e}")
    return out


def parity_legs() -> dict:
    """The two parity workloads, timed on the GPU with their accuracy stated.

    chain10: N=10 chain, 5 um, Omega = 4 pi, delta -10 -> 10 rad/us over 1000 ns
    (100 steps), |g..g>, occupations and energy at the end; complex128 and
    complex64 against the committed complex128 fixture tests/golden/chain10.npz
    (written by the oracle, cross-checked by dense eigendecomposition steps).
    square20: 4x5 lattice, 6 um, same pulse over 2000 ns (200 steps);
    complex64 (FP64 Lanczos scalars) against the complex128 GPU run."""
    import numpy as np
    import torch

    from emulators_amd import workloads
    from emulators_amd.emu_sv import Energy, Occupation, StateResult, SVBackend, SVConfig

    def run(seq, precision):
        cfg = SVConfig(observables=[StateResult(evaluation_times=[1.0]),
                                    Occupation(evaluation_times=[1.0]), Energy(evaluation_times=[1.0])],
                       log_level=40, precision=precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = SVBackend._run_from_sequence_data(seq, cfg)
        torch.cuda.synchronize()
        return res, time.perf_counter() - t0

    def compare(psi, occ, en, ref_psi, ref_occ, ref_en):
        a = psi.to(torch.complex128)
        b = ref_psi.to(torch.complex128)
        fid = float(torch.abs(torch.vdot(a, b)) ** 2 / (torch.vdot(a, a).real * torch.vdot(b, b).real))
        return {"fidelity_deficit": 1.0 - fid,
                "max_occupation_error": float(torch.max(torch.abs(occ - ref_occ))),
                "energy_error": abs(float(en) - float(ref_en))}

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "chain10.npz")
    with np.load(path, allow_pickle=False) as f:
        gold = {k: f[k] for k in f.files}
    seq = workloads.chain10()
    out = {"chain10": {}, "square20": {}}
    run(seq, "complex128")  # warm
    for prec in ("complex128", "complex64"):
        res, t = run(seq, prec)
        out["chain10"][prec] = {"seconds": t, **compare(
            res.state[-1].data.cpu(), res.occupation[-1].cpu(), res.energy[-1],
            torch.as_tensor(gold["final_state"]), torch.as_tensor(gold["occupation"]), gold["energy"])}
    seq = workloads.square20()
    r128, t128 = run(seq, "complex128")
    r64, t64 = run(seq, "complex64")
    out["square20"] = {"seconds_c128": t128, "seconds_c64": t64,
                       "c64_vs_c128": compare(r64.state[-1].data.cpu(), r64.occupation[-1].cpu(),
                                              r64.energy[-1], r128.state[-1].data.cpu(),
                                              r128.occupation[-1].cpu(), r128.energy[-1])}
    c = out["chain10"]["complex128"]
    log(f"parity: chain10 c128 {c['seconds']:.2f} s, 1-F {c['fidelity_deficit']:.1e}, "
        f"occupation err {c['max_occupation_error']:.1e}; square20 c64 vs c128 1-F "
        f"{out['square20']['c64_vs_c128']['fidelity_deficit']:.1e}")
    return out


def noise_legs() -> dict:
    """The noisy paths on the GPU.

    monte_carlo: 10-atom chain (the chain10 pulse, 100 steps) with relaxation
    (1 /us) and dephasing (0.5 /us); four seeded quantum-jump trajectories
    (Python ``random`` seeds 0..3).  lindblad: the same pulse and noise on 6
    atoms as a density matrix (2^12 elements).  The final states are kept
    (``_states``) for the CPU-baseline leg, which replays the same workloads
    on the oracle and reports the agreement."""
    import random

    import torch

    from emulators_amd import workloads
    from emulators_amd.emu_sv import Occupation, StateResult, SVBackend, SVConfig

    lind = workloads.relaxation_dephasing(1.0, 0.5)
    cfg = SVConfig(observables=[StateResult(evaluation_times=[1.0]),
                                Occupation(evaluation_times=[1.0])], log_level=40)
    out = {"_states": {}}
    seq = workloads.noisy_chain(10, lind, stochastic=True)
    times = []
    for seed in range(4):
        random.seed(seed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = SVBackend._run_from_sequence_data(seq, cfg)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        if seed == 0:
            out["_states"]["monte_carlo"] = res.state[-1].data.cpu()
    out["monte_carlo"] = {
        "workload": "10-atom chain, 5 um, Omega 4pi, delta -10 -> 10 rad/us over 1000 ns "
                    "(100 steps), relaxation 1/us + dephasing 0.5/us, complex128",
        "trajectories": len(times), "seconds_per_trajectory": sum(times[1:]) / (len(times) - 1),
        "first_trajectory_s": times[0]}
    # a 16-atom trajectory: the tuned noisy H.psi kernels and the Gram-form
    # Gram-Schmidt (seed 0; a short warm run first, which also tunes)
    seq16 = workloads.noisy_chain(16, lind, stochastic=True)
    random.seed(0)
    SVBackend._run_from_sequence_data(workloads.noisy_chain(16, lind, True, duration=50), cfg)
    random.seed(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = SVBackend._run_from_sequence_data(seq16, cfg)
    torch.cuda.synchronize()
    out["monte_carlo_n16"] = {
        "workload": "16-atom chain, the same pulse and noise (100 steps), one trajectory (seed 0), "
                    "complex128", "seconds_per_trajectory": time.perf_counter() - t0}
    n = 6
    seq = workloads.noisy_chain(n, lind, stochastic=False)
    SVBackend._run_from_sequence_data(workloads.noisy_chain(n, lind, False, duration=50), cfg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = SVBackend._run_from_sequence_data(seq, cfg)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    out["_states"]["lindblad"] = (res.state[-1].data.cpu(), torch.as_tensor(res.occupation[-1]))
    out["lindblad"] = {
        "workload": f"{n}-atom chain density matrix (2^{2 * n} elements), the same pulse and noise, "
                    "100 steps, complex128", "seconds": t_gpu}
    log(f"noise: {out['monte_carlo']['seconds_per_trajectory']:.2f} s per Monte-Carlo trajectory "
        f"(N=10), {out['monte_carlo_n16']['seconds_per_trajectory']:.2f} s (N=16), "
        f"{t_gpu:.2f} s for the N=6 density matrix")
    return out


def cpu_noise_baseline(noise: dict) -> dict:
    """The noisy workloads of ``noise_legs`` on the oracle (host cores):
    trajectory seed 0 (same draws, so the same jumps) and the density matrix,
    timed, with the agreement of the GPU results."""
    import random

    import torch

    from emulators_amd import workloads
    from oracle import emu_sv_oracle as orc

    lind = workloads.relaxation_dephasing(1.0, 0.5)

    def loop(seq, state, step_fn):
        tt = seq.target_times
        for k in range(seq.omega.shape[0]):
            dt = (tt[k + 1] - tt[k]) * 1e-3
            state = step_fn(dt, seq.omega[k].real, seq.delta[k].real, seq.phi[k].real,
                            seq.interaction_matrix(tt[k])[0], state)
        return state

    out = {}
    seq = workloads.noisy_chain(10, lind, stochastic=True)
    random.seed(0)
    mc = orc.OracleMonteCarlo()
    psi0 = torch.zeros(2**10, dtype=torch.complex128)
    psi0[0] = 1.0
    t0 = time.perf_counter()
    psi = loop(seq, psi0, lambda dt, om, de, ph, u, st: mc.step(dt, om, de, ph, u, st, 1e-10, lind))
    t_mc = time.perf_counter() - t0
    psi = psi / psi.norm()
    a = noise["_states"]["monte_carlo"]
    a = a / a.norm()
    out["monte_carlo"] = {"seconds_per_trajectory": t_mc, "jumps": len(mc.jumps),
                          "fidelity_deficit_gpu_vs_oracle": 1.0 - float(torch.abs(torch.vdot(a, psi)) ** 2)}
    n = 6
    seq = workloads.noisy_chain(n, lind, stochastic=False)
    rho0 = torch.zeros(2**n, 2**n, dtype=torch.complex128)
    rho0[0, 0] = 1.0

    def dm_step(dt, om, de, ph, u, rho):
        lo = orc.OracleLindbladian(om, de, ph, lind, u)
        return orc.krylov_exp(lambda x: -1j * dt * (lo @ x), rho, 1e-10, 1e-10, is_hermitian=False)

    t0 = time.perf_counter()
    rho = loop(seq, rho0, dm_step)
    t_dm = time.perf_counter() - t0
    rho = rho / rho.trace()
    p = rho.diagonal().real
    occ = torch.tensor([float(sum(p[x] for x in range(2**n) if (x >> (n - 1 - q)) & 1))
                        for q in range(n)], dtype=torch.float64)
    rho_gpu, occ_gpu = noise["_states"]["lindblad"]
    out["lindblad"] = {"seconds": t_dm,
                       "max_abs_rho_error_gpu_vs_oracle": float((rho_gpu / rho_gpu.trace() - rho).abs().max()),
                       "max_occupation_error_gpu_vs_oracle": float((occ_gpu - occ).abs().max())}
    log(f"CPU baseline, noisy paths: Monte-Carlo trajectory {t_mc:.1f} s ({len(mc.jumps)} jumps, "
        f"1-F vs GPU {out['monte_carlo']['fidelity_deficit_gpu_vs_oracle']:.1e}), density matrix "
        f"{t_dm:.1f} s (occupation error {out['lindblad']['max_occupation_error_gpu_vs_oracle']:.1e})")
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--hpsi-reps", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=0,
                    help="timed repeats of the window (default: 5, or 1 for the whole pulse)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c64", action="store_true")
    ap.add_argument("--duration", type=int, default=3000, help="pulse duration in ns")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-noise", action="store_true")
    ap.add_argument("--only-noise", action="store_true", help="time only the noisy paths")
    ap.add_argument("--only-batch", action="store_true", help="time only the batching workload")
    ap.add_argument("--only-adjoint", action="store_true", help="time only the gradient workload")
    ap.add_argument("--only-matvec", action="store_true",
                    help="time only the Lanczos matvec (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the single-GPU N=26 state fits in one MI355X)", file=sys.stderr)
        return 2

    import torch

    from emulators_amd import _native as native
    from emulators_amd import workloads

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    tol = 1e-10
    seq = workloads.triangular26(duration=args.duration)
    nsteps = seq.omega.shape[0]
    steps = min(args.steps, nsteps)
    n = seq.omega.shape[1]
    log(f"workload triangular26: N={n}, {nsteps} steps; timing {steps} after {args.warmup} warmup")

    if args.only_batch:
        print(json.dumps(batch_throughput()), flush=True)
        return 0
    if args.only_noise:
        noise = noise_legs()
        if not args.no_cpu_baseline:
            noise["cpu_baseline"] = cpu_noise_baseline(noise)
        noise.pop("_states")
        print(json.dumps(noise), flush=True)
        return 0
    if args.only_adjoint:
        print(json.dumps(adjoint_cost()), flush=True)
        return 0
    if args.only_matvec:
        for dt in (torch.complex128, torch.complex64):
            ms, nbytes, passes, sched = time_matvec(seq, dt, args.hpsi_reps, nsteps // 2)
            log(f"{dt}: Lanczos matvec {ms:.3f} ms, {nbytes / ms / 1e6:.0f} GB/s algorithmic, "
                f"{passes} passes, {sched}")
        # plain copy of one complex128 vector (2 * 16 * 2^n bytes): calibrates
        # the profiler's FETCH_SIZE / WRITE_SIZE counters
        src = torch.zeros(2**n, dtype=torch.complex128, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(args.hpsi_reps):
            native.convert(n, src, dst)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.hpsi_reps} copies of {16 * 2**n} bytes")
        gbs, ms = copy_bandwidth(n, args.hpsi_reps)
        log(f"copy kernel: {gbs:.0f} GB/s ({ms:.3f} ms per 2 x {16 * 2**n} bytes)")
        return 0
    whole = steps >= nsteps
    k0 = 0 if whole else (nsteps - steps) // 2
    repeats = args.repeats if args.repeats else (1 if whole else 5)
    psi128, start128, walls128, mv128 = timed_window(seq, torch.complex128, k0, steps, repeats,
                                                     args.warmup, tol)
    wall128 = median(walls128)
    log(f"c128: steps {k0}..{k0 + steps - 1}, median {wall128:.3f} s of {repeats}, {mv128} matvecs")
    start_host = start128.cpu()  # the window's first state, for the CPU leg
    del start128
    result64 = None
    if not args.no_c64:
        psi64, start64, walls64, mv64 = timed_window(seq, torch.complex64, k0, steps, repeats,
                                                     args.warmup, tol)
        del start64
        wall64 = median(walls64)
        log(f"c64: median {wall64:.3f} s of {repeats}, {mv64} matvecs")
        a = psi128
        b = native.convert(n, psi64, torch.empty_like(psi128))
        fid = float(torch.abs(torch.vdot(a, b)) ** 2 / (torch.vdot(a, a).real * torch.vdot(b, b).real))
        occ_a = native.correlation(n, a).diagonal()
        occ_b = native.correlation(n, b).diagonal()
        result64 = {"value": steps / wall64, "ms_per_step": 1e3 * wall64 / steps,
                    "walls_s": walls64,
                    "speedup_vs_c128": wall128 / wall64, "fidelity_vs_c128": fid,
                    "fidelity_deficit": 1.0 - fid,
                    "fidelity_over": f"steps 0..{k0 + steps - 1}, each precision from t = 0",
                    "max_occupation_error": float(torch.max(torch.abs(occ_a - occ_b))),
                    "mean_lanczos_iterations": mv64 / steps}
        del psi64, b
    del psi128
    torch.cuda.empty_cache()

    mid = min(nsteps // 2, nsteps - 1)
    ms128, bytes128, passes128, sched128 = time_matvec(seq, torch.complex128, args.hpsi_reps, mid)
    gbs128 = bytes128 / (ms128 * 1e-3) / 1e9
    hpsi = {"c128": {"ms": ms128, "alg_bytes": bytes128, "GB/s": gbs128, "passes": passes128,
                     "schedule": sched128.kwargs()}}
    if not args.no_c64:
        ms64, bytes64, passes64, sched64 = time_matvec(seq, torch.complex64, args.hpsi_reps, mid)
        hpsi["c64"] = {"ms": ms64, "alg_bytes": bytes64, "GB/s": bytes64 / (ms64 * 1e-3) / 1e9,
                       "passes": passes64, "schedule": sched64.kwargs()}
    log(f"Lanczos matvec c128 {ms128:.3f} ms ({gbs128:.0f} GB/s algorithmic)")
    copy_gbs, copy_ms = copy_bandwidth(n, args.hpsi_reps)
    traffic128 = measured_traffic("c128", sched128, n)
    log(f"copy kernel: {copy_gbs:.0f} GB/s ({copy_ms:.3f} ms per 2 x {16 * 2**n} bytes)")

    adjoint = None if args.no_adjoint else adjoint_cost()
    batch = None if args.no_batch else batch_throughput()
    parity = None if args.no_parity else parity_legs()
    noise = None if args.no_noise else noise_legs()
    mean_iters = mv128 / steps
    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle) ...")
        cpu = cpu_baseline(seq, k0, start_host, tol)
        # the same step on the GPU, for the agreement of the two
        from emulators_amd.emu_base.math.krylov_exp import KrylovStats
        gpu_step = run_steps(seq, start_host.to("cuda"), k0, 1, torch.complex128, tol,
                             KrylovStats).cpu()
        ref = cpu.pop("_state")
        cpu["gpu_step_vs_oracle"] = {
            "step": k0, "max_abs_error": float((gpu_step - ref).abs().max()),
            "fidelity_deficit": 1.0 - float(torch.abs(torch.vdot(gpu_step, ref)) ** 2 /
                                            (torch.vdot(gpu_step, gpu_step).real * torch.vdot(ref, ref).real))}
        del gpu_step, ref
        log(f"CPU baseline {cpu['value']:.4f} steps/s on {cpu['cores']} threads; GPU step vs oracle "
            f"max |diff| {cpu['gpu_step_vs_oracle']['max_abs_error']:.1e}")
        if noise is not None:
            cpu["noise"] = cpu_noise_baseline(noise)
    if noise is not None:
        noise.pop("_states")

    out = {
        "metric": METRIC,
        "value": steps / wall128,
        "unit": "Krylov steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall128 / steps,
        "walls_s": walls128,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": "triangular26: N=26 triangular lattice 5.5 um (rows 7/6/7/6), Omega 2pi "
                        "rad/us sin^2 rise/fall 200 ns, delta -20->15 rad/us, "
                        f"{args.duration} ns at dt=10 ns, |g..g>, Krylov tolerance 1e-10",
            "n_qubits": n,
            "time_steps": steps,
            "window": {"first_step": k0, "last_step": k0 + steps - 1, "of_steps": nsteps,
                       "repeats": repeats, "statistic": "median" if repeats > 1 else "single run",
                       "start_state": "evolved untimed from t = 0" if k0 else "|g..g> at t = 0"},
            "mean_lanczos_iterations": mean_iters,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": gbs128,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbs128 / HBM_PEAK_GBS,
            "attainable": copy_gbs,
            "frac_of_attainable": gbs128 / copy_gbs,
            "attainable_kernel": "plain device copy of one complex128 vector (emu_sv_convert), "
                                 "HIP events, same run",
            "traffic": traffic128,
            # the HBM rate the iteration actually sustains: counter bytes per call / call time
            "traffic_GBps": None if traffic128 is None else traffic128 / (ms128 * 1e-3) / 1e9,
            "traffic_frac_of_attainable": (None if traffic128 is None
                                           else traffic128 / (ms128 * 1e-3) / 1e9 / copy_gbs),
            "kernel": f"one Lanczos iteration (emu_sv_lanczos_matvec, complex128): "
                      f"orthogonalise + {passes128} H.psi passes ({sched128.kernel} kernel, "
                      f"{tile_text(sched128, 14)}) + reductions; algorithmic bytes 3*D*16 per call "
                      "(read v_j, v_{j-1}, write v_{j+1})",
            "schedule": sched128.kwargs(),
        },
        "cpu_baseline": cpu,
        "c64": result64,
        "hpsi": hpsi,
        "adjoint": adjoint,
        "batch": batch,
        "parity": parity,
        "noise": noise,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
