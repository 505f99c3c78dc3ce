"""Throughput of the Langevin (stochastic Heun) step on one MI355X.

Workload: pyrochlore lattice, Fd-3m conventional cell a = 10 A with 16 sites
per cell, 32x32x32 cells = 524,288 spins per replica, :dipole mode with
s = 3/2, g = 2; Heisenberg J1 = 1.0 and J2 = 0.1 meV, easy-axis
-0.05 (S.n_i)^2 along the local <111> axes; Langevin dt = 0.025,
damping = 0.2, kT = 0.3; random spins from the seed; device Philox noise.

Reported value: spin-updates/s of the FP32 batched step (R replicas advanced
by the same launches, inputs resident in HBM).  Beside it: the plain FP64
step on one replica, the dominant kernel's roofline from HIP events recorded
on the library's stream, the FP32-vs-FP64 energy deviation, and the CPU
baseline (the FP64 CPU port stepping independent replicas, one per thread).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--replicas R]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "sunny.jl_amd")]

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOAD = ("pyrochlore Fd-3m 16c a=10A, 32^3 cells x 16 = 524288 spins/replica, dipole s=3/2, "
            "J1=1.0 J2=0.1 meV, -0.05(S.n_111)^2, Langevin dt=0.025 damping=0.2 kT=0.3")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_time(handle, family):
    from sunnyhip import _lib
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    handle.call("snh_kernel_time", family.encode(), ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value


def timed_steps(obj, integ, steps, warmup, families=("dipole_predict", "dipole_correct"), kTs=None):
    """Warm up, then time exactly `steps` steps bracketed by synchronisation
    (whole chunks of steps replay the library's captured HIP graph).  Then
    the same number of steps again with HIP events recorded on the library's
    stream around every kernel (eager launches): per-kernel durations."""
    import sunnyhip as sh
    h = obj._device()
    if warmup:
        sh.step_(obj, integ, nsteps=warmup, kTs=kTs)
    h.call("snh_synchronize")
    t0 = time.perf_counter()
    sh.step_(obj, integ, nsteps=steps, kTs=kTs)
    h.call("snh_synchronize")
    t1 = time.perf_counter()
    h.call("snh_enable_timing", 1)
    sh.step_(obj, integ, nsteps=steps, kTs=kTs)
    fams = {f: kernel_time(h, f) for f in families}
    h.call("snh_enable_timing", 0)
    return (t1 - t0), fams


def measured_traffic(kernel, plan):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 FETCH_SIZE and WRITE_SIZE passes, each scaled by what the
    counter reads for the plain copy kernel of known size), or None when no
    summary for this device plan exists."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_pyrochlore32_r64_fp32.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    # the launch geometry of this kernel's plan must match (the autotuner's
    # ranks, and the other stage's plan, do not matter)
    if kernel == "dipole_predict":
        geom = ("layout", "tile_sp", "tile_T1", "tile_T2", "tile_threads", "tile_work", "tile_lds")
    else:
        geom = ("layout", "correct_sp", "correct_T1", "correct_T2", "correct_threads", "correct_work")
    have = d.get("plan") or {}
    if any(have.get(k) != plan.get(k) for k in geom) or kernel not in d.get("kernels", {}):
        return None
    k = d["kernels"][kernel]
    return k["fetch_bytes"] + k["write_bytes"]


def host_cpu():
    """Cores this process may use and the host CPU model."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    # a worker pool is sized to the lease's CPU share (OMP_NUM_THREADS where set)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"nproc": os.cpu_count(), "affinity": allowed, "threads": min(allowed, share) if share > 0 else allowed,
            "model": model}


def cpu_baseline(base_sys, seconds=15.0):
    """FP64 CPU port (the oracle) on the same 524,288-spin replica: one
    independent replica per thread, as Threads.@threads over clones.  A
    stand-in, not Sunny: Julia is not installed, so Sunny's own step! cannot
    run on this host."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O
    host = host_cpu()
    cores = max(1, host["threads"])
    spec = base_sys.spec()
    rng = np.random.default_rng(5)
    states = []
    for _ in range(cores):
        v = rng.standard_normal((base_sys.nsites, 3))
        states.append(1.5 * v / np.linalg.norm(v, axis=1, keepdims=True))
    oracles = [O.Oracle(spec) for _ in range(cores)]

    def one(k, nsteps):
        S = states[k]
        nz = np.random.default_rng(k).standard_normal((base_sys.nsites, 3))
        for _ in range(nsteps):
            S = oracles[k].step_langevin(S=S, noise=nz, dt=0.025, damping=0.2, kT=0.3)
        return nsteps

    t0 = time.perf_counter()
    one(0, 1)  # calibrate one step on one core
    t1 = time.perf_counter()
    per_step = t1 - t0
    nsteps = max(1, int(seconds / per_step))
    with ThreadPoolExecutor(cores) as ex:
        t0 = time.perf_counter()
        done = sum(ex.map(lambda k: one(k, nsteps), range(cores)))
        t1 = time.perf_counter()
    value = done * base_sys.nsites / (t1 - t0)
    return {"value": value, "unit": "spin-updates/s", "cores": cores, "kind": "port",
            "label": "stand-in, not Sunny (FP64 C restatement of step!; no Julia on this host)",
            "host": host,
            "sample": f"{cores} replicas x {nsteps} Langevin steps of the 524288-spin pyrochlore model, "
                      f"FP64 CPU port (oracle), one replica per thread, {t1 - t0:.1f} s"}


def fp32_vs_fp64(make, integ, checkpoints=(1, 10, 100), kappa=None, coherent=False, kTs=None):
    """Error of the FP32 device path against the FP64 device path (parity
    with the FP64 oracle to 1e-11 is what the GPU tests establish) on the
    same initial state and the same Philox noise: relative energy deviation,
    max | |S| - kappa | (or | |Z|^2 - kappa |) of the FP32 state, and the max
    per-site deviation, after each checkpoint's number of steps."""
    import sunnyhip as sh
    a, b = make(32), make(64)
    out, done = {}, 0
    for k in checkpoints:
        sh.step_(a, integ, nsteps=k - done, kTs=kTs)
        sh.step_(b, integ, nsteps=k - done, kTs=kTs)
        done = k
        Ea, Eb = np.atleast_1d(sh.energy(a)), np.atleast_1d(sh.energy(b))
        if hasattr(a, "sync_to_systems"):
            a.sync_to_systems()
            b.sync_to_systems()
            sa, sb = a.systems, b.systems
        else:
            sa, sb = [a], [b]
        if coherent:
            Za = np.concatenate([s.coherents_flat() for s in sa])
            Zb = np.concatenate([s.coherents_flat() for s in sb])
            kap = np.concatenate([np.repeat(s.kappas, s.ncells) for s in sa])
            norm = float(np.max(np.abs((np.abs(Za) ** 2).sum(1) - kap)))
            dev = float(np.max(np.abs(Za - Zb)))
        else:
            Sa = np.concatenate([s.dipoles_flat() for s in sa])
            Sb = np.concatenate([s.dipoles_flat() for s in sb])
            kap = np.concatenate([np.repeat(s.kappas, s.ncells) for s in sa])
            norm = float(np.max(np.abs(np.linalg.norm(Sa, axis=1) - kap)))
            dev = float(np.max(np.abs(Sa - Sb)))
        out[f"step_{k}"] = {"rel_energy": float(np.max(np.abs(Ea - Eb) / np.abs(Eb))), "max_norm_error": norm,
                            "max_site_deviation": dev}
    return out


def bench_sun3(args):
    """Auxiliary line: 64 SU(3) triangular replicas, L_r uniform in 32..256,
    kT geometric 0.1..2.0, Langevin dt = 0.02, damping 0.1, FP32 batch."""
    import sunnyhip as sh
    systems, kTs = sh.workloads.sun3_tempering_batch(seed=0)
    batch = sh.ReplicaBatch(systems, precision=32)
    integ = sh.Langevin(0.02, damping=0.1, kT=0.0)
    dt, fams = timed_steps(batch, integ, args.steps, args.warmup, ("sun_predict", "sun_correct"), kTs)
    ups = batch.nsites * args.steps / dt
    # Schroedinger midpoint (deterministic) on the same batch
    mp = sh.ImplicitMidpoint(0.02, tol=1e-5)
    h = batch._device()
    sh.step_(batch, mp, nsteps=2)
    h.call("snh_synchronize")
    msteps = 20
    t0 = time.perf_counter()
    its = sh.step_(batch, mp, nsteps=msteps)
    h.call("snh_synchronize")
    dtm = time.perf_counter() - t0
    mid = {"value": batch.nsites * msteps / dtm, "steps": msteps, "iterations": int(np.max(its)), "tol": 1e-5}
    log(f"SU(3) midpoint: {mid['value']:.3e} spin-updates/s, {mid['iterations']} iterations")
    per = {f: ms / max(n, 1) for f, (ms, n) in fams.items()}
    log(f"SU(3) batch: {batch.nsites} sites, {dt * 1e3 / args.steps:.3f} ms/step, {ups:.3e} spin-updates/s, {per}")
    # FP32 against FP64 on the first four replicas (uneven sizes), same noise
    sub = systems[:4]

    def sub_batch(prec):
        return sh.ReplicaBatch([sh.clone_system(x) for x in sub], precision=prec)

    err = fp32_vs_fp64(sub_batch, sh.Langevin(0.02, damping=0.1, kT=0.0), coherent=True, kTs=kTs[:4])
    log(f"SU(3) FP32 vs FP64: {err}")
    print(json.dumps({"metric": "spin-updates/s (auxiliary: SU(3) tempering batch)", "value": ups,
                      "unit": "spin-updates/s", "ms_per_step": dt * 1e3 / args.steps, "steps": args.steps,
                      "dtype": "f32", "kernels_ms": per, "implicit_midpoint": mid, "fp32_vs_fp64": err,
                      "config": {"workload": "64 SU(3) triangular replicas, L uniform 32..256 (seed 0), "
                                             f"{batch.nsites} sites, kT 0.1..2.0 geometric, dt 0.02, damping 0.1"}}),
          flush=True)


def bench_cubic16(args):
    """Auxiliary line: simple cubic 16^3 = 4096 spins, :dipole s = 1, g = 2,
    NN J = 1 and -0.1 Sz^2, plain FP64; Langevin(0.05; damping 0.1, kT 0.5)
    and ImplicitMidpoint(0.05; damping 0.1), 1000 steps each (launch- and
    latency-bound); beside them the FP64 CPU port on one core."""
    import sunnyhip as sh
    from oracle import oracle as O
    nst = 1000
    s = sh.workloads.cubic_model(L=16, seed=0, precision=64)
    lang = sh.Langevin(0.05, damping=0.1, kT=0.5)
    dt_l, _ = timed_steps(s, lang, nst, 50)
    mp = sh.ImplicitMidpoint(0.05, damping=0.1)
    sh.step_(s, mp, nsteps=20)
    h = s._device()
    h.call("snh_synchronize")
    t0 = time.perf_counter()
    its = sh.step_(s, mp, nsteps=nst)
    h.call("snh_synchronize")
    dt_m = time.perf_counter() - t0
    o = O.Oracle(s.spec())
    # FP64 device against the FP64 oracle from one state with the same draws
    c = sh.workloads.cubic_model(L=16, seed=5, precision=64)
    S_ref, done, vs_oracle = c.dipoles_flat(), 0, {}
    key = O.replica_key(c.noise_seed, 0)
    for k in (1, 10, 100):
        S_ref, _ = o.langevin_run(S=S_ref, dt=0.05, damping=0.1, kT=0.5, key=key, step0=done, nsteps=k - done,
                                  every=0)
        sh.step_(c, lang, nsteps=k - done)
        done = k
        Sd = c.dipoles_flat()
        vs_oracle[f"step_{k}"] = {"rel_energy": abs(sh.energy(c) - o.energy(S_ref)) / abs(o.energy(S_ref)),
                                  "max_norm_error": float(np.max(np.abs(np.linalg.norm(Sd, axis=1) - 1.0))),
                                  "max_site_deviation": float(np.max(np.abs(Sd - S_ref)))}
    log(f"cubic 16^3 FP64 device vs oracle: {vs_oracle}")
    S = s.dipoles_flat()
    nz = np.random.default_rng(1).standard_normal((s.nsites, 3))
    ncpu = 200
    t0 = time.perf_counter()
    for _ in range(ncpu):
        S = o.step_langevin(S=S, noise=nz, dt=0.05, damping=0.1, kT=0.5)
    dt_c = time.perf_counter() - t0
    out = {"langevin": s.nsites * nst / dt_l, "midpoint": s.nsites * nst / dt_m,
           "midpoint_iterations_last": int(its[0]), "cpu_port_langevin_1core": s.nsites * ncpu / dt_c,
           "fp64_vs_oracle": vs_oracle,
           "ms_per_step": {"langevin": dt_l * 1e3 / nst, "midpoint": dt_m * 1e3 / nst}}
    log(f"cubic 16^3 FP64: Langevin {out['langevin']:.3e}, midpoint {out['midpoint']:.3e} spin-updates/s "
        f"({out['ms_per_step']}); CPU port {out['cpu_port_langevin_1core']:.3e}")
    print(json.dumps({"metric": "spin-updates/s (auxiliary: cubic 16^3 plain FP64)", "unit": "spin-updates/s",
                      "dtype": "f64", "steps": nst, **out,
                      "config": {"workload": "simple cubic 16^3 = 4096 spins, dipole s=1, J=1, -0.1 Sz^2, "
                                             "Langevin dt 0.05 damping 0.1 kT 0.5; ImplicitMidpoint dt 0.05 "
                                             "damping 0.1"}}), flush=True)


def bench_adjoint(args):
    """Auxiliary line: simple cubic 64^3 = 262,144 spins, plain FP64, labelled
    parameters J (NN exchange) and D (D Sz^2) at (0.9, -0.08);
    ImplicitMidpoint(0.05) with damping 0 for 500 steps, checkpoints every
    50; loss = mean over the 8^3 grid q = 2 pi m / 8 of (S(q) - S_target(q))^2
    with the target from (1.0, -0.1).  Times the recorded forward pass, the
    host loss and the reverse sweep (which recomputes every segment)."""
    import ctypes

    import sunnyhip as sh
    L, nq, nsteps, dt, every = 64, 8, 500, 0.05, 50
    dims = (L, L, L)
    ref = sh.workloads.cubic_model(L=L, J=1.0, D=-0.1, seed=11, params=True)
    sh.step_(ref, sh.ImplicitMidpoint(dt), nsteps=20)
    target, _ = sh.structure_factor(ref.dipoles_flat(), dims, nq)
    s = sh.workloads.cubic_model(L=L, J=0.9, D=-0.08, seed=11, params=True)
    h = s._device()
    h.call("snh_synchronize")
    t0 = time.perf_counter()
    h.call("snh_midpoint_record", dt, 1e-12, 100, nsteps, every)
    h.call("snh_synchronize")
    t1 = time.perf_counter()
    s._dev_dirty = True
    loss, dLdS = sh.structure_factor_loss(s.dipoles_flat(), dims, target, nq)
    t2 = time.perf_counter()
    specs = [sh.param_spec(s, lab) for lab in ("J", "D")]
    from sunnyhip import _lib
    packed = (_lib.SnhSpec * 2)(*[_lib.pack_spec(sp) for sp in specs])
    grad = np.zeros(2)
    h.call("snh_midpoint_adjoint", 2, ctypes.cast(packed, ctypes.POINTER(_lib.SnhSpec)),
           _lib.dptr(np.ascontiguousarray(dLdS, float)), _lib.dptr(grad), None)
    h.call("snh_synchronize")
    t3 = time.perf_counter()
    out = {"forward_record_s": t1 - t0, "loss_host_s": t2 - t1, "reverse_s": t3 - t2,
           "forward_spin_updates_per_s": s.nsites * nsteps / (t1 - t0),
           "gradient_per_s": 1.0 / (t3 - t0), "loss": loss, "grad": grad.tolist()}
    log(f"adjoint 64^3 x {nsteps}: forward {t1 - t0:.3f} s, loss {t2 - t1:.3f} s, reverse {t3 - t2:.3f} s, "
        f"grad {grad}")
    print(json.dumps({"metric": "trajectory gradients/s (auxiliary: midpoint adjoint, cubic 64^3 x 500 steps)",
                      "unit": "gradients/s", "value": out["gradient_per_s"], "dtype": "f64", **out,
                      "config": {"workload": "simple cubic 64^3 = 262144 spins, J=0.9 D=-0.08 labelled, "
                                             "ImplicitMidpoint dt 0.05 damping 0, 500 steps, checkpoints every 50, "
                                             "structure-factor loss on the 8^3 q grid"}}), flush=True)


def bench_corr(args):
    """Auxiliary line: sampled dynamical correlations (SampledCorrelations,
    ssf observables, six pairs) of pyrochlore 8^3 x 16 = 8192 spins, plain
    FP64, energies 0..4 in 100 steps (198 time slots); one sample = the
    midpoint trajectory plus the correlation accumulation; beside it the
    numpy restatement of the accumulation on the host."""
    import sunnyhip as sh
    from oracle.correlations_ref import SampledCorrelationsRef
    s = sh.workloads.pyrochlore_model(L=8, seed=3, precision=64)
    sh.step_(s, sh.Langevin(0.02, damping=0.2, kT=0.3), nsteps=200)
    sc = sh.SampledCorrelations(s, energies=np.linspace(0.0, 4.0, 100), dt=0.05)
    h = s._device()
    sh.add_sample_(sc, s)  # warm-up sample
    h.call("snh_synchronize")
    h.call("snh_enable_timing", 1)
    t0 = time.perf_counter()
    sh.add_sample_(sc, s)
    h.call("snh_synchronize")
    t1 = time.perf_counter()
    traj_ms, _ = kernel_time(h, "midpoint_iter")
    h.call("snh_enable_timing", 0)
    nsteps = sc.measperiod * (sc.n_all_omega // 2)
    # host restatement of accum_sample! on the same sizes (zero trajectory)
    ref = SampledCorrelationsRef(s.dims, s.crystal.natoms, sc.observables, sc.corr_pairs, sc.n_all_omega)
    buf = np.zeros((3, *s.dims, s.crystal.natoms, sc.n_all_omega), complex)
    t2 = time.perf_counter()
    ref.accum_sample(buf)
    t3 = time.perf_counter()
    out = {"sample_s": t1 - t0, "trajectory_steps": nsteps, "measperiod": sc.measperiod,
           "n_all_omega": sc.n_all_omega, "midpoint_iter_kernel_ms": traj_ms,
           "host_numpy_accumulation_s": t3 - t2, "value": 1.0 / (t1 - t0)}
    log(f"correlations pyrochlore 8^3: sample {t1 - t0:.3f} s ({nsteps} midpoint steps), host numpy "
        f"accumulation alone {t3 - t2:.3f} s")
    print(json.dumps({"metric": "samples/s (auxiliary: sampled correlations, pyrochlore 8^3 x 16)",
                      "unit": "samples/s", "dtype": "f64", **out,
                      "config": {"workload": "pyrochlore 8^3 x 16 = 8192 spins, FP64, ssf observables, "
                                             "energies 0..4 (100), dt 0.05 adjusted, ImplicitMidpoint"}}), flush=True)


def bench_local(args):
    """Auxiliary line: LocalSampler (Metropolis trials, propose_delta(0.3),
    kT 0.3) on the headline pyrochlore: one replica in FP32 and in FP64, and a
    batch of FP32 replicas.  trials/s counts the expected nsweeps * nsites
    trials of a call.  Beside it the oracle's scalar restatement on a bounded
    sample (pyrochlore 8^3 x 16 = 8192 spins, same colour classes)."""
    import sunnyhip as sh
    from oracle import oracle as O
    out = {}
    cases = (("fp32_r1", 32, 1), ("fp64_r1", 64, 1), (f"fp32_r{args.local_replicas}", 32, args.local_replicas))
    for label, prec, R in cases:
        systems = [sh.workloads.pyrochlore_model(L=args.L, seed=50 + r, precision=prec) for r in range(R)]
        obj = systems[0] if R == 1 else sh.ReplicaBatch(systems, precision=prec)
        sampler = sh.LocalSampler(kT=0.3, propose=sh.propose_delta(0.3))
        h = obj._device()
        sh.step_(obj, sampler, nsteps=2)
        h.call("snh_synchronize")
        t0 = time.perf_counter()
        sh.step_(obj, sampler, nsteps=args.local_sweeps)
        h.call("snh_synchronize")
        t1 = time.perf_counter()
        h.call("snh_enable_timing", 1)
        sh.step_(obj, sampler, nsteps=args.local_sweeps)
        ms, n = kernel_time(h, "local_color")
        h.call("snh_enable_timing", 0)
        m, nc, _ = sh.local_coloring(obj)
        ns = sum(s.nsites for s in systems)
        rate = ns * args.local_sweeps / (t1 - t0)
        out[label] = {"sites": ns, "ms_per_sweep": (t1 - t0) * 1e3 / args.local_sweeps, "value": rate,
                      "colors": nc, "period": list(m), "color_kernel_ms": ms / max(n, 1),
                      "color_launches_per_sweep": n / args.local_sweeps}
        log(f"LocalSampler {label}: {ns} sites, {nc} colours, {(t1 - t0) * 1e3 / args.local_sweeps:.3f} ms/sweep, "
            f"{rate:.3e} trials/s, colour kernel {ms / max(n, 1):.4f} ms")
        del obj, systems
    # the Ewald form: sequential trials at random sites, chunks of 64 per launch
    for prec in (32, 64):
        e = sh.workloads.pyrochlore_model(L=16, J1=0.05, J2=0.0, A=0.0, s=3.5, seed=40, precision=prec,
                                          dipolar=True, mu0_muB2=0.6745817653324668, demag=1 / 3)
        h = e._device()
        trials = args.local_ewald_trials
        sampler = sh.LocalSampler(kT=0.3, nsweeps=trials / e.nsites, propose=sh.propose_delta(0.3))
        sh.step_(e, sampler)
        h.call("snh_synchronize")
        dE0 = sampler.dE
        t0 = time.perf_counter()
        sh.step_(e, sampler)
        h.call("snh_synchronize")
        dt = time.perf_counter() - t0
        out[f"ewald_fp{prec}_r1"] = {"sites": e.nsites, "trials": trials, "us_per_trial": dt * 1e6 / trials,
                                     "value": trials / dt, "dE": sampler.dE - dE0}
        log(f"LocalSampler Ewald fp{prec}: {e.nsites} sites, {trials} sequential trials, "
            f"{dt * 1e6 / trials:.3f} us/trial")
        del e, h
    # a batch of replicas (parallel tempering's shape), FP32
    Rb = args.ewald_replicas
    systems = [sh.workloads.pyrochlore_model(L=16, J1=0.05, J2=0.0, A=0.0, s=3.5, seed=40 + r, precision=32,
                                             dipolar=True, mu0_muB2=0.6745817653324668, demag=1 / 3)
               for r in range(Rb)]
    b = sh.ReplicaBatch(systems, precision=32)
    h = b._device()
    trials = args.local_ewald_trials // 4
    sampler = sh.LocalSampler(kT=0.3, nsweeps=trials / systems[0].nsites, propose=sh.propose_delta(0.3))
    kTs = np.geomspace(0.1, 2.0, Rb)
    sh.step_(b, sampler, kTs=kTs)
    h.call("snh_synchronize")
    t0 = time.perf_counter()
    sh.step_(b, sampler, kTs=kTs)
    h.call("snh_synchronize")
    dt = time.perf_counter() - t0
    out[f"ewald_fp32_r{Rb}"] = {"sites": Rb * systems[0].nsites, "trials": Rb * trials,
                                "us_per_trial": dt * 1e6 / (Rb * trials), "value": Rb * trials / dt}
    log(f"LocalSampler Ewald fp32 batch: {Rb} x {systems[0].nsites} sites, {trials} sequential trials each, "
        f"{dt * 1e6 / (Rb * trials):.3f} us per trial, {Rb * trials / dt:.3e} trials/s")
    del b, h, systems
    small = sh.workloads.pyrochlore_model(L=8, seed=1, precision=64)
    m, nc, col = sh.local_coloring(small)
    o = O.Oracle(small.spec())
    S = small.dipoles_flat()
    sweeps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < args.cpu_seconds / 3:
        S, _ = o.local_sweeps(S, m, col.reshape(-1), nc, [(2, 0.3, 1.0)], 0.3, 7, sweeps, 1.0)
        sweeps += 1
    cpu = small.nsites * sweeps / (time.perf_counter() - t0)
    log(f"LocalSampler CPU oracle: {cpu:.3e} trials/s (1 core, {sweeps} sweeps of {small.nsites})")
    print(json.dumps({"metric": "Metropolis trials/s (auxiliary: LocalSampler, pyrochlore 32^3 x 16)",
                      "unit": "trials/s", "sweeps": args.local_sweeps, "by_case": out,
                      "cpu_baseline": {"value": cpu, "unit": "trials/s", "cores": 1, "kind": "port",
                                       "sample": f"oracle colour passes, pyrochlore 8^3 x 16, {sweeps} sweeps"},
                      "config": {"workload": "pyrochlore 32^3 x 16 = 524288 spins/replica, dipole s=3/2, "
                                             "J1 1.0 J2 0.1 A -0.05, LocalSampler kT 0.3 propose_delta(0.3); "
                                             "Ewald cases: pyrochlore 16^3 x 16 = 65536 spins, s=7/2, J1=0.05, "
                                             "Ewald mu0muB2=0.6745817653324668 demag 1/3"}}),
          flush=True)


def bench_ewald(args):
    """Auxiliary line: pyrochlore 16^3 cells x 16 = 65,536 spins, s = 7/2,
    J1 = 0.05 meV, Ewald dipole-dipole (mu0 muB^2 = 0.6745817653324668,
    demag 1/3), FP32 Langevin dt = 0.025, damping 0.2, kT 0.3; one replica
    and a batch sharing the transformed Ewald tensor."""
    import sunnyhip as sh
    out = {}
    for R in (1, args.ewald_replicas):
        systems = [sh.workloads.pyrochlore_model(L=16, J1=0.05, J2=0.0, A=0.0, s=3.5, seed=40 + r, precision=32,
                                                 dipolar=True, mu0_muB2=0.6745817653324668, demag=1 / 3)
                   for r in range(R)]
        obj = systems[0] if R == 1 else sh.ReplicaBatch(systems, precision=32)
        integ = sh.Langevin(0.025, damping=0.2, kT=0.3)
        dt, fams = timed_steps(obj, integ, args.steps, args.warmup, ("dipole_predict", "dipole_correct", "ewald_field"))
        ns = sum(s.nsites for s in systems)
        per = {f: ms / max(n, 1) for f, (ms, n) in fams.items()}
        ups = ns * args.steps / dt
        log(f"Ewald R={R}: {ns} sites, {dt * 1e3 / args.steps:.3f} ms/step, {ups:.3e} spin-updates/s, {per}")
        out[R] = {"ms_per_step": dt * 1e3 / args.steps, "value": ups, "kernels_ms": per}
        del obj, systems

    def ewald_model(prec):
        return sh.workloads.pyrochlore_model(L=16, J1=0.05, J2=0.0, A=0.0, s=3.5, seed=40, precision=prec,
                                             dipolar=True, mu0_muB2=0.6745817653324668, demag=1 / 3)

    err = fp32_vs_fp64(ewald_model, sh.Langevin(0.025, damping=0.2, kT=0.3))
    # the gradient itself at the initial state: FP32 against FP64
    g32, g64 = sh.energy_grad_dipoles(ewald_model(32)), sh.energy_grad_dipoles(ewald_model(64))
    err["grad_initial_rel"] = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
    log(f"Ewald FP32 vs FP64: {err}")
    print(json.dumps({"metric": "spin-updates/s (auxiliary: Ewald pyrochlore 16^3)", "unit": "spin-updates/s",
                      "dtype": "f32", "steps": args.steps, "by_replicas": out, "fp32_vs_fp64": err,
                      "config": {"workload": "pyrochlore 16^3 cells x 16 = 65536 spins, s=7/2, J1=0.05, "
                                             "Ewald mu0muB2=0.6745817653324668 demag 1/3, Langevin dt 0.025"}}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--local-ewald-trials", type=int, default=65536,
                    help="sequential Metropolis trials timed on the Ewald case of --workload local")
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--layout", choices=["auto", "ghost"], default="auto",
                    help="device gather layout (auto: LDS tiles when they fit)")
    ap.add_argument("--fp32-only", action="store_true", help="time only the FP32 batch (tuning runs)")
    ap.add_argument("--fp64-only", action="store_true", help="time only the plain FP64 step (tuning runs)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the cost model's first tile plan")
    ap.add_argument("--tune-candidates", type=int, default=8, help="tile plans the autotuner times")
    ap.add_argument("--kT", type=float, default=0.3, help="temperature of the timed runs (tuning: 0 disables noise)")
    ap.add_argument("--workload", choices=["pyrochlore", "sun3", "ewald", "cubic16", "adjoint", "corr", "local"],
                    default="pyrochlore",
                    help="pyrochlore: the headline batch; auxiliary lines: sun3 (64 uneven SU(3) tempering "
                         "replicas), ewald (dipolar pyrochlore 16^3), cubic16 (plain FP64 16^3), adjoint "
                         "(midpoint trajectory gradient, cubic 64^3 x 500 steps), corr (sampled correlations), local "
                         "(LocalSampler Metropolis sweeps)")
    ap.add_argument("--ewald-replicas", type=int, default=32, help="batch size of the Ewald auxiliary line")
    ap.add_argument("--local-replicas", type=int, default=16, help="batch size of the LocalSampler line")
    ap.add_argument("--local-sweeps", type=int, default=20, help="timed sweeps of the LocalSampler line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (replicas shard across GPUs "
              "as independent processes; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import sunnyhip as sh
    if args.workload == "sun3":
        return bench_sun3(args)
    if args.workload == "ewald":
        return bench_ewald(args)
    if args.workload == "cubic16":
        return bench_cubic16(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "corr":
        return bench_corr(args)
    if args.workload == "local":
        return bench_local(args)
    t0 = time.perf_counter()
    base = sh.workloads.pyrochlore_model(L=args.L, seed=0, precision=32)
    log(f"model built in {time.perf_counter() - t0:.1f}s: {base.nsites} spins, {base.spec()['nbonds']} directed bonds")
    integ = sh.Langevin(0.025, damping=0.2, kT=args.kT)
    R = args.replicas

    layout = 1 if args.layout == "ghost" else 0
    if args.fp64_only:
        plain = sh.clone_system(base)
        plain.precision = 64
        plain.layout = layout
        h = plain._device()
        plan = {k: h.query(k) for k in ("layout", "tile_sp", "tile_T1", "tile_T2", "tile_threads", "tile_work")}
        steps64 = max(10, args.steps // 4)
        dt64, f64 = timed_steps(plain, integ, steps64, args.warmup)
        log(f"FP64 plain {plan}: {dt64 * 1e3 / steps64:.4f} ms/step, {plain.nsites * steps64 / dt64:.3e} spin-updates/s")
        return

    # FP32 batched
    reps = [base] + [sh.clone_system(base, seed=100 + r) for r in range(1, R)]
    for r in range(1, R):
        sh.randomize_spins_(reps[r])
    batch = sh.ReplicaBatch(reps, precision=32, layout=layout)
    hb = batch._device()
    if not args.no_autotune:
        tune_ms = hb.autotune(candidates=args.tune_candidates, launches=3)
        log(f"autotune: kept plan ranks {hb.query('tile_tuned')} (predictor), {hb.query('tile_tuned_correct')} "
            f"(corrector) of {hb.query('tile_candidates')}, {tune_ms:.4f} ms per predictor + corrector")
    plan = {k: hb.query(k) for k in ("layout", "tile_sp", "tile_T1", "tile_T2", "tile_threads", "tile_work",
                                     "tile_lds", "tile_tuned", "correct_sp", "correct_T1", "correct_T2",
                                     "correct_threads", "correct_work", "tile_tuned_correct", "pair_on")}
    log(f"FP32 batch: {R} replicas, {batch.nsites} spins, plan {plan}")
    dt32, fams32 = timed_steps(batch, integ, args.steps, args.warmup)
    ups32 = batch.nsites * args.steps / dt32
    log(f"FP32 batched: {dt32 * 1e3 / args.steps:.3f} ms/step, {ups32:.3e} spin-updates/s, "
        f"predict {fams32['dipole_predict'][0] / max(fams32['dipole_predict'][1], 1):.4f} ms, "
        f"correct {fams32['dipole_correct'][0] / max(fams32['dipole_correct'][1], 1):.4f} ms")
    if args.fp32_only:
        # the plain copy kernel runs here too, so that counter passes over this
        # short run see the calibration launches beside the stencil kernels
        nb_ = ctypes.c_int64()
        hb.call("snh_copy_calibration", 5, ctypes.byref(nb_))
        return

    # FP64 plain (one replica, same initial state as replica 0)
    plain = sh.clone_system(base)
    plain.precision = 64
    plain.layout = layout
    steps64 = max(10, args.steps // 4)
    dt64, fams64 = timed_steps(plain, integ, steps64, args.warmup)
    ups64 = plain.nsites * steps64 / dt64
    log(f"FP64 plain: {dt64 * 1e3 / steps64:.3f} ms/step, {ups64:.3e} spin-updates/s")

    # FP64 batched: the same replicas in FP64 (no longer cache-resident), to
    # separate the effect of precision from that of the working-set size
    batch64 = sh.ReplicaBatch(reps, precision=64, layout=layout)
    steps64b = max(10, args.steps // 8)
    dt64b, _ = timed_steps(batch64, integ, steps64b, min(args.warmup, 5))
    ups64b = batch64.nsites * steps64b / dt64b
    log(f"FP64 batched: {dt64b * 1e3 / steps64b:.3f} ms/step, {ups64b:.3e} spin-updates/s")
    del batch64

    # ImplicitMidpoint (deterministic, damping 0) on the same batch and plain
    # system: spin-updates/s = spins x steps / time; iterations reported
    mp = sh.ImplicitMidpoint(0.025, damping=0.0, tol=1e-5)
    msteps = max(4, args.steps // 5)
    sh.step_(batch, mp, nsteps=2)
    hb.call("snh_synchronize")
    t0 = time.perf_counter()
    it32 = sh.step_(batch, mp, nsteps=msteps)
    hb.call("snh_synchronize")
    dtm32 = time.perf_counter() - t0
    hb.call("snh_enable_timing", 1)
    sh.step_(batch, mp, nsteps=msteps)
    mi_ms, mi_n = kernel_time(hb, "midpoint_iter")
    hb.call("snh_enable_timing", 0)
    mp64 = sh.ImplicitMidpoint(0.025, damping=0.0, tol=1e-10)
    sh.step_(plain, mp64, nsteps=2)
    hp = plain._device()
    hp.call("snh_synchronize")
    t0 = time.perf_counter()
    it64 = sh.step_(plain, mp64, nsteps=msteps)
    hp.call("snh_synchronize")
    dtm64 = time.perf_counter() - t0
    midpoint = {"fp32_batched": batch.nsites * msteps / dtm32, "fp64_plain": plain.nsites * msteps / dtm64,
                "unit": "spin-updates/s", "steps": msteps, "iterations_fp32": int(np.max(it32)),
                "iterations_fp64": int(np.max(it64)), "tol_fp32": 1e-5, "tol_fp64": 1e-10,
                "group_reruns_fp32": int(hb.query("mid_reruns")), "group_reruns_fp64": int(hp.query("mid_reruns")),
                # one fixed-point sweep reads Shat (gathered), S and the previous mu;
                # writes the next Shat and mu (the iterate S'' = 2 mu Shat - S stays
                # implicit).  Time per working iteration: all launches (including
                # the budget's launches after convergence, which return at once)
                # over iterations x steps
                "iter_kernel_ms": mi_ms / max(int(np.max(it32)) * msteps, 1), "iter_kernel_launches": mi_n,
                "iter_kernel_bytes_per_spin": 44}
    midpoint["iter_kernel_GBs"] = 44 * batch.nsites / (midpoint["iter_kernel_ms"] * 1e-3) / 1e9
    midpoint["speedup"] = midpoint["fp32_batched"] / midpoint["fp64_plain"]
    log(f"midpoint: FP32 batched {midpoint['fp32_batched']:.3e} ({midpoint['iterations_fp32']} it), "
        f"FP64 plain {midpoint['fp64_plain']:.3e} ({midpoint['iterations_fp64']} it), "
        f"iter kernel {midpoint['iter_kernel_ms']:.4f} ms = {midpoint['iter_kernel_GBs']:.0f} GB/s")

    # FP32 vs FP64 on the same start and noise (one 524,288-spin replica)
    def one_replica(prec):
        c = sh.clone_system(base)
        c.precision = prec
        return c

    err32 = fp32_vs_fp64(one_replica, integ)
    log(f"FP32 vs FP64: {err32}")

    # plain copy of the batch's state planes: calibrates traffic counters
    # (profiles/) and gives the achievable streaming bandwidth
    hb.call("snh_enable_timing", 1)
    nbytes = ctypes.c_int64()
    hb.call("snh_copy_calibration", 20, ctypes.byref(nbytes))
    cms, cn = kernel_time(hb, "copy_planes")
    hb.call("snh_enable_timing", 0)
    copy_gbs = nbytes.value / (cms / cn * 1e-3) / 1e9
    log(f"copy kernel: {nbytes.value / 1e6:.1f} MB per launch, {copy_gbs:.0f} GB/s")

    # roofline of the batched Heun step (predictor + corrector launches):
    # compulsory HBM bytes per spin-update are read S + write S = 24 B in
    # FP32; the two-launch Heun moves 36 B per spin in each stage (the
    # predictor reads S, writes S' and H; the corrector reads S' and H,
    # writes S), reported per stage below
    ms_p, n_p = fams32["dipole_predict"]
    ms_c, n_c = fams32["dipole_correct"]
    avg_p, avg_c = ms_p / max(n_p, 1), ms_c / max(n_c, 1)
    pair_ms = avg_p + avg_c
    alg_bytes = 24 * batch.nsites
    achieved = alg_bytes / (pair_ms * 1e-3) / 1e9
    tr_p, tr_c = measured_traffic("dipole_predict", plan), measured_traffic("dipole_correct", plan)
    traffic = tr_p + tr_c if tr_p is not None and tr_c is not None else None
    step_ms = dt32 * 1e3 / args.steps
    step_achieved = 24 * batch.nsites / (step_ms * 1e-3) / 1e9
    stages = {k: {"avg_launch_ms": a, "bytes_per_spin": 36, "achieved": 36 * batch.nsites / (a * 1e-3) / 1e9,
                  "traffic": t}
              for k, a, t in (("dipole_predict", avg_p, tr_p), ("dipole_correct", avg_c, tr_c))}

    line = {
        "metric": "spin-updates/s, Langevin midpoint step, FP32 batched vs FP64 plain; % HBM peak",
        "value": ups32,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "replicas": R, "spins_per_replica": base.nsites,
                   "integrator": "Langevin (stochastic Heun), device Philox noise", "device_plan": plan},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_over_algorithmic": traffic / alg_bytes if traffic else None,
                     "kernel": "dipole_predict + dipole_correct (one Heun step)", "bytes_per_spin_update": 24,
                     "algorithmic_bytes": alg_bytes, "avg_launch_ms": pair_ms},
        "roofline_stages": stages,
        "roofline_wallclock": {"bytes_per_spin_update": 24, "achieved": step_achieved, "unit": "GB/s",
                               "frac": step_achieved / HBM_PEAK_GBS},
        "kernels_ms": {"dipole_predict": avg_p, "dipole_correct": avg_c},
        "copy_kernel": {"bytes_per_launch": nbytes.value, "GB/s": copy_gbs},
        "fp64_plain": {"value": ups64, "unit": "spin-updates/s", "ms_per_step": dt64 * 1e3 / steps64,
                       "steps": steps64, "dtype": "f64"},
        "fp64_batched": {"value": ups64b, "unit": "spin-updates/s", "ms_per_step": dt64b * 1e3 / steps64b,
                         "steps": steps64b, "dtype": "f64", "replicas": R},
        "speedup_fp32_batched_vs_fp64_plain": ups32 / ups64,
        "speedup_fp32_batched_vs_fp64_batched": ups32 / ups64b,
        "fp32_vs_fp64": err32,
        "implicit_midpoint": midpoint,
        "cpu_baseline": None,
    }
    if not args.skip_cpu:
        log("CPU baseline ...")
        line["cpu_baseline"] = cpu_baseline(base, args.cpu_seconds)
        log(f"CPU: {line['cpu_baseline']['value']:.3e} spin-updates/s on {line['cpu_baseline']['cores']} cores")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
