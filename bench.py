#!/usr/bin/env python3
"""Lanczos (SIL) steps/s of the MCTDHB configurational core at N=100, M=5.

One step is one short-iterative-Lanczos coefficient step (reference
coefficients_integration.c:59-160): m = 6 Krylov vectors, 6 H.C
applications, full reorthogonalisation, the tridiagonal exponential and the
recombination, real time dt = 0.01, on NC = C(104, 4) = 4,598,126
configurations with seeded complex-normal coefficients, a seeded Hermitian
one-body matrix (scale 1) and a seeded symmetrised two-body tensor (scale 0.1).
All inputs are resident in HBM before timing starts.

Prints one JSON line.  `value` is the mixed-precision form (complex64 H.C and
density sums, FP64 Lanczos scalars); the FP64 forms, their ratio, the error
of the mixed path against FP64, p50/p99 step latency, the H.C kernel's
roofline, the whole MCTDHB step at N=10, M=4, mixed against FP64 on
(10,4), (200,4), (30,7) and (100,5) with m = 8, and a CPU baseline (the
oracle's restatement of the reference, timed on this host) ride along.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

# More hardware queues than HIP's default 4 let the batched and served
# sections' independent solves (one stream each) run side by side; HIP reads
# this when it initialises, so it is set before torch is imported.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "1d-mctdhb_amd"))
sys.path.insert(0, ROOT)

METRIC = "Lanczos steps/s at N=100,M=5 (mixed vs FP64) at stated error; p50/p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def algorithmic_bytes(npar: int, norb: int, bc: int, lm: int) -> dict:
    """Per-launch algorithmic bytes: coefficient stream in and out, the
    occupation row, and each used jump-table entry once (the work-normalised
    figure of a table-driven H.C); vector passes of the SIL step on top."""
    N, M = npar, norb
    nc = math.comb(N + M - 1, M - 1)
    nc3 = math.comb(N + M - 3, M - 1)
    j1 = M * (M - 1) * math.comb(N + M - 2, M - 1)
    j2s = M * nc3 * M * (M - 1) // 2
    j2d = math.comb(M, 2) * nc3 * (M - 1) * (M - 2) // 2
    hc = 2 * nc * bc + 2 * M * nc + 4 * (j1 + j2s + j2d)
    V = 16 * nc
    sil = lm * hc + V * sum(2 * i + 5 for i in range(lm - 1)) + (lm + 1) * V
    return {"hc": hc, "sil": sil, "coef_only": 2 * nc * bc}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def physical_cores() -> int:
    """Physical cores among the CPUs this process may run on: distinct
    (package, core) pairs of /proc/cpuinfo over the affinity mask."""
    try:
        allowed = os.sched_getaffinity(0)
    except AttributeError:
        allowed = set(range(os.cpu_count() or 1))
    cores, cpu, pkg = set(), None, 0
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                k, _, v = ln.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    cpu, pkg = int(v), 0
                elif k == "physical id":
                    pkg = int(v)
                elif k == "core id" and cpu in allowed:
                    cores.add((pkg, int(v)))
    except OSError:
        pass
    return len(cores) or len(allowed)


def cpu_baseline(npar, norb, lm, dt, steps, seed):
    """The oracle (the reference's algorithm restated in C/OpenMP, its
    table-driven H.C) on this host, timed around each SIL step with a
    monotonic clock as the reference times its steps (driver.c:251, :284):
    first on every physical core of this process's CPU set, then on half of
    them, the reference's own default (apps/mctdhb_main.c:138)."""
    import oracle
    from mctdhb_amd import inputs

    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        oracle.build()
    phys = physical_cores()
    S = oracle.Space(npar, norb)
    c0 = inputs.coefficients(S.dim, seed)
    h, g = inputs.onebody(norb, seed), inputs.twobody(norb, seed, 0.1)
    runs = {}
    for label, threads in (("all", phys), ("half", max(1, phys // 2))):
        oracle.set_threads(threads)
        c = c0
        per = []
        for _ in range(steps):
            t0 = time.perf_counter()
            _, c = S.sil_step(c, h, g, lm, dt, False)
            per.append(time.perf_counter() - t0)
        runs[label] = {"threads": threads, "steps_per_s": 1.0 / float(np.median(per)),
                       "step_s": [float(v) for v in per]}
    return {"value": runs["all"]["steps_per_s"], "unit": "steps/s", "cores": phys, "kind": "port",
            "physical_cores": phys, "half_cores": runs["half"],
            "sample": f"{steps} real-time SIL steps (m={lm}, dt={dt}) at N={npar}, M={norb} with the "
                      f"reference's table-driven H.C, median step; {phys} OpenMP threads (all physical cores "
                      f"of this process's CPUs), then {runs['half']['threads']} (half, the reference's "
                      f"default); table build excluded"}


def hc_loads_per_config(M: int, c64: bool) -> int:
    """Vector-memory load instructions of one configuration's row in the
    fused H.C kernel when every occupation is >= 2 (every branch taken, the
    case of nearly every wave at N=100): the byte-packed occupation row, the
    row's own coefficient, and one gather per distinct jump target
    (M(M-1) single, M.C(M,2) same-orbital and C(M,2).C(M-1,2) two-orbital
    double jumps).  In the complex64 form, targets that differ only in
    whether an added particle lands in orbital 0 or 1 are adjacent and share
    one 16-byte load when the removed orbitals are >= 2."""
    singles = M * (M - 1)
    same = M * (M - 1) * M // 2
    diff = math.comb(M, 2) * (M - 2) * (M - 1) // 2
    if c64:
        singles -= M - 2
        same -= (M - 2) * (1 + (M - 3))
        diff -= math.comb(M - 2, 2) * (1 + max(M - 4, 0))
    return singles + same + diff + 2


# Texture-addresser cycles per 64-lane gather instruction and the unit
# count: every vector-memory instruction of a wave occupies its CU's
# texture addresser for ~16 cycles whatever its width (rocprofv3
# TA_TA_BUSY / SQ_INSTS_VMEM_RD on the H.C kernel, 256 CUs at 2.4 GHz).
TA_CYCLES_PER_LOAD = 16.0
TA_UNITS, CLOCK_GHZ = 256, 2.4


def measure_traffic(args) -> dict:
    """HBM bytes per launch of the mixed H.C kernel, measured now: two
    rocprofv3 counter runs of `bench.py --probe-hc` (FETCH_SIZE, then
    WRITE_SIZE), each scaled by what the same counter reports for a device
    copy of known size in the same run."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if prof is None:
        return {"bytes": None, "error": "rocprofv3 not found"}
    out = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="hc_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", counter, "-f", "csv", "-d", d, "-o", "p", "--", sys.executable,
               os.path.join(ROOT, "bench.py"), "--probe-hc", "--npar", str(args.npar), "--norb", str(args.norb),
               "--seed", str(args.seed)]
        try:
            subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           timeout=150, check=False)
        except subprocess.TimeoutExpired:
            return {"bytes": None, "error": f"{counter} run timed out"}
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") == counter:
                        rows.append((int(row.get("Dispatch_Id", 0) or 0), row.get("Kernel_Name", ""),
                                     float(row.get("Counter_Value", "nan"))))
        shutil.rmtree(d, ignore_errors=True)
        rows.sort()
        # the probe's last dispatches: 5 H.C launches, then 5 device-to-device
        # copies of COPY_PROBE_BYTES (the runtime's copyBuffer kernel)
        hc = [v for _, name, v in rows if "k_hc_fused" in name][-5:]
        cp = [v for _, name, v in rows if "copyBuffer" in name][-5:]
        if len(hc) < 5 or len(cp) < 5:
            return {"bytes": None, "error": f"{counter}: the probe's H.C or copy dispatches are missing"}
        scale = COPY_PROBE_BYTES / float(np.mean(cp))  # bytes per counter unit
        out[counter] = {"hc_counter_mean": float(np.mean(hc)), "copy_counter_mean": float(np.mean(cp)),
                        "scale_bytes_per_unit": scale, "hc_bytes": float(np.mean(hc)) * scale}
    total = out["FETCH_SIZE"]["hc_bytes"] + out["WRITE_SIZE"]["hc_bytes"]
    return {"bytes": int(round(total)), "counters": out}


COPY_PROBE_BYTES = 512 << 20


def probe_hc(args, space, C0, H, G, stream):
    """Five launches of the mixed on-the-fly H.C kernel, then five of a
    plain device copy of COPY_PROBE_BYTES (read and written once each), so a
    FETCH_SIZE / WRITE_SIZE counter run can be scaled by what the same
    counter reports for a copy of known size."""
    import torch
    import mctdhb_amd as amd

    space.hc_tiling(args.probe_tiling)  # default: the one-lane kernel (the SIL step's Lanczos forms run it)
    cf = torch.empty(space.coef_bytes(amd.MIXED), dtype=torch.uint8, device="cuda")
    space.hc_coefficients(H, G, cf, amd.MIXED, stream=stream)
    src = C0.to(torch.complex64)
    out = torch.empty_like(src)
    for _ in range(5):
        space.apply_hamiltonian_cf(src, cf, out, amd.MIXED, amd.INDEX_FLY, stream=stream)
    a = torch.empty(COPY_PROBE_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "hc_mixed_fly", "hc_launches": 5, "copy_launches": 5,
                      "copy_bytes_read": COPY_PROBE_BYTES, "copy_bytes_written": COPY_PROBE_BYTES}), flush=True)


def probe_density(args, space, C0, H, G, stream):
    """rho1 + rho2 three times in FP64, then three times in the mixed form
    (on-the-fly indices), for counter runs of the density kernels."""
    import torch
    import mctdhb_amd as amd

    M = args.norb
    r1 = torch.empty(M * M, dtype=torch.complex128, device="cuda")
    r2 = torch.empty(M ** 4, dtype=torch.complex128, device="cuda")
    c32 = C0.to(torch.complex64)
    for src, prec in ((C0, amd.FP64), (c32, amd.MIXED)):
        for _ in range(3):
            space.density(src, r1, r2, prec, amd.INDEX_FLY, stream=stream)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "density", "launches_per_form": 3}), flush=True)


def probe_sil(args, space, C0, H, G, stream):
    """Two mixed SIL steps (their H.C launches carry the Lanczos epilogue),
    then five device copies of COPY_PROBE_BYTES, for counter runs of the
    kernels that dominate the timed step."""
    import torch
    import mctdhb_amd as amd

    sil = amd.DeviceSil(space, args.lanczos, amd.MIXED, amd.INDEX_FLY)
    C = C0.clone()
    for _ in range(2):
        sil.step(C, H, G, args.dt, False, stream=stream)
    a = torch.empty(COPY_PROBE_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    sil.close()
    print(json.dumps({"probe": "sil_mixed", "steps": 2, "copy_launches": 5, "copy_bytes": COPY_PROBE_BYTES}),
          flush=True)


def full_step_bench(args, stream):
    """The whole MCTDHB time step (orbital split step, h/g rebuild, SIL,
    density sync; reference driver.c:159-202) on the reference's own
    ground-state case: N=10, M=4 (NC=286), 256 grid intervals on [-8, 8]
    (grid_size 257, the last point the periodic copy of the first, FFT length
    256), harmonic trap, g=1, imaginary time dt=0.005, Lanczos dimension 5,
    FP64.  Eager steps and
    replays of the captured step graph run from the same initial state; the
    oracle's full step is timed on a bounded sample of the same run."""
    import torch
    import mctdhb_amd as amd
    from mctdhb_amd import inputs

    N, M, grid, dt, lm, g = 10, 4, 257, 0.005, 5, 1.0
    x = np.linspace(-8.0, 8.0, grid)
    pot = 0.5 * x ** 2
    orb0 = inputs.hermite_gauss(M, x, seed=args.seed, noise=1e-3)
    orb0[:, -1] = orb0[:, 0]
    space = amd.DeviceSpace(N, M, tables=False)
    dim = math.comb(N + M - 1, M - 1)

    def state():
        w = amd.DeviceOrbital(M, grid, x[0], x[-1], dt, g, pot, imagtime=True)
        sil = amd.DeviceSil(space, lm, amd.FP64, amd.INDEX_FLY)
        d = {"orb": torch.from_numpy(np.ascontiguousarray(orb0.ravel())).cuda(),
             "coef": torch.from_numpy(inputs.coefficients(dim, args.seed)).cuda()}
        for k, n in (("hob", M * M), ("hint", M ** 4), ("r1", M * M), ("r2", M ** 4), ("rinv", M * M)):
            d[k] = torch.zeros(n, dtype=torch.complex128, device="cuda")
        w.orthonormalize(d["orb"], stream=stream)
        w.orbital_matrices(d["orb"], d["hob"], d["hint"], stream=stream)
        amd.density_sync(space, d["coef"], d["r1"], d["r2"], d["rinv"], stream=stream)
        torch.cuda.synchronize()
        return w, sil, d

    warm, timed = 20, 500

    def timeit(fn):
        for _ in range(warm):
            fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(timed):
            fn()
        b.record(stream)
        b.synchronize()
        return a.elapsed_time(b) / timed

    w_e, sil_e, de = state()
    host0 = {k: v.cpu().numpy() for k, v in de.items()}
    eager_ms = timeit(lambda: amd.propagate_step(space, sil_e, w_e, de["orb"], de["coef"], de["hob"], de["hint"],
                                                 de["r1"], de["r2"], de["rinv"], stream=stream))
    w_g, sil_g, dg = state()
    graph = amd.StepGraph(space, sil_g, w_g, dg["orb"], dg["coef"], dg["hob"], dg["hint"], dg["r1"], dg["r2"],
                          dg["rinv"])
    graph_ms = timeit(lambda: graph.launch(stream))
    same = all(torch.equal(de[k], dg[k]) for k in de)
    r1 = de["r1"].cpu().numpy().reshape(M, M)
    graph.close()
    out = {"workload": f"imaginary-time MCTDHB steps N={N} M={M} NC={dim} Ng={grid} dt={dt} m={lm} FP64",
           "eager_ms_per_step": eager_ms, "graph_ms_per_step": graph_ms,
           "graph_steps_per_s": 1e3 / graph_ms, "graph_vs_eager": eager_ms / graph_ms,
           "time_to_2000_steps_s": 2000 * graph_ms / 1e3, "graph_equals_eager_bitwise": bool(same),
           "after_steps": warm + timed,
           "natural_occupations": [float(v) for v in np.sort(np.linalg.eigvalsh(r1))[::-1] / N]}
    log(f"full step: eager {eager_ms:.3f} ms, graph {graph_ms:.3f} ms, bitwise equal {same}")
    if not args.no_cpu:
        import oracle

        S = oracle.Space(N, M)
        eq = oracle.make_orbeq(M, grid, x[0], x[-1], dt, g, pot, imagtime=True)
        h = {k: np.ascontiguousarray(v) for k, v in host0.items()}
        lvec = np.zeros(lm * dim, complex)
        steps = 200
        t0 = time.perf_counter()
        for _ in range(steps):
            oracle.lib().orc_propagate_step(S._p, eq, 1, 1, lm, h["orb"], h["coef"], h["hob"], h["hint"], h["r1"],
                                            h["r2"], h["rinv"], lvec)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": steps / el, "unit": "steps/s",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                               "kind": "port", "sample": f"{steps} full steps of the same run"}
    return out


MIXED_SHAPES = ((10, 4), (200, 4), (30, 7), (100, 5))


def mixed_shapes_bench(args, stream, steps=50, lm=8, dt=0.01):
    """Mixed precision against FP64 on several shapes: from the same seeded
    coefficients, h and g (scale 0.1), `steps` real-time SIL steps with
    Lanczos dimension `lm`, once with complex64 H.C (FP64 Lanczos scalars)
    and once all-FP64, both with on-the-fly jump indices.  Reports the
    per-step p50 of each and the relative error of the energy per particle,
    of the natural occupations and of rho1 (Frobenius) at the end."""
    import torch
    import mctdhb_amd as amd
    from mctdhb_amd import inputs

    out = []
    for N, M in MIXED_SHAPES:
        nc = math.comb(N + M - 1, M - 1)
        space = amd.DeviceSpace(N, M, tables=False)
        C0 = torch.from_numpy(inputs.coefficients(nc, args.seed)).cuda()
        H = torch.from_numpy(np.ascontiguousarray(inputs.onebody(M, args.seed))).cuda()
        G = torch.from_numpy(inputs.twobody(M, args.seed, 0.1)).cuda()
        res = {"npar": N, "norb": M, "dim": nc, "lanczos_iter": lm, "dt": dt, "steps": steps}
        finals = {}
        for name, prec in (("mixed", amd.MIXED), ("fp64", amd.FP64)):
            sil = amd.DeviceSil(space, lm, prec, amd.INDEX_FLY)
            C = C0.clone()
            evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
            evs[0].record(stream)
            for k in range(steps):
                sil.step(C, H, G, dt, False, stream=stream)
                evs[k + 1].record(stream)
            evs[-1].synchronize()
            per = [evs[k].elapsed_time(evs[k + 1]) for k in range(steps)]
            sil.close()
            res[f"{name}_p50_ms"] = percentile(per, 50)
            res[f"{name}_steps_per_s"] = 1e3 / percentile(per, 50)
            hc = torch.empty_like(C)
            space.apply_hamiltonian(C, H, G, hc, amd.FP64, amd.INDEX_FLY, stream=stream)
            r1 = torch.empty(M * M, dtype=torch.complex128, device="cuda")
            space.density(C, r1, None, amd.FP64, amd.INDEX_FLY, stream=stream)
            torch.cuda.synchronize()
            rho1 = r1.cpu().numpy().reshape(M, M)
            finals[name] = (torch.vdot(C, hc).real.item() / N, rho1, np.sort(np.linalg.eigvalsh(rho1)) / N)
        (e32, r32, o32), (e64, r64, o64) = finals["mixed"], finals["fp64"]
        res["energy_rel"] = abs(e32 - e64) / abs(e64)
        res["natural_occupations_rel"] = float(np.max(np.abs(o32 - o64) / np.maximum(np.abs(o64), 1e-300)))
        res["rho1_rel_fro"] = float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64))
        res["mixed_vs_fp64"] = res["mixed_steps_per_s"] / res["fp64_steps_per_s"]
        log(f"mixed vs FP64 N={N} M={M}: {res['mixed_p50_ms']:.3f} / {res['fp64_p50_ms']:.3f} ms, "
            f"E rel {res['energy_rel']:.2e}, rho1 rel {res['rho1_rel_fro']:.2e}")
        out.append(res)
        del space, C0
        torch.cuda.empty_cache()
    return out


def batched_bench(args):
    """Uneven imaginary-time ground states solved at once: N in [4, 64], M in
    {2..6} (NC <= 2e5), Ng in {256, 512, 1024} on [-L, L) with V(L) = 200,
    double-well traps, g log-uniform in [0.01, 10], dt = 0.005, m = 5
    (mctdhb_amd.batch).
    The instances of the one-workgroup FP64 step run packed (one launch per
    stage for each group of alike instances, mctdhb_pack_*); the others
    replay their own step graphs round-robin over a few streams.  For
    comparison the same instances are also stepped one after another, and
    together, for min(K, 10) steps each, and every instance by its own graph
    for K steps."""
    import torch
    from mctdhb_amd import batch

    draws = batch.draw_ground_states(args.batch_instances, seed=args.seed)
    t0 = time.perf_counter()
    insts = [batch.Instance(d) for d in draws]
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    b = batch.Batch(insts, nstreams=args.batch_streams, packed=not args.batch_graphs)
    K = args.batch_steps
    b.run(1)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(K)
    b.synchronize()
    wall = time.perf_counter() - t0
    log(f"batched: {len(insts)} instances x {K} steps in {wall:.2f} s (setup {setup_s:.1f} s)")
    # every mixed result against its FP64 twin after the same 1 + K steps
    obs = [inst.observables() for inst in insts]
    mixed_idx = [i for i, inst in enumerate(insts) if inst.prec != 0]
    twins = [batch.Instance(draws[i], fp64=True) for i in mixed_idx]
    tb = batch.Batch(twins, nstreams=args.batch_streams)
    tb.run(1 + K)
    tb.synchronize()
    twin_obs = [t.observables() for t in twins]
    errs = [batch.errors_vs_fp64(obs[i], o) for i, o in zip(mixed_idx, twin_obs)]
    worst_occ = None
    if errs:
        j = max(range(len(errs)), key=lambda j: errs[j]["natural_occupations_rel"] or 0.0)
        d = draws[mixed_idx[j]]
        worst_occ = {"npar": d.npar, "norb": d.norb, "dim": d.dim, "g": d.g, **errs[j],
                     "natural_occupations_fp64": [float(v) for v in twin_obs[j][2]],
                     "natural_occupations_mixed": [float(v) for v in obs[mixed_idx[j]][2]]}
    for t in twins:
        t.close()

    def worst(key):
        vals = [e[key] for e in errs if e[key] is not None]
        return max(vals) if vals else None
    # the same instances one after another on one stream, and together
    K2 = min(K, 10)
    st = b.streams[0]
    t0 = time.perf_counter()
    for inst in insts:
        for _ in range(K2):
            inst.step(st)
    st.synchronize()
    serial = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.run(K2)
    b.synchronize()
    together = time.perf_counter() - t0
    log(f"{K2} steps each: one after another {serial:.2f} s, together {together:.2f} s")
    # every instance by its own step graph (no pack), K steps
    gb = batch.Batch(insts, nstreams=args.batch_streams)
    gb.run(1)
    gb.synchronize()
    t0 = time.perf_counter()
    gb.run(K)
    gb.synchronize()
    graphs_wall = time.perf_counter() - t0
    log(f"graphs only: {len(insts)} x {K} steps in {graphs_wall:.2f} s")
    e = [inst.energy_per_particle() for inst in insts]
    finite = [v for v in e if math.isfinite(v)]
    dims = np.array([d.dim for d in draws])
    out = {"workload": f"{len(insts)} uneven imaginary-time ground states x {K} steps (dt=0.005, m=5)",
           "instances": len(insts), "steps_each": K, "streams": args.batch_streams,
           "instance_steps_per_s": len(insts) * K / wall, "wall_s": wall, "setup_s": setup_s,
           "packed_instances": b.packed_count, "pack_groups": b.pack.groups if b.pack is not None else 0,
           "graphs_only_instance_steps_per_s": len(insts) * K / graphs_wall,
           "comparison_steps_each": K2, "one_after_another_s": serial, "together_s": together,
           "concurrency_gain": serial / together,
           "dim_min_median_max": [int(dims.min()), int(np.median(dims)), int(dims.max())],
           "mixed_instances": int(sum(1 for i in insts if i.prec != 0)),
           "instances_with_finite_energy": len(finite),
           "energy_per_particle_first8": e[:8],
           "errors_vs_fp64": {"what": f"every mixed instance against its FP64 twin after {1 + K} steps; FP64 "
                                      f"instances are exact by construction",
                              "mixed_checked": len(errs), "energy_rel_max": worst("energy_rel"),
                              "natural_occupations_rel_max": worst("natural_occupations_rel"),
                              "natural_occupations_abs_max": worst("natural_occupations_abs"),
                              "rho1_rel_fro_max": worst("rho1_rel_fro"), "worst_occupations": worst_occ}}
    b.close()
    for inst in insts:
        inst.close()
    return out


def worst_requests(recs, by_id, n=5):
    """The n mixed solved requests with the largest energy error against
    their FP64 twins, with what characterises them."""
    rows = [r for r in recs if r["mixed"] and r["status"] == "ok" and r["errors_vs_fp64"]["energy_rel"] is not None]
    rows.sort(key=lambda r: -r["errors_vs_fp64"]["energy_rel"])
    out = []
    for r in rows[:n]:
        q = by_id[r["rid"]]
        d = q.draw
        out.append({"rid": r["rid"], "kind": r["kind"], "npar": d.npar, "norb": d.norb, "dim": d.dim,
                    "ngrid": d.ngrid, "g": d.g, "g_factor": q.g_factor, "imag_steps": q.imag_steps,
                    "real_steps": q.real_steps, "energy_per_particle": r["energy_per_particle"],
                    **r["errors_vs_fp64"]})
    return out


def serve_bench(args):
    """The request loop (mctdhb_amd.serve): Poisson arrivals at 200/s for a
    window, 70 % ground states and 30 % interaction quenches, NC
    log-uniform in [1e2, 5e6], admitted against a fixed HBM budget and run
    together; every step count is scaled by --serve-time-scale."""
    from mctdhb_amd import serve

    reqs = serve.draw_requests(args.serve_rate, args.serve_window, seed=args.seed,
                               time_scale=args.serve_time_scale)
    srv = serve.Server(int(args.serve_budget_gb * 1e9), nstreams=8, chunk=8,
                       max_active=args.serve_max_active or None,
                       reclaim_bytes=int(args.serve_reclaim_gb * 1e9) if args.serve_reclaim_gb > 0 else None)
    t0 = time.perf_counter()
    recs = srv.run(reqs)
    wall = time.perf_counter() - t0
    steps = sum(r["steps"] for r in recs)
    lat = serve.latency_summary(recs)
    t1 = time.perf_counter()
    serve.attach_fp64_errors(recs, reqs)  # every mixed result against its FP64 twin, run alone
    log(f"served: FP64 twins of the mixed results in {time.perf_counter() - t1:.1f} s")
    checked = [r["errors_vs_fp64"] for r in recs if r["mixed"] and r["status"] == "ok"]

    def worst(key):
        vals = [e[key] for e in checked if e[key] is not None]
        return max(vals) if vals else None
    log(f"served {len(recs)} requests in {wall:.2f} s ({lat['count']} solved): time to solution p50 "
        f"{lat['p50_s']} s, p99 {lat['p99_s']} s, at most {srv.peak_active} at once")
    dims = np.array([r["dim"] for r in recs])
    by_id = {q.rid: q for q in reqs}
    ok = [r for r in recs if r["status"] == "ok"]
    failed = []
    for r in recs:
        if r["status"] == "ok":
            continue
        q = by_id[r["rid"]]
        d = q.draw
        failed.append({"rid": r["rid"], "kind": r["kind"], "reason": r["reason"], "npar": d.npar, "norb": d.norb,
                       "dim": d.dim, "ngrid": d.ngrid, "g": d.g, "g_factor": q.g_factor,
                       "g_n": d.g * (d.npar - 1) * (q.g_factor if q.kind == "quench" else 1.0),
                       "imag_steps": q.imag_steps, "real_steps": q.real_steps})
    return {"workload": f"{len(recs)} requests, Poisson {args.serve_rate}/s for {args.serve_window} s, "
                        f"step counts x{args.serve_time_scale}",
            "requests": len(recs), "quenches": sum(r["kind"] == "quench" for r in recs),
            "wall_s": wall, "solves_per_s": len(ok) / wall, "solved": len(ok), "failed": len(failed),
            "instance_steps": steps, "instance_steps_per_s": steps / wall,
            "time_to_solution": lat, "peak_admitted": srv.peak_active,
            "budget_gb": args.serve_budget_gb, "max_active": args.serve_max_active or None,
            "reclaim_gb": args.serve_reclaim_gb, "bulk_releases": srv.releases,
            "dim_min_median_max": [int(dims.min()), int(np.median(dims)), int(dims.max())],
            "failed_requests": failed[:20],
            "errors_vs_fp64": {"what": "every solved mixed request against its FP64 twin run alone; FP64 requests "
                                       "are exact by construction",
                               "mixed_checked": len(checked), "energy_rel_max": worst("energy_rel"),
                               "natural_occupations_rel_max": worst("natural_occupations_rel"),
                               "natural_occupations_abs_max": worst("natural_occupations_abs"),
                               "rho1_rel_fro_max": worst("rho1_rel_fro"),
                               "worst5": worst_requests(recs, by_id)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--npar", type=int, default=100)
    ap.add_argument("--norb", type=int, default=5)
    ap.add_argument("--lanczos", type=int, default=6)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--hc-reps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--latency-steps", type=int, default=200,
                    help="further steps, each timed on its own, for the p50/p99 step latency")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter runs that measure the H.C kernel's HBM bytes")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-full-step", action="store_true",
                    help="skip the whole-step (orbitals + SIL + densities) eager/graph timing")
    ap.add_argument("--no-mixed-shapes", action="store_true",
                    help="skip the mixed-vs-FP64 comparison on (10,4), (200,4), (30,7), (100,5) with m=8")
    ap.add_argument("--batch-instances", type=int, default=128)
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--batch-streams", type=int, default=8)
    ap.add_argument("--batch-graphs", action="store_true",
                    help="batched workload: every instance by its own step graph (no packed steps)")
    ap.add_argument("--no-batched", action="store_true", help="skip the many-instance timing")
    ap.add_argument("--batched-only", action="store_true", help="only the many-instance timing (its JSON object)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-window", type=float, default=1.0)
    ap.add_argument("--serve-time-scale", type=float, default=0.02)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-max-active", type=int, default=0, help="admitted at once (0: budget only)")
    ap.add_argument("--serve-reclaim-gb", type=float, default=64.0,
                    help="retired requests' device memory is released in bulk once this much is held (0: at once)")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-loop timing")
    ap.add_argument("--serve-only", action="store_true", help="only the request loop (its JSON object)")
    ap.add_argument("--full-step-only", action="store_true",
                    help="only the whole-step timing (prints its JSON object)")
    ap.add_argument("--probe-density", action="store_true",
                    help="rho1 + rho2 three times per form (FP64, mixed) for counter runs")
    ap.add_argument("--probe-sil", action="store_true",
                    help="only two mixed SIL steps and a plain copy kernel of known size, for hardware-counter runs")
    ap.add_argument("--probe-tiling", type=int, default=0,
                    help="mixed H.C tiling of --probe-hc (0 one configuration per lane, 1 pairs, 2 windows)")
    ap.add_argument("--probe-hc", action="store_true",
                    help="only the mixed H.C kernel (5 launches) and a plain copy kernel of known size, "
                         "for hardware-counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)

    import torch
    import mctdhb_amd as amd
    from mctdhb_amd import inputs

    if args.full_step_only:
        print(json.dumps(full_step_bench(args, torch.cuda.current_stream())), flush=True)
        return
    if args.batched_only:
        print(json.dumps(batched_bench(args)), flush=True)
        return
    if args.serve_only:
        print(json.dumps(serve_bench(args)), flush=True)
        return

    N, M, lm, dt = args.npar, args.norb, args.lanczos, args.dt
    dev = "cuda"
    stream = torch.cuda.current_stream()
    nc = math.comb(N + M - 1, M - 1)
    c0 = inputs.coefficients(nc, args.seed)
    h = inputs.onebody(M, args.seed)
    g = inputs.twobody(M, args.seed, 0.1)
    C0 = torch.from_numpy(c0).to(dev)
    H = torch.from_numpy(np.ascontiguousarray(h)).to(dev)
    G = torch.from_numpy(g).to(dev)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    space = amd.DeviceSpace(N, M, tables=True)
    torch.cuda.synchronize()
    table_build_s = time.perf_counter() - t0
    log(f"space N={N} M={M} dim={nc} tables built in {table_build_s:.2f} s")

    def ev():
        return torch.cuda.Event(enable_timing=True)

    if args.probe_hc:
        probe_hc(args, space, C0, H, G, stream)
        return
    if args.probe_sil:
        probe_sil(args, space, C0, H, G, stream)
        return
    if args.probe_density:
        probe_density(args, space, C0, H, G, stream)
        return

    # ---- H.C kernel alone (coefficient block prebuilt): FP64 with tables and
    # on the fly, mixed with the tiling the space chose, and mixed with one
    # configuration per lane (the kernel of the SIL step's Lanczos forms,
    # whose roofline is reported)
    hc_ms = {}
    forms = {"fp64_tables": (amd.FP64, amd.INDEX_TABLES), "fp64_fly": (amd.FP64, amd.INDEX_FLY),
             "mixed_fly": (amd.MIXED, amd.INDEX_FLY), "mixed_fly_lane": (amd.MIXED, amd.INDEX_FLY)}
    tuned = None
    for name, (prec, index) in forms.items():
        if name == "mixed_fly_lane":
            tuned = space.hc_tiling()
            space.hc_tiling(0)
        cf = torch.empty(space.coef_bytes(prec), dtype=torch.uint8, device=dev)
        space.hc_coefficients(H, G, cf, prec, stream=stream)
        src = C0 if prec == amd.FP64 else C0.to(torch.complex64)
        out = torch.empty_like(src)
        for _ in range(3):
            space.apply_hamiltonian_cf(src, cf, out, prec, index, stream=stream)
        a, b = ev(), ev()
        a.record(stream)
        for _ in range(args.hc_reps):
            space.apply_hamiltonian_cf(src, cf, out, prec, index, stream=stream)
        b.record(stream)
        b.synchronize()
        hc_ms[name] = a.elapsed_time(b) / args.hc_reps
        log(f"H.C {name}: {hc_ms[name]:.3f} ms")
        del cf, out
    space.hc_tiling(tuned)

    # ---- density matrices (rho1 + rho2)
    dm_ms = {}
    for name, (prec, index) in {"fp64": (amd.FP64, amd.INDEX_FLY), "mixed": (amd.MIXED, amd.INDEX_FLY)}.items():
        src = C0 if prec == amd.FP64 else C0.to(torch.complex64)
        r1 = torch.empty(M * M, dtype=torch.complex128, device=dev)
        r2 = torch.empty(M ** 4, dtype=torch.complex128, device=dev)
        space.density(src, r1, r2, prec, index, stream=stream)
        a, b = ev(), ev()
        a.record(stream)
        for _ in range(5):
            space.density(src, r1, r2, prec, index, stream=stream)
        b.record(stream)
        b.synchronize()
        dm_ms[name] = a.elapsed_time(b) / 5
        log(f"density {name}: {dm_ms[name]:.3f} ms")

    # ---- SIL steps: K timed steps (the value), then `latency_steps` more,
    # each bracketed by its own events, for the per-step p50/p99
    def run_sil(prec, index, lat_steps):
        sil = amd.DeviceSil(space, lm, prec, index)
        C = C0.clone()
        for _ in range(args.warmup):
            sil.step(C, H, G, dt, False, stream=stream)
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record(stream)
        for k in range(args.steps):
            sil.step(C, H, G, dt, False, stream=stream)
        b.record(stream)
        b.synchronize()
        total_ms = a.elapsed_time(b)
        per = []
        if lat_steps:
            evs = [ev() for _ in range(lat_steps + 1)]
            evs[0].record(stream)
            for k in range(lat_steps):
                sil.step(C, H, G, dt, False, stream=stream)
                evs[k + 1].record(stream)
            evs[-1].synchronize()
            per = [evs[k].elapsed_time(evs[k + 1]) for k in range(lat_steps)]
        done, _, _ = sil.scalars(stream=stream)
        sil.close()
        log(f"SIL prec={prec} index={index}: {total_ms / args.steps:.3f} ms/step")
        return C, total_ms, per, done, args.warmup + args.steps + lat_steps

    C_mixed, t_mixed, per_mixed, done_mixed, n_mixed = run_sil(amd.MIXED, amd.INDEX_FLY, args.latency_steps)
    C_fly, t_fly, per_fly, _, n_fly = run_sil(amd.FP64, amd.INDEX_FLY, args.latency_steps)
    C_tab, t_tab, per_tab, _, _ = run_sil(amd.FP64, amd.INDEX_TABLES, 0)
    if n_fly != n_mixed:
        raise RuntimeError("the mixed and FP64 runs must end at the same time")

    # ---- error of the mixed path against FP64 at the same time t = n dt
    def observables(C):
        hc = torch.empty_like(C)
        space.apply_hamiltonian(C, H, G, hc, amd.FP64, amd.INDEX_FLY, stream=stream)
        r1 = torch.empty(M * M, dtype=torch.complex128, device=dev)
        space.density(C, r1, None, amd.FP64, amd.INDEX_FLY, stream=stream)
        torch.cuda.synchronize()
        e = torch.vdot(C, hc).real.item() / N
        rho1 = r1.cpu().numpy().reshape(M, M)
        occ = np.sort(np.linalg.eigvalsh(rho1)) / N
        return e, rho1, occ

    e64, rho64, occ64 = observables(C_fly)
    e32, rho32, occ32 = observables(C_mixed)
    errors = {
        "energy_rel": abs(e32 - e64) / abs(e64),
        "natural_occupations_rel": float(np.max(np.abs(occ32 - occ64) / np.abs(occ64))),
        "natural_occupations_abs": float(np.max(np.abs(occ32 - occ64))),
        "rho1_rel_fro": float(np.linalg.norm(rho32 - rho64) / np.linalg.norm(rho64)),
        "coef_norm_drift_mixed": abs(torch.linalg.vector_norm(C_mixed).item() - 1.0),
        "after_steps": n_mixed,
    }

    # ---- the same mixed step with C handed over in host memory, as the
    # reference-named host entry points take it: host -> device copy, step,
    # device -> host copy, host clock around each step.  Reported beside the
    # value, never as it.
    def host_boundary(nsteps=5):
        sil = amd.DeviceSil(space, lm, amd.MIXED, amd.INDEX_FLY)
        Cd = C0.clone()
        pinned = torch.empty(C0.shape, dtype=C0.dtype, pin_memory=True)
        pinned.copy_(C0)
        res = {"what": f"one mixed SIL step with C ({C0.numel() * 16 / 1e6:.1f} MB) copied from host memory "
                       f"and back per step, {nsteps} steps, host clock"}
        for name, hbuf in (("pinned", pinned), ("pageable", C0.cpu())):
            sil.step(Cd, H, G, dt, False, stream=stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(nsteps):
                Cd.copy_(hbuf, non_blocking=True)
                sil.step(Cd, H, G, dt, False, stream=stream)
                hbuf.copy_(Cd, non_blocking=True)
                stream.synchronize()
            el = (time.perf_counter() - t0) / nsteps
            res[name] = {"ms_per_step": el * 1e3, "steps_per_s": 1.0 / el}
        sil.close()
        log(f"SIL step across the host boundary: pinned {res['pinned']['ms_per_step']:.3f} ms, "
            f"pageable {res['pageable']['ms_per_step']:.3f} ms")
        return res

    host_bound = host_boundary()

    steps_mixed = args.steps / (t_mixed / 1e3)
    steps_fly = args.steps / (t_fly / 1e3)
    steps_tab = args.steps / (t_tab / 1e3)
    ab_c64 = algorithmic_bytes(N, M, 8, lm)
    ab_f64 = algorithmic_bytes(N, M, 16, lm)

    # ---- roofline of the dominant kernel (the mixed on-the-fly H.C).  Its
    # algorithmic bytes are what this algorithm must move: coefficients in
    # and out (8 B each) and the byte-packed occupation row (8 B); it reads
    # no jump tables.  The table-driven figure is kept as `effective_gbs`.
    t_hc = hc_ms["mixed_fly_lane"] / 1e3
    hc_bytes = 24 * nc
    traffic = None if args.no_traffic else measure_traffic(args)
    waves = -(-nc // 64)
    loads = hc_loads_per_config(M, True)
    floor_s = waves * loads * TA_CYCLES_PER_LOAD / (TA_UNITS * CLOCK_GHZ * 1e9)
    roofline = {
        "bound": "hbm",
        "kernel": f"k_hc_fused<{M}, float, fly, c64, EpiPlain> (mixed H.C, on-the-fly jump indices)",
        "achieved": hc_bytes / t_hc / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": hc_bytes / t_hc / 1e9 / HBM_PEAK_GBS,
        "traffic": traffic["bytes"] if traffic else None,
        "algorithmic_bytes_per_launch": hc_bytes,
        "algorithmic_bytes": "coefficients in + out (8 B each) + packed occupation row (8 B) per configuration",
        "launch_us": hc_ms["mixed_fly_lane"] * 1e3,
        "traffic_measurement": traffic,
        "gather_issue_roof": {
            "what": "vector-memory load instructions x texture-addresser cycles each / (CUs x clock)",
            "loads_per_configuration": loads, "waves": waves, "ta_cycles_per_load": TA_CYCLES_PER_LOAD,
            "floor_us": floor_s * 1e6, "frac": floor_s / t_hc,
        },
        "binding_roof": "gather issue (texture addresser), not HBM",
        "effective_gbs_table_driven_bytes": ab_c64["hc"] / t_hc / 1e9,
    }

    full = full_step_bench(args, stream) if not args.no_full_step else None
    shapes = mixed_shapes_bench(args, stream) if not args.no_mixed_shapes else None
    batched = batched_bench(args) if not args.no_batched else None
    served = serve_bench(args) if not args.no_serve else None

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        try:
            cpu = cpu_baseline(N, M, lm, dt, args.cpu_steps, args.seed)
        except Exception as exc:  # reported, never silently replaced
            cpu = {"value": None, "error": repr(exc)}

    line = {
        "metric": METRIC,
        "value": steps_mixed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 H.C/rho (Kahan) + f64 Lanczos",
        "data": "synthetic (seeded complex-normal C, random Hermitian h scale 1, symmetrised g scale 0.1)",
        "config": {"workload": f"real-time SIL steps N={N} M={M} NC={nc} m={lm} dt={dt}",
                   "npar": N, "norb": M, "dim": nc, "lanczos_iter": lm, "dt": dt},
        "p50_ms": percentile(per_mixed, 50) if per_mixed else None,
        "p99_ms": percentile(per_mixed, 99) if per_mixed else None,
        "latency_samples": len(per_mixed),
        "fp64_fly_steps_per_s": steps_fly,
        "fp64_tables_steps_per_s": steps_tab,
        "fp64_fly_p50_ms": percentile(per_fly, 50) if per_fly else None,
        "fp64_fly_p99_ms": percentile(per_fly, 99) if per_fly else None,
        "mixed_vs_fp64_fly": steps_mixed / steps_fly,
        "mixed_fly_vs_fp64_tables": steps_mixed / steps_tab,
        "lanczos_iter_done": done_mixed,
        "errors_mixed_vs_fp64": errors,
        "host_boundary": host_bound,
        "hc_ms": hc_ms,
        "hc_per_s_mixed": 1e3 / hc_ms["mixed_fly"],
        "hc_effective_gbs": {"fp64_tables": ab_f64["hc"] / hc_ms["fp64_tables"] / 1e6,
                             "fp64_fly": ab_f64["hc"] / hc_ms["fp64_fly"] / 1e6,
                             "mixed_fly": ab_c64["hc"] / hc_ms["mixed_fly"] / 1e6},
        "hc_mixed_tiling": {0: "one configuration per lane", 1: "configuration pairs",
                            2: "coefficient windows in LDS"}.get(space.hc_tiling(), "?"),
        "density_ms": dm_ms,
        "table_build_s": table_build_s,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "full_step": full,
        "mixed_vs_fp64_shapes": shapes,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
