"""Benchmark of the fused per-stage DG kernel on one MI355X.

Workload (the configuration the metric is quoted on): one GPU's block of 64^3 hexahedra, N=4 (125
LGL nodes each, 32.8M nodes x 4 FP64 fields), water with B/A=5, a 1 MHz 1e6 Pa tone burst entering
through z=0, 8-element sponge, LSRK(5,4) at CFL 0.3; the FP64 drop-in path. A "step" is one
LSRK54 step (5 fused stage kernels) over the whole block with state resident in HBM.

Printed: one JSON line. Besides the headline value it carries the dominant kernel's roofline
(algorithmic bytes per launch / HIP-event launch time), the CPU oracle timed on the host on a
bounded sample, and the mixed-precision comparisons (same mesh at N=4 and at N=7: speed-up over
FP64, rel L2 / peak pressure / second harmonic vs FP64 at t_end).
"""
import argparse
import importlib
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per node-stage of the fused LSRK54 stage, averaged over the 5 stages
# (stage 1 reads q, writes du and q: 96 B; stages 2-5 also read du: 128 B), FP64 / mixed
BYTES_FP64 = (96.0 + 4 * 128.0) / 5.0   # 121.6
BYTES_MIXED = (80.0 + 4 * 96.0) / 5.0   # 92.8

_T0 = time.time()


def progress(msg):
    """One line on stderr as each part of the run ends (stdout carries only the JSON line)."""
    print(f"bench [{time.time() - _T0:6.1f} s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--elements", type=int, default=64)
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed / N=7 comparison runs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--variant", type=int, default=-1,
                    help="pin the stage-kernel variant 0..7 (default: autotune), e.g. for profiled runs")
    ap.add_argument("--xcd", type=int, default=1, help="workgroup order 0..3 when --variant is pinned")
    ap.add_argument("--batch-steps", type=int, default=20, help="timed steps of the batched workload")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 256-solve workload")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched 256-solve workload")
    ap.add_argument("--case", choices=["c3m_valu", "c3m_mfma_f32", "c3m_mfma_bf16", "c5m_valu", "c5m_mfma_f32",
                                       "c5m_mfma_bf16", "c5_fp64", "c3_fp64", "n6_fp64", "n6m_valu", "c2_rk33",
                                       "c2_ldg"], default=None,
                    help="time one workload alone (for profiling): 64^3 N=7, 48^3 N=5 (stretched two-layer) or "
                         "48^3 N=6 (the burst workload at that order), FP64 or mixed with the VALU or a "
                         "matrix-core contraction; 64^3 N=4 with RK33 or LDG diffusion; prints one JSON line")
    ap.add_argument("--c5-steps", type=int, default=2000, help="steps of the combined 48^3 N=5 workload (0: skip)")
    ap.add_argument("--precision-steps", type=int, default=500,
                    help="least steps of the 64^3 N=4 and N=7 precision comparisons (they run on until the burst "
                         "has passed the probe)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    return ap.parse_args()


def calibrate_copy():
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel give the scale factor
    between the counters and HBM bytes (1 GiB read + 1 GiB written per dispatch)."""
    import torch
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        b.copy_(a)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / 20
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copy_ms": el * 1e3,
                      "copy_GBs": 2 * n * 8 / el / 1e9}))


STREAM_PROBE = os.path.join(ROOT, "3ddg_amd", "dg3d_stream_probe")


def stream_ceiling():
    """The device's streaming rate for the LSRK stage kernel's access pattern at the 64^3 N=4
    block's size (group-blocked FP64 state read and written, pair-major RK register read and
    written in place, workgroups over whole elements), from 3ddg_amd/dg3d_stream_probe run as a
    child process. None when the probe is not built or fails."""
    if not os.path.exists(STREAM_PROBE):
        return None
    try:
        r = subprocess.run([STREAM_PROBE], capture_output=True, text=True, timeout=120)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0 or not r.stdout.strip():
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def timed_run(pk, p, precision, steps, warmup, probe=None, variant=-1, xcd=1, contraction=0, scheme=None,
              with_q=False):
    scheme = pk.dg3d.LSRK54 if scheme is None else scheme
    ctx = pk.problem.make_context(p, precision=precision, contraction=contraction, with_q=with_q)
    dt = ctx.cfl_dt(p.cfl)
    if variant >= 0:
        ctx.set_variant(scheme, variant, xcd)
    else:
        variant, xcd = ctx.autotune(scheme)
    if probe is not None:
        ctx.set_probes([probe], steps + warmup)
    t = 0.0
    if warmup:
        t = ctx.step(scheme, dt, warmup, t)
    ctx.sync()
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t = ctx.step(scheme, dt, steps, t)
    ctx.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    stage_ms, launches = ctx.last_step_timing()
    progress(f"  {p.name} {'FP64' if precision == pk.dg3d.FP64 else 'mixed'} contraction {contraction}: "
             f"{steps} timed steps, stage {stage_ms:.4f} ms")
    return ctx, dict(elapsed=el, stage_ms=stage_ms, launches=launches, dt=dt, t_end=t, variant=variant, xcd=xcd)


def other_paths_run(pk, elements, steps, warmup):
    """The 64^3 N=4 FP64 workload through the reference's own scheme (RK33, Grid.cpp:642-835:
    64, 96 and 96 B per node for its three stages, 85.3 B on average) and through LSRK54 with LDG
    diffusion on (per stage the Q kernel, 32 B per node, and the stage, which also reads Q:
    121.6 + 24 B; 177.6 B per node-stage in all). Rates per stage launch group."""
    out = {}
    p = pk.problem.config_c2(num_elements=elements, steps=steps)
    ctx, r = timed_run(pk, p, pk.dg3d.FP64, steps, warmup, scheme=pk.dg3d.RK33)
    ctx.close()
    bpn = (64.0 + 96.0 + 96.0) / 3.0
    out["rk33_fp64"] = dict(stage_ms=r["stage_ms"], variant=[r["variant"], r["xcd"]], bytes_per_node_stage=bpn,
                            dof_stage_per_s=p.dof * 3 * steps / r["elapsed"],
                            hbm_frac=p.dof * bpn / (r["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS)
    p = pk.problem.config_c2(num_elements=elements, steps=steps)
    p.diffusion = np.full(p.E, 1.0e-6)
    ctx, r = timed_run(pk, p, pk.dg3d.FP64, steps, warmup, with_q=True)
    ctx.close()
    bpn = 32.0 + BYTES_FP64 + 24.0
    out["lsrk54_ldg_fp64"] = dict(stage_ms=r["stage_ms"], variant=[r["variant"], r["xcd"]], bytes_per_node_stage=bpn,
                                  dof_stage_per_s=p.dof * 5 * steps / r["elapsed"],
                                  hbm_frac=p.dof * bpn / (r["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS)
    return out


def small_block_run(pk, steps=2000, warmup=20):
    """The 8^3 N=3 LSRK54 block (32 768 nodes, Gaussian pulse, IC halos): kernels of a few
    microseconds, so the rate is set by launches. Times dg3d_step with its step graphs and with
    direct launches (wall clock over the timed steps)."""
    import torch
    p = pk.problem.config_c1(steps=steps)
    out = {"workload": "c1", "elements": list(p.num_elements), "order": p.order, "dof": p.dof, "steps": steps}
    for graphs in (True, False):
        ctx = pk.problem.make_context(p)
        ctx.set_graphs(graphs)
        dt = ctx.cfl_dt(p.cfl)
        v, x = ctx.autotune(pk.dg3d.LSRK54)
        t = ctx.step(pk.dg3d.LSRK54, dt, warmup, 0.0)
        ctx.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t = ctx.step(pk.dg3d.LSRK54, dt, steps, t)
        ctx.sync()
        el = time.perf_counter() - t0
        key = "graph" if graphs else "direct"
        out[f"{key}_dof_stage_per_s"] = p.dof * 5 * steps / el
        out[f"{key}_us_per_step"] = el / steps * 1e6
        out[f"{key}_variant"] = [v, x]
        ctx.close()
    out["graph_speedup"] = out["graph_dof_stage_per_s"] / out["direct_dof_stage_per_s"]
    return out


def probe_node(p, depth_frac=0.05):
    """Global node index on the axis (x, y at the centre) near the source face."""
    n = p.n
    nE = p.num_elements
    i, j = nE[0] // 2, nE[1] // 2
    zc = np.concatenate([[0.0], np.cumsum(p.widths[2])])
    k = int(np.searchsorted(zc, depth_frac * zc[-1])) - 1
    k = max(0, min(nE[2] - 1, k))
    e = (i * nE[1] + j) * nE[2] + k
    return e * n ** 3 + (0 * n + 0) * n + 0


def probe_depth(p, node):
    """z coordinate of a probe node (global node index into a field)."""
    n = p.n
    e, nd = divmod(int(node), n ** 3)
    k = e % p.num_elements[2]
    zc = np.concatenate([[0.0], np.cumsum(p.widths[2])])
    return float(zc[k] + p.widths[2][k] * (p.nodes[nd % n] / 2.0 + 0.5))


def burst_window_steps(p, dt, depth):
    """Steps until the whole tone burst (envelope e^-9 at t0 +- 3 tau) has passed a probe at
    `depth`: the probe trace then holds the burst from before its start to after its end."""
    s = p.source
    return int(np.ceil((s.t0 + 3.0 * s.tau + depth / float(np.min(p.c0))) / dt))


def hann(n):
    return 0.5 * (1.0 - np.cos(2.0 * np.pi * np.arange(n) / max(n - 1, 1)))


def dft_bin(trace, dt, freq):
    """|sum w(t) x(t) e^(-2 pi i f t)| * 2 / sum w, Hann window w over the trace (the amplitude of
    a steady sinusoid of that frequency)."""
    w = hann(len(trace))
    t = np.arange(len(trace)) * dt
    return float(np.abs(np.sum(w * trace * np.exp(-2j * np.pi * freq * t))) * 2.0 / max(np.sum(w), 1e-300))


def fubini_estimate(p, depth, dt, count):
    """The plane-wave second harmonic expected at `depth` in the weak (pre-shock) regime, where the
    Fubini solution reduces to its first term: p = p1(tau) + depth * beta/(2 rho0 c^3) *
    d(p1^2)/dtau, with p1 the source burst delayed by depth/c and beta = 1 + B/(2A) (the
    nonlinearity of this system's equation of state and mass flux, which also sets the Fubini
    shock distance rho0 c^3/(beta omega p0)). Returns the 2 f0 and f0 bins of that waveform under
    the same Hann window as the measured trace (so the burst's own spectral tail at 2 f0 enters
    both), and sigma = depth / shock distance at the burst's peak."""
    s = p.source
    c, r0, ba = float(p.c0[0]), float(p.rho0[0]), float(p.b_over_a[0])
    beta = 1.0 + ba / 2.0
    t = (np.arange(count) + 1) * dt
    tr = t - depth / c
    p1 = s.amplitude * np.sin(2 * np.pi * s.frequency * tr + s.phase) * np.exp(-((tr - s.t0) / s.tau) ** 2)
    dp1sq = np.gradient(p1 * p1, dt)
    pq = p1 + depth * beta / (2.0 * r0 * c ** 3) * dp1sq
    sigma = depth * beta * 2 * np.pi * s.frequency * s.amplitude / (r0 * c ** 3)
    return dft_bin(pq, dt, 2 * s.frequency), dft_bin(pq, dt, s.frequency), sigma


CONTRACTION_NAMES = ["valu", "mfma_f32", "mfma_bf16x3"]


def _mixed_vs(p, a, b, tr_a, tr_b, dt, f0):
    """Accuracy of a mixed state `a` (and probe trace) against the FP64 state `b`: rel L2 per
    field, peak pressure and second-harmonic errors."""
    rel = [float(np.linalg.norm(a[f] - b[f]) / max(np.linalg.norm(b[f]), 1e-300)) for f in range(4)]
    ba = (p.b_over_a / (2 * p.rho0))[:, None]
    c2 = (p.c0 ** 2)[:, None]
    ra = a[0].reshape(p.E, -1)
    rb = b[0].reshape(p.E, -1)
    pka = float(np.abs(c2 * (ra + ba * ra * ra)).max())
    pkb = float(np.abs(c2 * (rb + ba * rb * rb)).max())
    ha, hb = dft_bin(tr_a, dt, 2 * f0), dft_bin(tr_b, dt, 2 * f0)
    return dict(rel_l2=rel, rel_l2_max=max(rel), peak_p_rel_err=abs(pka - pkb) / max(pkb, 1e-300),
                second_harmonic_rel_err=abs(ha - hb) / max(hb, 1e-300), peak_p=pkb)


def compare_precisions(pk, p, steps, warmup, mixed_contraction=0, other_contractions=True, timing_steps=20,
                       until_burst_passed=False, full_length_others=()):
    """FP64 and mixed runs of the same problem from t = 0: speed, and accuracy of mixed against
    FP64 at t_end (the mixed run's derivative contraction: 0 VALU, 1 MFMA FP32, 2 MFMA
    split-BF16). With until_burst_passed the runs go on until the whole tone burst has passed
    the on-axis probe (at least `steps`), so that the probe trace holds it; the second harmonic
    is measured there and set beside the plane-wave Fubini estimate. The other two contractions
    run the same whole trajectory when named in full_length_others; otherwise they are timed over
    `timing_steps` and compared with an FP64 run of the same length."""
    import torch
    pr = probe_node(p)
    depth = probe_depth(p, pr)
    if until_burst_passed:
        steps = max(steps, burst_window_steps(p, pk.problem.cfl_dt_host(p), depth) - warmup)
    total = steps + warmup
    ctx64, r64 = timed_run(pk, p, pk.dg3d.FP64, steps, warmup, probe=pr)
    s64 = ctx64.get_state(0)
    tr64 = ctx64.get_probes()[:, 0]
    ctx64.close()
    ctx32, r32 = timed_run(pk, p, pk.dg3d.MIXED, steps, warmup, probe=pr, contraction=mixed_contraction)
    s32 = ctx32.get_state(0)
    tr32 = ctx32.get_probes()[:, 0]
    ctx32.close()
    a = s32.reshape(4, -1)
    b = s64.reshape(4, -1)
    del s32
    f0 = p.source.frequency
    dt = r64["dt"]
    acc = _mixed_vs(p, a, b, tr32, tr64, dt, f0)
    rel = acc["rel_l2"]
    del a
    h64 = dft_bin(tr64, dt, 2 * f0)
    f64 = dft_bin(tr64, dt, f0)
    rate64 = p.dof * 5 * steps / r64["elapsed"]
    rate32 = p.dof * 5 * steps / r32["elapsed"]
    out = dict(
        workload=p.name, elements=list(p.num_elements), order=p.order, dof=p.dof, steps=total, dt=dt,
        t_end=r64["t_end"], timed_steps=steps, mixed_contraction=CONTRACTION_NAMES[mixed_contraction],
        fp64_dof_stage_per_s=rate64, mixed_dof_stage_per_s=rate32, mixed_speedup=rate32 / rate64,
        fp64_stage_ms=r64["stage_ms"], mixed_stage_ms=r32["stage_ms"],
        fp64_variant=[r64["variant"], r64["xcd"]], mixed_variant=[r32["variant"], r32["xcd"]],
        fp64_hbm_frac=p.dof * BYTES_FP64 / (r64["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        mixed_hbm_frac=p.dof * BYTES_MIXED / (r32["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        mixed_rel_l2_max=max(rel), mixed_rel_l2=rel,
        peak_p_fp64=acc["peak_p"], peak_p_rel_err=acc["peak_p_rel_err"],
        probe_depth_m=depth, probe_trace_max_p_fp64=float(np.abs(tr64).max()),
        fundamental_fp64=f64, second_harmonic_fp64=h64, second_harmonic_ratio_fp64=h64 / max(f64, 1e-300),
        second_harmonic_rel_err=acc["second_harmonic_rel_err"],
        harmonic_window="Hann over the whole probe trace, t = dt .. t_end",
    )
    if until_burst_passed and p.source.lead is None:
        e2, e1, sigma = fubini_estimate(p, depth, dt, len(tr64))
        out.update(fubini_second_harmonic=e2, fubini_ratio=e2 / max(e1, 1e-300), fubini_sigma=sigma,
                   second_harmonic_vs_fubini=h64 / max(e2, 1e-300))
    others = {}
    for con in full_length_others:
        ctxm, rm = timed_run(pk, p, pk.dg3d.MIXED, steps, warmup, probe=pr, contraction=con)
        sm = ctxm.get_state(0).reshape(4, -1)
        trm = ctxm.get_probes()[:, 0]
        ctxm.close()
        accm = _mixed_vs(p, sm, b, trm, tr64, dt, f0)
        del sm
        others[CONTRACTION_NAMES[con]] = dict(stage_ms=rm["stage_ms"], variant=[rm["variant"], rm["xcd"]],
                                              speedup_vs_fp64=r64["stage_ms"] / rm["stage_ms"], steps=total,
                                              rel_l2_max=accm["rel_l2_max"], rel_l2=accm["rel_l2"],
                                              peak_p_rel_err=accm["peak_p_rel_err"],
                                              second_harmonic_rel_err=accm["second_harmonic_rel_err"])
    del b, s64
    if other_contractions:
        short64, rs = timed_run(pk, p, pk.dg3d.FP64, timing_steps, warmup, variant=r64["variant"], xcd=r64["xcd"])
        ref = short64.get_state(0).reshape(4, -1)
        short64.close()
        for con in range(len(CONTRACTION_NAMES)):
            if con == mixed_contraction or con in full_length_others:
                continue
            ctxm, rm = timed_run(pk, p, pk.dg3d.MIXED, timing_steps, warmup, contraction=con)
            sm = ctxm.get_state(0).reshape(4, -1)
            ctxm.close()
            relm = [float(np.linalg.norm(sm[f] - ref[f]) / max(np.linalg.norm(ref[f]), 1e-300)) for f in range(4)]
            others[CONTRACTION_NAMES[con]] = dict(stage_ms=rm["stage_ms"], variant=[rm["variant"], rm["xcd"]],
                                                  speedup_vs_fp64=r64["stage_ms"] / rm["stage_ms"],
                                                  steps=timing_steps + warmup, rel_l2_max=max(relm))
        del ref
    out["other_contractions"] = others
    # the fastest contraction among those run over the whole trajectory, by stage time
    best = (CONTRACTION_NAMES[mixed_contraction], r64["stage_ms"] / r32["stage_ms"], max(rel))
    for name, o in others.items():
        if o["steps"] == total and o["speedup_vs_fp64"] > best[1]:
            best = (name, o["speedup_vs_fp64"], o["rel_l2_max"])
    out.update(fastest_contraction=best[0], fastest_stage_speedup_vs_fp64=best[1], fastest_rel_l2_max=best[2])
    torch.cuda.synchronize()
    return out


def batch_run(pk, steps, warmup=1):
    """The batched workload: 256 seeded unit-cube solves (4..16 elements per side, N 2..6), FP32
    arithmetic / FP64 state, under the tuned stage schedule, replayed from step graphs; accuracy of
    every solve against the same batch in FP64 at the end. The stage time of every order in one
    persistent launch (schedule 4) is reported beside it (one_launch_stage_ms)."""
    import torch
    ps = pk.problem.config_c4(steps=steps)
    descs = [pk.problem.batch_solve(p) for p in ps]
    out = dict(workload="256 unit-cube solves, 4..16 elements per side, N 2..6 (mt19937_64 seed 0), "
                        "Gaussian pulse, 2-element sponge, LSRK54 CFL 0.3, FP32 arithmetic / FP64 state")
    runs = {}
    for name, prec in (("mixed", pk.dg3d.MIXED), ("fp64", pk.dg3d.FP64)):
        b = pk.dg3d.Batch(descs, precision=prec)
        b.autotune()
        b.step(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        el = time.perf_counter() - t0
        info = b.info()
        info["schedule"] = b.schedule()
        stage_ms, _ = b.last_step_timing()
        # every order in one persistent launch (schedule 4), timed the same way, for comparison;
        # both precisions take the same steps, so that their states compare at one time
        tuned = info["schedule"]
        b.set_schedule(4)
        b.step(steps)
        one_ms, _ = b.last_step_timing()
        b.set_schedule(tuned)
        b.step(steps)
        runs[name] = dict(b=b, elapsed=el, info=info, stage_ms=stage_ms, one_ms=one_ms)
    nodes = runs["mixed"]["info"]["total_nodes"]
    worst = 0.0
    for i in range(len(ps)):
        a = runs["mixed"]["b"].get_state(i).reshape(4, -1)
        r = runs["fp64"]["b"].get_state(i).reshape(4, -1)
        for f in range(4):
            nr = np.linalg.norm(r[f])
            if nr > 0:
                worst = max(worst, float(np.linalg.norm(a[f] - r[f]) / nr))
    for r in runs.values():
        r["b"].close()
    m, d = runs["mixed"], runs["fp64"]
    out.update(
        solves=len(ps), nodes=nodes, steps=steps,
        dof_stage_per_s=nodes * 5 * steps / m["elapsed"], stage_ms=m["stage_ms"],
        hbm_frac=nodes * BYTES_MIXED / (m["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        fp64_dof_stage_per_s=nodes * 5 * steps / d["elapsed"], fp64_stage_ms=d["stage_ms"],
        fp64_hbm_frac=nodes * BYTES_FP64 / (d["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        mixed_rel_l2_max_vs_fp64=worst,
        orders=m["info"]["orders"], variants=m["info"]["variants"], tasks=m["info"]["tasks"],
        grids=m["info"]["grids"], schedule=m["info"]["schedule"], fp64_variants=d["info"]["variants"],
        fp64_schedule=d["info"]["schedule"],
        one_launch_stage_ms=m["one_ms"], fp64_one_launch_stage_ms=d["one_ms"],
    )
    return out


def cpu_baseline(p, steps):
    """The CPU oracle (per-line dense solves, OpenMP over elements) on `steps` LSRK54 steps of the
    same workload, on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Oracle, load_problem_into
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = Oracle(p, literal=False, num_registers=3, threads=threads)
    load_problem_into(o, p)
    dt = importlib.import_module("3ddg_amd").problem.cfl_dt_host(p)
    o.set_time(0.0)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.lsrk54(dt, 0, 1, 2)
    el = time.perf_counter() - t0
    del o
    return dict(value=p.dof * 5 * steps / el, unit="DOF-stage/s", cores=threads, kind="port",
                sample=f"{steps} LSRK54 step(s) (5 stages) of the full {p.name} block "
                       f"({p.num_elements[0]}^3 elements, N={p.order}, {p.dof} nodes), FP64 oracle "
                       f"with per-line dense solves, OpenMP over elements, {el:.1f} s",
                host=host_cpu())


def host_cpu():
    """The host's CPU model (from /proc/cpuinfo) and the cores this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return dict(model=model, cpus_available=avail, cpus_online=os.cpu_count())


def _split_blocks(nE, count):
    """Blocks per axis (a product equal to `count`, each dividing nE) for one block per core:
    the axis with the longest block is halved first."""
    nb = [1, 1, 1]
    left = count
    while left > 1:
        for f in (2, 3, 5, 7):
            if left % f == 0:
                break
        else:
            f = left
        order = sorted(range(3), key=lambda a: -nE[a] // nb[a])
        for a in order:
            if (nE[a] // nb[a]) % f == 0:
                nb[a] *= f
                break
        else:
            break
        left //= f
    return nb


def cpu_literal_baseline(pk, nE_full, order, h, cores, steps):
    """The reference's own arithmetic and parallel model on the CPU: the block split into one
    block per core (the reference's NUM_BLOCKS, one per single-threaded MPI rank), each block on
    its own std::thread running the oracle in literal mode (FP32 state and flux constants,
    per-line dense solves, Grid::RK33 stages). The blocks exchange no halos, so the MPI exchange
    is left out of the time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    nb = _split_blocks(nE_full, cores)
    sub = tuple(nE_full[a] // nb[a] for a in range(3))
    blocks = []
    for bi in range(nb[0]):
        for bj in range(nb[1]):
            for bk in range(nb[2]):
                p = pk.problem._burst_problem("c2_block", order, sub, h, steps)
                p.sigma = pk.dg3d.sponge_sigma(sub, tuple(nb), (bi, bj, bk), (h, h, h), (8, 8, 8, 8, 0, 8),
                                               3 * 1500.0 / (8 * h), 1)
                if bk > 0:
                    p.source = pk.problem.Source()
                o = ob.Oracle(p, literal=True, num_registers=4)
                o.zero(0)
                o.set_time(0.0)
                blocks.append(o)
    dt = pk.problem.cfl_dt_host(p)
    t0 = time.perf_counter()
    ob.rk33_blocks(blocks, dt, steps)
    el = time.perf_counter() - t0
    nodes = int(np.prod(nE_full)) * (order + 1) ** 3
    del blocks
    return dict(value=nodes * 3 * steps / el, unit="DOF-stage/s", cores=int(np.prod(nb)), kind="port",
                sample=f"{steps} RK33 step(s) (3 stages) of the {nE_full[0]}x{nE_full[1]}x{nE_full[2]} N={order} "
                       f"block split into {nb[0]}x{nb[1]}x{nb[2]} blocks of {sub[0]}x{sub[1]}x{sub[2]} elements, one "
                       f"std::thread per block (the reference's one block per MPI rank, exchange left out), "
                       f"literal FP32 oracle (the reference's arithmetic), {el:.1f} s",
                host=host_cpu())


def traffic_from_profiles(name):
    path = os.path.join(ROOT, "profiles", f"{name}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("traffic_bytes_per_launch")
    except Exception:
        return None


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    pk = importlib.import_module("3ddg_amd")
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.batch_only:
        print(json.dumps({"batch_c4": batch_run(pk, args.batch_steps)}))
        return
    if args.case in ("c2_rk33", "c2_ldg"):
        p = pk.problem.config_c2(num_elements=args.elements, steps=args.steps)
        if args.case == "c2_ldg":
            p.diffusion = np.full(p.E, 1.0e-6)
        ctx, r = timed_run(pk, p, pk.dg3d.FP64, args.steps, args.warmup, variant=args.variant, xcd=args.xcd,
                           scheme=pk.dg3d.RK33 if args.case == "c2_rk33" else None, with_q=args.case == "c2_ldg")
        ctx.close()
        print(json.dumps(dict(case=args.case, dof=p.dof, steps=args.steps, stage_ms=r["stage_ms"],
                              variant=[r["variant"], r["xcd"]])))
        return
    if args.case:
        name, rest = args.case.split("_", 1)
        if name in ("c3m", "c3"):
            p = pk.problem.config_c3(steps=args.steps)
        elif name in ("n6", "n6m"):
            p = pk.problem.config_c2(num_elements=48, steps=args.steps, order=6)
        else:
            p = pk.problem.config_c5(steps=args.steps)
        if rest == "fp64":
            prec, con, bpn = pk.dg3d.FP64, 0, BYTES_FP64
        else:
            prec, bpn = pk.dg3d.MIXED, BYTES_MIXED
            con = {"valu": pk.dg3d.CONTRACT_VALU, "mfma_f32": pk.dg3d.CONTRACT_MFMA_F32,
                   "mfma_bf16": pk.dg3d.CONTRACT_MFMA_BF16}[rest]
        ctx, r = timed_run(pk, p, prec, args.steps, args.warmup, variant=args.variant, xcd=args.xcd, contraction=con)
        ctx.close()
        print(json.dumps(dict(case=args.case, elements=list(p.num_elements), order=p.order, dof=p.dof,
                              steps=args.steps, stage_ms=r["stage_ms"], variant=[r["variant"], r["xcd"]],
                              hbm_frac=p.dof * bpn / (r["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              contraction_flop_per_node_stage=12 * (p.order + 1))))
        return
    p = pk.problem.config_c2(num_elements=args.elements, steps=args.steps)
    ctx, r = timed_run(pk, p, pk.dg3d.FP64, args.steps, args.warmup, variant=args.variant, xcd=args.xcd)
    maxabs = ctx.reduce(pk.dg3d.REDUCE_MAXABS, 0)
    ctx.close()
    progress(f"headline 64^3 N={p.order} FP64: {args.steps} steps, stage {r['stage_ms']:.4f} ms")
    if not np.all(np.isfinite(maxabs[:4])):
        print("non-finite state after the run", file=sys.stderr)
        sys.exit(3)
    value = p.dof * 5 * args.steps / r["elapsed"]
    achieved = p.dof * BYTES_FP64 / (r["stage_ms"] * 1e-3) / 1e9
    out = {
        "metric": "DOF-stage/s (FP64 drop-in fused LSRK54 stage; % HBM peak)",
        "value": value,
        "unit": "DOF-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded tone burst, zero initial state)",
        "config": {
            "workload": f"{p.num_elements[0]}^3 hex elements, N={p.order} LGL, 4 FP64 fields, water B/A=5, "
                        f"z=0 1 MHz 1e6 Pa tone burst, 8-element sponge, LSRK54 CFL {p.cfl}",
            "elements": p.E, "order": p.order, "dof": p.dof, "scheme": "LSRK54", "dt": r["dt"],
            "t_end": r["t_end"], "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles("stage_fp64_c2"),
            "kernel": "k_rhs<NN=5, EPB, MINW, double, EPI_LSRK, double> (fused stage)",
            "variant": r["variant"], "xcd_remap": r["xcd"], "avg_launch_ms": r["stage_ms"],
            "bytes_per_node_stage": BYTES_FP64,
        },
    }
    if not args.no_extra:
        sc = stream_ceiling()
        if sc is not None:
            out["stream_ceiling"] = dict(
                GBs=sc["stream_GBs"], frac_of_peak=sc["stream_GBs"] / HBM_PEAK_GBS,
                stage_frac_of_ceiling=achieved / sc["stream_GBs"],
                pattern=sc["pattern"])
        progress("streaming probe done")
        # the precision comparisons run from t = 0 until the whole burst has passed the probe
        # (at least the workloads' 500 steps), whatever --steps is
        out["mixed_c2"] = compare_precisions(pk, pk.problem.config_c2(num_elements=args.elements),
                                             args.precision_steps, args.warmup, timing_steps=args.steps,
                                             until_burst_passed=True)
        progress("precision comparison 64^3 N=4 done")
        # the N=7 precision workload runs its derivative contraction as split-BF16 MFMA GEMMs;
        # the FP32 MFMA and VALU contractions run the same whole trajectory beside it
        out["mixed_c3"] = compare_precisions(pk, pk.problem.config_c3(num_elements=args.elements),
                                             args.precision_steps, args.warmup,
                                             mixed_contraction=pk.dg3d.CONTRACT_MFMA_BF16, timing_steps=args.steps,
                                             until_burst_passed=True,
                                             full_length_others=(pk.dg3d.CONTRACT_MFMA_F32, pk.dg3d.CONTRACT_VALU))
        progress("precision comparison 64^3 N=7 done")
    if not args.no_extra:
        out["small_c1"] = small_block_run(pk)
        progress("8^3 N=3 block done")
        out["other_paths_c2"] = other_paths_run(pk, args.elements, min(args.steps, 200), args.warmup)
        progress("RK33 and LDG paths done")
    if not args.no_batch and not args.no_extra:
        out["batch_c4"] = batch_run(pk, args.batch_steps)
        progress("256-solve batch done")
    if not args.no_extra and args.c5_steps > 0:
        # combined workload: stretched two-layer 48^3 N=5 with a focused burst; mixed with the
        # split-BF16 matrix-core contraction against FP64, and the FP32 matrix-core and VALU
        # contractions over the same trajectory beside it
        out["combined_c5"] = compare_precisions(pk, pk.problem.config_c5(steps=args.c5_steps),
                                                args.c5_steps - args.warmup, args.warmup,
                                                mixed_contraction=pk.dg3d.CONTRACT_MFMA_BF16,
                                                other_contractions=False,
                                                full_length_others=(pk.dg3d.CONTRACT_MFMA_F32,
                                                                    pk.dg3d.CONTRACT_VALU))
        progress("48^3 N=5 combined workload done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pk.problem.config_c2(num_elements=args.elements), args.cpu_steps)
        progress("CPU oracle baseline done")
        cores = out["cpu_baseline"]["cores"]
        out["cpu_literal_rk33"] = cpu_literal_baseline(pk, (args.elements,) * 3, 4, 0.75e-3, cores,
                                                       max(1, args.cpu_steps // 2))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
