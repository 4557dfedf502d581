"""Throughput of the MPLS interface-transport step on one MI355X.

Workload: the 3D LeVeque deformation at 256^3 -- sphere r = 0.15 at
(0.35, 0.35, 0.35), velocity x cos(pi t / T) with T = 3, dt = rdt = 0.1 dx,
ls_wid = 1.5 dx, reinitialisation every step, rho_12 = 1.  One step = the
hot path of one physical step: device velocity at t^{n+1}, CRWENO5 TVD-RK3
advection, reinitialisation (btn = 1), mass correction.  Inputs are resident
in HBM before the timed region starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--precision fp64|fp32]

Prints one JSON line (metric, value = cell-updates/s, roofline of the dominant
kernel from HIP-event timing, the chosen forms over two windows of steps, and
the CPU baseline: the oracle's OpenMP build on every host core).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mass-preserving-level-set-method_amd")
METRIC = "cell-updates/s per RK3 step (adv+reinit+mass fix) at 256^3; L1 shape err, |dM|/M"
HBM_PEAK_GBS = 8000.0
T2S = 3.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_lsmp():
    import importlib.util
    spec = importlib.util.spec_from_file_location("lsmp", os.path.join(PKG, "lsmp.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lsmp"] = mod
    spec.loader.exec_module(mod)
    return mod


def sphere_box(n):
    """phi0 = 0.15 - |x - (0.35, 0.35, 0.35)| on the ghosted box with the
    reference's single-precision literals (Vortex/src/setup/problem_init.f90:69)."""
    xb = (np.arange(-3, n + 3) + 0.5) / n
    c, r = float(np.float32(0.35)), float(np.float32(0.15))
    z, y, x = np.meshgrid(xb, xb, xb, indexing="ij", sparse=True)
    box = -np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2) + r
    # ghosts are copies of the nearest interior cell (zero-gradient fill)
    idx = np.clip(np.arange(n + 6) - 3, 0, n - 1) + 3
    return np.ascontiguousarray(box[np.ix_(idx, idx, idx)])


# Algorithmic bytes per interior cell per launch, in units of V = sizeof(real):
# the minimum each kernel must read and write (scratch of the line solves and
# halo re-reads are not counted).  The mass correction's kernels work on the
# cells of the lines that cross the band (BAND_CELLS); their figures count
# those cells, and the mass update's only its read of phi (the read of
# |grad phi| and the store happen on the band cells alone).  A reinitialisation
# stage counts the reference's stored-increment form, 16 V per pseudo-step
# over its three stages (stage 1: phi, S0 in, phi, l1, phi_start out; stage
# 2: phi, S0, l1 in, phi, l2 out; stage 3: phi, S0, l1, l2, phi_start in, phi
# out); the Shu-Osher form run here moves 11 V (REDIS_MOVED_V).
BYTES_V = {
    "adv_x": 3, "adv_y": 4,            # 3D y pass accumulates into R
    "redis_stage": 16.0 / 3.0,
    "ccd_x": 2, "ccd_y": 2, "ccd_z": 4, "mass_update": 1, "line_hsums": 1, "band_lines": 1,
    "velocity": 3, "sign_init": 2,
}
REDIS_MOVED_V = 11.0 / 3.0             # stages read phi, S0 [, phi0], write phi: 3, 4, 4
ADV_Z_V = (5 + 6 + 6) / 3.0            # last pass: R, w, phi (+ l1, l2) in, phi (+ l1/l2) out


def kernel_cells(name, n, lines):
    """Interior cells one launch of `name` covers: the listed lines of its
    direction for the mass correction's CCD passes and update (lines = the
    last correction's counts along x, y, z), the unlisted z lines for
    line_hsums, every cell otherwise."""
    d = {"ccd_x": 0, "ccd_y": 1, "ccd_z": 2, "mass_update": 2}.get(name)
    if d is not None and lines[d] >= 0:
        return lines[d] * n
    if name == "line_hsums":
        return (n * n - lines[2]) * n
    return n ** 3


# kernel names of the committed counter passes for each timed kernel
PMC_NAMES = {
    "redis_stage": "k_redis3d<", "adv_x": "k_adv_tile<double, 0,", "adv_y": "k_adv_tile<double, 1,",
    "adv_z": "k_adv_tile<double, 2,", "ccd_x": "k_ccd_part<double, 3, false", "ccd_y": "k_ccd_part<double, 3, false",
    "ccd_z": "k_ccd_part<double, 3, true", "mass_update": "k_mass_update_lines<", "line_hsums": "k_line_hsums<",
    "sign_init": "k_sign_init<", "velocity": "k_velocity<",
}
FP64_PEAK_TFS = 78.6
# the learned form advects and reinitialises within |phi| < 12 dx (the
# network's band is 6 dx, its stencil 2 dx more)
LEARNED_TUBE_CELLS = 12.0


def pmc_path(n, fp64):
    return os.path.join(ROOT, "profiles", f"pmc_traffic_{n}cube_{'fp64' if fp64 else 'fp32'}.json")


def pmc_figures(kernel, n, fp64):
    """(HBM bytes, FP64 flop) per launch of `kernel` from the committed
    rocprofv3 counter passes (FETCH_SIZE and WRITE_SIZE calibrated on a copy of
    known size; FP64 VALU instruction counts, FMA as two flop), or
    (None, None) when no pass covers this kernel, size and precision."""
    path = pmc_path(n, fp64)
    if not os.path.exists(path) or kernel not in PMC_NAMES:
        return None, None
    data = json.load(open(path))["kernels"]
    hits = [v for k, v in data.items() if PMC_NAMES[kernel] in k]
    if not hits:
        return None, None
    traffic = sum(v["traffic_bytes"] for v in hits) / len(hits)
    flop = sum(v.get("fp64_flop_per_launch", 0.0) for v in hits) / len(hits)
    return round(traffic), flop


def measured_traffic(kernel, n, fp64):
    return pmc_figures(kernel, n, fp64)[0]


def traffic_source(n, fp64):
    """Where the counter figures come from: the committed pass's file, the
    commit of the kernels it measured and its command."""
    path = pmc_path(n, fp64)
    if not os.path.exists(path):
        return None
    meta = json.load(open(path))
    return {"file": os.path.relpath(path, ROOT), "kernels_at_commit": meta.get("commit"),
            "command": meta.get("command")}


def stream_copy_gbs(nbytes=2 << 30, reps=10):
    """HBM bandwidth of a device-to-device copy (bytes read + written per
    second) through the HIP runtime: the peak measured on this box, reported
    beside the published one.  None if the buffers cannot be allocated."""
    import ctypes as C
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    hip.hipEventDestroy.argtypes = [C.c_void_p]
    a, b, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), nbytes) != 0:
        return None
    if hip.hipMalloc(C.byref(b), nbytes) != 0:
        hip.hipFree(a)
        return None
    try:
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        D2D = 3   # hipMemcpyDeviceToDevice
        hip.hipMemcpyAsync(b, a, nbytes, D2D, None)   # warm-up
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, nbytes, D2D, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        return round(2 * nbytes * reps / (ms.value * 1e-3) / 1e9, 1) if ms.value > 0 else None
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def run_steps(s, nsteps, t, imass, field, net=None, learned=False):
    """MPLS steps; learned: the learned reinitialisation with `net` (None: the
    sweeps-only control).  Returns (t, pseudo-steps, mass iterations,
    fallbacks)."""
    import lsmp as L
    npseudo = nmass = fallbacks = 0
    for _ in range(nsteps):
        t = t + s.params.dt
        if not learned:
            a, b = s.mpls_step(t, T2S, field, imass)
        else:
            s.set_velocity_analytic(field, t, T2S)
            s.rk3_advect()
            _, fb, a = s.reinit_learned(net, 6.0, 2, 1e-2, 1.5, L.FALLBACK_CONTINUE)
            fallbacks += fb
            _, b = s.mass_fix(imass)
        npseudo += a
        nmass += b
    return t, npseudo, nmass, fallbacks


FORMS = (  # name, precision, learned reinitialisation, network, narrow band half-width (cells)
    ("fp32-plain", "fp32", False, False, 0.0),
    ("fp32-learned", "fp32", True, True, LEARNED_TUBE_CELLS),
    ("fp32-sweeps", "fp32", True, False, LEARNED_TUBE_CELLS),
)


def form_window(lsmp, n, prec, steps, warmup, learned=False, net=None, narrow_band=0.0, start=None, t_start=0.0):
    """One form over a window of steps: from the initial sphere, or from the
    field `start` at time t_start (the L1 shape error still against the
    initial sphere, the mass against the initial mass)."""
    s = lsmp.Solver(lsmp.make_params(n, dim=3, precision=prec))
    s.set_narrow_band(narrow_band)
    s.upload_phi(sphere_box(n))
    s.store_phi0()
    imass, _ = s.mass_vol()
    if start is not None:
        s.upload_phi(start)
    t, _, _, _ = run_steps(s, warmup, t_start, imass, lsmp.VEL_LEVEQUE3D, net, learned)
    s.synchronize()
    t0 = time.perf_counter()
    t, npseudo, _, fallbacks = run_steps(s, steps, t, imass, lsmp.VEL_LEVEQUE3D, net, learned)
    s.synchronize()
    el = time.perf_counter() - t0
    mass, _ = s.mass_vol()
    out = {"value": n ** 3 * steps / el, "ms_per_step": el / steps * 1e3,
           "pseudo_steps_per_step": npseudo / steps, "fallbacks": fallbacks,
           "t_end": round(t, 6), "l1_shape_err": s.interface_l1(), "dM_over_M": abs(mass - imass) / imass,
           "narrow_band_cells": narrow_band}
    s.close()
    return out


def half_period_field(lsmp, n, net, warmup):
    """The field `warmup` steps before t = T/2, made by the FP32 + learned form
    with the narrow band from the initial sphere (the cheapest way there);
    every form's second window starts from it."""
    s = lsmp.Solver(lsmp.make_params(n, dim=3, precision=lsmp.FP32))
    s.set_narrow_band(LEARNED_TUBE_CELLS)
    s.upload_phi(sphere_box(n))
    s.store_phi0()
    imass, _ = s.mass_vol()
    nsteps = int(round(0.5 * T2S / s.params.dt)) - 12 - warmup
    t, _, _, _ = run_steps(s, nsteps, 0.0, imass, lsmp.VEL_LEVEQUE3D, net, True)
    field = s.download_phi()
    s.close()
    return field, t


CPU_SAMPLE = """
import json, sys, time
sys.path.insert(0, {tests!r})
import oracle_lib as O
o = O.Oracle(O.make_params({n}, dim=3))
o.initial_phi(2, 0)
o.store_initial()
t0 = time.perf_counter()
for _ in range({steps}):
    o.mpls_step(2, {t2s}, 1)
dt = time.perf_counter() - t0
print(json.dumps(dict(seconds=dt, pseudo=o.get("redis_iters"), mass=o.get("mass_iters"))))
"""


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(n_sample=128, steps=3):
    """The FP64 C oracle built with OpenMP (oracle/_build/liblsmp_oracle_omp.so,
    its line solves, stencils and ghost fills on every host thread; the sums
    sequential) on a bounded sample of the same workload: `steps` MPLS steps of
    the LeVeque sphere at n_sample^3, in a child process."""
    so = os.path.join(ROOT, "oracle", "_build", "liblsmp_oracle_omp.so")
    if not os.path.exists(so):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    env = dict(os.environ, LSMP_ORACLE_SO=so, OMP_NUM_THREADS=str(threads))
    code = CPU_SAMPLE.format(tests=os.path.join(ROOT, "tests"), n=n_sample, steps=steps, t2s=T2S)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log(f"bench: CPU baseline failed: {r.stderr[-400:]}")
        return None
    m = json.loads(r.stdout.strip().splitlines()[-1])
    return {
        "value": n_sample ** 3 * steps / m["seconds"],
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "cpu": cpu_model(),
        "sample": f"{steps} MPLS steps of the same LeVeque sphere case at {n_sample}^3 from t = 0 "
                  f"(the FP64 C oracle with OpenMP on {threads} threads, {m['seconds']:.1f} s; "
                  f"{int(m['pseudo'])} pseudo-steps and {int(m['mass'])} mass iterations in the last step)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the FP32 and learned-form windows")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path's sharding "
              "across GPUs is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)

    lsmp = load_lsmp()
    n = args.n
    prec = lsmp.FP64 if args.precision == "fp64" else lsmp.FP32
    s = lsmp.Solver(lsmp.make_params(n, dim=3, precision=prec))
    s.upload_phi(sphere_box(n))
    s.store_phi0()
    imass, _ = s.mass_vol()
    log(f"bench: {n}^3 {args.precision}, imass {imass:.9e}; warmup {args.warmup} steps")

    t = 0.0
    t, _, _, _ = run_steps(s, args.warmup, t, imass, lsmp.VEL_LEVEQUE3D)
    s.synchronize()
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    t, npseudo, nmass, _ = run_steps(s, args.steps, t, imass, lsmp.VEL_LEVEQUE3D)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = n ** 3 * args.steps / elapsed
    mass, _ = s.mass_vol()
    l1 = s.interface_l1()
    log(f"bench: {ms_per_step:.2f} ms/step, {value:.3e} cell-updates/s, "
        f"pseudo-steps/step {npseudo / args.steps:.1f}, mass iters/step {nmass / args.steps:.2f}")

    # per-kernel device time with HIP events on the library's stream
    s.set_profiling(True)
    t, pp, pm, _ = run_steps(s, args.profile_steps, t, imass, lsmp.VEL_LEVEQUE3D)
    kernels = {}
    for name in ("adv_x", "adv_y", "adv_z", "redis_stage", "redis_band", "sign_init", "band_lines", "line_hsums",
                 "ccd_x", "ccd_y", "ccd_z", "finalize", "mass_update", "velocity"):
        cnt, ms = s.kernel_time(name)
        if cnt:
            kernels[name] = {"launches": cnt, "total_ms": round(ms, 4), "avg_us": round(1e3 * ms / cnt, 2)}
    lines = s.band_line_counts()
    s.set_profiling(False)
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    V = 8 if prec == lsmp.FP64 else 4
    bv = ADV_Z_V if dom == "adv_z" else BYTES_V.get(dom, 0.0)
    bytes_per_launch = bv * V * kernel_cells(dom, n, lines)
    avg_s = kernels[dom]["avg_us"] * 1e-6
    achieved = bytes_per_launch / avg_s / 1e9
    moved = (REDIS_MOVED_V * V * kernel_cells(dom, n, lines) / avg_s / 1e9) if dom == "redis_stage" else None
    # a 3D reinit stage is two launches: every cell (the dominant kernel), then
    # the band cells; the same bytes over both launches' time
    stage_s = avg_s + (kernels["redis_band"]["avg_us"] * 1e-6 if dom == "redis_stage" and "redis_band" in kernels
                       else 0.0)
    for k, v in kernels.items():
        b = (ADV_Z_V if k == "adv_z" else BYTES_V.get(k, 0.0)) * V * kernel_cells(k, n, lines)
        v["algo_GBs"] = round(b / (v["avg_us"] * 1e-6) / 1e9, 1) if b else None
        tb, fl = pmc_figures(k, n, prec == lsmp.FP64)
        if tb:
            v["traffic_GBs"] = round(tb / (v["avg_us"] * 1e-6) / 1e9, 1)
        if fl:
            v["fp64_TFs"] = round(fl / (v["avg_us"] * 1e-6) / 1e12, 2)
    # on-box streaming ceiling: the runtime's device-to-device copy, or the
    # library's streaming kernels over every cell (sign init: phi in, S0 out;
    # velocity: u, v, w out) at their counter-measured bytes, whichever moves most
    copy_gbs = stream_copy_gbs()
    stream, stream_src = max([(copy_gbs or 0.0, "hipMemcpy device-to-device")] +
                             [(kernels.get(k, {}).get("traffic_GBs") or 0.0, f"k_{k}, counter-measured bytes")
                              for k in ("sign_init", "velocity")])
    stream = stream or None
    _, dom_flop = pmc_figures(dom, n, prec == lsmp.FP64)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == lsmp.FP64 else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"3D LeVeque deformation {n}^3, sphere r=0.15 at (0.35,0.35,0.35), T=3, "
                        f"dt=rdt=0.1dx, ls_wid=1.5dx, reinit every step, MPLS, "
                        f"{'FP64 plain' if prec == lsmp.FP64 else 'FP32 bulk + FP64 mass sums'}",
            "grid": [n, n, n],
            "timed_steps_from_t": round(t - (args.steps + args.profile_steps) * s.params.dt, 6),
        },
        "l1_shape_err": l1,
        "dM_over_M": abs(mass - imass) / imass,
        "pseudo_steps_per_step": npseudo / args.steps,
        "mass_iters_per_step": nmass / args.steps,
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "algo_bytes_per_launch": round(bytes_per_launch),
            # the bytes the Shu-Osher stages actually have to move (11 V per pseudo-step)
            "achieved_moved": round(moved, 1) if moved else None,
            "frac_moved": round(moved / HBM_PEAK_GBS, 4) if moved else None,
            "achieved_stage_with_band": round(bytes_per_launch / stage_s / 1e9, 1) if stage_s > avg_s else None,
            "frac_stage_with_band": round(bytes_per_launch / stage_s / 1e9 / HBM_PEAK_GBS, 4)
            if stage_s > avg_s else None,
            "traffic": measured_traffic(dom, n, prec == lsmp.FP64),
            "traffic_source": traffic_source(n, prec == lsmp.FP64),
            # measured on this box: a device-to-device copy, bytes read + written
            "stream_peak": stream,
            "stream_source": stream_src if stream else None,
            "copy_GBs": copy_gbs,
            "frac_of_stream": round(achieved / stream, 4) if stream else None,
            # the FP64 co-bound of the same kernel (FP64 VALU flop per launch
            # from the counter passes over its average launch time)
            "fp64": ({"achieved": round(dom_flop / avg_s / 1e12, 2), "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                      "frac": round(dom_flop / avg_s / 1e12 / FP64_PEAK_TFS, 4)} if dom_flop else None),
        },
        "kernels": kernels,
        # lines the mass correction solved along x, y, z (the lines crossing
        # the band |phi| < ls_wid; n^2 per direction in all)
        "band_lines": lines,
    }
    s.close()
    if not args.no_forms:
        # the chosen forms beside the plain one over the same windows of steps:
        # from t = 0, and around t = T/2 (maximum stretch) from a common field
        net = lsmp.Net()
        net.load(os.path.join(PKG, "weights", "reinit_mlp.npy"))
        forms = {"fp64-plain": {"from_t0": {"value": value, "ms_per_step": ms_per_step,
                                            "pseudo_steps_per_step": npseudo / args.steps, "fallbacks": 0,
                                            "l1_shape_err": l1, "dM_over_M": abs(mass - imass) / imass}}}
        for name, pr, learned, with_net, nb in FORMS:
            log(f"bench: {name} window from t = 0")
            forms[name] = {"from_t0": form_window(lsmp, n, lsmp.FP32 if pr == "fp32" else lsmp.FP64, args.steps,
                                                  args.warmup, learned, net if with_net else None, nb)}
        log("bench: the common field before t = T/2 (FP32 + learned, narrow band)")
        start, t_start = half_period_field(lsmp, n, net, args.warmup)
        forms["fp64-plain"]["around_T_half"] = form_window(lsmp, n, lsmp.FP64, args.steps, args.warmup,
                                                           start=start, t_start=t_start)
        for name, pr, learned, with_net, nb in FORMS:
            log(f"bench: {name} window around t = T/2")
            forms[name]["around_T_half"] = form_window(lsmp, n, lsmp.FP32 if pr == "fp32" else lsmp.FP64,
                                                       args.steps, args.warmup, learned,
                                                       net if with_net else None, nb, start=start, t_start=t_start)
        net.close()
        for name, f in forms.items():
            for wname, w in f.items():
                w["vs_fp64_plain"] = round(w["value"] / forms["fp64-plain"][wname]["value"], 3)
        out["forms"] = forms
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (the oracle with OpenMP on the host cores)")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
