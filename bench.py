"""Benchmark of the MI355X CTU+CT MHD step: FP64 cell-updates/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--variant exact|fast]

Workload (one step = one iteration of the reference main loop, main.c:221-318:
clamp dt, integrate_2d, advance the clock, new_dt, set_bvals): the MHD blast
wave of athinput.blast scaled to 4096 x 4096 zones on [-0.5,0.5]x[-0.75,0.75],
periodic, Roe flux, PLM, CFL 0.4, FP64, state resident in HBM. The exact
(no-contraction) variant is the default: it is the one that meets the parity
bar against the reference CPU integrator.

Prints one JSON line with the metric, a roofline object for the dominant kernel
(the step's algorithmic bytes, 168 B per cell-update, over that kernel's
HIP-event duration in a kernel-by-kernel pass; its own plane I/O, scratch
included, under separate keys), the step-level roofline at 168 algorithmic
bytes per cell-update, the FP64 work rate, the CPU baseline
(the reference's own CPU integrator timed on this host on a bounded sample),
the batched-ensemble rate and the adjoint/forward cost ratio.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X vector FP64
ALG_BYTES_PER_CELL = 168   # 10 FP64 read (U x8, B1i, B2i) + 11 written (U x8, B1i, B2i, B3i)

# Plane I/O of each kernel, scratch planes included: the planes it reads once
# and writes once (8 B each), times the units of one launch on an N1 x N2
# grid. Stencil neighbours served from cache are not counted again. Reported
# as roofline.kernel_io_*; the roofline's algorithmic bytes are the metric's
# 168 B per cell-update.
KERNEL_IO = {
    # name: (planes read, planes written, units(N1, N2))
    "k_pred_x1": (9, 22, lambda a, b: (a + 4) * (b + 4)),       # U, B1i -> Ul,Ur,F (21) + emf3_cc
    "k_pred_x2": (9, 21, lambda a, b: (a + 5) * (b + 4)),       # U, B2i -> Ul,Ur,F (21)
    "k_corner<pred>": (5, 1, lambda a, b: (a + 4) * (b + 4)),
    # corrector x1 faces: Ul1, Ur1, 6 x2-flux components, B1i, B1c..B3c, M3, d, emf -> F1c
    "k_corr<1>": (27, 7, lambda a, b: (a + 2) * (b + 3)),
    # corrector x2 faces + Step 7: Ul2, Ur2, 6 + 3 flux components, B1i, B2i, cell fields,
    # emf -> F2c, emf3 at t^{n+1/2}
    "k_corr<2>": (33, 8, lambda a, b: (a + 3) * (b + 3)),
    "k_corner<corr>": (5, 1, lambda a, b: (a + 2) * (b + 2)),
    "k_final": (26, 11, lambda a, b: (a + 1) * (b + 1)),        # U, B, 14 fluxes, emf -> U, B
    "k_ghost": (11, 11, lambda a, b: 8 * (a + 8) + 8 * (b + 1)),
    "k_begin": (0, 0, lambda a, b: 1),
    "k_end": (0, 0, lambda a, b: 1),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pmc_summary():
    """Per-kernel HBM traffic (FETCH_SIZE/WRITE_SIZE scaled by the copy-kernel
    calibration) and FP64 VALU instruction counts from the committed
    rocprofv3 --pmc summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", "step_kernels_blast4096_pmc.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)
    except (OSError, IndexError, ValueError):
        return None


def cpu_leg(overrides, steps):
    """The reference's own CPU integrator (oracle/_ref, built from
    /root/reference/src in the build container; the oracle restatement where
    that is absent) on the blast of athinput.blast with these overrides, for
    a bounded number of main-loop steps, one core, in a child process.
    Returns (kind, steps, seconds)."""
    from helpers import REF_DIR, ref_available
    ov = list(overrides)
    if ref_available("roe"):
        code = f"""
import ctypes, time, json
L = ctypes.CDLL({os.path.join(REF_DIR, 'libathref_roe.so')!r})
ov = [o.encode() for o in {ov!r}]
arr = (ctypes.c_char_p * len(ov))(*ov)
L.ref_open({os.path.join(ROOT, 'athena-cuda_amd', 'inputs', 'athinput.blast').encode()!r}, len(ov), arr)
L.ref_setup(0)
t = time.perf_counter()
n = L.ref_run(ctypes.c_double(0.2), {steps})
dt = time.perf_counter() - t
print(json.dumps(dict(steps=n, seconds=dt)))
"""
        kind = "reference"
    else:
        code = f"""
import sys, time, json
sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
from helpers import oracle_setup
g, tlim, _ = oracle_setup("blast", {ov!r})
t = time.perf_counter()
n = g.run(tlim, {steps})
dt = time.perf_counter() - t
print(json.dumps(dict(steps=n, seconds=dt)))
"""
        kind = "port"
    env = dict(os.environ, OMP_NUM_THREADS="1")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         check=True).stdout
    # the reference's own stdout shares the pipe: take the result line
    r = json.loads([x for x in out.splitlines() if x.startswith('{"steps"')][-1])
    return kind, r["steps"], r["seconds"]


def cpu_baseline(n=4096, steps=2):
    """The CPU baseline on a bounded sample of the bench workload itself
    (blast n^2, `steps` main-loop steps, one core), and beside it the
    reference's parity configuration (blast 400x600, 100 steps)."""
    kind, ns, sec = cpu_leg([f"grid/Nx1={n}", f"grid/Nx2={n}"], steps)
    rate = n * n * ns / sec
    kind2, ns2, sec2 = cpu_leg([], 100)
    cpu = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": rate, "unit": "cell-updates/s", "cores": 1, "kind": kind,
            "sample": f"blast {n}x{n}, {ns} main-loop steps, single thread ({sec:.1f} s; "
                      f"host {cpu}, nproc {os.cpu_count()})",
            "blast_400x600_100_steps": {"value": 400 * 600 * ns2 / sec2,
                                        "unit": "cell-updates/s", "cores": 1, "kind": kind2,
                                        "steps": ns2, "seconds": sec2}}


def batched_leg(pkg, args, plain_rate):
    """The batched uneven ensemble (seed 0): every instance runs its own CFL
    dt for the timed steps; aggregate cell-updates/s over all instances."""
    import importlib
    ens = importlib.import_module(pkg.__name__ + ".ensemble")
    params = ens.ensemble_params(seed=0, n=args.batch)
    ncell = ens.cells(params)
    log(f"[bench] batched ensemble: {args.batch} instances, {ncell} cells per step")
    gb = ens.build(params, nlim=args.batch_warmup + args.batch_steps, variant=args.variant)
    shape = None
    if not args.no_tune:
        shape = gb.autotune(2)
        shape = gb.launch_shape() + (shape[3], gb.x2_kind(), gb.fusion(), gb.strip(),
                                     gb.final_rows())
        log(f"[bench] batch shape x2_seg={shape[0]} persist={shape[1]} xcd_remap={shape[2]} "
            f"corr_rows={shape[3]} x2_form={shape[5]} fusion={shape[6]} strip={shape[7]} "
            f"final_rows={shape[8]} "
            f"({shape[4]:.3f} ms/step)")
    for _ in range(args.batch_warmup):
        gb.step()
    gb.sync()
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        gb.step()
    gb.sync()
    sec = time.perf_counter() - t0
    steps = [gb.clock(k)[2] for k in range(gb.n)]
    gb.close()
    if min(steps) != max(steps) or steps[0] != args.batch_warmup + args.batch_steps:
        raise RuntimeError(f"ensemble instances ran {min(steps)}..{max(steps)} steps")
    rate = ncell * args.batch_steps / sec
    log(f"[bench] batched {1e3 * sec / args.batch_steps:.3f} ms/step, {rate / 1e9:.3f} G "
        f"cell-updates/s, {rate / plain_rate:.3f} of plain")
    return {"value": rate, "unit": "cell-updates/s", "instances": args.batch,
            "cells_per_step": ncell, "steps": args.batch_steps,
            "ms_per_step": 1e3 * sec / args.batch_steps, "ratio_to_plain": rate / plain_rate,
            "workload": "seeded uneven blast ensemble (seed 0): nx, ny in {64..1024 step 32}, "
                        "off-centre blasts, Roe even / HLLE odd, per-instance CFL dt, "
                        "one HIP graph per step",
            "shape": None if shape is None else {"x2_seg": shape[0], "persist": shape[1],
                                                 "xcd_remap": shape[2], "corr_rows": shape[3],
                                                 "x2_form": shape[5], "fusion": shape[6],
                                                 "strip": shape[7], "final_rows": shape[8]}}


def adjoint_leg(pkg, args):
    """Differentiation form: loss = 0.5 ||rho(T) - rho_target||^2 after N
    fixed-dt steps at 512^2 (checkpoint every 16). Cost ratio = (forward with
    checkpoints + reverse with recomputation) / plain forward of N steps."""
    import importlib
    import numpy as np
    import torch
    A = importlib.import_module(pkg.__name__ + ".adjoint")
    n, N, every, dt = args.adj_size, args.adj_steps, 16, 2e-4
    dev = torch.device("cuda:0")
    desc = A.sensitivity_desc(n)
    s0_np, _ = A.sensitivity_state(1, n)
    _, target = A.sensitivity_state(2, n)
    s0 = torch.from_numpy(s0_np).to(dev)
    tgt = torch.from_numpy(target).to(dev)
    inner = (slice(4, n + 4), slice(4, n + 4))
    log(f"[bench] adjoint: {n}x{n}, N={N}, checkpoint every {every}")
    # plain forward: the same rollout through the plain step (no checkpoints)
    g = pkg.Grid(desc, variant="exact")
    gas = np.ascontiguousarray(np.moveaxis(s0_np[:8], 0, 2))
    g.upload(0, gas, s0_np[8], s0_np[9], s0_np[10])
    g.set_clock(0, dt, 0.0, 0)
    g.set_bvals()
    for _ in range(3):
        g.integrate()
        g.set_bvals()
    g.sync()
    t0 = time.perf_counter()
    for _ in range(N):
        g.integrate()
        g.set_bvals()
    g.sync()
    t_plain = time.perf_counter() - t0
    g.close()
    h = A.Rollout(desc, dt, N, every)
    x = s0.clone().requires_grad_(True)
    sN = A.rollout(x, h)  # warm-up of both passes
    (0.5 * ((sN[0][inner] - tgt) ** 2).sum()).backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sN = h.forward(s0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    gN = torch.zeros_like(sN)
    gN[0][inner] = sN[0][inner] - tgt
    t0 = time.perf_counter()
    g0 = h.backward(gN)
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    ratio = (t_fwd + t_bwd) / t_plain
    log(f"[bench] adjoint: plain {1e3 * t_plain:.1f} ms, forward+ckpt {1e3 * t_fwd:.1f} ms, "
        f"reverse {1e3 * t_bwd:.1f} ms, ratio {ratio:.2f} (reverse alone {t_bwd / t_plain:.2f})")
    full = full_trajectory(A, desc, dt, N, s0, gN, g0, t_plain)

    def fd_check(state, tg, handle, dirs, eps=1e-6):
        """max relative |FD - adjoint| over the directions (central, eps)."""
        t = torch.from_numpy(tg).to(dev)

        def loss_of(st):
            return 0.5 * ((handle.forward(st)[0][inner] - t) ** 2).sum().item()

        xx = state.clone().requires_grad_(True)
        (0.5 * ((A.rollout(xx, handle)[0][inner] - t) ** 2).sum()).backward()
        rels = []
        for v in dirs:
            v = torch.from_numpy(v).to(dev)
            fd = (loss_of(state + eps * v) - loss_of(state - eps * v)) / (2 * eps)
            ad = (xx.grad * v).sum().item()
            rels.append(abs(fd - ad) / abs(ad))
        return max(rels)

    rng = np.random.default_rng(7)
    dens = []
    for _ in range(args.adj_dirs):
        v = np.zeros_like(s0_np)
        v[0][inner] = rng.normal(size=(n, n))
        dens.append(v)
    rel_dens = fd_check(s0, target, h, dens)
    log(f"[bench] adjoint FD (density directions, {len(dens)}): max rel {rel_dens:.2e}")
    h.close()
    import sys as _sys
    _sys.path.insert(0, os.path.join(ROOT, "tests"))
    from adjoint_cases import rollout_state, stable_dt
    s_sm = rollout_state(21, n, n)
    h2 = A.Rollout(dict(desc), stable_dt(s_sm, n, n), args.adj_smooth_steps, every)
    dirs = [rng.normal(size=s_sm.shape) for _ in range(args.adj_smooth_dirs)]
    s_sm_t = torch.from_numpy(s_sm).to(dev)
    tg_sm = rollout_state(22, n, n)[0][inner]
    rel_sm = fd_check(s_sm_t, tg_sm, h2, dirs)
    rel_sm8 = fd_check(s_sm_t, tg_sm, h2, dirs, eps=1e-8)
    log(f"[bench] adjoint FD (smooth state, all fields, {len(dirs)}): max rel {rel_sm:.2e} "
        f"(eps 1e-6), {rel_sm8:.2e} (eps 1e-8)")
    h2.close()
    return {"cost_ratio": ratio, "reverse_to_plain": t_bwd / t_plain,
            "plain_forward_ms": 1e3 * t_plain,
            "forward_with_checkpoints_ms": 1e3 * t_fwd, "reverse_ms": 1e3 * t_bwd,
            "workload": f"{n}x{n} periodic, rho = 1 + 16 Fourier modes, P = 1, v = 0, "
                        f"B = (0.5, 0.2, 0), dt = {dt}, N = {N}, checkpoint every {every}",
            "fd_max_rel_density_dirs": rel_dens, "fd_density_dirs": len(dens),
            "fd_max_rel_smooth_all_fields": rel_sm, "fd_max_rel_smooth_all_fields_eps1e-8": rel_sm8,
            "fd_smooth_dirs": len(dirs),
            "fd_smooth_steps": args.adj_smooth_steps, "grad_norm": float(g0.norm().item()),
            "full_trajectory": full}


def full_trajectory(A, desc, dt, N, s0, gN, g_ref, t_plain):
    """The same rollout with one segment (checkpoint spacing N): the forward
    pass stores every step's forward scratch (88 planes, ~0.2 GB per step at
    512^2, ~50 GB in all) and the reverse re-runs nothing. Not the specified
    configuration (checkpoint every 16): reported beside it. None if the
    slots do not fit."""
    import torch
    try:
        hf = A.Rollout(desc, dt, N, N)
    except RuntimeError as e:
        log(f"[bench] adjoint, one segment: not run ({e})")
        return None
    hf.forward(s0)
    hf.backward(gN)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hf.forward(s0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = hf.backward(gN)
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    diff = ((g - g_ref).abs().max() / g_ref.abs().max()).item()
    hf.close()
    ratio = (t_fwd + t_bwd) / t_plain
    log(f"[bench] adjoint, one segment (every step stored): forward {1e3 * t_fwd:.1f} ms, "
        f"reverse {1e3 * t_bwd:.1f} ms, ratio {ratio:.2f}; gradient vs checkpointed: "
        f"max rel diff {diff:.1e}")
    return {"cost_ratio": ratio, "forward_ms": 1e3 * t_fwd, "reverse_ms": 1e3 * t_bwd,
            "checkpoint_every": N, "grad_max_rel_diff_vs_checkpointed": diff}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--variant", default="exact", choices=["exact", "fast"])
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--batch", type=int, default=256,
                    help="instances of the batched uneven ensemble (0: skip that leg)")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-warmup", type=int, default=3)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--shape", default=None,
                    help="x2_seg,persist,xcd_remap,corr_rows,x2_form,fusion[,strip[,final_rows]]: "
                         "use this "
                         "launch "
                         "shape for the plain step instead of autotuning (profiling runs)")
    ap.add_argument("--adjoint", type=int, default=1, help="0: skip the adjoint leg")
    ap.add_argument("--adj-size", type=int, default=512)
    ap.add_argument("--adj-steps", type=int, default=256)
    ap.add_argument("--adj-dirs", type=int, default=32)
    ap.add_argument("--adj-smooth-dirs", type=int, default=8)
    ap.add_argument("--adj-smooth-steps", type=int, default=64)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy kernel whose HBM "
                         "counters calibrate FETCH_SIZE/WRITE_SIZE under rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the library "
              "runs on one GPU (the multi-GPU halo decomposition is design-only).",
              file=sys.stderr)
        sys.exit(2)

    if args.adjoint:
        # torch ships its own HIP runtime under the same soname as the one
        # libathmi.so links; whichever loads first serves the process, and
        # torch needs its own
        import torch  # noqa: F401
        torch.cuda.init()
    from helpers import load_pkg
    pkg = load_pkg()
    L = pkg.lib()
    L.athmi_profile_steps.restype = ctypes.c_int
    L.athmi_profile_steps.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.athmi_step_kernel_name.restype = ctypes.c_char_p
    L.athmi_step_kernel_name.argtypes = [ctypes.c_int]

    n = args.size
    log(f"[bench] setup blast {n}x{n} ({args.variant})")
    g, par, tlim, _ = pkg.setup_problem("blast", [f"grid/Nx1={n}", f"grid/Nx2={n}"],
                                        variant=args.variant)
    g.set_limits(1e30, -1)  # fixed number of steps: the loop never reaches tlim
    shape = None
    if args.shape:
        v = [int(x) for x in args.shape.split(",")]
        g.set_shape(v[0], v[1])
        g.set_tile_order(v[2])
        g.set_corr_rows(v[3])
        g.set_x2_kind(v[4])
        g.set_fusion(v[5])
        g.set_strip(v[6] if len(v) > 6 else 0)
        g.set_final_rows(v[7] if len(v) > 7 else 8)
        shape = tuple(v[:4]) + (float("nan"), v[4], v[5], v[6] if len(v) > 6 else 0,
                                 v[7] if len(v) > 7 else 8)
        log(f"[bench] fixed shape {args.shape}")
    elif not args.no_tune:
        shape = g.autotune(3)
        shape = g.launch_shape() + (shape[3], g.x2_kind(), g.fusion(), g.strip(), g.final_rows())
        log(f"[bench] autotuned shape x2_seg={shape[0]} persist={shape[1]} xcd_remap={shape[2]} "
            f"corr_rows={shape[3]} x2_form={shape[5]} fusion={shape[6]} strip={shape[7]} "
            f"final_rows={shape[8]} "
            f"({shape[4]:.3f} ms/step)")
    for _ in range(args.warmup):
        g.step()
    g.sync()
    log(f"[bench] timing {args.steps} steps")
    n0 = g.clock(0)[2]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.step()
    g.sync()
    t1 = time.perf_counter()
    # every timed call advanced the instance (a stopped instance's step does no work)
    if g.clock(0)[2] - n0 != args.steps:
        raise RuntimeError(f"the timed region advanced {g.clock(0)[2] - n0} of {args.steps} steps; "
                           f"failure word {g.failure(0)}")
    sec = t1 - t0
    cells = n * n
    value = cells * args.steps / sec
    ms_step = 1e3 * sec / args.steps
    log(f"[bench] {ms_step:.3f} ms/step, {value / 1e9:.3f} G cell-updates/s")

    # constrained transport: max |div B| over the interior after the timed steps
    # (the exact variant reproduces the reference's roundoff; B is O(1) here)
    divb = g.divb_max(0)
    dx = min(1.0 / n, 1.5 / n)
    log(f"[bench] max|div B| {divb:.3e} (x dx {divb * dx:.3e}) after {args.warmup + args.steps} "
        "steps")

    # kernel-by-kernel pass with HIP events on the handle's stream
    avg = (ctypes.c_double * 16)()
    pkg._check(L.athmi_profile_steps(g._h, args.profile_steps, avg, 16), "profile")
    kern = {L.athmi_step_kernel_name(k).decode(): avg[k] for k in range(16)
            if L.athmi_step_kernel_name(k)}
    dom = max(kern, key=kern.get)
    # the metric's algorithmic bytes (168 B per cell-update: the state read
    # once and written once) over one launch of the dominant kernel: the whole
    # step's compulsory traffic, priced against that kernel's duration alone
    alg_bytes = ALG_BYTES_PER_CELL * n * n
    achieved = alg_bytes / (kern[dom] * 1e-3) / 1e9
    # the kernel's own plane I/O, scratch planes included (not algorithmic)
    rd, wr, units = KERNEL_IO[dom]
    io_bytes = (rd + wr) * 8 * units(n, n)
    pmc = pmc_summary() if n == 4096 else None
    traffic = None
    if pmc and dom in pmc["kernels"]:
        traffic = pmc["kernels"][dom]["traffic_bytes"]
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "kernel_ms": kern[dom], "alg_bytes_per_launch": alg_bytes,
                "alg_bytes_per_cell_update": ALG_BYTES_PER_CELL,
                "kernel_io_bytes_per_launch": io_bytes,
                "kernel_io_frac": io_bytes / (kern[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "traffic_source": None if not traffic else {
                    "file": "profiles/step_kernels_blast4096_pmc.md",
                    "commit": pmc.get("commit"), "workload": pmc.get("workload"),
                    "measured_in_this_run": False},
                "kernels_ms": kern}
    step_bw = ALG_BYTES_PER_CELL * value / 1e9
    out = {
        "metric": "FP64 cell-updates/s at 4096² plain vs batched; adjoint/forward step cost",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"MHD blast {n}x{n} (athinput.blast scaled), Roe + PLM, "
                               "CTU+CT, periodic, CFL 0.4; one step = integrate + new_dt + "
                               "set_bvals", "nx1": n, "nx2": n, "variant": args.variant,
                   "shape": None if shape is None else {"x2_seg": shape[0], "persist": shape[1],
                                                 "xcd_remap": shape[2], "corr_rows": shape[3],
                                                 "x2_form": shape[5], "fusion": shape[6],
                                                 "strip": shape[7], "final_rows": shape[8]}},
        "roofline": roofline,
        "divb": {"max_abs": divb, "max_abs_times_dx": divb * dx,
                 "steps": args.warmup + args.steps},
        "step_roofline": {"bound": "hbm", "alg_bytes_per_cell_update": ALG_BYTES_PER_CELL,
                          "achieved": step_bw, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_bw / HBM_PEAK_GBS},
        "fp64": None if not pmc else {
            "flop_per_cell_update": pmc["per_cell_update"]["fp64_flop_fma2"],
            "lane_instr_per_cell_update": pmc["per_cell_update"]["fp64_lane_instr"],
            "achieved": pmc["per_cell_update"]["fp64_flop_fma2"] * value / 1e12,
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": pmc["per_cell_update"]["fp64_flop_fma2"] * value / 1e12 / FP64_PEAK_TFLOPS,
            "hbm_bytes_per_cell_update": pmc["per_cell_update"]["hbm_bytes"],
            "source": "SQ_INSTS_VALU_*_F64 counts, profiles/step_kernels_blast4096_pmc.md"},
    }
    g.close()
    if args.batch > 0:
        out["batched"] = batched_leg(pkg, args, value)
    if args.adjoint:
        out["adjoint"] = adjoint_leg(pkg, args)
    if args.calibrate_copy:
        L.athmi_copy_calibration.argtypes = [ctypes.c_size_t, ctypes.c_int]
        pkg._check(L.athmi_copy_calibration(1 << 30, 3), "copy calibration")
        out["copy_calibration_bytes"] = {"read": 1 << 30, "write": 1 << 30, "reps": 3}
    if not args.no_cpu:
        log("[bench] CPU baseline (reference CPU integrator, 1 core): blast 4096^2 x 2 steps, "
            "400x600 x 100 steps")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
