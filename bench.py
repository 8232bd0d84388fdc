"""RK4 steps/s of the FFF Navier-Stokes hot path at 512^3 on one MI355X.

Workload (one "step" = one Lawson RK4 step of the pseudo-spectral
incompressible Navier-Stokes equations, all four nonlinear evaluations,
pressure projection, forcing and the strong reality condition):
  grid 512^3 (box 2 pi), FP64 state and transforms, reference corner
  dealiasing, nu = 1e-3, dt = 1e-3; seeded solenoidal initial field with
  E(k) ~ k^-5/3 on 2 <= k <= 170 (random phases, total energy 0.5);
  band-limited random forcing on 2 < k <= 3 injecting eps = 0.1 (phases
  redrawn every 0.01 time units).
Inputs are generated on the device before the timed region (synthetic data).

Besides the headline FP64 number the line reports:
  fp32       the same step with FP32 transforms/products (FP64 state), its
             steps/s and its relative energy error against FP64 after the
             same steps;
  diag       one evaluation of the energy flux over 20 log-spaced spheres
             and the 20x20 shell-to-shell transfer (plus its sum check);
  roofline   the dominant kernel's algorithmic HBM bytes per launch over its
             HIP-event-measured average duration, against 8 TB/s;
  step_roofline  the whole step's algorithmic bytes over the step time;
  cpu_baseline   the FP64 CPU oracle (restatement of the reference
             algorithm, own FFT, OpenMP) on one RK4 step of the same
             512^3 workload, on the host's cores;
  rbc        Rayleigh-Benard convection on the SFF basis (sin/cos along x):
             512 x 1024 x 1024 FP64, Ra = 1e8, Pr = 1 (nu = kappa = 1e-4),
             theta = white noise of amplitude 1e-3 in the first 32 modes per
             axis, u = 0, corner dealias, weak reality; RK4 steps/s, per-kernel
             times, step roofline (8 (4 + 8) S + 11 * 4 S64 bytes per step) and
             the Nusselt number;
  adjoint    the adjoint/forward cost ratio on its own workload: 256^3 FP64,
             200 unforced RK4 steps (no reality condition), checkpoints
             every 20 steps, loss = energy in the shell 8 <= |k| < 16 at the
             last step; ratio = (forward sweep with tape + backward sweep
             with recomputation) / (200 plain forward steps), measured time.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# band-limited random forcing at 2 < k <= 3 injecting eps = 0.1 (no helicity),
# phases redrawn every 0.01 time units (compute_force_random.cc)
HEADLINE_FORCE = dict(force_kind=2, force_eps=0.1, force_r_in=2.0, force_r_out=3.0, force_helicity=0.0,
                      force_tau=0.01, force_seed=7)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512, help="grid size per axis")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="headline FP64 number only")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost measurement")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint cost measurement")
    ap.add_argument("--no-rbc", action="store_true", help="skip the Rayleigh-Benard measurement")
    ap.add_argument("--rbc-only", action="store_true", help="run only the Rayleigh-Benard measurement")
    ap.add_argument("--no-mhd", action="store_true", help="skip the MHD precision measurement")
    ap.add_argument("--mhd-only", action="store_true", help="run only the MHD precision measurement")
    ap.add_argument("--copy-calib", action="store_true", help="run a plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--fp32-only", action="store_true",
                    help="run only the FP32-transform RK4 steps of the headline workload (for profiling)")
    return ap.parse_args()


def copy_bandwidth(torch, gib=4, reps=5):
    """Attainable HBM bandwidth: a device-to-device copy of `gib` GiB (read +
    write bytes over the event-timed average)."""
    a = torch.empty(gib << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / reps
    gbs = 2 * a.numel() * 8 / t / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def algorithmic_bytes(kernel, N, prec_bytes, nstate=3):
    """Algorithmic HBM bytes of one launch class per RK4 step (all launches)."""
    nc = N[0] * N[1] * (N[2] // 2 + 1)
    Sp = nc * 2 * prec_bytes      # one complex work field in the transform precision
    S64 = nc * 16                 # one complex FP64 state field
    # an RK stage transforms 3 stage inputs and 5 products (the trace-shifted
    # u_x^2 - u_z^2, u_y^2 - u_z^2, u_x u_y, u_y u_z, u_z u_x: the stage only uses the
    # nonlinear term projected, which removes the gradient the shift adds)
    if kernel == "y_fft":         # 4 stages x (3 inverse + 5 forward fields), read + write
        return 4 * 8 * 2 * Sp
    if kernel == "z_products":    # 4 stages x (read 3, write 5)
        return 4 * 8 * Sp
    if kernel == "x_nlin_rk":     # 4 stages x (read 5 products, write 3) + per component: V read 4x and
        # written once, the stage accumulator written 3x and read 3x (both FP64)
        return 4 * 8 * Sp + nstate * 11 * S64
    if kernel == "reality":       # two planes of V (read+write) and the patched columns
        return 2 * 2 * nstate * N[0] * N[1] * 16 + nstate * N[0] * N[1] * 2 * (16 + 2 * prec_bytes)
    if kernel == "x_inv":
        return nstate * (S64 + Sp)
    return 0


def step_bytes_model(N, prec_bytes, nstate=3, n_tr=8):
    """B_step = 8 (n_inv + n_fwd) S_p + 11 n_state S64 (two-sweep model: the FP64
    state read at every stage and written once, the FP64 stage accumulator written
    and read three times each)."""
    nc = N[0] * N[1] * (N[2] // 2 + 1)
    return 8 * n_tr * nc * 2 * prec_bytes + nstate * nc * 11 * 16


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU slab "
              "decomposition is a design, not part of this build.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from tarang_amd import api
    from tarang_amd._lib import KERNELS, check

    torch.cuda.init()
    dev = torch.device("cuda", 0)

    if args.copy_calib:
        a = torch.ones(1 << 27, dtype=torch.float64, device=dev)   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes": a.numel() * 8, "launches": 3}))
        return

    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_cost(api, torch, np)}))
        return
    if args.rbc_only:
        print(json.dumps({"rbc": rbc_rate(api, torch, np)}))
        return
    if args.mhd_only:
        print(json.dumps({"mhd": mhd_precision(api, torch, np)}))
        return
    if args.fp32_only:
        print(json.dumps({"fp32_only": fp32_steps(api, torch, args.n, args.steps, args.warmup)}))
        return

    hbm_copy = copy_bandwidth(torch)
    print(f"[bench] device copy: {hbm_copy:.0f} GB/s", file=sys.stderr, flush=True)
    N = (args.n, args.n, args.n)
    nu, dt = 1e-3, 1e-3
    force = HEADLINE_FORCE
    kmax_ic = 170.0 if args.n >= 512 else args.n / 3.0

    def make(prec):
        b = api.FFF_Basis(N, precision=prec, dealias=api.DEALIAS_REF_CORNER, strong_reality=True)
        U = api.FluidVF(b, nu)
        return b, U

    b64, U64 = make(api.PREC_FP64)
    U64.Init_cond_spectrum(kind=1, kmin=2.0, kmax=kmax_ic, energy=0.5, seed=2)
    V_init = [v.clone() for v in U64.cvf] if not args.no_extra else None
    ta = api.Time_advance_incompress(U64, dt, **force)

    for _ in range(args.warmup):
        ta.Time_advance_step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ta.Time_advance_step(1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / max(args.steps, 1)
    value = 1e3 / ms_step
    print(f"[bench] fp64 {args.steps} steps: {ms_step:.3f} ms/step", file=sys.stderr, flush=True)

    # per-kernel HIP-event timing on the plan's stream over 2 extra steps
    L = b64.L
    check("tg_profile_reset", L.tg_profile_reset(b64.plan))
    check("tg_profile_enable", L.tg_profile_enable(b64.plan, 1))
    prof_steps = 2
    ta.Time_advance_step(prof_steps)
    check("tg_profile_enable", L.tg_profile_enable(b64.plan, 0))
    import ctypes
    kern = {}
    for k, name in enumerate(KERNELS):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check("tg_profile_read", L.tg_profile_read(b64.plan, k, ctypes.byref(ms), ctypes.byref(n)))
        if n.value:
            kern[name] = (ms.value / prof_steps, n.value // prof_steps)  # per step: ms, launches
    dom = max(kern, key=lambda k: kern[k][0])
    dom_ms_step, dom_launch = kern[dom]
    dom_bytes_step = algorithmic_bytes(dom, N, 8)
    dom_achieved = dom_bytes_step / (dom_ms_step * 1e-3) / 1e9
    step_model = step_bytes_model(N, 8)
    roofline = {
        "kernel": dom, "bound": "hbm", "achieved": round(dom_achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(dom_achieved / HBM_PEAK_GBS, 4),
        "bytes_per_launch": int(dom_bytes_step / dom_launch), "avg_launch_ms": round(dom_ms_step / dom_launch, 4),
        "traffic": measured_traffic(dom, N, "fp64"),
        "attainable_copy_GBps": round(hbm_copy, 1),
        "frac_of_copy": round(dom_achieved / hbm_copy, 4),
    }
    per_kernel = {k: {"ms_per_step": round(v[0], 3), "launches_per_step": v[1],
                      "GBps": round(algorithmic_bytes(k, N, 8) / (v[0] * 1e-3) / 1e9, 1)
                      if algorithmic_bytes(k, N, 8) else None} for k, v in kern.items()}
    step_roofline = {"bytes_model": step_model, "achieved": round(step_model / (ms_step * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(step_model / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    extra = {}
    if not args.no_extra:
        # FP32 transforms, same initial field, same number of steps
        E64 = U64.Compute_total_energy()
        nsteps_total = args.warmup + args.steps + prof_steps
        del ta
        b32, U32 = make(api.PREC_FP32)
        U32.set_field(V_init)
        ta32 = api.Time_advance_incompress(U32, dt, **force)
        for _ in range(args.warmup):
            ta32.Time_advance_step(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ta32.Time_advance_step(1)
        torch.cuda.synchronize()
        ms32 = (time.perf_counter() - t0) * 1e3 / max(args.steps, 1)
        L32 = b32.L
        check("tg_profile_reset", L32.tg_profile_reset(b32.plan))
        check("tg_profile_enable", L32.tg_profile_enable(b32.plan, 1))
        ta32.Time_advance_step(prof_steps)
        check("tg_profile_enable", L32.tg_profile_enable(b32.plan, 0))
        kern32 = {}
        for k, name in enumerate(KERNELS):
            kms, cnt = ctypes.c_double(), ctypes.c_int64()
            check("tg_profile_read", L32.tg_profile_read(b32.plan, k, ctypes.byref(kms), ctypes.byref(cnt)))
            if cnt.value:
                kern32[name] = {"ms_per_step": round(kms.value / prof_steps, 3),
                                "launches_per_step": cnt.value // prof_steps,
                                "GBps": round(algorithmic_bytes(name, N, 4) / (kms.value / prof_steps * 1e-3) / 1e9, 1)
                                if algorithmic_bytes(name, N, 4) else None}
        E32 = U32.Compute_total_energy()
        # energy flux through the 20 default spheres, FP32 transforms vs FP64,
        # after the same steps from the same field
        radii = b64.default_radii(20)
        flux64 = api.EnergyTr(b64, radii, radii).Compute_flux(U64)
        flux32 = api.EnergyTr(b32, radii, radii).Compute_flux(U32)
        fin = np.isfinite(radii[1:])
        flux_err = float(np.max(np.abs(flux32[1:][fin] - flux64[1:][fin])) / np.max(np.abs(flux64[1:][fin])))
        print(f"[bench] fp32 {args.steps} steps: {ms32:.3f} ms/step", file=sys.stderr, flush=True)
        step_model32 = step_bytes_model(N, 4)
        extra["fp32"] = {"steps_per_s": round(1e3 / ms32, 3), "ms_per_step": round(ms32, 3),
                         "speedup_vs_fp64": round(ms_step / ms32, 3),
                         "energy_rel_err_vs_fp64": abs(E32 - E64) / abs(E64), "steps_compared": nsteps_total,
                         "flux_rel_err_vs_fp64": flux_err,
                         "step_roofline_frac": round(step_model32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "kernels": kern32}
        del ta32, U32, b32
        torch.cuda.empty_cache()
        # energy-transfer diagnostics on the FP64 state (20 spheres / 20 shells, the same radii): the
        # two reference calls, then both from one giver evaluation per shell
        radii = b64.default_radii(20)
        et = api.EnergyTr(b64, radii, radii)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        flux = et.Compute_flux(U64)
        S = et.Compute_shell_tr(U64)
        torch.cuda.synchronize()
        tsep = time.perf_counter() - t0
        t0 = time.perf_counter()
        flux_c, S_c = et.Compute_flux_and_shell_tr(U64)
        torch.cuda.synchronize()
        tdiag = time.perf_counter() - t0
        extra["diag"] = {"flux_and_shell_to_shell_ms": round(tdiag * 1e3, 1), "spheres": 20, "shells": 20,
                         "separate_calls_ms": round(tsep * 1e3, 1),
                         "combined_vs_separate_flux_rel": float(np.max(np.abs(flux_c - flux)) / np.max(np.abs(flux))),
                         "combined_vs_separate_S_rel": float(np.max(np.abs(S_c - S)) / np.max(np.abs(S))),
                         "shell_to_shell_sum": float(S.sum()), "shell_to_shell_abs_sum": float(np.abs(S).sum()),
                         "flux_max": float(np.max(flux)),
                         # one RK4 step followed by both diagnostics, as in a run that outputs them every step
                         "step_with_diagnostics_ms": round(ms_step + tdiag * 1e3, 1),
                         "steps_per_s_with_diagnostics": round(1e3 / (ms_step + tdiag * 1e3), 3)}
        print(f"[bench] diagnostics {tdiag * 1e3:.1f} ms (separate calls {tsep * 1e3:.1f} ms)", file=sys.stderr,
              flush=True)
        # the transfer identity at full size: 2/3 rule, the initial field (excited for K <= 170, inside the
        # largest inscribed sphere, so no receiver is cut): sum S = 0, S antisymmetric, to round-off
        del et
        torch.cuda.empty_cache()
        b23 = api.FFF_Basis(N, precision=api.PREC_FP64, dealias=api.DEALIAS_TWOTHIRDS, strong_reality=True)
        U23 = api.FluidVF(b23, nu)
        U23.set_field(V_init)
        f23, S23 = api.EnergyTr(b23, radii, radii).Compute_flux_and_shell_tr(U23)
        ns = len(radii) - 1
        Sm = S23[1:ns, 1:ns]
        extra["diag_twothirds"] = {
            "field": "initial k^-5/3 field, K <= 170", "dealias": "2/3 rule", "spheres": ns,
            "shell_to_shell_sum": float(S23.sum()), "shell_to_shell_abs_sum": float(np.abs(S23).sum()),
            "sum_rel": float(abs(S23.sum()) / np.abs(S23).sum()),
            "antisymmetry_rel": float(np.abs(Sm + Sm.T).max() / np.abs(S23).sum()),
            "flux_last_sphere": float(f23[-1]), "flux_max": float(np.abs(f23).max())}
        print(f"[bench] 2/3-rule transfer at {N}: |sum S|/sum|S| = {extra['diag_twothirds']['sum_rel']:.2e}",
              file=sys.stderr, flush=True)
        del U23, b23
        torch.cuda.empty_cache()
        extra["energy"] = E64
        if not args.no_adjoint:
            extra["adjoint"] = adjoint_cost(api, torch, np)
            torch.cuda.empty_cache()
        if not args.no_rbc:
            extra["rbc"] = rbc_rate(api, torch, np)
            torch.cuda.empty_cache()
        if not args.no_mhd:
            extra["mhd"] = mhd_precision(api, torch, np)
            torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        src = V_init if V_init is not None else U64.cvf
        cpu, V_cpu = cpu_baseline(N, nu, dt, force, src)
        # the same step on the GPU from the same field, against the oracle's step
        b1, U1 = make(api.PREC_FP64)
        U1.set_field(src)
        api.Time_advance_incompress(U1, dt, **force).Time_advance_step(1)
        num = sum(float(np.linalg.norm((U1.cvf[c].cpu().numpy() - V_cpu[c]).ravel()) ** 2) for c in range(3))
        den = sum(float(np.linalg.norm(V_cpu[c].ravel()) ** 2) for c in range(3))
        cpu["gpu_step_rel_l2_vs_oracle"] = (num / den) ** 0.5
        print(f"[bench] GPU step vs oracle step at {N}: rel L2 {cpu['gpu_step_rel_l2_vs_oracle']:.3e}",
              file=sys.stderr, flush=True)
        del U1, b1

    out = {
        "metric": "RK4 steps/s at 512^3 (FP64 vs FP32, % HBM roofline); adjoint/forward cost",
        "value": round(value, 4),
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded k^-5/3 solenoidal field generated on device)",
        "config": {"workload": f"FFF incompressible Navier-Stokes RK4 step, {N[0]}x{N[1]}x{N[2]}, FP64, "
                               "corner dealias, strong reality, band-limited random forcing eps=0.1 on 2<k<=3 "
                               "(random interval 0.01)",
                   "grid": list(N), "nu": nu, "dt": dt, "parallelism": "single GPU"},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "kernels": per_kernel,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


def fp32_steps(api, torch, n, steps, warmup, nu=1e-3, dt=1e-3):
    """The headline workload with FP32 transforms alone (same field, forcing
    and dealias), timed over `steps` RK4 steps: the run the FP32 profiles
    come from."""
    N = (n, n, n)
    b = api.FFF_Basis(N, precision=api.PREC_FP32, dealias=api.DEALIAS_REF_CORNER, strong_reality=True)
    U = api.FluidVF(b, nu)
    U.Init_cond_spectrum(kind=1, kmin=2.0, kmax=170.0 if n >= 512 else n / 3.0, energy=0.5, seed=2)
    ta = api.Time_advance_incompress(U, dt, **HEADLINE_FORCE)
    for _ in range(warmup):
        ta.Time_advance_step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ta.Time_advance_step(1)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / max(steps, 1)
    print(f"[bench] fp32 {steps} steps: {ms:.3f} ms/step", file=sys.stderr, flush=True)
    return {"grid": list(N), "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 4)}


def rbc_run(api, torch, np, prec, N, steps, warmup, Ra, Pr, dt, seed, profile):
    """warmup + steps (timed) + 1 (profiled) RK4 steps of the RBC workload in one precision mode."""
    import ctypes
    from tarang_amd._lib import KERNELS, check
    b = api.SFF_Basis(N, precision=prec, dealias=api.DEALIAS_REF_CORNER, reality=2)
    U, T = api.FluidVF(b), api.FluidSF(b)
    nx, ny, nzc = b.shape_complex_array
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    m = 32
    rows = torch.cat([torch.arange(0, m), torch.arange(ny - m + 1, ny)]).to("cuda")
    shape = (m - 1, rows.numel(), m)
    noise = torch.complex(torch.randn(shape, generator=gen, device="cuda", dtype=torch.float64),
                          torch.randn(shape, generator=gen, device="cuda", dtype=torch.float64))
    T.F[1:m, rows, :m] = noise
    b.Satisfy_strong_reality_condition_in_Array(T.F)
    # real-space RMS 1e-3: energy (1/2) <theta^2> = 5e-7 with the SFF weights
    T.F.mul_((5e-7 / T.Compute_total_energy()) ** 0.5)
    e0 = T.Compute_total_energy()
    ta = api.Time_advance_incompress_RBC(U, T, Ra, Pr, dt)
    ta.Time_advance_step(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ta.Time_advance_step(steps)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    L = b.L
    kern = {}
    if profile:
        check("tg_profile_reset", L.tg_profile_reset(b.plan))
        check("tg_profile_enable", L.tg_profile_enable(b.plan, 1))
    ta.Time_advance_step(1)
    if profile:
        check("tg_profile_enable", L.tg_profile_enable(b.plan, 0))
        for k, name in enumerate(KERNELS):
            kms, cnt = ctypes.c_double(), ctypes.c_int64()
            check("tg_profile_read", L.tg_profile_read(b.plan, k, ctypes.byref(kms), ctypes.byref(cnt)))
            if cnt.value:
                kern[name] = {"ms_per_step": round(kms.value, 3), "launches_per_step": cnt.value}
    res = {"ms": ms, "kernels": kern, "energy_theta_initial": e0, "nusselt": api.Correlation.Get_Nusselt_no(U, T, Ra, Pr),
           "energy_u": U.Compute_total_energy(), "energy_theta": T.Compute_total_energy(),
           "plan_GB": round(b.device_bytes() / 1e9, 1), "S": nx * ny * nzc * 16}
    del ta, U, T, b
    torch.cuda.empty_cache()
    return res


def rbc_rate(api, torch, np, N=(512, 1024, 1024), steps=3, warmup=1, Ra=1e8, Pr=1.0, dt=1e-3, seed=6):
    r = rbc_run(api, torch, np, api.PREC_FP64, N, steps, warmup, Ra, Pr, dt, seed, True)
    ms, S = r["ms"], r["S"]
    # a stage transforms 4 inputs and 8 products (the velocity products trace-shifted, zz not formed)
    model = 8 * (4 + 8) * S + 11 * 4 * S
    print(f"[bench] rbc {N}: {ms:.1f} ms/step, Nu {r['nusselt']:.12g}", file=sys.stderr, flush=True)
    # FP32 transforms from the same fields, the same steps: Nusselt number and energies against FP64
    r32 = rbc_run(api, torch, np, api.PREC_FP32, N, steps, warmup, Ra, Pr, dt, seed, False)
    rel = lambda a, b: abs(a - b) / abs(b)
    print(f"[bench] rbc FP32: {r32['ms']:.1f} ms/step, Nu {r32['nusselt']:.12g}", file=sys.stderr, flush=True)
    return {"workload": f"SFF Rayleigh-Benard {N[0]}x{N[1]}x{N[2]} FP64, Ra={Ra:g}, Pr={Pr:g}, corner dealias, "
                        "weak reality, u = 0, theta = white noise in the first 32 modes per axis scaled to a "
                        "real-space RMS of 1e-3",
            "grid_note": f"{N[0]} sin/cos points along x, not 513: the transforms take powers of two (the "
                         "reference's own SFF examples use even sizes, info/data/RBC_scaling/3d/in/para.yaml)",
            "energy_theta_initial": r["energy_theta_initial"],
            "steps_per_s": round(1e3 / ms, 4), "ms_per_step": round(ms, 2), "steps": steps,
            "step_roofline": {"bytes_model": model, "achieved": round(model / (ms * 1e-3) / 1e9, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(model / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "kernels": r["kernels"], "nusselt": r["nusselt"], "energy_u": r["energy_u"],
            "energy_theta": r["energy_theta"], "plan_GB": r["plan_GB"],
            "fp32": {"ms_per_step": round(r32["ms"], 2), "speedup_vs_fp64": round(ms / r32["ms"], 3),
                     "steps_compared": warmup + steps + 1,
                     "nusselt_rel_err_vs_fp64": rel(r32["nusselt"], r["nusselt"]),
                     "energy_u_rel_err_vs_fp64": rel(r32["energy_u"], r["energy_u"]),
                     "energy_theta_rel_err_vs_fp64": rel(r32["energy_theta"], r["energy_theta"]),
                     "plan_GB": r32["plan_GB"]}}


def mhd_run(api, torch, np, prec, N, steps, nu, eta, dt):
    b = api.FFF_Basis(N, precision=prec)
    U, B = api.FluidVF(b, nu), api.FluidVF(b, eta)
    U.Init_cond_spectrum(kind=1, kmin=2.0, kmax=170.0, energy=0.5, seed=11)
    B.Init_cond_spectrum(kind=1, kmin=2.0, kmax=170.0, energy=0.25, seed=12)
    ta = api.Time_advance_incompress_MHD(U, B, dt)
    ta.Time_advance_step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ta.Time_advance_step(steps - 1)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / (steps - 1)
    radii = b.default_radii(20)
    u2b, b2b = api.EnergyTr(b, radii, radii).Compute_flux_MHD(U, B)
    res = {"ms": ms, "E_u": U.Compute_total_energy(), "E_b": B.Compute_total_energy(),
           "Hc": api.Correlation.Get_cross_helicity(U, B),
           "Ek_u": sum(b.Compute_shell_spectrum(a) for a in U.cvf),
           "Ek_b": sum(b.Compute_shell_spectrum(a) for a in B.cvf),
           "u2b": u2b, "b2b": b2b, "fin": np.isfinite(radii[1:])}
    del ta, U, B, b
    torch.cuda.empty_cache()
    return res


def mhd_precision(api, torch, np, n=512, steps=100, nu=1e-3, eta=1e-3, dt=1e-3):
    """FFF MHD (Elsasser form) at n^3, nu = eta = 1e-3, seeded solenoidal u and b with E(k) ~ k^-5/3 on
    2 <= k <= 170 and E_b = 0.5 E_u: the same steps with FP64 and with FP32 transforms and products (FP64
    state, FP64 reductions); relative errors of the FP32 run against the FP64 run."""
    N = (n, n, n)
    r64 = mhd_run(api, torch, np, api.PREC_FP64, N, steps, nu, eta, dt)
    r32 = mhd_run(api, torch, np, api.PREC_FP32, N, steps, nu, eta, dt)
    rel = lambda a, b: abs(a - b) / abs(b)

    def rel_arr(a, b, mask=None):
        a, b = np.asarray(a), np.asarray(b)
        if mask is not None:
            a, b = a[1:][mask], b[1:][mask]
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    print(f"[bench] mhd {n}^3 x {steps}: FP64 {r64['ms']:.1f} ms/step, FP32 {r32['ms']:.1f} ms/step",
          file=sys.stderr, flush=True)
    return {"workload": f"FFF MHD {n}^3, nu = eta = {nu:g}, dt = {dt:g}, u and b seeded with E(k) ~ k^-5/3 on "
                        f"2 <= k <= 170, E_b = 0.5 E_u; {steps} steps",
            "fp64_ms_per_step": round(r64["ms"], 2), "fp32_ms_per_step": round(r32["ms"], 2),
            "fp32_speedup": round(r64["ms"] / r32["ms"], 3),
            "fp32_rel_err_vs_fp64": {
                "energy_u": rel(r32["E_u"], r64["E_u"]), "energy_b": rel(r32["E_b"], r64["E_b"]),
                "cross_helicity": rel(r32["Hc"], r64["Hc"]),
                "spectrum_u": rel_arr(r32["Ek_u"], r64["Ek_u"]), "spectrum_b": rel_arr(r32["Ek_b"], r64["Ek_b"]),
                "flux_u_to_b": rel_arr(r32["u2b"], r64["u2b"], r64["fin"]),
                "flux_b_to_b": rel_arr(r32["b2b"], r64["b2b"], r64["fin"])},
            "fp64": {"energy_u": r64["E_u"], "energy_b": r64["E_b"], "cross_helicity": r64["Hc"]}}


def measured_traffic(kernel_class, N, precision):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by the copy calibration), or None
    when no pass of this kernel, grid and precision is on record."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    for r in rec.values():
        if r.get("kernel_class") == kernel_class and r.get("grid") == list(N) and r.get("precision") == precision:
            return r["traffic_bytes_per_launch"]
    return None


def adjoint_cost(api, torch, np, n=256, nsteps=200, every=20, nu=1e-3, dt=1e-3, r_lo=8.0, r_hi=16.0,
                 hbm_sized=True, fd_dirs=16):
    N = (n, n, n)
    b = api.FFF_Basis(N, precision=api.PREC_FP64, dealias=api.DEALIAS_REF_CORNER, strong_reality=False)
    U = api.FluidVF(b, nu)
    U.Init_cond_spectrum(kind=1, kmin=2.0, kmax=n / 3.0, energy=0.5, seed=7)
    V0 = [v.clone() for v in U.cvf]
    ta = api.Time_advance_incompress(U, dt)
    ta.Time_advance_step(2, chain=False)
    U.set_field(V0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ta.Time_advance_step(nsteps, chain=False)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    kx = torch.fft.fftfreq(n, 1.0 / n, dtype=torch.float64, device="cuda")
    kz = torch.arange(n // 2 + 1, dtype=torch.float64, device="cuda")
    K = torch.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2)
    m = ((K >= r_lo) & (K < r_hi)).to(torch.float64)
    m[:, :, 0] *= 0.5
    adj = api.Adjoint_incompress(U, dt, checkpoint_every=every)
    # allocation of checkpoints and tape outside the timed region: a sweep of
    # the full length (the buffers grow with the number of segments)
    U.set_field(V0)
    adj.Forward(nsteps)
    U.set_field(V0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    adj.Forward(nsteps)
    lam = [2.0 * m * v for v in U.cvf]
    adj.Backward(lam)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    loss = float(sum((m * (v.real ** 2 + v.imag ** 2)).sum().item() for v in U.cvf))
    gnorm = float(sum((x.real ** 2 + x.imag ** 2).sum().item() for x in lam)) ** 0.5
    # per-kernel HIP-event times of one segment (forward with tape + backward), per step
    import ctypes
    from tarang_amd._lib import KERNELS, check
    L = b.L
    U.set_field(V0)
    check("tg_profile_reset", L.tg_profile_reset(b.plan))
    check("tg_profile_enable", L.tg_profile_enable(b.plan, 1))
    adj.Forward(every)
    lam = [2.0 * m * v for v in U.cvf]
    adj.Backward(lam)
    check("tg_profile_enable", L.tg_profile_enable(b.plan, 0))
    kern = {}
    for k, name in enumerate(KERNELS):
        ms, cnt = ctypes.c_double(), ctypes.c_int64()
        check("tg_profile_read", L.tg_profile_read(b.plan, k, ctypes.byref(ms), ctypes.byref(cnt)))
        if cnt.value:
            kern[name] = {"ms_per_step": round(ms.value / every, 3), "launches_per_step": cnt.value // every}
    nc = n * n * (n // 2 + 1)

    def tape_bytes(e):
        # checkpoints (FP64 spectral) + four stage tapes of three real fields per step of one segment
        return (nsteps + e - 1) // e * 3 * nc * 16 + e * 12 * n * n * (n // 2) * 16

    adj.Free()
    print(f"[bench] adjoint {n}^3 x {nsteps}: forward {t_fwd:.3f} s, forward+backward {t_adj:.3f} s",
          file=sys.stderr, flush=True)
    fd = None
    if fd_dirs:
        # the gradient of the timed sweep against central differences of the forward map along seeded
        # solenoidal directions (eps = 1e-6), at the full size and length
        U.set_field(V0)
        adj.Forward(nsteps)
        grad = [2.0 * m * v for v in U.cvf]
        adj.Backward(grad)
        adj.Free()
        Ud = api.FluidVF(b, nu)

        def loss_of(Vs):
            U.set_field(Vs)
            ta.Time_advance_step(nsteps, chain=False)
            return float(sum((m * (v.real ** 2 + v.imag ** 2)).sum().item() for v in U.cvf))

        eps, worst, t0 = 1e-6, 0.0, time.perf_counter()
        for seed in range(8, 8 + fd_dirs):
            Ud.Init_cond_spectrum(kind=1, kmin=1.0, kmax=n / 3.0, energy=0.5, seed=seed)
            d = Ud.cvf
            lp = loss_of([v + eps * x for v, x in zip(V0, d)])
            lm = loss_of([v - eps * x for v, x in zip(V0, d)])
            fdv = (lp - lm) / (2 * eps)
            adv = float(sum(torch.sum(torch.conj(g) * x).real.item() for g, x in zip(grad, d)))
            worst = max(worst, abs(fdv - adv) / max(abs(fdv), 1e-3 * loss))
        print(f"[bench] adjoint FD check: {fd_dirs} directions, worst rel mismatch {worst:.3e} "
              f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        fd = {"directions": fd_dirs, "eps": eps, "worst_rel_mismatch": worst, "target": 1e-6}
        del Ud
    out = {"workload": f"{n}^3 FP64, {nsteps} unforced RK4 steps, checkpoints every {every}, "
                       f"loss = E in {r_lo:g} <= |k| < {r_hi:g}",
           "forward_s": round(t_fwd, 4), "forward_and_backward_s": round(t_adj, 4),
           "cost_ratio": round(t_adj / t_fwd, 3), "target_ratio": 3.0, "loss": loss, "grad_norm": gnorm,
           "checkpoint_and_tape_GB": round(tape_bytes(every) / 1e9, 2), "segment_kernels": kern,
           "fd_check": fd}
    if hbm_sized:
        # the same sweep with the checkpoint interval sized to the device memory: the largest divisor of
        # nsteps whose checkpoints and tape fit in 60% of it (fewer recomputed steps, the same gradient)
        hbm = torch.cuda.get_device_properties(0).total_memory
        fits = [e for e in range(every, nsteps + 1) if nsteps % e == 0 and tape_bytes(e) <= 0.6 * hbm]
        e2 = max(fits) if fits else every
        if e2 != every:
            adj2 = api.Adjoint_incompress(U, dt, checkpoint_every=e2)
            U.set_field(V0)
            adj2.Forward(nsteps)  # allocation outside the timed region
            U.set_field(V0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            adj2.Forward(nsteps)
            lam2 = [2.0 * m * v for v in U.cvf]
            adj2.Backward(lam2)
            torch.cuda.synchronize()
            t_adj2 = time.perf_counter() - t0
            gnorm2 = float(sum((x.real ** 2 + x.imag ** 2).sum().item() for x in lam2)) ** 0.5
            adj2.Free()
            print(f"[bench] adjoint, checkpoints every {e2}: forward+backward {t_adj2:.3f} s", file=sys.stderr,
                  flush=True)
            out["hbm_sized_checkpoints"] = {
                "checkpoint_every": e2, "forward_and_backward_s": round(t_adj2, 4),
                "cost_ratio": round(t_adj2 / t_fwd, 3), "checkpoint_and_tape_GB": round(tape_bytes(e2) / 1e9, 2),
                "grad_norm_rel_diff_vs_every_%d" % every: abs(gnorm2 - gnorm) / gnorm}
    return out


def _cpu_model():
    """lscpu's model name (falls back to /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_baseline(N, nu, dt, force, V_dev):
    """One full Lawson RK4 step (four stages: force, nonlinear term, pressure,
    RHS, the stage algebra and the strong reality condition) of the same
    workload from the same field on the FP64 CPU oracle (own FFT, OpenMP over
    the host threads), timed."""
    import numpy as np
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import tg_oracle as to

    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", nproc))
    o = to.Oracle(N, dealias=to.DEALIAS_REF_CORNER, strong_reality=True)
    V = [np.ascontiguousarray(v.cpu().numpy()) for v in V_dev]
    fl = to.Fluid(nu, dt, force["force_kind"], force["force_eps"], force["force_r_in"], force["force_r_out"], 0.0, 0,
                  force["force_helicity"], force["force_tau"], force["force_seed"])
    print(f"[bench] cpu baseline: one RK4 step of {N} on {threads} threads", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    o.rk4_step(fl, V)
    t = time.perf_counter() - t0
    print(f"[bench] cpu baseline: {t:.2f} s/step", file=sys.stderr, flush=True)
    return {"value": round(1.0 / t, 5), "unit": "RK4 steps/s", "cores": threads, "nproc": nproc, "kind": "port",
            "sample": f"one full RK4 step (4 stages, 36 FP64 3D transforms) of the same {N[0]}^3 workload from "
                      f"the same field on the FP64 oracle (own FFT, OpenMP, {threads} threads): {t:.1f} s",
            "cpu_model": _cpu_model()}, V


if __name__ == "__main__":
    main()
