#!/usr/bin/env python3
"""Throughput of the Lare3d hot path on one MI355X.

Workload (one "step" = one iteration of the lare3d.f90:82-88 loop body:
lagrangian_step + eulerian_remap + eta_calc, with the energy_account that
output_routines makes every step): the resistive Harris-sheet
configuration at 256^3 in FP64 -- Bx = tanh(y/0.5), rho = 1 + sech^2(y/0.5),
pressure balance, eta = 1e-3 plus 1e-2 above |J| = 1, seeded 1% velocity
noise, shock viscosity on, periodic x/z and BC_USER walls in y.

Prints one JSON line: cell-updates/s of the FP64 path over the timed steps
(inputs resident in HBM), the roofline of the dominant kernel from HIP events
on the library stream, the same workload in the FP32 mode (FP32 bulk storage
and arithmetic; "fp32": throughput, per-kernel times and the relative L2
error of every field against the FP64 run after the same steps), and the CPU
restatement timed on a bounded sample of the same workload as a reported
baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]

Also the kink-unstable twisted flux tube at 512 x 512 x 1024 (resistive
with conduction) in the FP32 mode against FP64 ("kink"), and the conduction
super-step with the learned surrogate ("conduction").
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per processed point (FP64 = 8 B per field access) for
# the kernels that can dominate; a point is one index of the kernel's launch
# box.  Counting: each distinct field read once, each written field once.
KERNEL_BYTES_PER_POINT = {
    "mom_remap": 8 * (3 + 6 + 3),                   # v x3, v1, db1, rho_v, cv1v, cv2v, dmv -> v x3 (new buffer)
    "ct_sweep": 8 * (3 + 3),                        # t1, t2, a -> t1, t2, a (new buffers); + v_sweep, db1 below
    "geom": 8 * (3 + 3),                            # v1 x3 -> cv1, cv2, db1 (full sweeps)
    "vertex": 8 * (4 + 4),                          # rho1, cv1, cv2, dm -> rho_v, cv1v, cv2v, dmv
    "alpha": 8 * (3 + 2 + 3),                       # v x3, rho_v, cs_v -> alpha1..3
    "visc_dt": 8 * (3 + 2 + 3 + 1 + 3 + 2),         # v x3, rho, e, B x3, eta, alpha x3 -> visc_heat, p0
    "vertex_prep": 8 * (5 + 2),                     # rho, e, B x3 -> rho_v, cs_v
    "force": 8 * (3 + 1 + 3 + 1 + 3 + 3),           # v x3, pp, Bpred x3, rho_v, alpha x3 -> v1 x3
    "shock_corr": 8 * (6 + 3 + 1 + 3 + 3 + 3),      # v, v1, alpha, pp, e, rho, visc_dep (rw) -> v_new x3
    "rho_energy": 8 * (6 + 3),                      # e, rho1, v_sweep, db1, cv1, cv2 -> dm, rho_new, e_new
    "rkl2_stage": 8 * (9 + 1),                      # Y2,Y1,E0,rho,B x3,larsen,Lc -> Y3
}


def box_points(name, n):
    """Index points in one launch box of the named kernel (n^3 grid)."""
    sizes = {
        # mom_remap: the updated nodes 0..n (the copies of the rest of the
        # padded box, 4.7% more points at 48 B, are not counted)
        "mom_remap": (n + 1) ** 3, "ct_sweep": (n + 5) ** 3, "geom": (n + 4) ** 3,
        "vertex": (n + 3) * (n + 1) ** 2, "alpha": (n + 3) * (n + 4) * (n + 3),
        "visc_dt": (n + 2) ** 3, "vertex_prep": (n + 3) ** 3, "force": (n + 1) ** 3,
        "shock_corr": n ** 3, "rho_energy": (n + 1) * (n + 2) ** 2,
        "rkl2_stage": n ** 3,
    }
    return sizes.get(name)


def launch_bytes(name, n):
    """Algorithmic bytes of one launch of the named kernel (n^3 grid)."""
    b = KERNEL_BYTES_PER_POINT[name] * box_points(name, n)
    if name == "ct_sweep":
        # the sweep velocity and db1, read over the flux box
        b += 8 * 2 * (n + 3) ** 3
    return b


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "harris256_fp64_hbm_traffic.json")


def csrc_sha1():
    """Hash of the kernel sources, to match counter data to the code it measured."""
    import glob
    import hashlib
    h = hashlib.sha1()
    d = os.path.join(ROOT, "lare3d_amd", "csrc")
    for p in sorted(glob.glob(os.path.join(d, "*.hpp")) + glob.glob(os.path.join(d, "*.hip"))):
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the FETCH_SIZE / WRITE_SIZE counter
    runs recorded under profiles/ (scaled by a plain copy of known size), or
    None when they were taken on different kernel sources."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except OSError:
        return None
    if rec.get("csrc_sha1") != csrc_sha1() or kernel not in rec["kernels"]:
        return None
    k = rec["kernels"][kernel]
    return k["fetch_bytes"] + k["write_bytes"]


def step_bytes_per_cell(n_sub, s_stages, conduction):
    """Whole-step algorithmic bytes per cell-update (FP64): 113 field accesses
    for ideal MHD, 20 per resistive substep, 8s + 2 for conduction."""
    acc = 113 + 20 * n_sub + ((8 * s_stages + 2) if conduction else 0)
    return 8.0 * acc


def cpu_baseline(n_sample, steps):
    """The CPU restatement (single thread) on a bounded sample of the same
    workload.  Reported baseline only."""
    from lare3d_amd import configs
    from oracle.oracle import Oracle, build
    build()
    grid, params, fields = configs.harris(n_sample)
    orc = Oracle(grid, params, fields)
    orc.boundary_conditions()
    orc.eta_calc()
    orc.step(1)  # untimed first step (allocation warm-up)
    t0 = time.perf_counter()
    for s in range(2, 2 + steps):
        orc.step(s)
    dt = time.perf_counter() - t0
    orc.close()
    return n_sample ** 3 * steps / dt, dt


def copy_probe(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of known size (torch's copy kernel), so
    that a hardware-counter run can scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def timed_run(dev, warmup, steps):
    """warmup untimed steps, then `steps` timed ones; (CU/s, ms/step, last step).
    A step is the lare3d.f90:84-88 loop body with the energy_account of
    output_routines (diagnostics.F90:46,76): lare_dev_step_info with the
    energies, which also returns the step's scalars."""
    n = dev.n
    dev.boundary_conditions()
    dev.eta_calc()
    step = 0
    for _ in range(warmup):
        step += 1
        dev.step_info(step, with_energy=True)
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step += 1
        dev.step_info(step, with_energy=True)
    dev.synchronize()
    elapsed = time.perf_counter() - t0
    return n[0] * n[1] * n[2] * steps / elapsed, 1e3 * elapsed / steps, step


def fp32_run(n, warmup, steps, f64_end):
    """The same workload with FP32 bulk storage: throughput, and the relative
    L2 error of every field against the FP64 run after the same steps."""
    from lare3d_amd import configs
    from lare3d_amd.abi import PREC_FP32
    from lare3d_amd.device import Device
    grid, params, fields = configs.harris(n)
    params.precision = PREC_FP32
    dev = Device(grid, params, fields)
    cups, ms, step = timed_run(dev, warmup, steps)
    f32_end = dev.fields()
    _, divb = dev.energy_account(with_divb=True)
    # per-kernel durations of three more steps (HIP events on the library stream)
    dev.profile(True)
    for _ in range(3):
        step += 1
        dev.step(step)
    stats = dev.kernel_stats()
    dev.profile(False)
    dev.close()
    ranked = sorted(stats.items(), key=lambda kv: -kv[1][1])
    top = {k: {"launches": v[0], "ms": round(v[1], 3)} for k, v in ranked[:12]}
    errs = {}
    # scalar fields, and the velocity and magnetic field as vectors
    for name, comps in (("rho", ("rho",)), ("energy", ("energy",)), ("v", ("vx", "vy", "vz")),
                        ("B", ("bx", "by", "bz"))):
        a = np.concatenate([f32_end.interior(c).ravel() for c in comps])
        b = np.concatenate([f64_end.interior(c).ravel() for c in comps])
        errs[name] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {"value": cups, "unit": "cell-updates/s", "ms_per_step": ms,
            "storage": "FP32 bulk fields and FP32 bulk arithmetic; FP64 face B and its CT path (quotients, "
                       "fluxes, update), Van Leer limiter comparisons, metrics and reductions",
            "field_rel_l2_vs_fp64": errs, "steps_compared": warmup + steps, "max_divb_dx": divb,
            "kernels": top}


def _vector_errors(fa, fb):
    """Relative L2 of fa against fb over the interior: rho and energy per
    field, v and B as vectors."""
    errs = {}
    for name, comps in (("rho", ("rho",)), ("energy", ("energy",)), ("v", ("vx", "vy", "vz")),
                        ("B", ("bx", "by", "bz"))):
        num = sum(float(np.linalg.norm(fa.interior(c) - fb.interior(c))) ** 2 for c in comps)
        den = sum(float(np.linalg.norm(fb.interior(c))) ** 2 for c in comps)
        errs[name] = (num / den) ** 0.5
    return errs


def kink_leg(n, warmup, steps):
    """The kink-unstable twisted flux tube (n x n x 2n; resistive with
    Braginskii conduction) in the FP32 mode and in FP64: throughput, the
    energy budget and div B of each run, and the field errors of the FP32
    run against the FP64 run after the same steps."""
    from lare3d_amd import configs
    from lare3d_amd.abi import PREC_FP32
    from lare3d_amd.device import Device
    out = {"workload": f"kink-unstable twisted flux tube {n}x{n}x{2 * n}, B_theta = 2r/(1+r^2), B_z = 1, rho = 1, "
                       "beta = 0.05 outside the tube, eta = 1e-3, Braginskii conduction (kappa_0 = 0.1)",
           "steps_timed": steps, "steps_total": warmup + steps}
    ends = {}
    for prec in ("fp64", "fp32"):
        grid, params, f = configs.kink_tube(n)
        params.precision = PREC_FP32 if prec == "fp32" else 0
        t0 = time.perf_counter()
        dev = Device(grid, params, f)
        del f
        t_up = time.perf_counter() - t0
        dev.boundary_conditions()
        dev.eta_calc()
        en0, divb0 = dev.energy_account(with_divb=True)
        heat0 = 0.0
        step = 0
        for _ in range(warmup):
            step += 1
            dev.step_info(step, with_energy=True)
        dev.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step += 1
            info = dev.step_info(step, with_energy=True)
        dev.synchronize()
        el = time.perf_counter() - t0
        ncell = grid.n[0] * grid.n[1] * grid.n[2]
        e1 = (info.en_b, info.en_ke, info.en_int)
        out[prec] = {"value": ncell * steps / el, "unit": "cell-updates/s", "ms_per_step": 1e3 * el / steps,
                     "rkl2_stages": info.n_s_stages, "resistive_substeps": info.substeps, "upload_s": round(t_up, 1),
                     "energy_start": {"magnetic": en0[0], "kinetic": en0[1], "internal": en0[2]},
                     "energy_end": {"magnetic": e1[0], "kinetic": e1[1], "internal": e1[2]},
                     "total_energy_drift": (sum(e1) - sum(en0)) / sum(en0),
                     "visc_heating": info.total_visc_heating - heat0, "ohmic_heating": info.total_ohmic_heating,
                     "max_divb_dx": info.max_divb, "max_divb_dx_start": divb0}
        ends[prec] = dev.fields()
        dev.close()
    out["fp32"]["field_rel_l2_vs_fp64"] = _vector_errors(ends["fp32"], ends["fp64"])
    out["fp32"]["speedup_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    return out


def conduction_leg(n, seed=1, kappa_0=0.5, s_target=7, rows_per_state=17000, train_steps=6000):
    """Braginskii conduction super-steps on the Harris sheet with seeded hot
    spots (n^3): the plain RKL2 solver against the learned surrogate, in FP64
    and the FP32 mode.  The surrogate is trained here, on the device, from
    three seeded states (rows and RKL2 targets collected by the library),
    and evaluated on a fourth.  Reports super-step times, the temperature
    error of the surrogate against RKL2, and its residual-check outcome."""
    from lare3d_amd import configs
    from lare3d_amd.abi import PREC_FP32
    from lare3d_amd.device import Device
    from lare3d_amd.surrogate import Surrogate, initial_params

    def state(sd):
        return configs.harris_conduction(n, seed=sd, s_target=s_target, kappa_0=kappa_0)

    grid, params, f0, dt0 = state(seed)
    held = state(seed + 99)
    out = {"workload": f"resistive Harris sheet {n}^3 with Braginskii conduction (kappa_0 = {kappa_0}, "
                       f"12 seeded hot spots), one conduction super-step of about {s_target} RKL2 stages"}
    res = {}
    for prec in ("fp64", "fp32"):
        params.precision = PREC_FP32 if prec == "fp32" else 0
        dev = Device(grid, params, f0)
        sur = Surrogate(dev)
        if prec == "fp64":
            sur.set_params(initial_params(0))
            t0 = time.perf_counter()
            for q in range(3):
                # training states span about 5, 7 and 9 stages so that the
                # held-out step (about 7) lies inside the sampled dt range
                g_, p_, f_, dt_ = configs.harris_conduction(n, seed=seed + 1 + q, s_target=s_target - 2 + 2 * q,
                                                            kappa_0=kappa_0)
                dev.upload(f_)
                dev.boundary_conditions()
                dev.set_scalars(dt=dt_)
                sur.collect(rows_per_state, 100 + q)
            sur.fit_norm()
            t_collect = time.perf_counter() - t0
            t0 = time.perf_counter()
            # Adam with a stepped learning rate; the moments carry over
            loss = np.concatenate([sur.train(train_steps // 3, batch=512, lr=lr, seed=7 + q)
                                   for q, lr in enumerate((1e-3, 3e-4, 1e-4))])
            t_train = time.perf_counter() - t0
            trained = (sur.params(), sur.norm())
            out["training"] = {"rows": 3 * rows_per_state, "steps": int(loss.size), "batch": 512,
                               "learning_rate": "1e-3, 3e-4, 1e-4 (thirds)",
                               "collect_s": round(t_collect, 2), "train_s": round(t_train, 2),
                               "loss_first50": float(loss[:50].mean()), "loss_last50": float(loss[-50:].mean())}
        else:
            sur.set_params(trained[0])
            sur.set_norm(*trained[1])
        _, _, fh, dth = held

        def load():
            dev.upload(fh)
            dev.boundary_conditions()
            dev.set_scalars(dt=dth)

        def plain():
            load()
            dev.synchronize()
            t0 = time.perf_counter()
            dev.conduct_heat()
            dev.synchronize()
            return time.perf_counter() - t0

        def surrogate(tol):
            load()
            sur.enable(False, tol)
            dev.synchronize()
            t0 = time.perf_counter()
            r, acc = sur.conduct()
            dev.synchronize()
            return time.perf_counter() - t0, r, acc

        tp = min(plain() for _ in range(3)) * 1e3
        stages = dev.scalars().n_s_stages
        e_rkl2 = dev.fields().interior("energy")
        ts = min(surrogate(1e30)[0] for _ in range(3)) * 1e3
        _, r_forced, _ = surrogate(1e30)
        e_sur = dev.fields().interior("energy")
        _, r_tol, acc_tol = surrogate(1e-4)
        e0 = fh.interior("energy")
        load()
        r_plain = sur.plain_residual()
        # per-kernel times of one plain and one surrogate super-step
        dev.profile(True)
        plain()
        k_plain = dev.kernel_stats()
        dev.profile(False)
        dev.profile(True)
        surrogate(1e30)
        k_sur = dev.kernel_stats()
        dev.profile(False)
        res[prec] = {"rkl2_ms": round(tp, 3), "rkl2_stages": stages, "surrogate_ms": round(ts, 3),
                     "surrogate_speedup": round(tp / ts, 2),
                     "T_rel_l2_vs_rkl2": float(np.linalg.norm(e_sur - e_rkl2) / np.linalg.norm(e_rkl2)),
                     "dT_rel_l2_vs_rkl2": float(np.linalg.norm((e_sur - e0) - (e_rkl2 - e0)) /
                                                np.linalg.norm(e_rkl2 - e0)),
                     "residual": r_forced, "residual_of_plain_rkl2": r_plain, "accepted_at_1e-4": bool(acc_tol),
                     "cell_updates_per_s_rkl2": n ** 3 / (tp * 1e-3),
                     "cell_updates_per_s_surrogate": n ** 3 / (ts * 1e-3),
                     "kernels_rkl2_ms": {k: round(v[1], 3) for k, v in k_plain.items()},
                     "kernels_surrogate_ms": {k: round(v[1], 3) for k, v in k_sur.items()}}
        dev.close()
    out.update(res)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--cpu-n", type=int, default=96)
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-mode run")
    ap.add_argument("--no-conduction", action="store_true",
                    help="skip the conduction leg (RKL2 vs the learned surrogate)")
    ap.add_argument("--no-kink", action="store_true", help="skip the twisted flux tube leg")
    ap.add_argument("--cond-stages", type=int, default=7,
                    help="RKL2 stage count the conduction leg's super-step is sized for")
    ap.add_argument("--kink-n", type=int, default=512, help="flux tube size n (grid n x n x 2n)")
    ap.add_argument("--kink-steps", type=int, default=8)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy and exit (counter calibration for HBM traffic; "
                         "run it as its own process under the same counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(multi-GPU decomposition is a design, not a built path)"}))
        sys.exit(2)

    if args.copy_probe:
        copy_probe()
        return
    from lare3d_amd import configs
    from lare3d_amd.device import Device

    n = args.n
    grid, params, fields = configs.harris(n)
    dev = Device(grid, params, fields)
    cups, ms_per_step, step = timed_run(dev, args.warmup, args.steps)
    sc = dev.scalars()
    en, divb = dev.energy_account(with_divb=True)
    f64_end = dev.fields() if not args.no_fp32 else None

    # per-kernel durations from HIP events on the library stream
    dev.profile(True)
    for _ in range(args.profile_steps):
        step += 1
        dev.step(step)
    stats = dev.kernel_stats()
    dev.profile(False)
    total_ms = sum(v[1] for v in stats.values())
    ranked = sorted(stats.items(), key=lambda kv: -kv[1][1])
    dom = next((k for k, _ in ranked if k in KERNEL_BYTES_PER_POINT), ranked[0][0])
    cnt, ms = stats[dom]
    avg_s = ms / cnt / 1e3
    bytes_launch = launch_bytes(dom, n)
    achieved = bytes_launch / avg_s / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                "bytes_per_launch": bytes_launch, "avg_launch_us": round(avg_s * 1e6, 2),
                "share_of_step": round(ms / total_ms, 3)}
    bpc = step_bytes_per_cell(max(sc.last_substeps, 1), sc.n_s_stages, bool(params.conduction))
    step_roof = {"bytes_per_cell": bpc, "achieved": round(cups * bpc / 1e9, 1), "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": round(cups * bpc / 1e9 / HBM_PEAK_GBS, 4)}
    top = {k: {"launches": v[0], "ms": round(v[1], 3)} for k, v in ranked[:12]}
    dev.close()
    fp32 = None
    if not args.no_fp32:
        fp32 = fp32_run(n, args.warmup, args.steps, f64_end)
        fp32["speedup_vs_fp64"] = fp32["value"] / cups

    cond = None if args.no_conduction else conduction_leg(n, s_target=args.cond_stages)
    kink = None if args.no_kink else kink_leg(args.kink_n, 2, args.kink_steps)

    cpu = None
    if not args.no_cpu:
        rate, secs = cpu_baseline(args.cpu_n, args.cpu_steps)
        cpu = {"value": rate, "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"this repository's C restatement of the reference (oracle/, one thread; the reference's "
                         f"own MPI build needs the absent SDF library), Harris sheet {args.cpu_n}^3 FP64, "
                         f"{args.cpu_steps} steps after one untimed step, {secs:.1f} s"}

    out = {
        "metric": "cell-updates/s at 256³ (FP64 plain vs FP32+surrogate), % of HBM peak",
        "value": cups,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Harris-sheet initial condition)",
        "config": {"workload": f"resistive Harris-sheet reconnection {n}^3 FP64, shock viscosity on, "
                               "eta = 1e-3 + 1e-2 above |J| = 1 (lagrangian_step + eulerian_remap + eta_calc "
                               "+ energy_account per step)",
                   "n": [n, n, n], "resistive_substeps": sc.last_substeps, "rkl2_stages": sc.n_s_stages,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "step_roofline": step_roof,
        "cpu_baseline": cpu,
        "fp32": fp32,
        "conduction": cond,
        "kink": kink,
        "kernels": top,
        "diagnostics": {"energy": en, "max_divb_dx": divb, "dt": sc.dt},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
