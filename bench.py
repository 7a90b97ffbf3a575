"""Benchmark: H_AC applies/s on the D = 1024 Hubbard U(1)xU(1) workload, FP64 and mixed precision
on the same GPU, with the VUMPS iteration, the transfer-matrix microbenchmark and the 32-momentum
quasiparticle-excitation run beside it.

Workload: Fermi-Hubbard chain (t=1, U=4, half filling), (N, M) representation, two-site unit cell,
d = 4, D = 1024 over 121 charge sectors of sizes 1..60.  One step = one grouped H_AC application
(VUMPS_helpers.hpp:38-88) to a block of `nvec` Krylov vectors resident in HBM.  Operator data are
seeded synthetic (isometry-scale normal entries, reduced operators computed on the GPU by
ReducedOp).

Prints ONE JSON line on stdout.  `value` is vector-applies/s of the FP64 path (the reference's
arithmetic); the mixed-precision rates (FP32 and split-BF16 MFMA products, FP32 accumulation) and
their ratio to FP64 are extra keys.  `roofline` is the FP64 panel GEMM kernel's (its own launch time
from HIP events on the launch stream), `roofline_mixed` the faster mixed one's.  `cpu_baseline` is
the same H_AC workload on the host with the FP64 CPU restatement, all cores of the affinity mask and
one core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 157.3}  # bf16x3 is priced against FP32 MFMA
HBM_PEAK_GBS = 8000.0


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "panel_gemm_kernel_hac_hubbard_d1024.md")


def committed_traffic(kernel: str, nvec: int):
    """HBM bytes per launch of a panel GEMM kernel variant from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE summary (copy-calibrated), for this batch; None if not recorded."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    key = f"traffic {kernel} nvec {nvec}:"
    with open(TRAFFIC_FILE) as f:
        for line in f:
            if line.startswith(key):
                return float(line.split(":")[1].split()[0])
    return None


def lieb_wu(U):
    """Ground-state energy per site of the half-filled Hubbard chain (Lieb-Wu) for H = -t sum c^+c +
    U sum n_up n_dn, t = 1: -4 int_0^inf J0(w) J1(w) / (w (1 + exp(w U / 2))) dw."""
    from scipy import integrate, special
    f = lambda w: special.j0(w) * special.j1(w) / (w * (1.0 + np.exp(min(w * U / 2.0, 700.0))))  # noqa: E731
    edges = np.linspace(0.0, 400.0, 4001)
    return -4.0 * sum(integrate.quad(f, a, b)[0] for a, b in zip(edges[:-1], edges[1:]))


def settle():
    """Collect what earlier runs left: left to the cyclic garbage collector, their plans' device
    frees (each a device synchronisation) would land inside a later timed region."""
    import gc
    import torch
    gc.collect()
    torch.cuda.synchronize()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nvec", type=int, default=16, help="Krylov vectors per batched apply")
    ap.add_argument("--cpu-seconds", type=float, default=16.0, help="bound on the CPU baseline samples (half on all cores, half on one)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-vumps", action="store_true")
    ap.add_argument("--no-xxz", action="store_true", help="skip the D = 64 Heisenberg VUMPS runs (keep the D = 1024 Hubbard one)")
    ap.add_argument("--no-exc", action="store_true")
    ap.add_argument("--no-tm", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--exc-maxit", type=int, default=1600, help="bound on the batched H_eff applications of the ground-state eigensolve")
    ap.add_argument("--exc-krylov", type=int, default=160, help="Lanczos basis size of the ground-state excitation eigensolve")
    ap.add_argument("--exc-mixed-eigs", action="store_true",
                    help="also run the ground-state eigensolve with FP32-refined TM chains (otherwise their applies only)")
    ap.add_argument("--gs-iters", type=int, default=1200, help="bound on the VUMPS iterations of the D = 512 ground state")
    ap.add_argument("--d1024-iters", type=int, default=1000, help="bound on the D = 1024 VUMPS iterations to F <= 1e-10")
    ap.add_argument("--exc-synthetic-only", action="store_true", help="excitations on the synthetic 81-sector state only (no ground state)")
    ap.add_argument("--precs", default="fp64,fp32,bf16x3")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU sharding is described, not built).", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_mib:
        n = args.copy_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.copy_mib} MiB x 5")
        return
    import imps_cpp_tools_amd as im
    from imps_cpp_tools_amd import workloads as wl

    cfg = wl.HUBBARD_1024
    L = wl.hac_workload_layouts(cfg)
    rng = np.random.default_rng(1003)
    H = im.model_ham(cfg["model"], cfg["params"])
    # one default-priority stream shared by torch (uploads, events) and the library's launches
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx = im.Context(0, stream)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()

    # operands: AL, AR at isometry scale, HL/HR by the GPU ReducedOp, EHL/EHR symmetric
    AL = dev(wl.random_data(L["AL"], rng, scale_by_rows=True))
    AR = dev(wl.random_data(L["AR"], rng, scale_by_rows=True))
    rl = im.RedOpPlan(ctx, H, 4, L["AL"], im.LEFT)
    rr = im.RedOpPlan(ctx, H, 4, L["AR"], im.RIGHT)
    HLl, HRl = rl.out_layout, rr.out_layout
    HL = torch.empty(HLl.nelem, dtype=torch.float64, device="cuda")
    HR = torch.empty(HRl.nelem, dtype=torch.float64, device="cuda")
    rl.apply(AL, HL)
    rr.apply(AR, HR)
    EHL = dev(wl.random_data(L["EHL"], rng, symmetric=True))
    EHR = dev(wl.random_data(L["EHR"], rng, symmetric=True))
    n = L["AC"].nelem
    NV = args.nvec
    X = dev(rng.standard_normal(n * NV))
    torch.cuda.synchronize()

    results = {}
    precs = {"fp64": im.FP64, "fp32": im.FP32, "bf16x3": im.BF16X3}
    for name in args.precs.split(","):
        prec = precs[name]
        p = im.HACPlan(ctx, L["AC"], L["EHL"], L["EHR"], HLl, HRl, prec)
        p.set_operator(EHL, EHR, HL, HR)
        flops, bytes_ = p.cost()
        Y = torch.empty(n * NV, dtype=torch.float64, device="cuda")
        r = {}
        for label, nv in [("batched", NV), ("single", 1)]:
            for _ in range(args.warmup):
                p.apply(X, Y, nv)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            e0.record(stream)
            for _ in range(args.steps):
                p.apply(X, Y, nv)
            e1.record(stream)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            ms = e0.elapsed_time(e1) / args.steps
            # the GEMM kernels alone: HIP events recorded by the library on this stream around the
            # panel-GEMM launches of each apply (the split copy of the input stays outside)
            p.timing(True)
            for _ in range(args.steps):
                p.apply(X, Y, nv)
            kms, kcount = p.timing(False)
            r[label] = dict(nvec=nv, ms_per_launch=ms, applies_per_s=nv / (ms * 1e-3), wall_s=wall,
                            tflops=nv * flops / (ms * 1e-3) / 1e12, kernel_us=kms / kcount * 1e3,
                            kernel_tflops=nv * flops / (kms / kcount * 1e-3) / 1e12)
        r["flops_per_apply"] = flops
        r["bytes_per_apply"] = bytes_
        r["schedule"] = {"batched": p.schedule(NV), "single": p.schedule(1)}  # (K-split, shared panels), autotuned
        r["y"] = Y[:n].cpu().numpy()
        results[name] = r
        log(f"{name}: batched {r['batched']['applies_per_s']:.4g} applies/s ({r['batched']['tflops']:.3g} TF/s), "
            f"single {r['single']['applies_per_s']:.4g} applies/s, {r['batched']['ms_per_launch'] * 1e3:.1f} us/apply, "
            f"GEMM kernels {r['batched']['kernel_us']:.1f} us/apply")

    # headline: the FP64 path (the reference's arithmetic); the mixed-precision paths (FP32 or
    # split-BF16 MFMA products with FP32 accumulation, refined in FP64 by the solvers) beside it
    rate = lambda k, lab="batched": results[k][lab]["applies_per_s"] if k in results else None  # noqa: E731
    mixed_names = [k for k in ("fp32", "bf16x3") if k in results]
    mixed = max(mixed_names, key=lambda k: results[k]["batched"]["applies_per_s"]) if mixed_names else None
    head = "fp64" if "fp64" in results else (mixed or next(iter(results)))
    hres0 = results[head]
    pidx = {"fp64": 0, "fp32": 1, "bf16x3": 2}

    def roofline(name):
        r = results[name]
        ks, sh = r["schedule"]["batched"]
        # rocprof's template names; batched applies read the split copy (direct input reads: false)
        kname = f"panel_gemm_shared_kernel<{pidx[name]}, false>" if sh else f"panel_gemm_kernel<{pidx[name]}, {ks}, false>"
        ach = r["batched"]["kernel_tflops"]
        return {"bound": "mfma", "achieved": ach, "peak": PEAK_TFLOPS[name], "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS[name],
                "traffic": committed_traffic(kname, NV), "kernel": kname,
                "per_launch": {"flops": NV * r["flops_per_apply"], "us": r["batched"]["kernel_us"],
                               "apply_us_incl_split_copy": r["batched"]["ms_per_launch"] * 1e3,
                               "algorithmic_bytes": algorithmic_bytes(name, r["bytes_per_apply"], NV)}}

    def algorithmic_bytes(name, single_bytes, nv):
        """Bytes of one launch of nv vectors with each operand read once: the operator store once,
        input and output vectors once per vector (imps_plan_cost counts one vector's apply)."""
        eb = 8 if name == "fp64" else 4
        vec = 2 * n * eb
        return (single_bytes - vec) + nv * vec

    flops = hres0["flops_per_apply"]
    out = {
        "metric": "H_AC applies/s & VUMPS iter time, D=1024 Hubbard: mixed vs FP64, same GPU",
        "value": hres0["batched"]["applies_per_s"],
        "unit": "H_AC vector-applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hres0["batched"]["ms_per_launch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3 (split-BF16 MFMA, FP32 accumulate; FP64 in/out)",
                  "fp32": "f32 (FP32 MFMA, FP32 accumulate; FP64 in/out)", "fp64": "f64"}[head],
        "data": "synthetic (seeded)",
        "config": {"workload": "H_AC apply, Fermi-Hubbard U(1)xU(1) (N,M), d=4, D=1024, 121 sectors of size 1..60",
                   "nvec_per_step": NV, "flops_per_apply": flops, "ac_elements": n},
        "mixed_path": mixed,
        "mixed_vs_fp64": (rate(mixed) / rate("fp64")) if mixed and "fp64" in results else None,
        "fp64_applies_per_s": rate("fp64"),
        "fp32_applies_per_s": rate("fp32"),
        "bf16x3_applies_per_s": rate("bf16x3"),
        "single_vector_applies_per_s": {k: results[k]["single"]["applies_per_s"] for k in results},
        "panel_schedules": {k: results[k]["schedule"] for k in results},
        "vumps_iter_s": None,
        "roofline": roofline(head),
    }
    if mixed:
        out["roofline_mixed"] = roofline(mixed)

    # VUMPS sweep time on the two-site Heisenberg chain, D = 64 (9 U(1) sectors, seed 1001).  The VUMPS
    # runs use the fixed panel-schedule rule: autotuned schedules (chosen by timing) change the
    # rounding, which changes the trajectory and the sweep count between runs.
    if not args.no_vumps:
        # VUMPS on torch's default stream: its iterations measured ~20% faster there than with the
        # main context on a separate stream (the side streams of the concurrent solves are separate
        # either way)
        settle()
        torch.cuda.set_stream(torch.cuda.default_stream())
        vctx = im.Context(0)
        prev_tune = os.environ.get("IMPS_AUTOTUNE")
        os.environ["IMPS_AUTOTUNE"] = "0"
        from imps_cpp_tools_amd.vumps import VUMPS
        keys0 = np.array([[2 * (i - 4)] for i in range(9)], np.int32)
        sizes0 = np.array([1, 3, 7, 12, 18, 12, 7, 3, 1], np.int32)
        vres = {}
        for name in [k for k in ("fp64", "fp32", "bf16x3") if k in results and not args.no_xxz]:
            v = VUMPS(vctx, im.XXZ, 2, [0], keys0, sizes0, [1.0, -1.0, 0.0], prec=precs[name], seed=1001)
            t0 = time.perf_counter()
            sweeps, st = 0, None
            while sweeps < 600:
                st = v.sweep()
                sweeps += 1
                if max(st.F) <= 1e-10:
                    break
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            vres[name] = dict(sweeps=sweeps, s_per_sweep=dt / sweeps, seconds_to_F=dt, F=max(st.F), energy=st.energy)
            v.close()
            del v
            settle()
            log(f"vumps xxz D=64 {name}: {sweeps} sweeps, {dt / sweeps * 1e3:.1f} ms/sweep, F {max(st.F):.2e}, E {st.energy:.15f}")
        out["vumps_xxz_d64"] = vres
        # VUMPS on the D = 1024 Hubbard chain, as the reference reaches it: from a D = 64 bond (25
        # sectors, seeded randMPS_LR state) the bonds grow by ExpandFromH (at most 256 states per bond
        # and iteration) to D = 1024, in FP64; that state goes through a state file.  From it, FP64
        # and FP32 (operators in FP32 under FP64 refinement) iterate at fixed structure to max F <=
        # 1e-10 (bounded by --d1024-iters); the iteration time is the median over the run (plans
        # are built in the first).  BF16X3 runs five iterations from the same state for its timing.
        import tempfile
        k64, s64 = wl.fhub_profile(5, 8, 1, 64)

        def hub(prec, **kw):
            return VUMPS(vctx, cfg["model"], cfg["N"], cfg["QN"], k64, s64, cfg["params"], prec=prec, seed=1003, krylov=30,
                         env_krylov=300, **kw)

        hres = {}
        names = [k for k in ("fp64", "fp32", "bf16x3") if k in results]
        if names:
            t0 = time.perf_counter()
            v = hub(im.FP64, mmax=1024, dm_max=256, expand_pause=1, expand_thresh=1.0)
            grow = 0
            while min(v.bond_dims()) < 1024 and grow < 20:
                v.sweep()
                grow += 1
            t_grow = time.perf_counter() - t0
            state = os.path.join(tempfile.mkdtemp(), "hub_d1024.imps")
            v.save(state)
            dims, secs = v.bond_dims(), [len(b[1]) for b in v.bonds]
            v.close()
            del v
            settle()
            log(f"vumps hubbard D=1024: grown in {grow} FP64 iterations ({t_grow:.1f} s) to {dims} ({secs} sectors)")
        for name in names:
            v = hub(precs[name])
            v.load(state)
            conv = name in ("fp64", "fp32")
            cap = args.d1024_iters if conv else 5
            sts, its = [], []
            t0 = time.perf_counter()
            while len(sts) < cap:
                t1 = time.perf_counter()
                sts.append(v.sweep())
                its.append(time.perf_counter() - t1)
                if len(sts) % 100 == 0:
                    log(f"  {name}: iteration {len(sts)}, F {max(sts[-1].F):.2e}")
                if conv and max(sts[-1].F) <= 1e-10:
                    break
            torch.cuda.synchronize()
            total = time.perf_counter() - t0
            st = sts[-1]
            dt = float(np.median(its[1:] if len(its) > 1 else its))
            r = dict(s_per_iter=dt, iterations=len(sts), growth_iters=grow, bond_dims=v.bond_dims(), sectors=[len(b[1]) for b in v.bonds],
                     energy=st.energy, F=max(st.F), eig_iters_per_iter=sum(x.eig_iters for x in sts) / len(sts),
                     env_iters_per_iter=sum(x.env_iters for x in sts) / len(sts))
            if conv:
                r.update(seconds_to_F=total, converged=bool(max(st.F) <= 1e-10))
            hres[name] = r
            log(f"vumps hubbard D=1024 {name}: {len(sts)} iterations" + (f" to F {max(st.F):.2e} in {total:.1f} s" if conv else "") +
                f", {dt:.3f} s/iteration (median; {r['eig_iters_per_iter']:.0f} Lanczos, {r['env_iters_per_iter']:.0f} GMRES "
                f"steps per iteration), E {st.energy:.15f}")
            v.close()
            del v
            settle()
        out["vumps_hubbard_d1024"] = hres
        torch.cuda.synchronize()
        torch.cuda.set_stream(stream)
        if prev_tune is None:
            os.environ.pop("IMPS_AUTOTUNE", None)
        else:
            os.environ["IMPS_AUTOTUNE"] = prev_tune
        if "fp64" in hres:
            mx = [k for k in ("fp32", "bf16x3") if k in hres and "seconds_to_F" in hres[k]]  # converged runs only
            best = min(mx, key=lambda k: hres[k]["s_per_iter"]) if mx else None
            out["vumps_iter_s"] = {**{k: v["s_per_iter"] for k, v in hres.items()},
                                   "mixed_vs_fp64": hres["fp64"]["s_per_iter"] / hres[best]["s_per_iter"] if best else None,
                                   "mixed_path": best}
            out["vumps_d1024_seconds_to_F1e-10"] = {k: v["seconds_to_F"] for k, v in hres.items() if "seconds_to_F" in v}
            lw = lieb_wu(cfg["params"][1]) - cfg["params"][1] / 4.0
            out["vumps_hubbard_d1024_energy_vs_lieb_wu"] = {k: v["energy"] - lw for k, v in hres.items()}
            out["vumps_hubbard_d1024_energy_vs_fp64"] = {k: v["energy"] - hres["fp64"]["energy"] for k, v in hres.items()}

    # Transfer-matrix microbenchmark: left and right TM apply of one site, anisotropic Heisenberg
    # U(1), d = 2, D = 512 over 17 sectors of sizes 2..64 (Gaussian in charge), seeded normal
    # entries; FP32 and BF16X3 against FP64, 1 and `nvec` vectors per launch.
    if not args.no_tm:
        tc = wl.TM_512
        tk, ts = wl.xxz_profile(tc["nsec"], tc["peak"], tc["lo"], tc["total"])
        tb = im.make_dims(tc["model"], tc["N"], tc["QN"], tk, ts)
        tl = im.mps_structure(tc["model"], tc["N"], tc["QN"], tb[0], tb[1])
        tA = dev(wl.random_data(tl, np.random.default_rng(1002), scale_by_rows=True))
        tres = {}
        for direction, dname in ((im.LEFT, "left"), (im.RIGHT, "right")):
            xl = wl.diag_layout(1, tb[0] if direction == im.LEFT else tb[1])
            yl = wl.diag_layout(1, tb[1] if direction == im.LEFT else tb[0])
            tx = dev(np.random.default_rng(1003).standard_normal(xl.nelem * NV))
            ref = None
            for name in ("fp64", "fp32", "bf16x3"):
                tp = im.TMPlan(ctx, direction, tl, None, xl, yl, precs[name])
                tp.set_mps(tA)
                tflops, _ = tp.cost()
                ty = torch.empty(yl.nelem * NV, dtype=torch.float64, device="cuda")
                r = {}
                for label, nv in (("batched", NV), ("single", 1)):
                    for _ in range(args.warmup):
                        tp.apply(tx, ty, nv)
                    torch.cuda.synchronize()
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    for _ in range(args.steps):
                        tp.apply(tx, ty, nv)
                    e1.record(stream)
                    torch.cuda.synchronize()
                    ms = e0.elapsed_time(e1) / args.steps
                    tp.timing(True)
                    for _ in range(args.steps):
                        tp.apply(tx, ty, nv)
                    kms, kn = tp.timing(False)
                    r[label] = {"applies_per_s": nv / (ms * 1e-3), "us_per_apply": ms * 1e3, "gemm_us": kms / kn * 1e3,
                                "gemm_tflops": nv * tflops / (kms / kn * 1e-3) / 1e12}
                y1 = ty[:yl.nelem].cpu().numpy()
                if ref is None:
                    ref = y1
                r["relerr_vs_fp64"] = float(np.linalg.norm(y1 - ref) / np.linalg.norm(ref))
                r["schedule"] = {"batched": tp.schedule(NV), "single": tp.schedule(1)}
                r["flops_per_apply"] = tflops
                tres.setdefault(dname, {})[name] = r
            log(f"tm D=512 {dname}: " + ", ".join(f"{k} {v['batched']['applies_per_s']:.3g}/s (x{NV}), {v['single']['applies_per_s']:.3g}/s (x1)"
                                                  for k, v in tres[dname].items()))
        out["tm_xxz_d512"] = tres

    # Adjoints (differentiation workload): anisotropic Heisenberg (reference parameters Jxy = 1,
    # Delta = -1.5), U(1), d = 2, D = 256 over 17 sectors, seeded normal entries.  Per FP64 call: the
    # left TM apply and its vector-Jacobian product (x-bar and A-bar), the H_AC apply and its VJP
    # (x-bar, EHL-, EHR-, HL-, HR-bar); each VJP is checked against a central difference of
    # <y-bar, f(inputs + h d)> along a random direction d: h = 1e-6 for the TM (cubic along the
    # line), h = 1e-2 for H_AC (bilinear in x and the operators, so quadratic along the line and
    # its central difference exact up to rounding).
    if not args.no_adjoint:
        ak, asz = wl.xxz_profile(17, 28, 2, 256)
        ab = im.make_dims(im.XXZ, 2, [0], ak, asz)
        arng = np.random.default_rng(1004)
        aA = im.mps_structure(im.XXZ, 2, [0], ab[1], ab[0])
        aAC = im.mps_structure(im.XXZ, 2, [0], ab[0], ab[1])
        ax0, ax1 = wl.diag_layout(1, ab[1]), wl.diag_layout(1, ab[0])
        nsteps = min(args.steps, 100)

        def timed(fn):
            for _ in range(3):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(nsteps):
                fn()
            e1.record(stream)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / nsteps * 1e3

        def rnd(n):
            return dev(arng.standard_normal(n))

        ares = {}
        # left TM: x on bond ab[1] -> y on bond ab[0] through A: ab[1] -> ab[0]
        tpa = im.TMPlan(ctx, im.LEFT, aA, None, ax0, ax1, im.FP64)
        A_, x_, yb = dev(wl.random_data(aA, arng, scale_by_rows=True)), rnd(ax0.nelem), rnd(ax1.nelem)
        tpa.set_mps(A_)
        y_ = torch.empty(ax1.nelem, dtype=torch.float64, device="cuda")
        xb_, Ab_ = torch.empty_like(x_), torch.empty_like(A_)
        fwd = timed(lambda: tpa.apply(x_, y_, 1))
        bwd = timed(lambda: tpa.vjp(A_, None, x_, yb, xb_, Ab_))
        dx, dA, hh = rnd(x_.numel()), rnd(A_.numel()), 1e-6

        def tm_f(x, A):
            tpa.set_mps(A)
            tpa.apply(x, y_, 1)
            return torch.dot(yb, y_).item()
        fd = (tm_f(x_ + hh * dx, A_ + hh * dA) - tm_f(x_ - hh * dx, A_ - hh * dA)) / (2 * hh)
        an = (torch.dot(xb_, dx) + torch.dot(Ab_, dA)).item()
        ares["tm_left"] = {"forward_us": fwd, "vjp_us": bwd, "vjp_per_s": 1e6 / bwd, "vjp_over_forward": bwd / fwd,
                           "fd_relerr": abs(fd - an) / abs(an)}
        # H_AC at the same bonds, HL / HR by the GPU ReducedOp of a random A
        aH = im.model_ham(im.XXZ, [1.0, -1.5, 0.0])
        arl, arr = im.RedOpPlan(ctx, aH, 2, aA, im.LEFT), im.RedOpPlan(ctx, aH, 2, aA, im.RIGHT)
        aHL = torch.empty(arl.out_layout.nelem, dtype=torch.float64, device="cuda")
        aHR = torch.empty(arr.out_layout.nelem, dtype=torch.float64, device="cuda")
        arl.apply(A_, aHL)
        arr.apply(A_, aHR)
        aEHL, aEHR = wl.diag_layout(1, ab[0]), wl.diag_layout(1, ab[1])
        ops = [dev(wl.random_data(aEHL, arng, symmetric=True)), dev(wl.random_data(aEHR, arng, symmetric=True)), aHL, aHR]
        hp = im.HACPlan(ctx, aAC, aEHL, aEHR, arl.out_layout, arr.out_layout, im.FP64)
        hx, hyb = rnd(aAC.nelem), rnd(aAC.nelem)
        hy, hxb = torch.empty_like(hx), torch.empty_like(hx)
        obars = [torch.empty_like(o) for o in ops]
        hp.set_operator(*ops)
        fwd = timed(lambda: hp.apply(hx, hy, 1))
        bwd = timed(lambda: hp.vjp(ops, hx, hyb, hxb, obars))
        dirs = [rnd(t.numel()) for t in [hx] + ops]

        hh = 1e-2

        def hac_f(sign):
            p = [t + sign * hh * d for t, d in zip([hx] + ops, dirs)]
            hp.set_operator(*p[1:])
            hp.apply(p[0], hy, 1)
            return torch.dot(hyb, hy).item()
        fd = (hac_f(1.0) - hac_f(-1.0)) / (2 * hh)
        an = sum(torch.dot(g, d).item() for g, d in zip([hxb] + obars, dirs))
        ares["hac"] = {"forward_us": fwd, "vjp_us": bwd, "vjp_per_s": 1e6 / bwd, "vjp_over_forward": bwd / fwd,
                       "fd_relerr": abs(fd - an) / abs(an)}
        ares["config"] = "XXZ (Jxy 1, Delta -1.5), U(1), d = 2, D = 256 over 17 sectors, FP64, one vector per call"
        log("adjoints D=256: " + ", ".join(f"{k} fwd {v['forward_us']:.1f} us, vjp {v['vjp_us']:.1f} us, fd relerr {v['fd_relerr']:.1e}"
                                          for k, v in ares.items() if isinstance(v, dict)))
        out["adjoints_xxz_d256"] = ares

    # Quasiparticle excitations (config 5): the Fermi-Hubbard ground state at D = 512, reached as the
    # reference does (bond growth from D = 64, then VUMPS iterations to F <= 1e-10), and the 32
    # momenta p = j / 31, j = 0..31 (units of pi per site), K = (0, 0), all in one batched ApplyHeff
    # (imps_heff_apply) and one batched eigensolve for the 4 lowest modes (imps_heff_eigs, tol 1e-10,
    # geometric sums to InvETol = 1e-11: the reference's defaults, Excitations_main.cpp:32,65).
    # The clustered low-energy spectrum of this gapless chain needs a large Lanczos basis
    # (--exc-krylov 160): about 1300 batched applications to converge every mode, bounded by
    # --exc-maxit.  The FP32-refined path is timed and compared on the applies (its eigensolve with
    # --exc-mixed-eigs).  The seeded random mixed-canonical state of the synthetic 81-sector D = 512
    # structure is solved to convergence beside it in both precisions.
    if not args.no_exc:
        from imps_cpp_tools_amd.nullspace import left_nullspace
        from imps_cpp_tools_amd.vumps import VUMPS
        settle()
        torch.cuda.set_stream(torch.cuda.default_stream())
        # the ground state's VUMPS on torch's default stream; the excitation engine on a stream of its
        # own (default priority), so that its kernels are the only ones there; every hand-over between
        # the two is a device synchronisation
        gctx = im.Context(0)
        xstream = torch.cuda.Stream()
        ectx = im.Context(0, xstream)
        prev_tune = os.environ.get("IMPS_AUTOTUNE")
        os.environ["IMPS_AUTOTUNE"] = "0"
        ec = wl.HUBBARD_512
        ps = [j / 31 for j in range(32)]
        exc = {}

        def heff_run(gs, label, maxit, check_conv, prec=im.FP64, gs_state=True):
            E0 = gs.energy()
            nls = [left_nullspace(gctx, gs.AL[n], gs.slay[n], seed=n) for n in range(ec["N"])]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            P = im.HeffPlan(ectx, gs.slay, gs.slay, [l for _, l in nls], [gs.blay[(n + 1) % ec["N"]] for n in range(ec["N"])], gs.H, E0,
                            (0, 0), len(ps), prec=prec)
            P.set_state(gs.AL, gs.AR, [a for a, _ in nls], gs.C, inv_tol=1e-11)
            torch.cuda.synchronize()
            setup = time.perf_counter() - t0
            trl, trr, _ = P.check()
            zr = np.random.default_rng(7)
            x = torch.from_numpy(zr.standard_normal((len(ps), P.mtot)) + 1j * zr.standard_normal((len(ps), P.mtot))).cuda()
            y = torch.empty_like(x)
            torch.cuda.synchronize()
            P.apply(ps, x, y)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                steps, rr = P.apply(ps, x, y)
            torch.cuda.synchronize()
            t_apply = (time.perf_counter() - t0) / 3
            # Hermiticity of H_eff(k) at every momentum: <x_t, H y_t> = conj(<y_t, H x_t>)
            y2 = torch.empty_like(x)
            w = torch.from_numpy(zr.standard_normal((len(ps), P.mtot)) + 1j * zr.standard_normal((len(ps), P.mtot))).cuda()
            torch.cuda.synchronize()
            P.apply(ps, w, y2)
            torch.cuda.synchronize()
            hx, hy, hw, hy2 = (a.cpu().numpy() for a in (x, y, w, y2))  # the check on the host
            herm = max(abs(np.vdot(hw[t], hy[t]) - np.conj(np.vdot(hx[t], hy2[t]))) / abs(np.vdot(hw[t], hy[t])) for t in range(len(ps)))
            r = {"state": label, "D": gs.bond_dims(), "sectors": [len(b[1]) for b in gs.bonds], "x_dim_complex": P.mtot,
                 "momenta": len(ps), "setup_s": setup, "heffconstants_trace_check": [trl, trr],
                 "batched_apply_s": t_apply, "momentum_applies_per_s": len(ps) / t_apply, "gmres_steps_per_apply": steps,
                 "gmres_worst_relres": rr, "hermiticity_worst_rel": herm, "_y": hy}
            label = label + ("" if prec == im.FP64 else f" [{ {im.FP32: 'fp32', im.BF16X3: 'bf16x3'}[prec]} TM chains, FP64 refinement]")
            msg = f"excitations {label}: D {gs.bond_dims()}, mtot {P.mtot}, batched apply {t_apply * 1e3:.1f} ms for {len(ps)} momenta ({steps} GMRES steps)"
            if maxit > 0:
                t0 = time.perf_counter()
                kry = args.exc_krylov if gs_state else 40
                ev, _, res, na = P.eigs(ps, nev=4, tol=1e-10, krylov=kry, maxit=maxit, check=False)
                torch.cuda.synchronize()
                t_eig = time.perf_counter() - t0
                rel = res / np.maximum(1.0, np.abs(ev))
                r.update({"eigs_krylov": kry, "eigs_s_32_momenta": t_eig, "eigs_batched_applies": na, "eigs_worst_rel_residual": float(rel.max()),
                          "eigs_worst_rel_residual_per_level": [float(v) for v in rel.max(axis=0)],
                          "eigs_converged": bool(rel.max() <= 1e-10), "lowest_energy_per_momentum": [float(v) for v in ev[:, 0]],
                          "four_lowest_p0": [float(v) for v in ev[0]], "min_energy": float(ev.min()), "_ev": ev})
                msg += f", eigs {t_eig:.2f} s / {na} applies, worst residual {rel.max():.1e}, min E {ev.min():.6f}"
            log(msg)
            del P
            return r

        if not args.exc_synthetic_only:
            # ground state
            k64, s64 = wl.fhub_profile(5, 8, 1, 64)
            t0 = time.perf_counter()
            gs = VUMPS(gctx, ec["model"], ec["N"], ec["QN"], k64, s64, ec["params"], prec=im.FP64, seed=1005, krylov=30, env_krylov=300,
                       mmax=ec["total"], dm_max=128, expand_pause=1, expand_thresh=1.0)
            it = 0
            while min(gs.bond_dims()) < ec["total"] and it < 40:
                st = gs.sweep()
                it += 1
            gs.mmax = None
            while it < args.gs_iters:
                st = gs.sweep()
                it += 1
                if max(st.F) <= 1e-10:
                    break
            torch.cuda.synchronize()
            t_gs = time.perf_counter() - t0
            exc["ground_state"] = {"vumps_iterations": it, "seconds": t_gs, "F": max(st.F), "energy": st.energy,
                                   "energy_vs_lieb_wu": st.energy - (lieb_wu(ec["params"][1]) - ec["params"][1] / 4.0)}
            log(f"ground state D=512: {it} VUMPS iterations in {t_gs:.1f} s, F {max(st.F):.2e}, E {st.energy:.14f}")
            exc["hubbard_d512_ground_state"] = heff_run(gs, "VUMPS ground state (grown from D=64, F<=1e-10)", args.exc_maxit, False)
            exc["hubbard_d512_ground_state_mixed"] = heff_run(gs, "VUMPS ground state (grown from D=64, F<=1e-10)",
                                                              args.exc_maxit if args.exc_mixed_eigs else 0, False, prec=im.FP32)
            gs.close()
            del gs
            settle()
        # the synthetic structure (81 sectors) with its seeded random mixed-canonical state
        k5, s5 = wl.fhub_profile(ec["nside"], ec["peak"], 1, ec["total"])
        gr = VUMPS(gctx, ec["model"], ec["N"], ec["QN"], k5, s5, ec["params"], prec=im.FP64, seed=1005, env_krylov=300)
        exc["hubbard_d512_random_81_sectors"] = heff_run(gr, "seeded random mixed-canonical state, 81 sectors", 2000, True, gs_state=False)
        exc["hubbard_d512_random_81_sectors_mixed"] = heff_run(gr, "seeded random mixed-canonical state, 81 sectors", 2000, True, prec=im.FP32,
                                                               gs_state=False)
        # the mixed runs against FP64: energies (all 32 momenta x 4 levels) and speed
        for key in ("hubbard_d512_ground_state", "hubbard_d512_random_81_sectors"):
            if key in exc and key + "_mixed" in exc:
                a, b = exc[key], exc[key + "_mixed"]
                # the same seeded input vectors: H_eff(p) x of the refined path against FP64
                b["apply_relerr_vs_fp64"] = float(np.linalg.norm(b["_y"] - a["_y"]) / np.linalg.norm(a["_y"]))
                b["apply_speedup_vs_fp64"] = a["batched_apply_s"] / b["batched_apply_s"]
                msg = f"excitations {key}: FP32 TM chains vs FP64: apply relerr {b['apply_relerr_vs_fp64']:.1e}, x{b['apply_speedup_vs_fp64']:.2f}"
                if "_ev" in a and "_ev" in b:
                    b["energy_err_vs_fp64"] = float(np.abs(b["_ev"] - a["_ev"]).max())
                    b["eigs_speedup_vs_fp64"] = a["eigs_s_32_momenta"] / b["eigs_s_32_momenta"]
                    msg += f", energies {b['energy_err_vs_fp64']:.1e}, eigs x{b['eigs_speedup_vs_fp64']:.2f}"
                log(msg)
        for r in exc.values():
            r.pop("_ev", None)
            r.pop("_y", None)
        gr.close()
        del gr
        settle()
        out["excitations_hubbard_d512"] = exc
        torch.cuda.set_stream(stream)
        if prev_tune is None:
            os.environ.pop("IMPS_AUTOTUNE", None)
        else:
            os.environ["IMPS_AUTOTUNE"] = prev_tune

    # CPU baseline: the same H_AC workload on the host, FP64 CPU restatement of the reference's
    # products (compiled once, OpenMP over output blocks), on every core of the affinity mask and on
    # one core, each over a bounded sample; the first apply is checked against the oracle's H_AC
    if not args.no_cpu:
        from oracle import pyoracle as orc
        lay_ob = lambda l, d: orc.OB.like(l, d)  # noqa: E731
        ob = dict(EHL=lay_ob(L["EHL"], EHL.cpu().numpy()), EHR=lay_ob(L["EHR"], EHR.cpu().numpy()),
                  HL=lay_ob(HLl, HL.cpu().numpy()), HR=lay_ob(HRl, HR.cpu().numpy()), AC=lay_ob(L["AC"], X[:n].cpu().numpy()))
        cp = orc.HacPlanCPU(ob["EHL"], ob["EHR"], ob["HL"], ob["HR"], ob["AC"])
        xh = np.ascontiguousarray(ob["AC"].data)
        yh = np.empty_like(xh)
        cores = len(os.sched_getaffinity(0))
        timed = {}
        for th in (cores, 1):
            cp.apply(xh, yh, th)
            reps, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < args.cpu_seconds / 2:
                cp.apply(xh, yh, th)
                reps += 1
            dt = time.perf_counter() - t0
            timed[th] = (reps / dt, reps, dt)
        yref = orc.hac(ob["EHL"], ob["EHR"], ob["HL"], ob["HR"], ob["AC"]).data
        v_all, r_all, d_all = timed[cores]
        out["cpu_baseline"] = {"value": v_all, "unit": "H_AC vector-applies/s", "cores": cores, "kind": "port",
                               "sample": f"{r_all} single-vector H_AC applies of the same workload on {cores} threads ({d_all:.1f} s); "
                                         f"1 thread: {timed[1][1]} applies ({timed[1][2]:.1f} s)",
                               "single_core_value": timed[1][0],
                               "cpu_vs_oracle_relerr": float(np.linalg.norm(yh - yref) / np.linalg.norm(yref))}
        nr = np.linalg.norm(yref)
        out["err_vs_oracle"] = {k: float(np.linalg.norm(results[k]["y"] - yref) / nr) for k in results}
        log(f"cpu: {v_all:.3g} applies/s on {cores} cores, {timed[1][0]:.3g} on 1 core; errors {out['err_vs_oracle']}")
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
