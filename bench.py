"""Benchmark: H_AC applies/s on the D = 1024 Hubbard U(1)xU(1) workload, mixed precision vs FP64
on the same GPU.

Workload: Fermi-Hubbard chain (t=1, U=4, half filling), (N, M) representation, two-site unit cell,
d = 4, D = 1024 over 121 charge sectors of sizes 1..60.  One step = one grouped H_AC application
(VUMPS_helpers.hpp:38-88) to a block of `nvec` Krylov vectors resident in HBM (one kernel launch).
Operator data are seeded synthetic (isometry-scale normal entries, reduced operators computed on
the GPU by ReducedOp).

Prints ONE JSON line on stdout.  The value is vector-applies/s of the mixed-precision path
(split-BF16 on MFMA, FP32 accumulation); FP64 and FP32 rates, single-vector rates, the roofline of
the panel GEMM kernel (its own launch time from HIP events on the launch stream), VUMPS iteration
times (Heisenberg D = 64 to convergence; Hubbard D = 1024 first iteration) and the CPU oracle timed
on the host (cpu_baseline) are reported alongside.
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
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound on the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-vumps", action="store_true")
    ap.add_argument("--no-xxz", action="store_true", help="skip the D = 64 Heisenberg VUMPS runs (keep the D = 1024 Hubbard one)")
    ap.add_argument("--no-exc", action="store_true")
    ap.add_argument("--no-tm", action="store_true")
    ap.add_argument("--exc-momenta", type=int, default=4, help="momenta of the excitation sample (of 32)")
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

    # the mixed-precision path: FP32 or split-BF16 (BF16X3) MFMA products with FP32 accumulation,
    # refined in FP64 by the solvers; the faster of the two in this run is the headline value
    mixed_names = [k for k in ("fp32", "bf16x3") if k in results]
    head = (max(mixed_names, key=lambda k: results[k]["batched"]["applies_per_s"]) if mixed_names else next(iter(results)))
    mixed = results[head]
    rate = lambda k, lab="batched": results[k][lab]["applies_per_s"] if k in results else None  # noqa: E731
    flops = mixed["flops_per_apply"]
    achieved_tf = mixed["batched"]["kernel_tflops"]
    ks, sh = results[head]["schedule"]["batched"]
    pidx = {"fp64": 0, "fp32": 1, "bf16x3": 2}[head]
    kname = f"panel_gemm_shared_kernel<{pidx}>" if sh else f"panel_gemm_kernel<{pidx}, {ks}>"  # rocprof's template names
    out = {
        "metric": "H_AC applies/s & VUMPS iter time, D=1024 Hubbard: mixed vs FP64, same GPU",
        "value": mixed["batched"]["applies_per_s"],
        "unit": "H_AC vector-applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["batched"]["ms_per_launch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16x3": "bf16x3 (split-BF16 MFMA, FP32 accumulate; FP64 in/out)",
                  "fp32": "f32 (FP32 MFMA, FP32 accumulate; FP64 in/out)", "fp64": "f64"}[head],
        "data": "synthetic (seeded)",
        "config": {"workload": "H_AC apply, Fermi-Hubbard U(1)xU(1) (N,M), d=4, D=1024, 121 sectors of size 1..60",
                   "nvec_per_step": NV, "flops_per_apply": flops, "ac_elements": n},
        "mixed_vs_fp64": (mixed["batched"]["applies_per_s"] / rate("fp64")) if "fp64" in results else None,
        "bf16x3_applies_per_s": rate("bf16x3"),
        "fp64_applies_per_s": rate("fp64"),
        "fp32_applies_per_s": rate("fp32"),
        "single_vector_applies_per_s": {k: results[k]["single"]["applies_per_s"] for k in results},
        "panel_schedules": {k: results[k]["schedule"] for k in results},
        "vumps_iter_s": None,
        "roofline": {"bound": "mfma", "achieved": achieved_tf, "peak": PEAK_TFLOPS[head], "unit": "TFLOP/s",
                     "frac": achieved_tf / PEAK_TFLOPS[head], "traffic": committed_traffic(kname, NV), "kernel": kname,
                     "per_launch": {"flops": NV * flops, "us": mixed["batched"]["kernel_us"],
                                    "apply_us_incl_split_copy": mixed["batched"]["ms_per_launch"] * 1e3}},
    }
    if "fp64" in results:
        out["roofline"]["fp64_frac_of_fp64_peak"] = results["fp64"]["batched"]["kernel_tflops"] / PEAK_TFLOPS["fp64"]

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
        # VUMPS outer-iteration time on the D = 1024 Hubbard chain, as the reference reaches it: from a
        # D = 64 bond (25 sectors, seeded randMPS_LR state) the bonds grow by ExpandFromH (at most 256
        # states per bond and iteration) to D = 1024; after one more iteration on the final structure
        # (plans built), the median wall time of the next five iterations is the iteration time.
        k64, s64 = wl.fhub_profile(5, 8, 1, 64)
        hres = {}
        for name in [k for k in ("fp64", "fp32", "bf16x3") if k in results]:
            v = VUMPS(vctx, cfg["model"], cfg["N"], cfg["QN"], k64, s64, cfg["params"], prec=precs[name], seed=1003, krylov=30,
                      env_krylov=300, mmax=1024, dm_max=256, expand_pause=1, expand_thresh=1.0)
            grow = 0
            while min(v.bond_dims()) < 1024 and grow < 20:
                v.sweep()
                grow += 1
            v.sweep()
            settle()
            sts, its = [], []
            for _ in range(5):
                t0 = time.perf_counter()
                sts.append(v.sweep())
                torch.cuda.synchronize()
                its.append(time.perf_counter() - t0)
            dt = float(np.median(its))
            st = sts[-1]
            hres[name] = dict(s_per_iter=dt, iter_s=its, growth_iters=grow, bond_dims=v.bond_dims(), sectors=[len(b[1]) for b in v.bonds],
                              energy=st.energy, F=max(st.F), eig_iters_per_iter=sum(x.eig_iters for x in sts) / len(sts),
                              env_iters_per_iter=sum(x.env_iters for x in sts) / len(sts))
            log(f"vumps hubbard D=1024 {name}: grown in {grow} iterations to {v.bond_dims()} ({hres[name]['sectors']} sectors), "
                f"{dt:.3f} s/iteration (median of {' '.join(f'{t:.3f}' for t in its)}; {hres[name]['eig_iters_per_iter']:.0f} Lanczos, {hres[name]['env_iters_per_iter']:.0f} GMRES "
                f"steps), E {st.energy:.12f}, F {max(st.F):.2e}")
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
        if "fp64" in hres and head in hres:
            out["vumps_iter_s"] = {**{k: v["s_per_iter"] for k, v in hres.items()},
                                   "mixed_vs_fp64": hres["fp64"]["s_per_iter"] / hres[head]["s_per_iter"]}

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

    # Quasiparticle excitations on the D = 512 Hubbard structure (81 (N, M) sectors): ApplyHeff
    # (two nested geometric sums per apply) and the 4 lowest modes per momentum on a sample of the
    # 32 momenta p = j / 31 (units of pi per site), K = (0, 0).  The state is the seeded random
    # mixed-canonical MPS of that structure (a converged ground state at this D is reached by bond
    # growth, not built here); the cost per apply depends on the structure and on the geometric
    # sums' iteration counts.
    if not args.no_exc:
        from imps_cpp_tools_amd.excitations import ExcitationEngine
        from imps_cpp_tools_amd.nullspace import left_nullspace
        from imps_cpp_tools_amd.vumps import VUMPS
        ec = wl.HUBBARD_512
        k5, s5 = wl.fhub_profile(ec["nside"], ec["peak"], 1, ec["total"])
        t0 = time.perf_counter()
        gs = VUMPS(ctx, ec["model"], ec["N"], ec["QN"], k5, s5, ec["params"], prec=im.FP64, seed=1005, env_krylov=300)
        _, ecell = gs._left_env(0, 1e-14)
        nls = [left_nullspace(ctx, gs.AL[n], gs.slay[n], seed=n) for n in range(ec["N"])]
        eng = ExcitationEngine(ctx, gs.AL, gs.slay, gs.AR, gs.slay, [a for a, _ in nls], [l for _, l in nls], gs.C,
                               [gs.blay[(n + 1) % ec["N"]] for n in range(ec["N"])], gs.H, gs.d, ecell / ec["N"], (0, 0))
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        zr = np.random.default_rng(7)
        z = torch.from_numpy(zr.standard_normal(eng.mtot) + 1j * zr.standard_normal(eng.mtot)).cuda()
        eng.matvec(z, 0.5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            eng.matvec(z, 0.5)
        torch.cuda.synchronize()
        t_apply = (time.perf_counter() - t0) / 5
        ps = [j / 31 for j in np.linspace(1, 30, args.exc_momenta).round().astype(int)]
        per = []
        for p in ps:
            t0 = time.perf_counter()
            w, _, nmv = eng.lowest(p, nev=4, tol=1e-8, krylov=40)
            torch.cuda.synchronize()
            per.append((p, time.perf_counter() - t0, nmv))
        s_mom = float(np.mean([x[1] for x in per]))
        out["excitations_hubbard_d512"] = {
            "precision": "fp64", "x_dim_complex": eng.mtot, "setup_s": setup,
            "applyheff_per_s": 1.0 / t_apply, "applyheff_s_at_p0.5": t_apply,
            "momenta_sample": [round(p, 4) for p in ps], "s_per_momentum_4_lowest": s_mom,
            "applyheff_per_momentum": [x[2] for x in per], "est_s_32_momenta": 32 * s_mom}
        log(f"excitations hubbard D=512: ApplyHeff {t_apply * 1e3:.1f} ms, 4 lowest {s_mom:.2f} s per momentum "
            f"({[x[2] for x in per]} applies), setup {setup:.2f} s")

    # CPU baseline: the oracle (clean-room FP64 restatement of ApplyHA) on the host, bounded sample
    if not args.no_cpu:
        from oracle import pyoracle as orc
        lay_ob = lambda l, d: orc.OB.like(l, d)  # noqa: E731
        ob = dict(EHL=lay_ob(L["EHL"], EHL.cpu().numpy()), EHR=lay_ob(L["EHR"], EHR.cpu().numpy()),
                  HL=lay_ob(HLl, HL.cpu().numpy()), HR=lay_ob(HRl, HR.cpu().numpy()), AC=lay_ob(L["AC"], X[:n].cpu().numpy()))
        hs = {k: v.handle() for k, v in ob.items()}
        lib = orc.lib()
        reps, t0 = 0, time.perf_counter()
        yref = None
        while time.perf_counter() - t0 < args.cpu_seconds:
            h = lib.orc_hac(hs["EHL"].h, hs["EHR"].h, hs["HL"].h, hs["HR"].h, hs["AC"].h)
            if yref is None:
                yref = orc._H(h).get().data
            else:
                lib.orc_free(h)
            reps += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": reps / dt, "unit": "H_AC vector-applies/s", "cores": 1, "kind": "port",
                               "sample": f"{reps} single-vector H_AC applies of the same workload ({dt:.1f} s)"}
        nr = np.linalg.norm(yref)
        out["err_vs_oracle"] = {k: float(np.linalg.norm(results[k]["y"] - yref) / nr) for k in results}
        log(f"cpu oracle: {reps / dt:.3g} applies/s; errors {out['err_vs_oracle']}")
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
