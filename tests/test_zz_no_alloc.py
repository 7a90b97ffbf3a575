"""The allocation-free H_AC entry point (imps_hac_apply_static) against imps_hac_apply.

The D = 1024 Hubbard H_AC workload (121 sectors) and a small Hubbard cell, FP64 / FP32 / BF16X3,
autotuning off so that both entry points run the same schedule: the results must be bitwise equal
(same kernels, same summation order) on two calls in a row, the second with fewer vectors and then
on the smaller plan (the fixed scratch keeps the previous call's contents, which must not leak into
the pads); sizes beyond the 64 MiB scratch are refused with IMPS_INVALID_ARG."""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
from imps_cpp_tools_amd import workloads as wl

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def split_bytes_per_vector(prec, n):
    plane = (n + 64 + 63) // 64 * 64
    return {im.FP64: 16, im.FP32: 8, im.BF16X3: 8}[prec] * plane


def _plans(ctx, prec):
    out = []
    for cfg, seed in ((wl.HUBBARD_1024, 1003), (dict(wl.HUBBARD_1024, nside=5, peak=6, total=None), 7)):
        if cfg["total"] is None:
            k0, s0 = wl.fhub_profile(cfg["nside"], cfg["peak"], 1)
            bonds = im.make_dims(cfg["model"], cfg["N"], cfg["QN"], k0, s0)
            AC = im.mps_structure(cfg["model"], cfg["N"], cfg["QN"], bonds[0], bonds[1])
            A10 = im.mps_structure(cfg["model"], cfg["N"], cfg["QN"], bonds[1], bonds[0])
            L = dict(AC=AC, AL=A10, AR=A10, EHL=wl.diag_layout(2, bonds[0]), EHR=wl.diag_layout(2, bonds[1]))
        else:
            L = wl.hac_workload_layouts(cfg)
        rng = np.random.default_rng(seed)
        H = im.model_ham(cfg["model"], cfg["params"])
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()  # noqa: E731
        AL = dev(wl.random_data(L["AL"], rng, scale_by_rows=True))
        AR = dev(wl.random_data(L["AR"], rng, scale_by_rows=True))
        rl = im.RedOpPlan(ctx, H, 4, L["AL"], im.LEFT)
        rr = im.RedOpPlan(ctx, H, 4, L["AR"], im.RIGHT)
        HL = torch.empty(rl.out_layout.nelem, dtype=torch.float64, device="cuda")
        HR = torch.empty(rr.out_layout.nelem, dtype=torch.float64, device="cuda")
        rl.apply(AL, HL)
        rr.apply(AR, HR)
        p = im.HACPlan(ctx, L["AC"], L["EHL"], L["EHR"], rl.out_layout, rr.out_layout, prec)
        p.set_operator(dev(wl.random_data(L["EHL"], rng, symmetric=True)), dev(wl.random_data(L["EHR"], rng, symmetric=True)), HL, HR)
        out.append((p, L["AC"].nelem, rng))
    return out


@pytest.mark.parametrize("prec", [im.FP64, im.FP32, im.BF16X3], ids=["fp64", "fp32", "bf16x3"])
def test_static_apply_equals_apply(prec, monkeypatch):
    monkeypatch.setenv("IMPS_AUTOTUNE", "0")
    ctx = im.Context(0)
    (big, nb, rng), (small, ns, _) = _plans(ctx, prec)
    print(f"static apply prec {im.PREC_NAMES[prec]}: n {nb} / {ns}", flush=True)
    for p, n, nvec in ((big, nb, 8), (big, nb, 3), (small, ns, 5), (big, nb, 1)):
        x = torch.from_numpy(rng.standard_normal(n * nvec)).cuda()
        y1 = torch.full((n * nvec,), float("nan"), dtype=torch.float64, device="cuda")
        y2 = torch.full((n * nvec,), float("nan"), dtype=torch.float64, device="cuda")
        p.apply(x, y1, nvec)
        p.apply_static(x, y2, nvec)
        torch.cuda.synchronize()
        assert torch.isfinite(y2).all()
        rel = (torch.linalg.vector_norm(y2 - y1) / torch.linalg.vector_norm(y1)).item()
        print(f"  n {n} nvec {nvec}: rel diff {rel:.1e}", flush=True)
        assert torch.equal(y1, y2)
    # beyond the fixed scratch: refused, nothing written
    nmax = SCRATCH // split_bytes_per_vector(prec, nb)
    x = torch.zeros(nb * (nmax + 1), dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    with pytest.raises(im.ImpsError) as e:
        big.apply_static(x, y, nmax + 1)
    assert e.value.status == 1 and "scratch" in str(e.value)
    big.apply_static(x, y, nmax)  # the largest batch that fits is accepted
    torch.cuda.synchronize()
    ctx.close()
