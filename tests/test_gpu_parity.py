"""GPU parity of the grouped HIP kernels against the FP64 oracle, through the C ABI.

Tolerances (norm-wise relative error ||y - y_oracle|| / ||y_oracle||):
  FP64   : 1e-12  (the target is 1e-10 relative)
  FP32   : 1e-5   (FP32 MFMA, FP32 accumulation over K <= ~500)
  BF16X3 : 2e-4   (split-BF16 hi*hi + hi*lo + lo*hi, each product accurate to ~2^-16 relative)
Block structures are exact by construction: every output block of the layout is compared.
Cells: XXZ U(1), Fermi-Hubbard U(1)xU(1), and the Z(q) / Bose-Hubbard structures (quantum Ising
Z(2), 3-state Potts Z(3), whose blocks wrap modulo q, Bose-Hubbard d = 4 U(1) with an empty slice).
"""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
from states import BHUB, FHUB, QIM, QPOTTS, XXZ, Cell, fhub_bond0, random_like, xxz_bond0
from oracle import pyoracle as orc

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

TOL = {im.FP64: 1e-12, im.FP32: 1e-5, im.BF16X3: 2e-4}
PRECS = [im.FP64, im.FP32, im.BF16X3]


@pytest.fixture(scope="module")
def ctx():
    c = im.Context(0)
    yield c
    c.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def lay(ob):
    return im.BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols, ob.periods)


def relerr(got, ref):
    n = np.linalg.norm(ref)
    return np.linalg.norm(got - ref) / (n if n > 0 else 1.0)


CELLS = {}


def cell(name):
    if name not in CELLS:
        if name == "xxz":
            k, s = xxz_bond0(9, 12)
            CELLS[name] = Cell(XXZ, 2, [0], k, s, seed=101)
        elif name == "fhub":
            k, s = fhub_bond0(5, 10)
            CELLS[name] = Cell(FHUB, 2, [2, 0], k, s, seed=102)
        elif name == "xxz_big":  # TM microbenchmark shapes: 17 sectors, sizes up to 64
            k, s = xxz_bond0(17, 64, 2)
            CELLS[name] = Cell(XXZ, 2, [0], k, s, seed=1002)
        elif name in ZCELLS:
            model, QN, dims = ZCELLS[name]
            b = orc.make_dims_model(model, dims, 2, QN)
            CELLS[name] = Cell(model, 2, QN, b[0][0], b[0][1], seed=103)
    return CELLS[name]


ZCELLS = {"qim": (QIM, [], [40, 36]), "qpotts3": (QPOTTS(3), [], [20, 26, 17]), "bhub4": (BHUB(4), [2], [9, 14, 11, 6])}
HAM = {XXZ: [1.0, -1.0, 0.1], FHUB: [1.0, 4.0, 0.0, 0.0, 0.0], QIM: [1.0, 0.7], QPOTTS(3): [0.8], BHUB(4): [1.0, 3.0, 0.4]}
NAMES = ["xxz", "fhub", "qim", "qpotts3", "bhub4"]


def ham(c):
    return orc.model_ham(c.model, HAM[c.model])


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", NAMES + ["xxz_big"])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("mixed", [False, True])
def test_tm_parity(ctx, prec, name, direction, mixed):
    c = cell(name)
    rng = np.random.default_rng(7)
    A = random_like(c.struct[0], rng)
    B = random_like(c.struct[0], rng) if mixed else None
    xo = random_like(c.diag(0 if direction == 0 else 1), rng)
    yl = c.diag(1 if direction == 0 else 0)
    ref = orc.tm_mixed(direction, A, B, xo, yl).data
    p = im.TMPlan(ctx, direction, lay(A), lay(B) if mixed else None, lay(xo), lay(yl), prec)
    p.set_mps(dev(A.data), dev(B.data) if mixed else None)
    nvec = 3
    X = np.stack([xo.data, 2 * xo.data, -xo.data])
    y = torch.full((nvec * yl.nelem,), float("nan"), dtype=torch.float64, device="cuda")
    p.apply(dev(X), y, nvec)
    got = y.cpu().numpy().reshape(nvec, -1)
    for v, f in enumerate([1.0, 2.0, -1.0]):
        assert relerr(got[v], f * ref) < TOL[prec], (v, relerr(got[v], f * ref))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", NAMES)
def test_hac_parity(ctx, prec, name):
    c = cell(name)
    H = ham(c)
    rng = np.random.default_rng(8)
    HL = orc.reduced_op(H, c.d, c.AL[1], 0)
    HR = orc.reduced_op(H, c.d, c.AR[1], 1)
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(1), rng, sym=True)
    xs = [random_like(c.struct[0], rng) for _ in range(5)]
    refs = [orc.hac(EHL, EHR, HL, HR, x).data for x in xs]
    p = im.HACPlan(ctx, lay(xs[0]), lay(EHL), lay(EHR), lay(HL), lay(HR), prec)
    p.set_operator(dev(EHL.data), dev(EHR.data), dev(HL.data), dev(HR.data))
    for nvec in (1, 2, 3, 4, 5):  # K-split 4, 2, 2, 1, 1 (partly filled workgroups for 3 and 5)
        X = np.concatenate([x.data for x in xs[:nvec]])
        y = torch.full((X.size,), float("nan"), dtype=torch.float64, device="cuda")
        p.apply(dev(X), y, nvec)
        got = y.cpu().numpy().reshape(nvec, -1)
        for v in range(nvec):
            assert relerr(got[v], refs[v]) < TOL[prec], (nvec, v, relerr(got[v], refs[v]))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", NAMES)
def test_hc_parity(ctx, prec, name):
    """H_C through the reduced form equals the reference's per-nonzero 5-factor chain."""
    c = cell(name)
    H = ham(c)
    rng = np.random.default_rng(9)
    AL, AR = c.AL[1], c.AR[0]
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(0), rng, sym=True)
    C = random_like(c.diag(0), rng)
    ref = orc.hc(EHL, EHR, H, c.d, AL, AR, C).data
    HL = orc.reduced_op(H, c.d, AL, 0)
    p = im.HCPlan(ctx, lay(C), lay(EHL), lay(EHR), lay(HL), lay(AR), prec)
    p.set_operator(dev(EHL.data), dev(EHR.data), dev(HL.data), dev(AR.data))
    y = torch.full((C.nelem,), float("nan"), dtype=torch.float64, device="cuda")
    p.apply(dev(C.data), y)
    assert relerr(y.cpu().numpy(), ref) < TOL[prec]
    # three vectors at once (the third a scaled copy of the first: H_C is linear)
    C2 = random_like(c.diag(0), rng)
    ref2 = orc.hc(EHL, EHR, H, c.d, AL, AR, C2).data
    X = np.concatenate([C.data, C2.data, -2.0 * C.data])
    y3 = torch.full((X.size,), float("nan"), dtype=torch.float64, device="cuda")
    p.apply(dev(X), y3, 3)
    got = y3.cpu().numpy().reshape(3, -1)
    for g, r in zip(got, [ref, ref2, -2.0 * ref]):
        assert relerr(g, r) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("ksplit", ["1", "2", "4"])
def test_hac_ksplit_variants(ctx, prec, ksplit, monkeypatch):
    """Every K-split of the panel kernel (waves per vector) gives the oracle's result, for a
    vector count that leaves workgroups partly filled."""
    monkeypatch.setenv("IMPS_KSPLIT", ksplit)
    c = cell("fhub")
    H = ham(c)
    rng = np.random.default_rng(18)
    HL = orc.reduced_op(H, c.d, c.AL[1], 0)
    HR = orc.reduced_op(H, c.d, c.AR[1], 1)
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(1), rng, sym=True)
    xs = [random_like(c.struct[0], rng) for _ in range(3)]
    p = im.HACPlan(ctx, lay(xs[0]), lay(EHL), lay(EHR), lay(HL), lay(HR), prec)
    p.set_operator(dev(EHL.data), dev(EHR.data), dev(HL.data), dev(HR.data))
    X = np.concatenate([x.data for x in xs])
    y = torch.full((X.size,), float("nan"), dtype=torch.float64, device="cuda")
    p.apply(dev(X), y, 3)
    got = y.cpu().numpy().reshape(3, -1)
    for v, x in enumerate(xs):
        assert relerr(got[v], orc.hac(EHL, EHR, HL, HR, x).data) < TOL[prec]


@pytest.mark.parametrize("prec", [im.FP64, im.FP32])
@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("direction", [0, 1])
def test_redop_parity(ctx, prec, name, direction):
    c = cell(name)
    H = ham(c)
    A = c.AL[1] if direction == 0 else c.AR[1]
    ref = orc.reduced_op(H, c.d, A, direction)
    p = im.RedOpPlan(ctx, H, c.d, lay(A), direction, prec)
    assert p.out_layout == lay(ref)
    y = torch.full((ref.nelem,), float("nan"), dtype=torch.float64, device="cuda")
    p.apply(dev(A.data), y)
    assert relerr(y.cpu().numpy(), ref.data) < TOL[prec]


def test_full_size_hubbard_hac(ctx):
    """D = 1024 Hubbard U(1)xU(1), d = 4: every precision against the oracle at full size."""
    import synth_full
    st = synth_full.hubbard_1024(seed=1003)
    ref = orc.hac(st["EHL"], st["EHR"], st["HL"], st["HR"], st["AC"]).data
    for prec in PRECS:
        p = im.HACPlan(ctx, lay(st["AC"]), lay(st["EHL"]), lay(st["EHR"]), lay(st["HL"]), lay(st["HR"]), prec)
        p.set_operator(dev(st["EHL"].data), dev(st["EHR"].data), dev(st["HL"].data), dev(st["HR"].data))
        y = torch.full((ref.size,), float("nan"), dtype=torch.float64, device="cuda")
        p.apply(dev(st["AC"].data), y)
        assert relerr(y.cpu().numpy(), ref) < TOL[prec], (prec, relerr(y.cpu().numpy(), ref))
        # batched: 16 vectors (v-th = (v + 1) * AC; H_AC is linear), as in the benchmark
        X = np.concatenate([(v + 1.0) * st["AC"].data for v in range(16)])
        y16 = torch.full((X.size,), float("nan"), dtype=torch.float64, device="cuda")
        p.apply(dev(X), y16, 16)
        got = y16.cpu().numpy().reshape(16, -1)
        for v in range(16):
            assert relerr(got[v], (v + 1.0) * ref) < TOL[prec], (prec, v)


@pytest.mark.parametrize("sched", [0, 1, 2, 3], ids=["ks1", "ks1-shared", "ks2", "ks4"])
def test_every_schedule_every_precision(ctx, sched, monkeypatch):
    """The autotuner may pick any panel schedule for any precision: each one (forced through
    IMPS_SCHEDULE) against the oracle, full-size D = 1024 Hubbard H_AC at 1, 3 and 16 vectors and the
    Heisenberg TM (both directions) at 5 vectors."""
    import synth_full
    monkeypatch.setenv("IMPS_SCHEDULE", str(sched))
    st = synth_full.hubbard_1024(seed=1003)
    ref = orc.hac(st["EHL"], st["EHR"], st["HL"], st["HR"], st["AC"]).data
    for prec in PRECS:
        p = im.HACPlan(ctx, lay(st["AC"]), lay(st["EHL"]), lay(st["EHR"]), lay(st["HL"]), lay(st["HR"]), prec)
        p.set_operator(dev(st["EHL"].data), dev(st["EHR"].data), dev(st["HL"].data), dev(st["HR"].data))
        assert p.schedule(16) == ({0: 1, 1: 1, 2: 2, 3: 4}[sched], sched == 1)
        for nv in (1, 3, 16):
            X = np.concatenate([(v + 1.0) * st["AC"].data for v in range(nv)])
            y = torch.full((X.size,), float("nan"), dtype=torch.float64, device="cuda")
            p.apply(dev(X), y, nv)
            got = y.cpu().numpy().reshape(nv, -1)
            for v in range(nv):
                assert relerr(got[v], (v + 1.0) * ref) < TOL[prec], (prec, nv, v)
    c = cell("xxz")
    rng = np.random.default_rng(31)
    for direction in (0, 1):
        xl, yl = (c.diag(0), c.diag(1)) if direction == 0 else (c.diag(1), c.diag(0))
        xs = [random_like(xl, rng) for _ in range(5)]
        refs = [orc.tm_mixed(direction, c.AL[0], None, x, yl).data for x in xs]
        for prec in PRECS:
            p = im.TMPlan(ctx, direction, lay(c.AL[0]), None, lay(xl), lay(yl), prec)
            p.set_mps(dev(c.AL[0].data))
            y = torch.full((yl.nelem * 5,), float("nan"), dtype=torch.float64, device="cuda")
            p.apply(dev(np.concatenate([x.data for x in xs])), y, 5)
            got = y.cpu().numpy().reshape(5, -1)
            for v in range(5):
                assert relerr(got[v], refs[v]) < TOL[prec], (direction, prec, v)


def test_tuned_schedules_cached(ctx, tmp_path, monkeypatch):
    """IMPS_TUNE_CACHE: the first plan times the four schedules and records its choice under the
    program's key; a second plan of the same program takes the recorded schedule (rewritten here to
    another one) without timing or appending, and still matches the oracle."""
    for v in ("IMPS_SCHEDULE", "IMPS_KSPLIT"):
        monkeypatch.delenv(v, raising=False)
    monkeypatch.setenv("IMPS_AUTOTUNE", "1")
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("IMPS_TUNE_CACHE", str(path))
    table = {0: (1, False), 1: (1, True), 2: (2, False), 3: (4, False)}
    c = cell("fhub")
    H = ham(c)
    rng = np.random.default_rng(19)
    HL = orc.reduced_op(H, c.d, c.AL[1], 0)
    HR = orc.reduced_op(H, c.d, c.AR[1], 1)
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(1), rng, sym=True)
    xs = [random_like(c.struct[0], rng) for _ in range(4)]
    X = np.concatenate([x.data for x in xs])
    refs = [orc.hac(EHL, EHR, HL, HR, x).data for x in xs]

    def run():
        p = im.HACPlan(ctx, lay(xs[0]), lay(EHL), lay(EHR), lay(HL), lay(HR), im.FP32)
        p.set_operator(dev(EHL.data), dev(EHR.data), dev(HL.data), dev(HR.data))
        y = torch.full((X.size,), float("nan"), dtype=torch.float64, device="cuda")
        p.apply(dev(X), y, 4)
        got = y.cpu().numpy().reshape(4, -1)
        for v in range(4):
            assert relerr(got[v], refs[v]) < TOL[im.FP32]
        return p
    p1 = run()
    words = path.read_text().split()
    assert len(words) == 2
    key, sc = words[0], int(words[1])
    assert p1.schedule(4) == table[sc]
    other = (sc + 1) % 4
    path.write_text(f"{key} {other}\n")
    p2 = run()
    assert p2.schedule(4) == table[other]
    assert path.read_text() == f"{key} {other}\n"
