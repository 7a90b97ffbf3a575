"""The excitation program's front end (imps_cpp_tools_amd.excitations_main, Excitations_main.cpp) on
the GPU against the oracle.

The D = 16 XXZ and Hubbard states of excitation_state.py are written to IMPS state files
(write_imps, the saveIMPS format) and run through the front end's steps: load, I2K check,
ApplyTransformation / shift, LRoverlaps and sign, E0 from both sides, null spaces, XDims,
HeffConstants, one batched eigensolve over the momentum grid.  References: the oracle's dense H_eff
of the same (transformed) state (tests/excitation_state.py).  Tolerances: E0 1e-12; energies 1e-9;
MeasureExcitations' per-site values 1e-10 against the oracle's ApplyOPeff on the same random X
(S^z in K = 2, whose per-site values sum to the excitation's S^z change, -1 in the charge convention)."""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
import excitation_state as es
from oracle import excitations as ex

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

_CTX = []


def ctx():
    if not _CTX:
        _CTX.append(im.Context(0))
    return _CTX[0]


def _lay(ob):
    return im.BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def save_state(path, s):
    from imps_cpp_tools_amd.imps_io import write_imps
    lam = [{q: np.linalg.svd(m, compute_uv=False) for q, (_, m) in c.items()} for c in s["C"]]
    write_imps(str(path), lam, [(_lay(c), c.data) for c in s["Cob"]], [(_lay(a), a.data) for a in s["ALob"]],
               [(_lay(a), a.data) for a in s["ARob"]])


def run(tmp_path, which, K, op="", shift=0, extra=()):
    from imps_cpp_tools_amd import excitations_main as em
    from imps_cpp_tools_amd import vumps_main as vm
    s = es.state() if which == "xxz" else es.hubbard_state()
    path = tmp_path / f"{which}.bin"
    save_state(path, s)
    model = ["--model=XXZ", "--Delta=-1.5", "--QN=[0]"] if which == "xxz" else ["--model=FHUB", "--U=4", "--QN=[2,0]"]
    args = model + [f"--state={path}", f"--K=[{','.join(str(k) for k in K)}]", "--pmin=0", "--pmax=0.5", "--np=3", "--nbands=2",
                    f"--op={op}", f"--shift={shift}"] + list(extra)
    print(f"excitations_main {which} K={K} op={op!r} shift={shift}", flush=True)
    return em.run(em.settings(vm.parse_lines(args)), ctx=ctx(), log=lambda *a: print(*a, flush=True)), s


@pytest.mark.parametrize("which,K,op,shift", [("xxz", (0,), "", 0), ("xxz", (2,), "", 0), ("xxz", (0,), "FS", 0), ("xxz", (-1,), "", 1),
                                              ("hub", (0, 0), "FS", 0)],
                         ids=["xxz-K0", "xxz-K2", "xxz-FS-K0", "xxz-shift1-K-1", "hub-FS-K00"])
def test_energies_match_oracle(tmp_path, which, K, op, shift):
    r, s = run(tmp_path, which, K, op, shift)
    assert abs(r["E0"] - s["E0"]) < 1e-12
    assert r["converged"]
    ref = es.transformed_state(which, op, K, shift)
    assert abs(r["OL"] - ref["OL"]) < 1e-10
    for t, p in enumerate(r["p"]):
        M = es.oracle_heff_matrix(K, p, ref, proj_at_unit_phase=True)
        w = np.sort(np.linalg.eigvalsh(0.5 * (M + M.conj().T)))[:2]
        print(f"p={p}: {r['dE'][t]} oracle {w}", flush=True)
        assert np.abs(r["dE"][t] - w).max() < 1e-9


def test_measure_excitations_match_oracle(tmp_path):
    r, s = run(tmp_path, "xxz", (2,), extra=["--obs=[sz]", "--test=1", "--p=0.3"])
    assert "dE" not in r  # test = 1 stops after the observables (Excitations_main.cpp:218)
    O = np.diag([0.5, -0.5])
    _, _, (tl, _) = ex.opeff_constants(s["AL"], s["AR"], s["L"], s["R"], O)
    O0 = O - float(np.real(tl)) / s["N"] * np.eye(2)
    ref_s = es.transformed_state("xxz", "", (2,), 0)  # AR[0] with the sign of the K = 2 overlap
    xd = ex.xdims(s["NL"], ref_s["AR"], (2,))
    n = sum(rr * c for dv in xd for *_, rr, c in dv)
    rng = np.random.default_rng(0)
    z = rng.standard_normal((1, n)) + 1j * rng.standard_normal((1, n))
    z = z[0] / np.linalg.norm(z)
    X = ex.from_vec_list(z, xd)
    OPLtot, OPRtot, _ = ex.opeff_constants(s["AL"], ref_s["AR"], s["L"], s["R"], O0)
    Y = ex.apply_opeff(X, np.exp(1j * 0.3 * np.pi * s["N"]), s["AL"], ref_s["AR"], s["NL"], None, None, O0, OPLtot, OPRtot)
    ref = np.array([np.vdot(ex.vectorize(x), ex.vectorize(y)) for x, y in zip(X, Y)])
    got = r["obs_sz"]
    print(f"sz per site {got} oracle {ref}", flush=True)
    assert np.abs(got - ref).max() < 1e-10
    assert abs(got.sum() + 1.0) < 1e-10  # a normalised K = 2 (2 S^z) excitation changes S^z by -1 in this charge convention


def _read_xvec(path):
    """The excitation-vector file (CBLOCKMATARRAYVEC of CBLOCKMATARRAY of CBLOCKMAT) as a list of
    per-vector lists of per-site {qin: (qout, complex matrix)}."""
    import struct
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0

    def line():
        nonlocal pos
        e = buf.index(b"\n", pos)
        out, pos = buf[pos:e], e + 1
        return out

    def u32():
        nonlocal pos
        v = struct.unpack_from("<I", buf, pos)[0]
        pos += 4
        return v

    def key():
        nonlocal pos
        n = u32()
        k = struct.unpack_from("<%di" % n, buf, pos)
        pos += 4 * n
        return tuple(k)
    assert line() == b"CBLOCKMATARRAYVEC"
    vecs = []
    for _ in range(u32()):
        assert line() == b"CBLOCKMATARRAY"
        sites = []
        for _ in range(u32()):
            assert line() == b"CBLOCKMAT"
            bm = {}
            for _ in range(u32()):
                qi, qo = key(), key()
                assert line() == b"ARMA_MAT_BIN_FC016"
                r, c = (int(v) for v in line().split())
                m = np.frombuffer(buf, "<c16", r * c, pos).reshape(r, c, order="F")
                pos += 16 * r * c
                bm[qi] = (qo, m)
            sites.append(bm)
        vecs.append(sites)
    assert pos == len(buf)
    return vecs


def test_saved_files(tmp_path):
    """saveE / saveX: the energies as an arma matrix, per momentum the excitation vectors (each an
    eigenvector of the oracle's H_eff with its energy), and the ALRN state file."""
    from imps_cpp_tools_amd.imps_io import _r_mat
    r, s = run(tmp_path, "xxz", (2,), extra=["--saveE=1", "--saveX=1", f"--datafolder={tmp_path}"])
    files = sorted(p for p in tmp_path.rglob("*_dE.bin"))
    assert len(files) == 1
    with open(files[0], "rb") as f:
        dE = _r_mat(f)
    assert np.array_equal(dE, r["dE"])
    assert len(list(tmp_path.rglob("*_ALRN.bin"))) == 1
    ref = es.transformed_state("xxz", "", (2,), 0)
    xd = ex.xdims(ref["NL"], ref["AR"], (2,))
    # X is stored in the basis of the run's own null spaces (the ALRN file's NL), a rotation per
    # sector of the oracle's: X_oracle[n] = (sum_s NL_oracle[n][s]^T NL_run[n][s]) X_run[n]
    from oracle import pyoracle as orc
    rot = [ex.tml_unity(ex.from_ob(orc.OB.like(lay, t.cpu().numpy())), ref["NL"][n]) for n, (lay, t) in enumerate(r["NL"])]
    for t, p in enumerate(r["p"]):
        (xf,) = tmp_path.rglob(f"*_X{t}.bin")
        vecs = _read_xvec(xf)
        assert len(vecs) == 2
        M = es.oracle_heff_matrix((2,), p, ref, proj_at_unit_phase=True)
        for j, sites in enumerate(vecs):
            assert np.array_equal(ex.vectorize_list(sites), r["X"][t, j].cpu().numpy())
            z = ex.vectorize_list([ex.bm_mul(rot[n], x) for n, x in enumerate(sites)])
            assert z.size == sum(rr * c for dv in xd for *_, rr, c in dv)
            assert abs(np.linalg.norm(z) - 1.0) < 1e-12
            res = np.linalg.norm(M @ z - r["dE"][t, j] * z)
            assert res < 1e-8
