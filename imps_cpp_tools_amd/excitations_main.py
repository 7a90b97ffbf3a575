"""Quasiparticle excitations from the reference's configuration files: the keys, defaults and steps
of Excitations_main.cpp, run on this build's GPU excitation engine.

    python -m imps_cpp_tools_amd.excitations_main exc.cfg --state=gs.bin --K=[0,0] --np=32

Format: that of vumps_main (parser.cpp:4-46).  Keys (Excitations_main.cpp:26-61): state (an IMPS
state file, loadIMPS), QN, K (excitation charge), pmin, pmax, np (momenta linspace(pmin, pmax, np) in
units of pi per site), tol (1e-10), InvETol (1e-14, raised to tol / 10), nbands (4), shift
(relative shift of AR, 0), glob_shift (0), op (ApplyTransformation: FS for XXZ, FS / FC for FHUB),
obs (one-site operators for MeasureExcitations at momentum p, 0.5), saveE, datafolder (.), test,
verbose, and the model keys of vumps_main.  Extra keys of this runner: prec (fp64 | fp32: the
geometric sums' transfer matrices in FP32 under FP64 refinement), maxit (bound on batched H_eff
applications of the eigensolver, 0 = the engine's default), krylov (0 = max(20, 2 nbands + 1)).

Steps (Excitations_main.cpp:80-297): load the state, check it against the model's I2K, shift it
globally, L = C^T C and R = C C^T on the last bond, ApplyTransformation and the relative shift of AR
and L, LRoverlaps in K with AR[0] times the overlap's sign, E0 from AL with R and from AR with L
(they must agree to tol), H - E0, the null spaces of AL, XDims, HeffConstants, then all momenta at
once: one batched eigensolve of H_eff(p) (the reference loops over the momenta with one eigs_n
each).  saveE / saveX write, under datafolder/<run tag>/, the momenta and the state actually
excited (saveALRN: C, L, R, AL, AR, NL with the shifts and transformation), then the energies
(np x nbands) or, per momentum, the nbands excitation vectors (Excitations_main.cpp:221-239, 285-309).
"""
from __future__ import annotations

import os
import sys
import time
from typing import Dict, List, Optional

import numpy as np

from .vumps_main import _bool, _get, _vector, model_id, model_settings, parse


def settings(values: Dict[str, str]) -> Dict:
    """The run settings with the reference's defaults (Excitations_main.cpp:28-64)."""
    model, dim, params = model_settings(values)
    tol = _get(values, "tol", float, 1e-10)
    cfg = dict(
        model=model, dim=dim, params=params, state=_get(values, "state", str, required=True),
        QN=_get(values, "QN", lambda s: _vector(s, int), required=True), K=_get(values, "K", lambda s: _vector(s, int), required=True),
        pmin=_get(values, "pmin", float, required=True), pmax=_get(values, "pmax", float, required=True),
        np=_get(values, "np", int, required=True), tol=tol, InvETol=max(tol / 10.0, _get(values, "InvETol", float, 1e-14)),
        nbands=_get(values, "nbands", int, 4), shift=_get(values, "shift", int, 0), glob_shift=_get(values, "glob_shift", int, 0),
        op=_get(values, "op", str, ""), obs=_get(values, "obs", lambda s: _vector(s, str.strip), []), p=_get(values, "p", float, 0.5),
        saveE=_get(values, "saveE", _bool, False), saveX=_get(values, "saveX", _bool, False),
        datafolder=_get(values, "datafolder", str, "."), test=_get(values, "test", _bool, False),
        verbose=_get(values, "verbose", _bool, False), prec=_get(values, "prec", str, "fp64"), maxit=_get(values, "maxit", int, 0),
        krylov=_get(values, "krylov", int, 0))
    if cfg["np"] < 1:
        raise ValueError("np must be at least 1")
    if cfg["prec"] not in ("fp64", "fp32"):
        raise ValueError("prec: fp64 or fp32")
    return cfg


def _shift(arr: List, s: int) -> List:
    """shift(arr, s) (MPSBlockUtilities.hpp:1153-1168): site n takes site n - s."""
    N = len(arr)
    return [arr[(n - s) % N] for n in range(N)]


def check_i2k(model: int, N: int, QN, lays) -> None:
    """The loaded tensors against the model's I2K (Excitations_main.cpp:100-101): every block of
    physical index s changes the charge by I2K[s] (modulo the charge periods)."""
    import imps_cpp_tools_amd as im
    keys = im.i2k(model, N, QN).astype(np.int64)
    for lay in lays:
        if lay.nlist != keys.shape[0]:
            raise ValueError("state not compatible with I2K from model (physical dimension)")
        per = np.asarray(lay.periods if lay.periods is not None else np.zeros(lay.nsym), np.int64)
        b = 0
        for s, nb in enumerate(lay.nblk):
            for _ in range(int(nb)):
                dk = lay.qout[b].astype(np.int64) - lay.qin[b].astype(np.int64) - keys[s]
                if np.any(np.where(per > 1, dk % np.maximum(per, 1), dk) != 0):
                    raise ValueError("state not compatible with I2K from model")
                b += 1


def bond_energies(ops, H: np.ndarray, A, Alay, env, side: str) -> np.ndarray:
    """MeasureObservables of a two-site H (MPSBlockUtilities.hpp:1017-1073) on one canonical side:
    side "L": AL with the right fixed points env[n] = C_n C_n^T (bond right of site n), values
    tr(OpTM_left(AL[n] AL[n+1]) R_{n+1}); side "R": AR with the left fixed point env (bond N - 1),
    propagated through AR's transfer matrix, values tr(L_{n-1} OpTM_right(AR[n] AR[n+1])).  The
    bond energies of the cell, computed with the library's block products."""
    from .excitations import DObj
    N = len(A)
    As = [DObj(l, t) for l, t in zip(Alay, A)]
    out = np.zeros(N)
    if side == "L":
        for n in range(N):
            AA = ops.mps_mps(As[n], As[(n + 1) % N])
            X = ops.diag_part(ops.tml_unity(ops.apply_operator(AA, H), AA))
            out[n] = ops._tr(X, env[(n + 1) % N]).real
    else:
        Lb = env
        for n in range(N):
            AA = ops.mps_mps(As[n], As[(n + 1) % N])
            Y = ops.diag_part(ops.tmr_unity(ops.apply_operator(AA, H), AA))
            out[n] = ops._tr(Lb, Y).real
            Lb = ops.diag_part(ops.tml(As[n], As[n], Lb))
    return out


def site_values(ops, O: np.ndarray, A, Alay, env, side: str) -> np.ndarray:
    """MeasureObservables of a one-site operator on one canonical side (as bond_energies)."""
    from .excitations import DObj
    N = len(A)
    As = [DObj(l, t) for l, t in zip(Alay, A)]
    out = np.zeros(N)
    Lb = env
    for n in range(N):
        if side == "L":
            X = ops.diag_part(ops.tml_unity(ops.apply_operator(As[n], O), As[n]))
            out[n] = ops._tr(X, env[n]).real
        else:
            Y = ops.diag_part(ops.tmr_unity(ops.apply_operator(As[n], O), As[n]))
            out[n] = ops._tr(Lb, Y).real
            Lb = ops.diag_part(ops.tml(As[n], As[n], Lb))
    return out


def _save_arma(path: str, m: np.ndarray) -> None:
    from .imps_io import _w_mat
    with open(path, "wb") as f:
        _w_mat(f, m)


def run(cfg: Dict, ctx=None, log=print):
    """Excitations_main on the GPU; returns dict(p, dE (np x nbands), residuals, E0, OL, ...)."""
    import torch

    import imps_cpp_tools_amd as im
    from .excitations import BlockOps, DObj
    from .imps_io import read_imps
    from .nullspace import left_nullspace
    from .transform import transformed_state
    from .vumps import VUMPS

    model = model_id(cfg["model"], cfg.get("dim"))
    nsym = im.model_nsym(model)
    periods = im.model_periods(model)
    ctx = im.Context(0) if ctx is None else ctx
    K = tuple(int(k) for k in cfg["K"])
    if len(K) != nsym:
        raise ValueError("K has wrong number of quantum numbers")
    H = np.asarray(im.model_ham(model, cfg["params"]), np.float64)
    d = im.model_d(model)
    _, C, AL, AR = read_imps(cfg["state"], nsym, periods)
    N = len(AL)
    check_i2k(model, N, cfg["QN"], [l for l, _ in AL])
    check_i2k(model, N, cfg["QN"], [l for l, _ in AR])
    g = cfg["glob_shift"] % N
    AL, AR, C = _shift(AL, g), _shift(AR, g), _shift(C, g)
    dev = torch.device(f"cuda:{ctx.device}")

    def up(x):
        return torch.from_numpy(np.ascontiguousarray(x, np.float64)).to(dev)
    ALl, ALd = [l for l, _ in AL], [up(x) for _, x in AL]
    ARl, ARd = [l for l, _ in AR], [up(x) for _, x in AR]
    Cl, Cd = [l for l, _ in C], [up(x) for _, x in C]
    t0 = time.perf_counter()
    st = transformed_state(ctx, model, N, cfg["QN"], ALd, ALl, ARd, ARl, Cd, Cl, K, op=cfg["op"], rel_shift=cfg["shift"],
                           nev=cfg["nbands"], tol=cfg["InvETol"])
    log(f"mixed TM AL and AR: overlap {st['OL']:.12g}" + (" (AR[0] sign flipped)" if st["OL"] < 0 else ""))
    # ground-state energy from both sides (Excitations_main.cpp:182-191)
    ops = BlockOps(ctx, nsym, d)
    Rs = []
    for n in range(N):
        P = im.ProdPlan(ctx, im.MPS_MATT, Cl[n], Cl[n])
        y = torch.zeros(P.out_layout.nelem, dtype=torch.float64, device=dev)
        P.apply(Cd[n], Cd[n], y)
        Rs.append(DObj(P.out_layout, y))
    Lobj = DObj(*st["L"])
    e_l = bond_energies(ops, H, ALd, ALl, Rs, "L")
    e_r = bond_energies(ops, H, st["AR"], st["ARl"], Lobj, "R")
    E0L, E0R = float(e_l.mean()), float(e_r.mean())
    log(f"E0 from AL: {E0L:.14f}, from AR: {E0R:.14f}")
    if abs(E0L - E0R) > cfg["tol"]:
        raise ValueError(f"ground state energies from AL and AR differ by {E0L - E0R}")
    E0 = 0.5 * (E0L + E0R)
    nls = [left_nullspace(ctx, ALd[n], ALl[n], seed=n) for n in range(N)]
    NLd, NLl = [a for a, _ in nls], [l for _, l in nls]
    prec = im.FP64 if cfg["prec"] == "fp64" else im.FP32
    pvec = np.linspace(cfg["pmin"], cfg["pmax"], cfg["np"])
    out = dict(p=pvec, E0=E0, E0L=E0L, E0R=E0R, OL=st["OL"])
    if cfg["obs"]:
        # MeasureExcitations (Excitations_helpers.hpp:774-826) at momentum p on a random excitation
        table = VUMPS.model_operators(model)
        proj_p = cfg["p"]
        for name in cfg["obs"]:
            if name not in table:
                raise ValueError(f"operator {name} not defined (operators of this model: {', '.join(table)})")
            O = table[name]
            eobs = 0.5 * (site_values(ops, O, ALd, ALl, Rs, "L").mean() + site_values(ops, O, st["AR"], st["ARl"], Lobj, "R").mean())
            Pm = im.HeffPlan(ctx, ALl, st["ARl"], NLl, Cl, None, eobs, K, 1, observable=O)
            Pm.set_state(ALd, st["AR"], NLd, Cd, inv_tol=cfg["InvETol"], L=st["L"], LM=st["LM"], RM=st["RM"])
            rng = np.random.default_rng(0)
            z = rng.standard_normal((1, Pm.mtot)) + 1j * rng.standard_normal((1, Pm.mtot))
            x = torch.from_numpy(z / np.linalg.norm(z)).to(dev)
            vals = Pm.measure([proj_p], x)[0]
            log(f"{name} (subtracting {eobs:.12g}): " + " ".join(f"{v.real:.12g}" for v in vals) + f"  total: {vals.sum().real:.12g}")
            out[f"obs_{name}"] = vals
            del Pm
    if cfg["test"]:
        return out
    P = im.HeffPlan(ctx, ALl, st["ARl"], NLl, Cl, H, E0, K, cfg["np"], prec=prec)
    P.set_state(ALd, st["AR"], NLd, Cd, inv_tol=cfg["InvETol"], L=st["L"], LM=st["LM"], RM=st["RM"])
    trl, trr, _ = P.check()
    if cfg["verbose"]:
        log(f"HeffConstants check: tr(HL R) = {trl:.3e}, tr(L HR) = {trr:.3e}; mtot = {P.mtot}")
    t1 = time.perf_counter()
    ev, vecs, res, napp = P.eigs(pvec, nev=cfg["nbands"], tol=cfg["tol"], maxit=cfg["maxit"], krylov=cfg["krylov"],
                                 vectors=cfg["saveX"], check=False)
    torch.cuda.synchronize(dev)
    t_eig = time.perf_counter() - t1
    for t, p in enumerate(pvec):
        log(f"p = {p:g} {N} pi: \t" + "\t".join(f"{e:.12g}" for e in ev[t]))
    rel = res / np.maximum(1.0, np.abs(ev))
    conv = bool(rel.max() <= cfg["tol"])
    log(f"{napp} batched H_eff applications, worst relative residual {rel.max():.2e}" + ("" if conv else " (not converged)"))
    tel = time.perf_counter() - t0
    log(f"elapsed time: {tel / 60:.3f} min." if tel > 60 else f"elapsed time: {tel:.3f} s.")
    out.update(dE=ev, residuals=res, applies=napp, converged=conv, seconds_eigs=t_eig, seconds=tel, mtot=P.mtot)
    if cfg["saveX"]:
        out["X"] = vecs  # in the basis of the null spaces NL (saved in the ALRN file)
        out["NL"] = list(zip(NLl, NLd))
    if cfg["saveE"] or cfg["saveX"]:
        from .imps_io import write_alrn, write_excitations
        tag = f"{cfg['model']}_QN{''.join(str(q) for q in cfg['QN'])}_N{N}_p{cfg['pmin']:g}_{cfg['pmax']:g}_{cfg['np']}_nb{cfg['nbands']}"
        tag += f"_shift{cfg['shift']}" + (f"_op{cfg['op']}" if cfg["op"] else "") + f"_K{''.join(str(k) for k in K)}"
        folder = os.path.join(cfg["datafolder"], tag)
        os.makedirs(folder, exist_ok=True)
        _save_arma(os.path.join(folder, tag + "_pvec.bin"), pvec.reshape(-1, 1))
        host = lambda lt: (lt[0], lt[1].cpu().numpy())  # noqa: E731
        write_alrn(os.path.join(folder, tag + "_ALRN.bin"), [host(x) for x in zip(Cl, Cd)], [host(x) for x in st["Lvec"]],
                   [host((r.lay, r.data)) for r in Rs], [host(x) for x in zip(ALl, ALd)], [host(x) for x in zip(st["ARl"], st["AR"])],
                   [host(x) for x in zip(NLl, NLd)])
        if cfg["saveE"]:
            _save_arma(os.path.join(folder, tag + "_dE.bin"), ev)
        if cfg["saveX"]:
            xv = vecs.cpu().numpy()
            for t in range(cfg["np"]):
                write_excitations(os.path.join(folder, tag + f"_X{t}.bin"), xv[t], P.xdims)
        log(f"saved pvec, ALRN{', dE' if cfg['saveE'] else ''}{', X' if cfg['saveX'] else ''} under {folder}")
    return out


def main(argv: Optional[List[str]] = None) -> int:
    cfg = settings(parse(sys.argv[1:] if argv is None else argv))
    r = run(cfg)
    return 0 if r.get("converged", True) else 1


if __name__ == "__main__":
    sys.exit(main())
