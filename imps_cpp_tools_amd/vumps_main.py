"""VUMPS from the reference's configuration files: the `--key=value` format of parser.cpp and the
keys and defaults of VUMPS_main.cpp:25-100, run on this build's GPU VUMPS.

    python -m imps_cpp_tools_amd.vumps_main my.cfg --U=4 --mmax=[256] --save=1

Format (parser.cpp:4-46): every line or argument of the form `--key=value` sets key (anything else,
comments included, is ignored); files are read in order, then the command-line arguments, later
settings replacing earlier ones.  Vectors are written `[a,b,c]`.

Keys (VUMPS_main.cpp:51-69, Models.cpp:881-1114): model (XXZ | QIM | QPOTTS | BHUB | FHUB), N (>= 2),
QN, m0 (initial dims per sector, GetDimMap), mmax (one value or one per bond), thresh (stop when max F
is below), dmmax, exthresh, lamthresh, expand_pause, InvETol, InvEfac, save, folder, obs (one-site
operators of the model's GetObservables); XXZ: J (1), Delta, hz (0); QIM: J (1), hz; QPOTTS: q, h
(0); BHUB: d, t (1), U, mu (0); FHUB: t (1), U, V (0), mu (0), hz (0).  Extra keys of this runner:
maxit (outer iterations, 1000), prec (fp64 | fp32 | bf16x3), seed.
"""
from __future__ import annotations

import os
import sys
from typing import Dict, Iterable, List, Optional

MODELS = ("XXZ", "FHUB", "QIM", "QPOTTS", "BHUB")
PARAM_NAMES = {"XXZ": ("J", "Delta", "hz"), "FHUB": ("t", "U", "V", "mu", "hz"), "QIM": ("J", "h"), "QPOTTS": ("h",),
               "BHUB": ("t", "U", "mu")}


def model_id(name: str, dim: Optional[int] = None) -> int:
    """The library's model id of a CreateModel name (the local dimension q or d for QPOTTS / BHUB)."""
    import imps_cpp_tools_amd as im
    ids = {"XXZ": im.XXZ, "FHUB": im.FHUB, "QIM": im.QIM}
    if name in ids:
        return ids[name]
    if dim is None:
        raise ValueError(f"{name} needs its local dimension")
    return im.QPOTTS(dim) if name == "QPOTTS" else im.BHUB(dim)


def parse_lines(lines: Iterable[str], values: Optional[Dict[str, str]] = None) -> Dict[str, str]:
    """parser::parse_argument on each line: `--key=value` with a non-empty key; others ignored."""
    values = {} if values is None else values
    for line in lines:
        line = line.rstrip("\r\n")
        e = line.find("=", 2)
        if line[:2] == "--" and e > 2:
            values[line[2:e]] = line[e + 1:]
    return values


def parse(argv: List[str]) -> Dict[str, str]:
    """Files (arguments ending in .cfg) in order, then the other arguments."""
    values: Dict[str, str] = {}
    files = [a for a in argv if a.endswith(".cfg")]
    for f in files:
        if not os.path.isfile(f):
            raise FileNotFoundError(f"file {f} not found")
        with open(f) as fh:
            parse_lines(fh, values)
    parse_lines([a for a in argv if not a.endswith(".cfg")], values)
    return values


def _vector(s: str, conv=float) -> List:
    s = s.strip()
    if not (s.startswith("[") and s.endswith("]")):
        raise ValueError(f"expected a vector [a,b,...], got {s!r}")
    body = s[1:-1].strip()
    return [conv(x) for x in body.split(",")] if body else []


def _get(values, key, conv, default=None, required=False):
    if key not in values:
        if required:
            raise ValueError(f'argument "{key}" not found')
        return default
    return conv(values[key])


def _bool(s: str) -> bool:
    return s.strip().lower() in ("1", "true", "yes", "on")


def model_settings(values: Dict[str, str]):
    """(model name, local dimension or None, Hamiltonian parameters) of CreateModel
    (Models.cpp:881-1114) from the parsed keys."""
    model = _get(values, "model", str, required=True)
    if model not in MODELS:
        raise ValueError(f"{model} not found (models: {', '.join(MODELS)})")
    dim = None
    if model == "XXZ":
        params = [_get(values, "J", float, 1.0), _get(values, "Delta", float, required=True), _get(values, "hz", float, 0.0)]
        if _get(values, "d", int, 2) != 2:
            raise ValueError("XXZ: only spin-1/2 (d = 2) is built")
    elif model == "QIM":
        params = [_get(values, "J", float, 1.0), _get(values, "hz", float, required=True)]
    elif model == "QPOTTS":
        dim = _get(values, "q", int, required=True)
        params = [_get(values, "h", float, 0.0)]
    elif model == "BHUB":
        dim = _get(values, "d", int, required=True)
        params = [_get(values, "t", float, 1.0), _get(values, "U", float, required=True), _get(values, "mu", float, 0.0)]
    else:
        params = [_get(values, "t", float, 1.0), _get(values, "U", float, required=True), _get(values, "V", float, 0.0),
                  _get(values, "mu", float, 0.0), _get(values, "hz", float, 0.0)]
    return model, dim, params


def settings(values: Dict[str, str]) -> Dict:
    """The run settings with the reference's defaults (VUMPS_main.cpp:30-46)."""
    model, dim, params = model_settings(values)
    N = _get(values, "N", int, required=True)
    if N < 2:
        raise ValueError("unit cell size must be at least 2")
    mmax = _get(values, "mmax", lambda s: _vector(s, int), required=True)
    if len(mmax) == 1:
        mmax = mmax * N
    elif len(mmax) != N:
        raise ValueError("mmax needs one value or one per bond")
    return dict(
        model=model, dim=dim, N=N, params=params,
        QN=_get(values, "QN", lambda s: _vector(s, int), required=model not in ("QIM", "QPOTTS"), default=[]),
        m0=_get(values, "m0", lambda s: _vector(s, int), required=True), mmax=mmax,
        thresh=_get(values, "thresh", float, 1e-14), dmmax=_get(values, "dmmax", int, 20),
        exthresh=_get(values, "exthresh", float, 1e-5), lamthresh=_get(values, "lamthresh", float, 1e-12),
        expand_pause=_get(values, "expand_pause", int, 5), InvETol=_get(values, "InvETol", float, 5e-15),
        InvEfac=_get(values, "InvEfac", float, 2.0), save=_get(values, "save", _bool, False),
        folder=_get(values, "folder", str, "data"), maxit=_get(values, "maxit", int, 1000),
        prec=_get(values, "prec", str, "fp64"), seed=_get(values, "seed", int, 0),
        obs=_get(values, "obs", lambda s: _vector(s, str.strip), []))


def run(cfg: Dict, ctx=None, log=print):
    """VUMPS_main's outer loop on the GPU: iterate until max F < thresh (or maxit iterations) with
    bond growth up to mmax; returns the VUMPS object and the list of per-iteration stats."""
    import imps_cpp_tools_amd as im
    from .vumps import VUMPS
    model = model_id(cfg["model"], cfg.get("dim"))
    keys0, sizes0 = im.dimmap(model, cfg["N"], cfg["QN"], cfg["m0"])
    prec = {"fp64": im.FP64, "fp32": im.FP32, "bf16x3": im.BF16X3}[cfg["prec"]]
    ctx = im.Context(0) if ctx is None else ctx
    v = VUMPS(ctx, model, cfg["N"], cfg["QN"], keys0, sizes0, cfg["params"], prec=prec, seed=cfg["seed"], env_tol=cfg["InvETol"],
              env_fac=cfg["InvEfac"], mmax=cfg["mmax"], dm_max=cfg["dmmax"], expand_pause=cfg["expand_pause"],
              expand_thresh=cfg["exthresh"], lamthresh=cfg["lamthresh"], env_krylov=300, krylov=30)
    hist = []
    for it in range(cfg["maxit"]):
        st = v.sweep()
        hist.append(st)
        log(f"{it + 1}: E = {st.energy:.14f}  max F = {max(st.F):.3e}  D = {v.bond_dims()}  {st.seconds:.3f} s")
        if max(st.F) < cfg["thresh"]:
            break
    if cfg["obs"]:
        for name, vals in v.observables(cfg["obs"]).items():
            log(f"{name}: {' '.join(f'{x:.12f}' for x in vals)}  mean: {vals.mean():.12f}")
    if cfg["save"]:
        os.makedirs(cfg["folder"], exist_ok=True)
        dimtag = [f"{'q' if cfg['model'] == 'QPOTTS' else 'd'}{cfg['dim']}"] if cfg.get("dim") else []
        name = "_".join([f"UMPS_{cfg['model']}"] + dimtag + [f"{k}{p:g}" for k, p in zip(PARAM_NAMES[cfg["model"]], cfg["params"])]
                        + [f"N{cfg['N']}", f"chi{max(v.bond_dims())}"]) + ".bin"
        path = os.path.join(cfg["folder"], name)
        v.save(path)
        log(f"saved UMPS to {path}")
    return v, hist


def main(argv: Optional[List[str]] = None) -> int:
    cfg = settings(parse(sys.argv[1:] if argv is None else argv))
    run(cfg)
    return 0


if __name__ == "__main__":
    sys.exit(main())
