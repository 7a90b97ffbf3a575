"""Full-size synthetic states built with the oracle (isometric AL/AR by per-sector QR, reduced
operators by the oracle's ReducedOp) on the product's sector profiles."""
import numpy as np

from imps_cpp_tools_amd.workloads import HUBBARD_1024, fhub_profile
from oracle import pyoracle as orc
from states import FHUB, Cell, random_like


def hubbard_1024(seed=1003):
    cfg = HUBBARD_1024
    k0, s0 = fhub_profile(cfg["nside"], cfg["peak"], 1, cfg["total"])
    c = Cell(FHUB, cfg["N"], cfg["QN"], k0, s0, seed)
    H = orc.model_ham(FHUB, cfg["params"])
    rng = np.random.default_rng(seed + 1)
    return dict(cell=c, H=H,
                HL=orc.reduced_op(H, 4, c.AL[1], 0), HR=orc.reduced_op(H, 4, c.AR[1], 1),
                EHL=random_like(c.diag(0), rng, sym=True), EHR=random_like(c.diag(1), rng, sym=True),
                AC=random_like(c.struct[0], rng))
