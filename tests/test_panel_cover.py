"""Host check of the panel program builder on the Fermi-Hubbard H_AC structure (CPU only): every
task's chunks cover each of its block products' k exactly once in total, products with K up to one
lane group's k-run are packed at most four to a chunk, and packed chunks keep to the task's rows
(tests/panel_cover.cpp, compiled here with g++ against the builder's headers).  The structure:
U(1) x U(1) Hubbard bonds of 49 sectors of sizes 1..40 (products with K from 1 to 40, so single-run,
multi-run and multi-chunk products all occur), HL / HR block structures from the oracle's
ReducedOp of a seeded tensor."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write(f, ob):
    nb = np.asarray(ob.nblk)
    qin = np.asarray(ob.qin).reshape(-1, ob.nsym)
    qout = np.asarray(ob.qout).reshape(-1, ob.nsym)
    f.write(f"{ob.nsym} {len(nb)}\n" + " ".join(map(str, nb)) + "\n")
    for a, b, r, c in zip(qin, qout, np.asarray(ob.rows), np.asarray(ob.cols)):
        f.write(" ".join(map(str, a)) + " " + " ".join(map(str, b)) + f" {r} {c}\n")


def test_panel_chunks_cover_every_product(tmp_path):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    import imps_cpp_tools_amd as im
    from imps_cpp_tools_amd import workloads as wl
    from oracle import pyoracle as orc
    from states import random_like
    k0, s0 = wl.fhub_profile(4, 40, 1)
    bonds = im.make_dims(im.FHUB, 2, [2, 0], k0, s0)
    AC = im.mps_structure(im.FHUB, 2, [2, 0], bonds[0], bonds[1])
    A10 = orc.mps_structure(im.FHUB, 2, [2, 0], bonds[1], bonds[0])
    A = random_like(A10, np.random.default_rng(7))
    H = orc.model_ham(im.FHUB, [1.0, 4.0, 0, 0, 0])
    HL, HR = orc.reduced_op(H, 4, A, 0), orc.reduced_op(H, 4, A, 1)
    lay = tmp_path / "hac.txt"
    with open(lay, "w") as f:
        for ob in (AC, wl.diag_layout(2, bonds[0]), wl.diag_layout(2, bonds[1]), HL, HR):
            _write(f, ob)
    exe = str(tmp_path / "panel_cover")
    src = os.path.join(ROOT, "tests", "panel_cover.cpp")
    res = subprocess.run([gxx, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "imps_cpp_tools_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                          src, "-o", exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    run = subprocess.run([exe, str(lay)], capture_output=True, text=True, timeout=120)
    print(run.stdout)
    assert run.returncode == 0 and run.stdout.strip().endswith("ok"), run.stdout + run.stderr
