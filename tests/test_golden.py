"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must reproduce them (CPU), and so must the HIP path (GPU, <= 1e-12
relative L2, through the C ABI)."""
import glob
import os

import numpy as np
import pytest

from conftest import ROOT, rel_l2

FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz")))


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _ids():
    return [os.path.basename(p)[:-4] for p in FIXTURES]


def test_fixtures_present():
    assert len(FIXTURES) >= 3


@pytest.mark.parametrize("path", FIXTURES, ids=_ids())
def test_oracle_reproduces_fixture(oracle_mod, path):
    d = _load(path)
    shape, left, right = tuple(int(s) for s in d["shape"]), tuple(d["left"]), tuple(d["right"])
    nc = d["psi0"].shape[3]
    kind = {"lhy": oracle_mod.OR_LHY, "gp": oracle_mod.OR_GP}[str(d["kind"])]
    F = oracle_mod.Functional(kind=kind, couplings=d.get("couplings"), V=d.get("V"),
                              masses=tuple(d["masses"]) if "masses" in d else (1.0, 1.0))
    renorm = list(d["renorm"]) if d["renorm"].size else None
    out = oracle_mod.advance(F, d["psi0"], oracle_mod.Geometry(shape, left, right, nc), stepper=str(d["stepper"]),
                             dt=float(d["dt"]), imaginary=bool(d["imaginary"]), nsteps=int(d["nsteps"]),
                             renormalize=renorm)
    assert rel_l2(out, d["out"]) < 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=_ids())
def test_gpu_reproduces_fixture(path):
    from gp5_amd import capi

    d = _load(path)
    print("golden", os.path.basename(path), flush=True)
    shape, left, right = tuple(int(s) for s in d["shape"]), tuple(d["left"]), tuple(d["right"])
    nc = d["psi0"].shape[3]
    ctx = capi.Context(shape, left, right, nc)
    vb = ctx.real(d["V"]) if "V" in d else None
    kind = {"lhy": capi.GP5_FUNC_LHY, "gp": capi.GP5_FUNC_GP}[str(d["kind"])]
    f = capi.functional_desc(kind, couplings=d.get("couplings"), potential=vb,
                             masses=tuple(d["masses"]) if "masses" in d else (1.0, 1.0))
    renorm = list(d["renorm"]) if d["renorm"].size else None
    st = capi.GP5_STEP_RK4 if str(d["stepper"]) == "RK4" else capi.GP5_STEP_EULERO
    s = capi.stepper_desc(st, dt=float(d["dt"]), imaginary=bool(d["imaginary"]), renormalize=renorm)
    psi = ctx.field(d["psi0"])
    capi.advance(ctx, f, s, psi, int(d["nsteps"]))
    assert rel_l2(psi.download_complex(), d["out"]) <= 1e-12
