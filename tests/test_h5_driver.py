"""HDF5 field files (src/io.cpp:32-254) and the YAML driver (src/main.cpp:14-249).

The file layout is pinned with HDF5's own command-line tools, not with this
build's reader: h5dump prints single elements and the element type of a file
the build wrote, and h5import writes a plain-double file the build must read.
The driver's parsing and potentials run on the CPU; its end-to-end runs go
through the HIP path and are compared with the oracle."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import rel_l2, seeded_field

capi = pytest.importorskip("gp5_amd.capi")
driver = pytest.importorskip("gp5_amd.driver")


def _tool(name):
    return shutil.which(name) or (f"/opt/conda/bin/{name}" if os.path.exists(f"/opt/conda/bin/{name}") else None)


needs_h5 = pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="libgp5.so not built")


def _rand(shape, nc, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal(tuple(shape) + (nc,)) + 1j * rng.standard_normal(tuple(shape) + (nc,)))


# ---------------------------------------------------------------- files ----------

@needs_h5
@pytest.mark.parametrize("shape,nc", [((5, 6, 7), 1), ((4, 3, 2), 2), ((1, 1, 1), 1), ((16, 8, 4), 2)])
def test_h5_round_trip_bit_exact(tmp_path, shape, nc):
    a = _rand(shape, nc, seed=sum(shape) + nc)
    fn = str(tmp_path / "f.hdf5")
    capi.h5_write(fn, a)
    assert capi.h5_shape(fn) == (tuple(shape), nc)
    b = capi.h5_read(fn)
    assert b.shape == a.shape and b.flags.f_contiguous
    assert np.array_equal(a, b)
    assert np.array_equal(capi.h5_read(fn, shape, nc), a)


@needs_h5
def test_h5_three_d_array_is_one_component(tmp_path):
    a = _rand((3, 4, 5), 1, 7)[..., 0]
    fn = str(tmp_path / "f.h5")
    capi.h5_write(fn, a)
    assert np.array_equal(capi.h5_read(fn)[..., 0], a)


@needs_h5
@pytest.mark.skipif(_tool("h5dump") is None, reason="h5dump not installed")
def test_h5_layout_pinned_by_h5dump(tmp_path):
    """The reference's layout: dataset "field", H5T_ARRAY double[2] elements, dims
    (nC, Nz, Ny, Nx); the element at file index (c, k, j, i) is psi(i, j, k, c)."""
    shape, nc = (5, 6, 7), 2
    a = _rand(shape, nc, 11)
    fn = str(tmp_path / "f.hdf5")
    capi.h5_write(fn, a)
    h5dump = _tool("h5dump")
    head = subprocess.run([h5dump, "-H", fn], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "field"' in head
    assert re.search(r"H5T_ARRAY \{ \[2\] H5T_IEEE_F64LE \}", head)
    assert "( 2, 7, 6, 5 )" in head
    for (i, j, k, c) in [(0, 0, 0, 0), (4, 3, 2, 1), (1, 5, 6, 0), (2, 0, 3, 1)]:
        out = subprocess.run([h5dump, "-m", "%.17g", "-d", "field", "-s", f"{c},{k},{j},{i}", "-c", "1,1,1,1", fn],
                             capture_output=True, text=True, check=True).stdout
        m = re.search(r"\(\d+,\d+,\d+,\d+\): \[ ([-0-9.eE+]+), ([-0-9.eE+]+) \]", out)
        assert m, out
        assert float(m.group(1)) == a[i, j, k, c].real
        assert float(m.group(2)) == a[i, j, k, c].imag


@needs_h5
@pytest.mark.skipif(_tool("h5import") is None, reason="h5import not installed")
def test_h5_reads_plain_double_pairs_written_by_h5import(tmp_path):
    """Files with a (nC, Nz, Ny, Nx, 2) double dataset (the pair layout of
    python/gp/old/field.py:5-14) read as complex fields."""
    nc, nz, ny, nx = 2, 2, 3, 4
    vals = np.arange(nc * nz * ny * nx * 2, dtype=float) * 0.5 - 3.0
    txt = tmp_path / "in.txt"
    np.savetxt(txt, vals[None, :], fmt="%.17g")
    cfg = tmp_path / "cfg.txt"
    cfg.write_text("PATH field\nINPUT-CLASS TEXTFP\nRANK 5\n"
                   f"DIMENSION-SIZES {nc} {nz} {ny} {nx} 2\n"
                   "OUTPUT-CLASS FP\nOUTPUT-SIZE 64\nOUTPUT-ARCHITECTURE IEEE\nOUTPUT-BYTE-ORDER LE\n")
    fn = str(tmp_path / "imp.h5")
    subprocess.run([_tool("h5import"), str(txt), "-c", str(cfg), "-o", fn], check=True, capture_output=True)
    assert capi.h5_shape(fn) == ((nx, ny, nz), nc)
    a = capi.h5_read(fn)
    for (i, j, k, c) in [(0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (3, 2, 1, 1)]:
        flat = (((c * nz + k) * ny + j) * nx + i) * 2
        assert a[i, j, k, c] == vals[flat] + 1j * vals[flat + 1]


@needs_h5
def test_h5_errors(tmp_path):
    fn = str(tmp_path / "f.hdf5")
    capi.h5_write(fn, _rand((4, 4, 4), 1, 0))
    with pytest.raises(capi.GP5Error, match=r"expected \(1,4,4,5\)"):
        capi.h5_read(fn, (5, 4, 4), 1)
    with pytest.raises(capi.GP5Error, match=r"expected \(2,4,4,4\)"):
        capi.h5_read(fn, (4, 4, 4), 2)
    with pytest.raises(capi.GP5Error, match="cannot open"):
        capi.h5_read(str(tmp_path / "missing.hdf5"))
    not_h5 = tmp_path / "x.hdf5"
    not_h5.write_text("not an hdf5 file")
    with pytest.raises(capi.GP5Error, match="cannot open"):
        capi.h5_shape(str(not_h5))
    with pytest.raises(capi.GP5Error, match="cannot create"):
        capi.h5_write(str(tmp_path / "no" / "such" / "dir.hdf5"), _rand((2, 2, 2), 1, 0))


# ---------------------------------------------------------------- driver (CPU) ---

def _config(tmp_path, functional, *, nc=1, shape=(16, 16, 16), box=8.0, dt=1e-3, imaginary=True, renorm=True,
            nit=4, max_time=None, stepper="RK4", norms=None, constraint=None):
    ev = {"timeStep": dt, "imaginaryTime": imaginary, "stepper": stepper,
          "maxTime": (2.5 * nit * dt if max_time is None else max_time), "reNormalize": renorm}
    if constraint:
        ev["constraint"] = constraint
    return {
        "mesh": {"shape": list(shape)},
        "domain": {"left": [-box] * 3, "right": [box] * 3},
        "parallel": {"processorGrid": [1, 1, 1], "fftOrdering": [2, 0, 1]},
        "wavefunction": {"nComponents": nc, "normalization": norms or [50.0] * nc},
        "functional": functional,
        "evolution": ev,
        "initialCondition": {"file": str(tmp_path / "init.hdf5")},
        "output": {"folder": str(tmp_path / "output"), "nIterations": nit},
    }


def test_output_schedule_follows_float_accumulation():
    """t grows by timeStep once per step (src/main.cpp:196-201): 0.1 + 0.1 + 0.1 is
    0.30000000000000004 > 0.3, so a run to maxTime 0.3 with one step per output
    writes 3 files, not 4."""
    assert list(driver.output_times(0.1, 1, 0.3)) == [0.0, 0.1, 0.2]
    assert list(driver.output_times(0.25, 2, 1.0)) == [0.0, 0.5, 1.0]
    assert list(driver.output_times(1.0, 5, 0.0)) == [0.0]
    with pytest.raises(ValueError):
        list(driver.output_times(0.1, 0, 1.0))


def test_parse_reads_main_cpp_sections(tmp_path):
    cfg = _config(tmp_path, {"d": {"name": "LHYDroplet"}})
    rc = driver.parse(cfg)
    assert rc.shape == [16, 16, 16] and rc.left == [-8.0] * 3 and rc.right == [8.0] * 3
    assert rc.nc == 1 and rc.stepper == "RK4" and rc.imaginary and rc.renormalize
    assert rc.n_iterations == 4 and rc.out_dir.endswith("output")
    del cfg["evolution"]["reNormalize"]
    del cfg["parallel"]["fftOrdering"]
    assert driver.parse(cfg).renormalize is False
    cfg["parallel"]["processorGrid"] = [2, 1, 1]
    with pytest.raises(ValueError, match="one GPU"):
        driver.parse(cfg)
    cfg = _config(tmp_path, {"d": {"name": "LHYDroplet"}})
    del cfg["mesh"]
    with pytest.raises(ValueError, match="missing 'mesh'"):
        driver.parse(cfg)


def test_parse_constraint_takes_phase_file(tmp_path):
    cfg = _config(tmp_path, {"d": {"name": "LHYDroplet"}},
                  constraint={"other": {"name": "somethingElse"}, "v": {"name": "phaseConstraint", "file": "phi.hdf5"}})
    assert driver.parse(cfg).phase_file == "phi.hdf5"


def test_potentials_closed_forms():
    """harmonic: (wx^2 x^2 + wy^2 y^2 + wz^2 z^2) / 2 per component; vortex: 1 / (2 r_perp^2),
    on the cell-centred grid x_i = left + (i + 1/2) dx."""
    shape, left, right = (4, 4, 2), [-2.0, -2.0, -1.0], [2.0, 2.0, 1.0]
    V = driver.potential({"h": {"name": "harmonic", "omegas": [[1, 2, 3], [0.5, 0, 0]]}}, shape, left, right, 2)
    # (i, j, k) = (0, 3, 1): x = -1.5, y = 1.5, z = 0.5
    assert V[0, 3, 1, 0] == pytest.approx(0.5 * (2.25 + 4 * 2.25 + 9 * 0.25), rel=1e-15)
    assert V[0, 3, 1, 1] == pytest.approx(0.5 * 0.25 * 2.25, rel=1e-15)
    W = driver.potential({"v": {"name": "vortex"}}, shape, left, right, 1)
    assert W[1, 2, 0, 0] == pytest.approx(0.5 / (0.25 + 0.25), rel=1e-15)   # x = -0.5, y = 0.5
    assert W[0, 0, 1, 0] == pytest.approx(0.5 / (2.25 + 2.25), rel=1e-15)   # x = y = -1.5
    S = driver.potential({"h": {"name": "harmonic", "omegas": [[1, 1, 1]]}, "v": {"name": "vortex"}},
                         shape, left, right, 1)
    H = driver.potential({"h": {"name": "harmonic", "omegas": [[1, 1, 1]]}}, shape, left, right, 1)
    assert np.array_equal(S, H + W)
    with pytest.raises(ValueError, match="Unkown potential"):
        driver.potential({"x": {"name": "box"}}, shape, left, right, 1)


@needs_h5
def test_potential_from_file_real_and_complex(tmp_path):
    shape = (4, 3, 2)
    V = np.random.default_rng(1).standard_normal(shape + (1,))
    fn = str(tmp_path / "V.hdf5")
    capi.h5_write(fn, V.astype(complex))
    got = driver.potential({"f": {"name": "potentialFromFile", "filename": fn}}, shape, [-1] * 3, [1] * 3, 1)
    assert np.array_equal(got, V)
    capi.h5_write(fn, V + 1e-3j)
    with pytest.raises(ValueError, match="imaginary"):
        driver.potential({"f": {"name": "potentialFromFile", "filename": fn}}, shape, [-1] * 3, [1] * 3, 1)


def test_functional_section():
    shape, left, right = (4, 4, 4), [-1.0] * 3, [1.0] * 3
    desc, V = driver.functional({"gp": {"name": "gpFunctional", "masses": [2.0], "coupling": [[1.5, 0.25], [0.5, 3.0]]}},
                                shape, left, right, 2)
    assert desc["kind"] == capi.GP5_FUNC_GP and V is None
    # couplings(i, j) = coupling[j][i]
    assert desc["couplings"][0, 1] == 0.5 and desc["couplings"][1, 0] == 0.25
    assert desc["masses"] == (2.0, 1.0)
    desc, V = driver.functional({"a": {"name": "LHYDroplet"},
                                 "V": {"name": "externalPotential", "V": {"h": {"name": "harmonic", "omegas": [[1, 1, 1]]}}}},
                                shape, left, right, 1)
    assert desc["kind"] == capi.GP5_FUNC_LHY and V.shape == (4, 4, 4, 1)
    desc, _ = driver.functional({"a": {"name": "LHYDroplet"},
                                 "b": {"name": "UnlockedLHYDroplet", "alpha": 0.8, "beta": 1.0, "eta": -4}},
                                shape, left, right, 2)
    assert desc["kind"] == capi.GP5_FUNC_LHY_UNLOCKED and desc["eta"] == -4.0   # last entry wins
    with pytest.raises(ValueError, match="Unkown functional"):
        driver.functional({"a": {"name": "dipolar"}}, shape, left, right, 1)
    with pytest.raises(ValueError, match="Incompatible"):
        driver.functional({"a": {"name": "gpFunctional", "coupling": [[1.0]]}}, shape, left, right, 2)
    with pytest.raises(ValueError, match="needs gpFunctional"):
        driver.functional({"V": {"name": "externalPotential", "V": {}}}, shape, left, right, 1)


def test_driver_cli_usage():
    assert driver.main([]) == 2


# ---------------------------------------------------------------- driver (GPU) ---

def _write_yaml(tmp_path, cfg):
    import yaml

    fn = tmp_path / "input.yaml"
    fn.write_text(yaml.safe_dump(cfg))
    return str(fn)


@pytest.mark.gpu
def test_driver_lhy_imaginary_matches_oracle(tmp_path, oracle_mod, capsys):
    """LHY droplet, imaginary-time RK4 with renormalisation: every out_<k>.hdf5
    equals the oracle advanced k * nIterations steps (relative L2 <= 1e-12)."""
    O = oracle_mod
    shape, box, nit, dt = (16, 16, 16), 8.0, 4, 2e-3
    cfg = _config(tmp_path, {"d": {"name": "LHYDroplet"}}, shape=shape, box=box, dt=dt, nit=nit,
                  max_time=2 * nit * dt + dt / 2, norms=[50.0])
    geo = O.Geometry(shape, [-box] * 3, [box] * 3, 1)
    X, Y, Z = geo.grid()
    psi0 = np.asfortranarray((np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 6.0) * (1 + 0.1j * X))[..., None])
    capi.h5_write(cfg["initialCondition"]["file"], psi0)
    files = driver.run(_write_yaml(tmp_path, cfg))
    assert [os.path.basename(f) for f in files] == ["out_0.hdf5", "out_1.hdf5", "out_2.hdf5"]
    out = capsys.readouterr().out
    assert out.splitlines() == ["Time: 0", "saved", "Time: 0.008", "saved", "Time: 0.016", "saved"]
    f = O.Functional(kind=O.OR_LHY)
    ref = psi0
    for k, fn in enumerate(files):
        got = capi.h5_read(fn, shape, 1)
        if k == 0:
            assert np.array_equal(got, psi0)
        else:
            assert rel_l2(got, ref) <= 1e-12, (k, rel_l2(got, ref))
        ref = O.advance(f, ref, geo, stepper="RK4", dt=dt, imaginary=True, nsteps=nit, renormalize=[50.0])


@pytest.mark.gpu
def test_driver_two_component_gp_with_potentials_real_time(tmp_path, oracle_mod):
    """Two-component gpFunctional with couplings and masses, harmonic + from-file
    potential, real-time eulero without renormalisation (the
    exampleInputFiles/singleComponentGP.yaml pattern, widened to two components)."""
    O = oracle_mod
    shape, box, nit, dt = (12, 10, 8), 6.0, 3, 1e-3
    Vf = np.random.default_rng(5).uniform(0, 0.5, shape + (2,))
    capi.h5_write(str(tmp_path / "V.hdf5"), Vf.astype(complex))
    functional = {
        "gp": {"name": "gpFunctional", "masses": [1.0, 2.0], "coupling": [[1.0, 0.3], [0.3, 0.7]]},
        "V": {"name": "externalPotential",
              "V": {"h11": {"name": "harmonic", "omegas": [[1, 1, 1], [0.5, 1, 2]]},
                    "file": {"name": "potentialFromFile", "filename": str(tmp_path / "V.hdf5")}}},
    }
    cfg = _config(tmp_path, functional, nc=2, shape=shape, box=box, dt=dt, nit=nit, imaginary=False, renorm=False,
                  max_time=nit * dt * 1.5, stepper="eulero", norms=[10.0, 20.0])
    psi0 = seeded_field(shape, 2, seed=21, scale=0.3)
    capi.h5_write(cfg["initialCondition"]["file"], psi0)
    files = driver.run(_write_yaml(tmp_path, cfg))
    assert len(files) == 2
    geo = O.Geometry(shape, [-box] * 3, [box] * 3, 2)
    V = driver.potential(functional["V"]["V"], shape, [-box] * 3, [box] * 3, 2)
    X, Y, Z = geo.grid()
    Vh = np.stack([0.5 * (X ** 2 + Y ** 2 + Z ** 2), 0.5 * (0.25 * X ** 2 + Y ** 2 + 4 * Z ** 2)], axis=3)
    assert np.allclose(V, Vh + Vf, rtol=1e-14, atol=1e-14)
    f = O.Functional(kind=O.OR_GP, couplings=np.array([[1.0, 0.3], [0.3, 0.7]]), masses=(1.0, 2.0), V=V)
    ref = O.advance(f, psi0, geo, stepper="eulero", dt=dt, imaginary=False, nsteps=nit)
    got = capi.h5_read(files[1], shape, 2)
    assert rel_l2(got, ref) <= 1e-12


@pytest.mark.gpu
def test_field_save_load_hdf5(tmp_path):
    """gp.field.save/.load (python/gp/gp.py:98-160) and gpCpp field.save/load
    (src/gpCpp.cpp:105-113) write and read the reference's file."""
    from gp5_amd import gp

    geo = gp.geometry((16, 8, 12), left=[-4, -2, -3], right=[4, 2, 3])
    psi = gp.field(geo, nComponents=2)
    a = seeded_field((16, 8, 12), 2, seed=3)
    psi.array = a
    fn = str(tmp_path / "sub" / "psi.hdf5")
    psi.save(fn)
    assert np.array_equal(capi.h5_read(fn), a)
    psi2 = gp.field(geo, nComponents=2)
    psi2.load(fn)
    assert np.array_equal(psi2.array, a)
    fn2 = str(tmp_path / "cpp.h5")
    psi._fieldCpp.save(fn2)
    assert np.array_equal(capi.h5_read(fn2), a)
    psi3 = gp.field(geo, nComponents=2)
    psi3._fieldCpp.load(fn2)
    assert np.array_equal(psi3.array, a)
    bad = gp.field(gp.geometry((8, 8, 12), left=[-4, -2, -3], right=[4, 2, 3]), nComponents=2)
    with pytest.raises(Exception, match="expected"):
        bad.load(fn)


@pytest.mark.gpu
def test_driver_unlocked_droplet_two_components(tmp_path, oracle_mod):
    """The exampleInputFiles/dropletUnlocked.yaml pattern (UnlockedLHYDroplet,
    eta -4, alpha 0.8838834764831843, beta 1, imaginary-time RK4, two
    components) on a small grid, with renormalisation, against the oracle."""
    O = oracle_mod
    shape, box, nit, dt = (16, 12, 14), 6.0, 3, 2e-3
    functional = {"droplet": {"name": "UnlockedLHYDroplet", "eta": -4, "alpha": 0.8838834764831843, "beta": 1}}
    cfg = _config(tmp_path, functional, nc=2, shape=shape, box=box, dt=dt, nit=nit,
                  max_time=nit * dt * 1.5, norms=[10.0, 100.0])
    psi0 = seeded_field(shape, 2, seed=31, scale=0.5)
    capi.h5_write(cfg["initialCondition"]["file"], psi0)
    files = driver.run(_write_yaml(tmp_path, cfg))
    assert len(files) == 2
    geo = O.Geometry(shape, [-box] * 3, [box] * 3, 2)
    f = O.Functional(kind=O.OR_UNLOCKED, alpha=0.8838834764831843, beta=1.0, eta=-4.0)
    ref = O.advance(f, psi0, geo, stepper="RK4", dt=dt, imaginary=True, nsteps=nit, renormalize=[10.0, 100.0])
    assert rel_l2(capi.h5_read(files[1], shape, 2), ref) <= 1e-12


@pytest.mark.gpu
def test_driver_vortex_with_phase_constraint(tmp_path, oracle_mod):
    """gpFunctional with harmonic + vortex potentials and a phaseConstraint read from
    an HDF5 phase file (imprinted winding atan2(y, x)), imaginary-time eulero with
    renormalisation, against the oracle."""
    O = oracle_mod
    shape, box, nit, dt = (16, 16, 8), 4.0, 4, 1e-3
    geo = O.Geometry(shape, [-box] * 3, [box] * 3, 1)
    X, Y, Z = geo.grid()
    phase = np.arctan2(Y, X)[..., None]
    capi.h5_write(str(tmp_path / "phi.hdf5"), phase.astype(complex))
    functional = {
        "gp": {"name": "gpFunctional", "masses": [1], "coupling": [[1.5]]},
        "V": {"name": "externalPotential",
              "V": {"h11": {"name": "harmonic", "omegas": [[1, 1, 2]]}, "vortex": {"name": "vortex"}}},
    }
    cfg = _config(tmp_path, functional, shape=shape, box=box, dt=dt, nit=nit, max_time=nit * dt * 1.5,
                  stepper="eulero", norms=[5.0],
                  constraint={"vortexPhase": {"file": str(tmp_path / "phi.hdf5"), "name": "phaseConstraint"}})
    psi0 = np.asfortranarray((np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 4.0) + 0j)[..., None])
    capi.h5_write(cfg["initialCondition"]["file"], psi0)
    files = driver.run(_write_yaml(tmp_path, cfg))
    assert len(files) == 2
    V = 0.5 * (X ** 2 + Y ** 2 + 4 * Z ** 2) + 0.5 / (X ** 2 + Y ** 2)
    f = O.Functional(kind=O.OR_GP, couplings=np.array([[1.5]]), masses=(1.0,), V=V[..., None])
    ref = O.advance(f, psi0, geo, stepper="eulero", dt=dt, imaginary=True, nsteps=nit, renormalize=[5.0],
                    phase=phase)
    got = capi.h5_read(files[1], shape, 1)
    assert rel_l2(got, ref) <= 1e-12
    # the imprinted winding survives: the phase of the state is atan2(y, x) where it is not tiny
    mask = np.abs(got[..., 0]) > 1e-3 * np.abs(got).max()
    d = np.angle(got[..., 0][mask] * np.exp(-1j * phase[..., 0][mask]))
    assert np.abs(d).max() < 1e-12
