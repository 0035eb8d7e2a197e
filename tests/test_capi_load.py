"""CPU-side checks of the C ABI library: it loads and exports every symbol
include/gp5.h declares.  No compute call is made (no GPU here)."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "gp5.h")).read()
    return sorted(set(re.findall(r"\b(gp5_[a-z_0-9]+)\s*\(", src)))


def test_header_and_binding_agree():
    from gp5_amd import capi

    assert sorted(capi.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    from gp5_amd import capi

    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_library_reports_version_and_error_text():
    from gp5_amd import capi

    L = capi.load()
    assert L.gp5_version() >= 1
    # NULL context is rejected on the host without touching a device
    assert L.gp5_ctx_sync(None) == capi.GP5_ERR_ARG
    assert b"NULL" in L.gp5_last_error()


def test_entry_points_reject_null_arguments_on_the_host():
    """Argument checks of the labelling, resampling, batch-resampling and HDF5 entry
    points run before any device work (no GPU needed)."""
    import numpy as np

    from gp5_amd import capi

    L = capi.load()
    n = ctypes.c_int()
    assert L.gp5_decompose(None, None, 0, 0.5, None, ctypes.byref(n)) == capi.GP5_ERR_ARG
    assert L.gp5_resample(None, None, None, None) == capi.GP5_ERR_ARG
    assert L.gp5_batch_resample(None, 0, None, 0) == capi.GP5_ERR_ARG
    shape = (ctypes.c_int * 3)(2, 2, 2)
    buf = np.zeros(16)
    assert L.gp5_h5_write(None, buf.ctypes.data, shape, 1) == capi.GP5_ERR_ARG
    assert L.gp5_h5_read(b"x.h5", None, shape, 1) == capi.GP5_ERR_ARG
    assert L.gp5_h5_shape(None, shape, ctypes.byref(n)) == capi.GP5_ERR_ARG
    bad = (ctypes.c_int * 3)(2, 0, 2)
    assert L.gp5_h5_write(b"x.h5", buf.ctypes.data, bad, 1) == capi.GP5_ERR_ARG
    assert b"bad shape" in L.gp5_last_error()


def _build_facade_example(out_dir):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = os.path.join(str(out_dir), "facade_example")
    cmd = [gxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "gp5_amd", "csrc"),
           os.path.join(ROOT, "tests", "cpp", "facade_example.cpp"), "-L", os.path.join(ROOT, "gp5_amd"), "-lgp5",
           "-Wl,-rpath," + os.path.join(ROOT, "gp5_amd"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


def test_cpp_facade_example_compiles_and_links(tmp_path):
    """The reference-style C++ API (gp5pp.hpp) builds against libgp5.so."""
    assert os.path.exists(_build_facade_example(tmp_path))


@pytest.mark.gpu
def test_cpp_facade_example_runs(tmp_path, oracle_mod):
    """Runs the C++ example (64^3 LHY Gaussian, N = 500, 10 real-time RK4 steps) and checks
    its printed norm and energy against the oracle."""
    exe = _build_facade_example(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    parts = r.stdout.split()
    norm, energy = float(parts[1]), float(parts[3])

    O = oracle_mod
    geo = O.Geometry((64, 64, 64), (-10, -10, -10), (10, 10, 10), 1)
    X, Y, Z = geo.grid()
    a = (np.exp(-(X**2 + Y**2 + Z**2) / 8.0) + 0j)[..., None]
    a = O.normalize(a, geo, 0, 500.0)
    F = O.Functional(kind=O.OR_LHY)
    a = O.advance(F, a, geo, stepper="RK4", dt=1e-3, imaginary=False, nsteps=10)
    assert norm == pytest.approx(O.norm(a, geo, 0), rel=1e-11)
    assert energy == pytest.approx(O.energy(F, a, geo), rel=1e-10)


def test_library_exports_every_nn_symbol():
    src = open(os.path.join(ROOT, "include", "gp5_nn.h")).read()
    names = sorted(set(re.findall(r"\b(gp5_nn_[a-z_0-9]+)\s*\(", src)))
    assert len(names) >= 18
    from gp5_amd import capi

    lib = ctypes.CDLL(capi.LIB_PATH)
    assert not [s for s in names if not hasattr(lib, s)]


def test_network_entry_points_reject_bad_shapes_on_the_host():
    """The in-place-upsampling convolution, its weight gradient, the fused
    activation-backward and the fused loss check their shapes before any device work
    (no GPU needed; the pointers are never dereferenced)."""
    from gp5_amd import capi, nn

    L = nn._lib()
    fake = ctypes.c_void_p(0x1000)
    loss = ctypes.c_double()
    # conv3d_up: the cube side must be a power of two in [8, 128]
    assert L.gp5_nn_conv3d_up(fake, fake, None, None, None, None, fake, 1, 6, 16, 16, 0, 1) == capi.GP5_ERR_ARG
    assert b"power of two" in L.gp5_last_error()
    # wgrad_up: the cube side must be a multiple of 8
    assert L.gp5_nn_conv3d_wgrad_up(fake, fake, fake, fake, 1, 12, 16, 16) == capi.GP5_ERR_ARG
    assert b"multiple of 8" in L.gp5_last_error()
    # act_bwd_cast_bias: C / 4 must divide 256
    assert L.gp5_nn_act_bwd_cast_bias(fake, fake, fake, fake, 64, 48) == capi.GP5_ERR_ARG
    assert L.gp5_nn_act_bwd_cast_bias(None, fake, fake, fake, 64, 16) == capi.GP5_ERR_ARG
    # up2_add_act_bwd_cast_bias: G required, D even
    assert L.gp5_nn_up2_add_act_bwd_cast_bias(fake, fake, None, 0.125, fake, fake, 1, 8, 16) == capi.GP5_ERR_ARG
    assert L.gp5_nn_up2_add_act_bwd_cast_bias(fake, fake, fake, 0.125, fake, fake, 1, 7, 16) == capi.GP5_ERR_ARG
    assert b"even" in L.gp5_last_error()
    # conv3d_pool: 16 or 32 gradient channels, side 32, 64 or 128 (every call below is
    # refused before any launch: the pointers are fake)
    assert L.gp5_nn_conv3d_pool(fake, fake, fake, 1, 16, 16, 32, 1) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_pool(fake, fake, fake, 1, 16, 16, 16, 1) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_pool(fake, fake, fake, 1, 64, 64, 32, 1) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_pool(fake, fake, None, 1, 64, 16, 32, 1) == capi.GP5_ERR_ARG
    # conv3d_act_bwd: every pointer, and Cout / 4 dividing 256
    assert L.gp5_nn_conv3d_act_bwd(fake, fake, None, fake, fake, fake, 1, 16, 16, 16) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_act_bwd(fake, fake, fake, fake, fake, fake, 1, 16, 16, 48) == capi.GP5_ERR_ARG
    # conv3d_act_bwd_out0: every pointer, D a power of two in [16, 128], Cout a multiple of 16
    assert L.gp5_nn_conv3d_act_bwd_out0(fake, fake, fake, fake, None, fake, 1, 32, 16, 16) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_act_bwd_out0(fake, fake, fake, fake, fake, fake, 1, 24, 16, 16) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_act_bwd_out0(fake, fake, fake, fake, fake, fake, 1, 256, 16, 16) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_act_bwd_out0(fake, fake, fake, fake, fake, fake, 1, 32, 16, 8) == capi.GP5_ERR_ARG
    # wgrad_out0c: no full-row fallback, so shapes outside the kernel's are refused
    assert L.gp5_nn_conv3d_wgrad_out0c(fake, fake, fake, fake, 1, 16, 16, 16) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_wgrad_out0c(fake, fake, fake, fake, 1, 32, 32, 16) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_wgrad_out0c(fake, None, fake, fake, 1, 32, 16, 16) == capi.GP5_ERR_ARG
    # act_bwd_part_bf16 needs G and its partial buffer
    assert L.gp5_nn_act_bwd_part_bf16(fake, fake, None, 0.125, fake, fake, 1, 8, 16) == capi.GP5_ERR_ARG
    assert L.gp5_nn_act_bwd_part_bf16(fake, fake, fake, 0.125, fake, None, 1, 8, 16) == capi.GP5_ERR_ARG
    assert L.gp5_nn_act_bwd_part_bf16(fake, fake, fake, 0.125, fake, fake, 1, 7, 16) == capi.GP5_ERR_ARG
    # act_bwd_part needs its partial buffer; bias_final_multi 1..16 complete segments
    assert L.gp5_nn_act_bwd_part(fake, fake, None, 0.0, fake, None, 1, 8, 16) == capi.GP5_ERR_ARG
    bseg = nn.BiasSeg(fake, None, 16)
    assert L.gp5_nn_bias_final_multi(ctypes.byref(bseg), 1) == capi.GP5_ERR_ARG
    assert L.gp5_nn_bias_final_multi(ctypes.byref(bseg), 0) == capi.GP5_ERR_ARG
    # conv3d: act 0 or 1 only; out0 needs its output
    assert L.gp5_nn_conv3d(fake, fake, None, None, None, None, fake, 1, 8, 16, 16, 0, 2) == capi.GP5_ERR_ARG
    assert L.gp5_nn_conv3d_out0(fake, fake, None, None, 1, 8, 16, 16) == capi.GP5_ERR_ARG
    # mse_out0: H required, C a multiple of 8
    assert L.gp5_nn_mse_out0(fake, fake, fake, None, 16, None, 64, ctypes.byref(loss)) == capi.GP5_ERR_ARG
    assert L.gp5_nn_mse_out0(fake, fake, fake, fake, 12, None, 64, ctypes.byref(loss)) == capi.GP5_ERR_ARG
    assert L.gp5_nn_mse_out0_acc(fake, fake, fake, fake, 16, None, 64, None) == capi.GP5_ERR_ARG
    # mse_out0c: H0 required, and a loss destination
    assert L.gp5_nn_mse_out0c(fake, fake, fake, None, 16, None, 64, ctypes.byref(loss), None) == capi.GP5_ERR_ARG
    assert L.gp5_nn_mse_out0c(fake, fake, fake, fake, 16, None, 64, None, None) == capi.GP5_ERR_ARG
    # streams and events: NULL handles out and events in
    assert L.gp5_nn_stream_create(None) == capi.GP5_ERR_ARG
    assert L.gp5_nn_event_create(None) == capi.GP5_ERR_ARG
    assert L.gp5_nn_event_record(None, None) == capi.GP5_ERR_ARG
    assert L.gp5_nn_stream_wait(None, None) == capi.GP5_ERR_ARG
    assert L.gp5_nn_set_stream(None) == capi.GP5_OK
    assert L.gp5_nn_mse_out0_acc(fake, fake, fake, fake, 12, None, 64, fake) == capi.GP5_ERR_ARG
    # mse_bf16: C a multiple of 8, H required
    assert L.gp5_nn_mse_bf16(fake, 12, fake, fake, None, fake, None, 64, ctypes.byref(loss)) == capi.GP5_ERR_ARG
    assert L.gp5_nn_mse_bf16(fake, 16, fake, fake, None, None, None, 64, ctypes.byref(loss)) == capi.GP5_ERR_ARG
    # adam_multi: 1..16 tensors, complete ones, weight copies only of 27 x cin x cout tensors
    seg = nn.AdamSeg(fake, fake, fake, fake, None, None, 10, 0, 0)
    assert L.gp5_nn_adam_multi(ctypes.byref(seg), 0, 1e-3, 0.9, 0.999, 1e-8, 1) == capi.GP5_ERR_ARG
    assert L.gp5_nn_adam_multi(ctypes.byref(seg), 17, 1e-3, 0.9, 0.999, 1e-8, 1) == capi.GP5_ERR_ARG
    assert L.gp5_nn_adam_multi(ctypes.byref(seg), 1, 1e-3, 0.9, 0.999, 1e-8, 0) == capi.GP5_ERR_ARG
    bad = nn.AdamSeg(fake, fake, fake, fake, fake, None, 27 * 16 * 16 + 1, 16, 16)
    assert L.gp5_nn_adam_multi(ctypes.byref(bad), 1, 1e-3, 0.9, 0.999, 1e-8, 1) == capi.GP5_ERR_ARG
    assert b"27 x 16 x 16" in L.gp5_last_error()
    partial = nn.AdamSeg(fake, None, fake, fake, None, None, 10, 0, 0)
    assert L.gp5_nn_adam_multi(ctypes.byref(partial), 1, 1e-3, 0.9, 0.999, 1e-8, 1) == capi.GP5_ERR_ARG
