"""GPU kernels of the warm-start network (include/gp5_nn.h) against the CPU oracle
oracle/nn_ref.py: convolution forward, data and weight gradients on bf16 inputs
(exactly representable, so only the float accumulation order differs), pooling,
Adam, and the whole network's loss and gradients (bf16 activations: agreement to
the bf16 rounding level)."""

import ctypes

import numpy as np
import pytest

from oracle import nn_ref as R

nn = pytest.importorskip("gp5_amd.nn")


def test_bf16_round_trip_on_host():
    x = np.array([1.0, -2.5, 3.14159, 1e-3, 65504.0], dtype=np.float32)
    y = nn.from_bf16(nn.to_bf16(x))
    assert np.all(np.abs(y - x) <= np.abs(x) * 2.0 ** -8)
    assert nn.from_bf16(nn.to_bf16(np.float32([1.0 + 2.0 ** -9])))[0] == 1.0  # ties to even


def _rb(a):
    return nn.from_bf16(nn.to_bf16(a)).astype(np.float64)


def _buf(a, bf16=False):
    arr = nn.to_bf16(a) if bf16 else np.ascontiguousarray(a, dtype=np.float32)
    b = nn.DevBuf(arr.nbytes)
    b.upload(arr)
    return b


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B", [(8, 16, 32, 2), (4, 128, 64, 1), (16, 32, 16, 1), (8, 16, 144, 1)])
def test_conv_forward_dgrad_wgrad(D, Cin, Cout, B):
    rng = np.random.default_rng(D + Cin + Cout)
    X = _rb(rng.standard_normal((B, D, D, D, Cin)))
    W = _rb(rng.standard_normal((27, Cin, Cout)) * 0.1)
    b = rng.standard_normal(Cout).astype(np.float32)
    G = _rb(rng.standard_normal((B, D, D, D, Cout)))
    L = nn._lib()
    xb, gb = _buf(X, True), _buf(G, True)
    Wf = np.ascontiguousarray(W.transpose(0, 2, 1))
    wf, wb, bb = _buf(Wf, True), _buf(W, True), _buf(b)
    Yf = nn.DevBuf(B * D ** 3 * Cout * 4)
    A = nn.DevBuf(B * D ** 3 * Cout * 2)
    nn._ck(L.gp5_nn_conv3d(xb.ptr, wf.ptr, bb.ptr, A.ptr, None, None, Yf.ptr, B, D, Cin, Cout, 0, 1))
    ref = R.lrelu(R.conv(X, W, b.astype(np.float64)))
    got = Yf.download(np.float32, ref.shape)
    assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(ref))
    assert np.array_equal(A.download(np.uint16, ref.shape), nn.to_bf16(got))
    # data gradient through the mirrored taps of the [27][ci][co] copy
    dX = nn.DevBuf(B * D ** 3 * Cin * 4)
    nn._ck(L.gp5_nn_conv3d(gb.ptr, wb.ptr, None, None, None, None, dX.ptr, B, D, Cout, Cin, 1, 0))
    ref = R.conv_dgrad(G, W)
    got = dX.download(np.float32, ref.shape)
    assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(ref))
    # weight gradient
    ws = nn.DevBuf(L.gp5_nn_wgrad_ws(B, D, Cin, Cout))
    dW = nn.DevBuf(27 * Cin * Cout * 4)
    nn._ck(L.gp5_nn_conv3d_wgrad(xb.ptr, gb.ptr, dW.ptr, ws.ptr, B, D, Cin, Cout))
    ref = R.conv_wgrad(X, G)
    got = dW.download(np.float32, ref.shape)
    assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_pool_up_act_adam():
    rng = np.random.default_rng(7)
    L = nn._lib()
    B, D, C = 2, 8, 16
    X = _rb(rng.standard_normal((B, D, D, D, C)))
    xb = _buf(X, True)
    y = nn.DevBuf(B * (D // 2) ** 3 * C * 2)
    nn._ck(L.gp5_nn_pool2_bf16(xb.ptr, y.ptr, B, D, C, 1.0))
    assert np.array_equal(y.download(np.uint16, (B, 4, 4, 4, C)), nn.to_bf16(R.pool2(X).astype(np.float32)))
    u = nn.DevBuf(B * (2 * D) ** 3 * C * 2)
    nn._ck(L.gp5_nn_up2_bf16(xb.ptr, u.ptr, B, D, C, 1.0))
    assert np.array_equal(nn.from_bf16(u.download(np.uint16, (B, 16, 16, 16, C))), R.up2(X).astype(np.float32))
    # act backward
    G = rng.standard_normal(X.shape).astype(np.float32)
    gb, out = _buf(G), nn.DevBuf(G.nbytes)
    nn._ck(L.gp5_nn_act_bwd(xb.ptr, gb.ptr, out.ptr, G.size))
    assert np.allclose(out.download(np.float32, G.shape), G * R.lrelu_grad(X), rtol=1e-7, atol=0)
    # Adam, two steps
    p = rng.standard_normal(1000).astype(np.float32)
    g = rng.standard_normal(1000).astype(np.float32)
    pb, gb2, mb, vb = _buf(p), _buf(g), nn.DevBuf(4000), nn.DevBuf(4000)
    mb.zero()
    vb.zero()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in (1, 2):
        nn._ck(L.gp5_nn_adam(pb.ptr, gb2.ptr, mb.ptr, vb.ptr, p.size, 1e-2, 0.9, 0.999, 1e-8, t))
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - 1e-2 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    assert np.allclose(pb.download(np.float32, p.shape), p, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("B,D,C", [(2, 16, 16), (1, 8, 64), (3, 4, 32)])
def test_up2_add_act_bwd_equals_the_two_passes(B, D, C):
    """gp5_nn_up2_add_act_bwd_cast_bias gives bit for bit what gp5_nn_up2_add_f32
    into dA followed by gp5_nn_act_bwd_cast_bias gives (H and the bias gradient),
    and H matches bf16((dA + G/8 at the parent) * act'(A)) computed in numpy."""
    rng = np.random.default_rng(B * 100 + D + C)
    L = nn._lib()
    A = _rb(rng.standard_normal((B, D, D, D, C)))
    dA = rng.standard_normal((B, D, D, D, C)).astype(np.float32)
    G = rng.standard_normal((B, D // 2, D // 2, D // 2, C)).astype(np.float32)
    ab, gb = _buf(A, True), _buf(G)
    da1, da2 = _buf(dA), _buf(dA)
    H1, H2 = nn.DevBuf(dA.size * 2), nn.DevBuf(dA.size * 2)
    db1, db2 = nn.DevBuf(C * 4), nn.DevBuf(C * 4)
    nn._ck(L.gp5_nn_up2_add_act_bwd_cast_bias(ab.ptr, da1.ptr, gb.ptr, 0.125, H1.ptr, db1.ptr, B, D, C))
    nn._ck(L.gp5_nn_up2_add_f32(gb.ptr, da2.ptr, B, D // 2, C, 0.125))
    nn._ck(L.gp5_nn_act_bwd_cast_bias(ab.ptr, da2.ptr, H2.ptr, db2.ptr, B * D ** 3, C))
    h1 = H1.download(np.uint16, dA.shape)
    assert np.array_equal(h1, H2.download(np.uint16, dA.shape))
    assert np.array_equal(db1.download(np.float32, (C,)), db2.download(np.float32, (C,)))
    assert np.array_equal(da1.download(np.float32, dA.shape), dA)  # dA is not written
    # the same with the bias partials kept and finished in one multi-segment launch
    H3, db3 = nn.DevBuf(dA.size * 2), nn.DevBuf(C * 4)
    part = nn.DevBuf(L.gp5_nn_bias_part_floats(C) * 4)
    nn._ck(L.gp5_nn_act_bwd_part(ab.ptr, da1.ptr, gb.ptr, 0.125, H3.ptr, part.ptr, B, D, C))
    seg = (nn.BiasSeg * 1)(nn.BiasSeg(part.ptr, db3.ptr, C))
    nn._ck(L.gp5_nn_bias_final_multi(ctypes.cast(seg, ctypes.c_void_p), 1))
    assert np.array_equal(H3.download(np.uint16, dA.shape), h1)
    assert np.array_equal(db3.download(np.float32, (C,)), db1.download(np.float32, (C,)))
    tot = dA.astype(np.float64) + 0.125 * R.up2(G.astype(np.float64))
    ref = tot * R.lrelu_grad(A)
    assert np.max(np.abs(nn.from_bf16(h1) - ref)) <= 2.0 ** -8 * np.max(np.abs(ref))
    assert np.allclose(db1.download(np.float32, (C,)), ref.sum(axis=(0, 1, 2, 3)), rtol=1e-4, atol=1e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B", [(64, 16, 32, 2), (128, 16, 32, 1), (64, 16, 64, 1), (32, 32, 64, 4),
                                          (64, 32, 32, 1)])
def test_conv_pool_equals_conv_then_pool(D, Cin, Cout, B, monkeypatch):
    """gp5_nn_conv3d_pool (the data gradient summed over each cell's 2x2x2
    children, mirrored taps) in the z-sliding kernel's pooled epilogue equals
    gp5_nn_conv3d into a float tensor followed by gp5_nn_pool2_f32 with scale 8, bit
    for bit, and agrees with the oracle (the stride-2 form, which sums in another
    order, is off here and has its own test)."""
    monkeypatch.setenv("GP5_CONV_POOL2P", "0")
    rng = np.random.default_rng(D + Cin + Cout + B)
    L = nn._lib()
    G = _rb(rng.standard_normal((B, D, D, D, Cin)))
    W = _rb(rng.standard_normal((27, Cout, Cin)) * 0.1)  # [27][ci][co] of the layer: its co is Cin here
    gb, wb = _buf(G, True), _buf(W, True)
    full = nn.DevBuf(B * D ** 3 * Cout * 4)
    ref_h, got_h = nn.DevBuf(B * D ** 3 * Cout // 2), nn.DevBuf(B * D ** 3 * Cout // 2)
    nn._ck(L.gp5_nn_conv3d(gb.ptr, wb.ptr, None, None, None, None, full.ptr, B, D, Cin, Cout, 1, 0))
    nn._ck(L.gp5_nn_pool2_f32(full.ptr, ref_h.ptr, B, D, Cout, 8.0))
    nn._ck(L.gp5_nn_conv3d_pool(gb.ptr, wb.ptr, got_h.ptr, B, D, Cin, Cout, 1))
    shape = (B, D // 2, D // 2, D // 2, Cout)
    got = got_h.download(np.float32, shape)
    assert np.array_equal(got, ref_h.download(np.float32, shape))
    if D <= 64:  # the oracle at 128^3 takes minutes; there the two GPU paths are compared
        ref = 8.0 * R.pool2(R.conv_dgrad(G, W))
        assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("D,B", [(64, 2), (16, 1), (4, 2)])
def test_output_layer_channel0_and_its_loss(D, B):
    """gp5_nn_conv3d_out0 writes, bit for bit, channel 0 of gp5_nn_conv3d's float
    output (z-sliding, LDS-staged and direct kernels by cube side), and
    gp5_nn_mse_out0 on it gives the loss, the bf16 gradient rows and the bias
    gradient of gp5_nn_mse_bf16 on the 16-channel output, bit for bit."""
    rng = np.random.default_rng(D + B)
    L = nn._lib()
    C = 16
    X = _rb(rng.standard_normal((B, D, D, D, C)))
    W = _rb(rng.standard_normal((27, C, C)) * 0.1)
    b = rng.standard_normal(C).astype(np.float32)
    xb, wf, bb = _buf(X, True), _buf(np.ascontiguousarray(W.transpose(0, 2, 1)), True), _buf(b)
    P = B * D ** 3
    Yf, y0 = nn.DevBuf(P * C * 4), nn.DevBuf(P * 4)
    nn._ck(L.gp5_nn_conv3d(xb.ptr, wf.ptr, bb.ptr, None, None, None, Yf.ptr, B, D, C, C, 0, 0))
    nn._ck(L.gp5_nn_conv3d_out0(xb.ptr, wf.ptr, bb.ptr, y0.ptr, B, D, C, C))
    full = Yf.download(np.float32, (P, C))
    assert np.array_equal(y0.download(np.float32, (P,)), full[:, 0])
    base = rng.standard_normal(P).astype(np.float32)
    tgt = (base + 0.1 * rng.standard_normal(P)).astype(np.float32)
    baseb, tb = _buf(base), _buf(tgt)
    H1, H2 = nn.DevBuf(P * C * 2), nn.DevBuf(P * C * 2)
    db1, db2 = nn.DevBuf(C * 4), nn.DevBuf(C * 4)
    l1, l2 = ctypes.c_double(), ctypes.c_double()
    nn._ck(L.gp5_nn_mse_bf16(Yf.ptr, C, baseb.ptr, tb.ptr, None, H1.ptr, db1.ptr, P, ctypes.byref(l1)))
    nn._ck(L.gp5_nn_mse_out0(y0.ptr, baseb.ptr, tb.ptr, H2.ptr, C, db2.ptr, P, ctypes.byref(l2)))
    assert l1.value == l2.value
    assert np.array_equal(H1.download(np.uint16, (P, C)), H2.download(np.uint16, (P, C)))
    assert np.array_equal(db1.download(np.float32, (C,)), db2.download(np.float32, (C,)))
    e = (base.astype(np.float64) + full[:, 0]) - tgt
    assert abs(l2.value - np.mean(e * e)) <= 1e-6 * np.mean(e * e)


@pytest.mark.gpu
def test_adam_multi_equals_per_tensor_adam_and_relayout():
    """One launch over weight tensors (with both bf16 copies, one copy or none) and
    a bias equals gp5_nn_adam + gp5_nn_relayout per tensor, bit for bit, over three
    steps."""
    rng = np.random.default_rng(13)
    L = nn._lib()
    shapes = [(16, 32, "both"), (32, 16, "wf"), (16, 16, "none"), (None, 48, "bias")]
    ref, got = [], []
    for ci, co, kind in shapes:
        n = 27 * ci * co if ci else co
        p = rng.standard_normal(n).astype(np.float32)
        for dst in (ref, got):
            t = {"p": _buf(p), "m": nn.DevBuf(4 * n), "v": nn.DevBuf(4 * n), "g": nn.DevBuf(4 * n), "n": n,
                 "ci": ci or 0, "co": co, "kind": kind,
                 "wf": nn.DevBuf(2 * n) if kind in ("both", "wf") else None,
                 "wb": nn.DevBuf(2 * n) if kind == "both" else None}
            t["m"].zero()
            t["v"].zero()
            for k in ("wf", "wb"):
                if t[k] is not None:
                    t[k].zero()
            dst.append(t)
    for step in (1, 2, 3):
        for a, b in zip(ref, got):
            g = rng.standard_normal(a["n"]).astype(np.float32)
            a["g"].upload(g)
            b["g"].upload(g)
        for a in ref:
            nn._ck(L.gp5_nn_adam(a["p"].ptr, a["g"].ptr, a["m"].ptr, a["v"].ptr, a["n"], 3e-3, 0.9, 0.999, 1e-8, step))
            if a["kind"] != "bias" and a["kind"] != "none":
                nn._ck(L.gp5_nn_relayout(a["p"].ptr, a["wf"].ptr if a["wf"] else None, a["wb"].ptr if a["wb"] else None,
                                         a["ci"], a["co"]))
        segs = (nn.AdamSeg * len(got))(*[
            nn.AdamSeg(b["p"].ptr, b["g"].ptr, b["m"].ptr, b["v"].ptr, b["wf"].ptr if b["wf"] else None,
                       b["wb"].ptr if b["wb"] else None, b["n"], b["ci"], b["co"]) for b in got])
        nn._ck(L.gp5_nn_adam_multi(ctypes.cast(segs, ctypes.c_void_p), len(got), 3e-3, 0.9, 0.999, 1e-8, step))
    for a, b in zip(ref, got):
        for k in ("p", "m", "v"):
            np.testing.assert_array_equal(a[k].download(np.uint32, (a["n"],)), b[k].download(np.uint32, (a["n"],)))
        for k in ("wf", "wb"):
            if a[k] is not None:
                np.testing.assert_array_equal(a[k].download(np.uint16, (a["n"],)), b[k].download(np.uint16, (a["n"],)))


def _cos(a, b):
    a, b = a.ravel().astype(np.float64), b.ravel().astype(np.float64)
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))


@pytest.mark.gpu
def test_network_loss_and_gradients_match_oracle():
    rng = np.random.default_rng(11)
    B, D = 2, 16
    params = nn.init_params(seed=3)
    params["c8.W"] *= 10.0  # a non-trivial output for the check
    net = nn.UNet(B, D, params)
    X = rng.standard_normal((B, D, D, D, 16)).astype(np.float32) * 0.5
    base = rng.standard_normal((B, D, D, D)).astype(np.float32) * 0.1
    target = rng.standard_normal((B, D, D, D)).astype(np.float32)
    net.set_batch(X, base, target)
    net.forward()
    loss = net.backward()
    gp = net.grads()
    # oracle on the same bf16-rounded weights and inputs, float64 activations
    p64 = {k: (_rb(v) if k.endswith(".W") else v.astype(np.float64)) for k, v in params.items()}
    Lr, gr, yr = R.loss_and_grads(p64, _rb(X), base.astype(np.float64), target.astype(np.float64))
    y = net.output()
    assert np.linalg.norm(y - yr) <= 2e-2 * np.linalg.norm(yr)
    assert loss == pytest.approx(Lr, rel=2e-2)
    for name in gr:
        c = _cos(gp[name], gr[name])
        rel = np.linalg.norm(gp[name] - gr[name]) / (np.linalg.norm(gr[name]) + 1e-30)
        assert c > 0.995 and rel < 0.1, (name, c, rel)


@pytest.mark.gpu
def test_training_steps_reduce_the_loss():
    rng = np.random.default_rng(12)
    B, D = 2, 16
    net = nn.UNet(B, D, seed=5)
    X = rng.standard_normal((B, D, D, D, 16)).astype(np.float32) * 0.5
    base = np.zeros((B, D, D, D), dtype=np.float32)
    x = np.linspace(-1, 1, D)
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    target = np.broadcast_to(np.exp(-3 * r2), (B, D, D, D)).astype(np.float32)
    net.set_batch(X, base, target)
    losses = []
    for _ in range(40):
        net.forward()
        losses.append(net.backward())
        net.step(5e-4)
    assert min(losses[-5:]) < 0.3 * losses[0], losses


@pytest.mark.gpu
@pytest.mark.parametrize("B,D", [(1, 64), (2, 16)])
def test_side_stream_weight_gradients_change_nothing(B, D, monkeypatch):
    """The weight gradients on the side stream (GP5_NN_OVERLAP, on by default)
    give, bit for bit, the gradients and weights of the all-null-stream order over
    two training steps (64^3 takes the pooled data gradients, 16^3 the LDS-staged
    and side-8 kernels)."""
    rng = np.random.default_rng(B + D)
    X = rng.standard_normal((B, D, D, D, 16)).astype(np.float32) * 0.5
    base = rng.standard_normal((B, D, D, D)).astype(np.float32)
    target = (base + 0.1 * rng.standard_normal((B, D, D, D))).astype(np.float32)
    res = []
    for flag in ("0", "1"):
        monkeypatch.setenv("GP5_NN_OVERLAP", flag)
        net = nn.UNet(B, D, seed=4)
        net.set_batch(X, base, target)
        losses = []
        for _ in range(2):
            net.forward()
            losses.append(net.backward())
            net.step(5e-4)
        assert (getattr(net, "_aux", None) is not None) == (flag == "1")
        res.append((losses, net.grads(), net.params()))
        net._side_free()
    assert res[0][0] == res[1][0]
    for name in res[0][1]:
        assert np.array_equal(res[0][1][name], res[1][1][name]), name
    for key in res[0][2]:
        assert np.array_equal(res[0][2][key], res[1][2][key]), key


@pytest.mark.gpu
def test_device_loss_sum_equals_the_returned_losses():
    """backward(read_loss=False) (gp5_nn_mse_out0_acc) adds each loss to a device
    double in call order: after three training steps the sum equals, bit for bit,
    the Python sum of the losses that backward() returns for the same steps, and
    the weights come out the same."""
    rng = np.random.default_rng(21)
    B, D = 2, 16
    X = rng.standard_normal((B, D, D, D, 16)).astype(np.float32) * 0.5
    base = rng.standard_normal((B, D, D, D)).astype(np.float32)
    target = (base + 0.1 * rng.standard_normal((B, D, D, D))).astype(np.float32)
    nets = [nn.UNet(B, D, seed=3), nn.UNet(B, D, seed=3)]
    for net in nets:
        net.set_batch(X, base, target)
    total = 0.0
    for _ in range(3):
        for k, net in enumerate(nets):
            net.forward()
            if k == 0:
                total += net.backward()
            else:
                assert net.backward(read_loss=False) is None
            net.step(5e-4)
    assert nets[1].take_loss_sum() == total
    assert nets[1].take_loss_sum() == 0.0
    p0, p1 = nets[0].params(), nets[1].params()
    for key in p0:
        assert np.array_equal(p0[key], p1[key]), key


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B,exact", [(64, 16, 16, 1, True), (32, 32, 32, 1, True), (16, 32, 16, 2, True),
                                                (16, 128, 64, 2, False), (8, 64, 128, 1, False)])
def test_conv_lds_staged_path(D, Cin, Cout, B, exact, monkeypatch):
    """Cubes of side 8..128 (a power of two) run the kernel that stages each
    workgroup's input region in LDS.  Where both kernels take the same channel chunk
    (Cin = 16, or Cin = 32 on a cube of side 32 or less) they sum in the same order
    and agree bit for bit with the direct kernel (GP5_CONV_LDS=0), forward and
    mirrored (data gradient); every case agrees with the oracle."""
    rng = np.random.default_rng(D * Cin + Cout)
    X = _rb(rng.standard_normal((B, D, D, D, Cin)))
    W = _rb(rng.standard_normal((27, Cin, Cout)) * 0.1)
    b = rng.standard_normal(Cout).astype(np.float32)
    L = nn._lib()
    xb = _buf(X, True)
    wf, bb = _buf(np.ascontiguousarray(W.transpose(0, 2, 1)), True), _buf(b)
    shape = (B, D, D, D, Cout)
    monkeypatch.setenv("GP5_CONV_ZS", "0")
    outs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("GP5_CONV_LDS", flag)
        Yf = nn.DevBuf(B * D ** 3 * Cout * 4)
        A = nn.DevBuf(B * D ** 3 * Cout * 2)
        nn._ck(L.gp5_nn_conv3d(xb.ptr, wf.ptr, bb.ptr, A.ptr, None, None, Yf.ptr, B, D, Cin, Cout, 0, 1))
        outs.append((Yf.download(np.float32, shape), A.download(np.uint16, shape)))
    ref = R.lrelu(R.conv(X, W, b.astype(np.float64)))
    for got, Ab in outs:
        assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(ref))
        assert np.array_equal(Ab, nn.to_bf16(got))
    if exact:
        assert np.array_equal(outs[0][0], outs[1][0])
    # mirrored taps (data gradient through the [27][ci][co] copy), Cout -> Cin
    G = _rb(rng.standard_normal((B, D, D, D, Cout)))
    gb, wb = _buf(G, True), _buf(W, True)
    dref = R.conv_dgrad(G, W)
    douts = []
    for flag in ("1", "0"):
        monkeypatch.setenv("GP5_CONV_LDS", flag)
        dX = nn.DevBuf(B * D ** 3 * Cin * 4)
        nn._ck(L.gp5_nn_conv3d(gb.ptr, wb.ptr, None, None, None, None, dX.ptr, B, D, Cout, Cin, 1, 0))
        douts.append(dX.download(np.float32, dref.shape))
    for got in douts:
        assert np.max(np.abs(got - dref)) <= 1e-4 * np.max(np.abs(dref))
    if exact and Cout == Cin:
        assert np.array_equal(douts[0], douts[1])


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B", [(64, 16, 16, 1), (64, 32, 16, 1), (64, 16, 32, 1), (32, 32, 16, 8),
                                          (128, 16, 16, 1)])
def test_conv_z_sliding_path(D, Cin, Cout, B, monkeypatch):
    """16 or 32 input channels on cubes of side 16..128 with enough workgroups run
    the kernel that walks plane runs in z with a three-slab LDS ring and weights in
    registers.  It sums in the same order as the direct kernel with a chunk of Cin
    channels (GP5_CONV_ZS=0, GP5_CONV_LDS=0), so the two agree bit for bit, forward
    and mirrored; both agree with the oracle."""
    rng = np.random.default_rng(D + Cin * Cout + B)
    X = _rb(rng.standard_normal((B, D, D, D, Cin)))
    W = _rb(rng.standard_normal((27, Cin, Cout)) * 0.1)
    b = rng.standard_normal(Cout).astype(np.float32)
    G = _rb(rng.standard_normal((B, D, D, D, Cout)))
    L = nn._lib()
    xb, gb = _buf(X, True), _buf(G, True)
    wf, wb, bb = _buf(np.ascontiguousarray(W.transpose(0, 2, 1)), True), _buf(W, True), _buf(b)
    shape = (B, D, D, D, Cout)
    fwd, bwd = [], []
    for zs in ("1", "0"):
        monkeypatch.setenv("GP5_CONV_ZS", zs)
        monkeypatch.setenv("GP5_CONV_LDS", zs)
        Yf = nn.DevBuf(B * D ** 3 * Cout * 4)
        A = nn.DevBuf(B * D ** 3 * Cout * 2)
        nn._ck(L.gp5_nn_conv3d(xb.ptr, wf.ptr, bb.ptr, A.ptr, None, None, Yf.ptr, B, D, Cin, Cout, 0, 1))
        fwd.append((Yf.download(np.float32, shape), A.download(np.uint16, shape)))
        dX = nn.DevBuf(B * D ** 3 * Cin * 4)
        nn._ck(L.gp5_nn_conv3d(gb.ptr, wb.ptr, None, None, None, None, dX.ptr, B, D, Cout, Cin, 1, 0))
        bwd.append(dX.download(np.float32, (B, D, D, D, Cin)))
    assert np.array_equal(fwd[0][0], fwd[1][0]) and np.array_equal(fwd[0][1], fwd[1][1])
    ref = R.lrelu(R.conv(X, W, b.astype(np.float64)))
    assert np.max(np.abs(fwd[0][0] - ref)) <= 1e-4 * np.max(np.abs(ref))
    assert np.array_equal(fwd[0][1], nn.to_bf16(fwd[0][0]))
    dref = R.conv_dgrad(G, W)
    assert np.max(np.abs(bwd[0] - dref)) <= 1e-4 * np.max(np.abs(dref))
    if Cout in (16, 32):  # the mirrored call has Cout input channels: the same kernels and order
        assert np.array_equal(bwd[0], bwd[1])


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cout,B,up", [(64, 16, 1, 0), (16, 16, 2, 0), (32, 32, 1, 0), (64, 16, 1, 1)])
def test_conv_z_sliding_variants_agree(D, Cout, B, up, monkeypatch):
    """Every variant of the z-sliding kernel with 16 input channels (GP5_CONV_ZSV: bit
    0 wave-private epilogue barriers, bit 1 fragment reads one tap ahead) gives bit
    for bit what the direct kernel gives (forward with bias, activation and both
    outputs, and mirrored; up = 1 reads the half-side input upsampled, against the
    direct kernel on the upsampled input)."""
    Cin = 16
    rng = np.random.default_rng(D + Cout + B + up)
    Ds = D // 2 if up else D
    X = _rb(rng.standard_normal((B, Ds, Ds, Ds, Cin)))
    W = _rb(rng.standard_normal((27, Cin, Cout)) * 0.1)
    b = rng.standard_normal(Cout).astype(np.float32)
    L = nn._lib()
    U = X.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3) if up else X
    xb, ub = _buf(X, True), _buf(U, True)
    wf, wb, bb = _buf(np.ascontiguousarray(W.transpose(0, 2, 1)), True), _buf(W, True), _buf(b)
    shape = (B, D, D, D, Cout)
    outs = []
    for zs, v in (("0", "3"), ("1", "0"), ("1", "1"), ("1", "2"), ("1", "3")):
        monkeypatch.setenv("GP5_CONV_ZS", zs)
        monkeypatch.setenv("GP5_CONV_LDS", zs)
        monkeypatch.setenv("GP5_CONV_UP2P", "0")
        monkeypatch.setenv("GP5_CONV_ZSV", v)
        # the direct kernel (zs = 0) on the upsampled input itself
        f, x = (L.gp5_nn_conv3d_up, xb) if up and zs == "1" else (L.gp5_nn_conv3d, ub)
        Yf, A = nn.DevBuf(B * D ** 3 * Cout * 4), nn.DevBuf(B * D ** 3 * Cout * 2)
        nn._ck(f(x.ptr, wf.ptr, bb.ptr, A.ptr, None, None, Yf.ptr, B, D, Cin, Cout, 0, 1))
        got = [Yf.download(np.float32, shape), A.download(np.uint16, shape)]
        if not up and Cout == 16:  # mirrored: the data gradient of a 16 -> 16 layer
            dX = nn.DevBuf(B * D ** 3 * Cin * 4)
            nn._ck(L.gp5_nn_conv3d(xb.ptr, wb.ptr, None, None, None, None, dX.ptr, B, D, Cin, Cin, 1, 0))
            got.append(dX.download(np.float32, (B, D, D, D, Cin)))
        outs.append(got)
    for o in outs[1:]:
        for a, c in zip(outs[0], o):
            assert np.array_equal(a, c)
    ref = R.lrelu(R.conv(U, W, b.astype(np.float64)))
    assert np.max(np.abs(outs[1][0] - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B,up,mirror", [(8, 64, 128, 8, 0, 0), (8, 128, 64, 8, 0, 1),
                                                    (16, 64, 32, 2, 0, 1), (16, 128, 64, 1, 1, 0),
                                                    (8, 64, 32, 2, 0, 0)])
def test_conv_split_over_channel_chunks(D, Cin, Cout, B, up, mirror, monkeypatch):
    """Small cubes split the LDS-staged convolution over channel chunks (the splits'
    float sums added in order by a finishing pass, with bias, activation, skip and
    the bf16 outputs there).  Same results as the unsplit kernel (GP5_CONV_SPLIT=0)
    to float rounding, and as the oracle."""
    rng = np.random.default_rng(D * 5 + Cin + Cout + B + up)
    L = nn._lib()
    Ds = D // 2 if up else D
    X = _rb(rng.standard_normal((B, Ds, Ds, Ds, Cin)))
    W = _rb(rng.standard_normal((27, Cin, Cout)) * 0.1)
    b = rng.standard_normal(Cout).astype(np.float32)
    skip = _rb(rng.standard_normal((B, D, D, D, Cout)))
    xb, bb, sb = _buf(X, True), _buf(b), _buf(skip, True)
    # forward: the [27][co][ci] copy of W; data gradient (mirror): the [27][ci][co] layout of the layer
    # Wl = W^T (ci = this call's Cout), which is the same array
    wk = _buf(np.ascontiguousarray(W.transpose(0, 2, 1)), True)
    f = L.gp5_nn_conv3d_up if up else L.gp5_nn_conv3d
    shape = (B, D, D, D, Cout)
    outs = []
    for split in ("1", "0"):
        monkeypatch.setenv("GP5_CONV_SPLIT", split)
        Yf, A, Y = nn.DevBuf(B * D ** 3 * Cout * 4), nn.DevBuf(B * D ** 3 * Cout * 2), nn.DevBuf(B * D ** 3 * Cout * 2)
        if mirror:
            nn._ck(f(xb.ptr, wk.ptr, None, None, None, None, Yf.ptr, B, D, Cin, Cout, 1, 0))
        else:
            nn._ck(f(xb.ptr, wk.ptr, bb.ptr, A.ptr, sb.ptr, Y.ptr, Yf.ptr, B, D, Cin, Cout, 0, 1))
        outs.append((Yf.download(np.float32, shape), A.download(np.uint16, shape), Y.download(np.uint16, shape)))
    scale = np.max(np.abs(outs[1][0]))
    assert np.max(np.abs(outs[0][0] - outs[1][0])) <= 1e-6 * scale
    if mirror:
        # data gradient of a layer with Cout input channels: the oracle's conv_dgrad with W^T
        ref = R.conv_dgrad(X, np.ascontiguousarray(W.transpose(0, 2, 1)))
    else:
        U = X.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3) if up else X
        ref = R.lrelu(R.conv(U, W, b.astype(np.float64)))
        for k in (1, 2):  # bf16 outputs: within one rounding step of each other
            assert np.max(np.abs(nn.from_bf16(outs[0][k]) - nn.from_bf16(outs[1][k]))) <= 2.0 ** -7 * scale
    assert np.max(np.abs(outs[0][0] - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B", [(64, 16, 16, 2), (64, 32, 16, 1), (32, 64, 32, 1), (128, 16, 16, 1),
                                          (16, 128, 64, 2), (16, 16, 32, 1)])
def test_wgrad_z_sliding_path(D, Cin, Cout, B, monkeypatch):
    """Cube sides 16..128 take the weight gradient that stages channel-last X and dY
    into LDS while it walks plane runs (16x16x32 MFMAs).  It agrees with the oracle
    and with the kernel on the blocked channel-major copies (GP5_WGRAD_ZS=0), which
    sums in another order (float accumulation: agreement to rounding)."""
    rng = np.random.default_rng(D * 3 + Cin + 5 * Cout + B)
    X = _rb(rng.standard_normal((B, D, D, D, Cin)))
    G = _rb(rng.standard_normal((B, D, D, D, Cout)))
    L = nn._lib()
    xb, gb = _buf(X, True), _buf(G, True)
    grads = []
    for zs in ("1", "0"):
        monkeypatch.setenv("GP5_WGRAD_ZS", zs)
        ws = nn.DevBuf(L.gp5_nn_wgrad_ws(B, D, Cin, Cout))
        dW = nn.DevBuf(27 * Cin * Cout * 4)
        nn._ck(L.gp5_nn_conv3d_wgrad(xb.ptr, gb.ptr, dW.ptr, ws.ptr, B, D, Cin, Cout))
        grads.append(dW.download(np.float32, (27, Cin, Cout)))
    ref = R.conv_wgrad(X, G)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(grads[0] - ref)) <= 1e-4 * scale
    assert np.max(np.abs(grads[0] - grads[1])) <= 1e-5 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B", [(64, 32, 16, 1), (32, 64, 32, 1), (16, 128, 64, 2), (8, 64, 32, 2)])
def test_conv_reads_upsampled_input_in_place(D, Cin, Cout, B, monkeypatch):
    """gp5_nn_conv3d_up / gp5_nn_conv3d_wgrad_up on a cube of side D/2 give bit for
    bit what gp5_nn_conv3d / gp5_nn_conv3d_wgrad give on its nearest upsampling
    (the kernels that read the parent cell; the output-parity form, which sums in
    another order, is off here and has its own test)."""
    monkeypatch.setenv("GP5_CONV_UP2P", "0")
    monkeypatch.setenv("GP5_WGRAD_UP2", "0")
    rng = np.random.default_rng(D * 7 + Cin)
    Xh = _rb(rng.standard_normal((B, D // 2, D // 2, D // 2, Cin)))
    U = Xh.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)
    W = _rb(rng.standard_normal((27, Cin, Cout)) * 0.1)
    b = rng.standard_normal(Cout).astype(np.float32)
    G = _rb(rng.standard_normal((B, D, D, D, Cout)))
    L = nn._lib()
    xh, u, gb = _buf(Xh, True), _buf(U, True), _buf(G, True)
    wf, bb = _buf(np.ascontiguousarray(W.transpose(0, 2, 1)), True), _buf(b)
    shape = (B, D, D, D, Cout)
    outs = []
    for f, x in ((L.gp5_nn_conv3d_up, xh), (L.gp5_nn_conv3d, u)):
        Yf = nn.DevBuf(B * D ** 3 * Cout * 4)
        nn._ck(f(x.ptr, wf.ptr, bb.ptr, None, None, None, Yf.ptr, B, D, Cin, Cout, 0, 1))
        outs.append(Yf.download(np.float32, shape))
    assert np.array_equal(outs[0], outs[1])
    ref = R.lrelu(R.conv(U, W, b.astype(np.float64)))
    assert np.max(np.abs(outs[0] - ref)) <= 1e-4 * np.max(np.abs(ref))
    ws = nn.DevBuf(L.gp5_nn_wgrad_ws(B, D, Cin, Cout))
    grads = []
    for f, x in ((L.gp5_nn_conv3d_wgrad_up, xh), (L.gp5_nn_conv3d_wgrad, u)):
        dW = nn.DevBuf(27 * Cin * Cout * 4)
        nn._ck(f(x.ptr, gb.ptr, dW.ptr, ws.ptr, B, D, Cin, Cout))
        grads.append(dW.download(np.float32, (27, Cin, Cout)))
    assert np.array_equal(grads[0], grads[1])
    gref = R.conv_wgrad(U, G)
    assert np.max(np.abs(grads[0] - gref)) <= 1e-4 * np.max(np.abs(gref))


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B", [(64, 32, 16, 1), (32, 64, 32, 2), (16, 128, 64, 2), (16, 16, 32, 1),
                                          (128, 16, 16, 1), (32, 16, 48, 1)])
def test_upsampled_wgrad_by_output_parity(D, Cin, Cout, B, monkeypatch):
    """gp5_nn_conv3d_wgrad_up at D in {16, 32, 64, 128} sums, per fine parity a and
    coarse offset o, G_a[o] = sum_q X(q + o) dY(2q + a) and forms each tap from 8 of
    them (k_wgrad_up2).  It agrees with the oracle on the nearest upsampling to 1e-4,
    and to float rounding with the kernel that reads each fine position's parent
    (GP5_WGRAD_UP2=0) and with a single coarse plane run per cube."""
    rng = np.random.default_rng(D * 11 + Cin + Cout + B)
    Xh = _rb(rng.standard_normal((B, D // 2, D // 2, D // 2, Cin)))
    G = _rb(rng.standard_normal((B, D, D, D, Cout)))
    L = nn._lib()
    xh, gb = _buf(Xh, True), _buf(G, True)
    grads = []
    for up2, runs in (("1", None), ("0", None), ("1", "1")):
        monkeypatch.setenv("GP5_WGRAD_UP2", up2)
        if runs:
            monkeypatch.setenv("GP5_WU2_RUNS", runs)
        ws = nn.DevBuf(L.gp5_nn_wgrad_ws(B, D, Cin, Cout))
        dW = nn.DevBuf(27 * Cin * Cout * 4)
        nn._ck(L.gp5_nn_conv3d_wgrad_up(xh.ptr, gb.ptr, dW.ptr, ws.ptr, B, D, Cin, Cout))
        grads.append(dW.download(np.float32, (27, Cin, Cout)))
    scale = np.max(np.abs(grads[1]))
    assert np.max(np.abs(grads[0] - grads[1])) <= 1e-5 * scale
    assert np.max(np.abs(grads[2] - grads[1])) <= 1e-5 * scale
    if D <= 64:
        U = Xh.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)
        ref = R.conv_wgrad(U, G)
        assert np.max(np.abs(grads[0] - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("C", [16, 64])
def test_bias_gradient_entries(C):
    """gp5_nn_bias_grad, and gp5_nn_act_bwd_cast_bias (the leaky-ReLU backward with
    its bf16 copy and the bias gradient in one pass), against numpy."""
    rng = np.random.default_rng(C)
    P = 5000
    dA = rng.standard_normal((P, C)).astype(np.float32)
    A = nn.to_bf16(rng.standard_normal((P, C)).astype(np.float32))
    L = nn._lib()
    db = nn.DevBuf(C * 4)
    nn._ck(L.gp5_nn_bias_grad(_buf(dA).ptr, db.ptr, P, C))
    ref = dA.astype(np.float64).sum(axis=0)
    assert np.allclose(db.download(np.float32, (C,)), ref, rtol=1e-5, atol=1e-4)
    ab, gb = nn.DevBuf(A.nbytes), _buf(dA)
    ab.upload(A)
    H = nn.DevBuf(P * C * 2)
    nn._ck(L.gp5_nn_act_bwd_cast_bias(ab.ptr, gb.ptr, H.ptr, db.ptr, P, C))
    dP = dA * np.where(nn.from_bf16(A) > 0, np.float32(1.0), np.float32(0.1))
    assert np.array_equal(H.download(np.uint16, (P, C)), nn.to_bf16(dP))
    assert np.allclose(db.download(np.float32, (C,)), dP.astype(np.float64).sum(axis=0), rtol=1e-5, atol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B,up", [(64, 16, 16, 1, 0), (64, 32, 16, 1, 1), (32, 16, 32, 2, 0)])
def test_plane_run_counts_leave_results_unchanged(D, Cin, Cout, B, up, monkeypatch):
    """The number of plane runs per cube side of the z-sliding kernels (chosen from
    the occupancy, overridden by GP5_ZS_RUNS / GP5_WZ_RUNS) includes counts that do
    not divide the side (runs of unequal length).  The convolution sums each output
    in the same order whatever the count: bit for bit the same outputs; the pooled
    data gradient likewise over counts that pair planes; the weight gradient sums
    per-run partials (float: agreement to rounding, and to the oracle)."""
    rng = np.random.default_rng(D + Cin + 3 * Cout + B + up)
    L = nn._lib()
    S = D // 2 if up else D
    X = _rb(rng.standard_normal((B, S, S, S, Cin)))
    W = _rb(rng.standard_normal((27, Cin, Cout)) * 0.1)
    G = _rb(rng.standard_normal((B, D, D, D, Cout)))
    b = rng.standard_normal(Cout).astype(np.float32)
    xb, gb, bb = _buf(X, True), _buf(G, True), _buf(b)
    wf = _buf(np.ascontiguousarray(W.transpose(0, 2, 1)), True)
    conv = L.gp5_nn_conv3d_up if up else L.gp5_nn_conv3d
    shape = (B, D, D, D, Cout)
    outs, grads = [], []
    for runs in (None, "3", "5", "7", str(D // 2)):
        if runs is None:
            monkeypatch.delenv("GP5_ZS_RUNS", raising=False)
            monkeypatch.delenv("GP5_WZ_RUNS", raising=False)
        else:
            monkeypatch.setenv("GP5_ZS_RUNS", runs)
            monkeypatch.setenv("GP5_WZ_RUNS", runs)
        Yf = nn.DevBuf(B * D ** 3 * Cout * 4)
        nn._ck(conv(xb.ptr, wf.ptr, bb.ptr, None, None, None, Yf.ptr, B, D, Cin, Cout, 0, 1))
        outs.append(Yf.download(np.float32, shape))
        ws = nn.DevBuf(L.gp5_nn_wgrad_ws(B, D, Cin, Cout))
        dW = nn.DevBuf(27 * Cin * Cout * 4)
        wgrad = L.gp5_nn_conv3d_wgrad_up if up else L.gp5_nn_conv3d_wgrad
        nn._ck(wgrad(xb.ptr, gb.ptr, dW.ptr, ws.ptr, B, D, Cin, Cout))
        grads.append(dW.download(np.float32, (27, Cin, Cout)))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    U = X.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3) if up else X
    ref = R.conv_wgrad(U, G)
    scale = np.max(np.abs(ref))
    for g in grads:
        assert np.max(np.abs(g - ref)) <= 1e-4 * scale
    if Cin == 16 and Cout == 16 and not up:  # pooled data gradient (16 -> 32 channels of its layer)
        Wp = _rb(rng.standard_normal((27, 32, 16)) * 0.1)
        wp = _buf(Wp, True)
        pooled = []
        for runs in (None, "2", "4", "8"):
            if runs is None:
                monkeypatch.delenv("GP5_ZS_RUNS", raising=False)
            else:
                monkeypatch.setenv("GP5_ZS_RUNS", runs)
            h = nn.DevBuf(B * D ** 3 * 32 // 2)
            nn._ck(L.gp5_nn_conv3d_pool(gb.ptr, wp.ptr, h.ptr, B, D, 16, 32, 1))
            pooled.append(h.download(np.float32, (B, D // 2, D // 2, D // 2, 32)))
        for p in pooled[1:]:
            assert np.array_equal(p, pooled[0])


@pytest.mark.gpu
@pytest.mark.parametrize("B,D,Cin,Cout", [(2, 64, 16, 16), (1, 64, 32, 16), (2, 16, 16, 16), (1, 8, 64, 32)])
def test_conv_act_bwd_equals_the_two_passes(B, D, Cin, Cout):
    """gp5_nn_conv3d_act_bwd (a data gradient with the activation backward of the
    layer below in the z-sliding epilogue; the first two shapes take it, the others
    run the two passes) gives Yf and H bit for bit as gp5_nn_conv3d (mirrored)
    followed by gp5_nn_act_bwd_part, and the bias gradient from its partials to
    float rounding."""
    rng = np.random.default_rng(B + D + Cin + Cout)
    L = nn._lib()
    dY = _rb(rng.standard_normal((B, D, D, D, Cin)))
    W = _rb(rng.standard_normal((27, Cout, Cin)) * 0.1)  # [27][ci][co] of the layer: its co is Cin here
    Ab = _rb(rng.standard_normal((B, D, D, D, Cout)))
    gb, wb, ab = _buf(dY, True), _buf(W, True), _buf(Ab, True)
    P = B * D ** 3
    out = []
    for fused in (False, True):
        Yf, H = nn.DevBuf(P * Cout * 4), nn.DevBuf(P * Cout * 2)
        part, db = nn.DevBuf(L.gp5_nn_bias_part_floats(Cout) * 4), nn.DevBuf(Cout * 4)
        if fused:
            nn._ck(L.gp5_nn_conv3d_act_bwd(gb.ptr, wb.ptr, ab.ptr, Yf.ptr, H.ptr, part.ptr, B, D, Cin, Cout))
        else:
            nn._ck(L.gp5_nn_conv3d(gb.ptr, wb.ptr, None, None, None, None, Yf.ptr, B, D, Cin, Cout, 1, 0))
            nn._ck(L.gp5_nn_act_bwd_part(ab.ptr, Yf.ptr, None, 0.0, H.ptr, part.ptr, B, D, Cout))
        seg = (nn.BiasSeg * 1)(nn.BiasSeg(part.ptr, db.ptr, Cout))
        nn._ck(L.gp5_nn_bias_final_multi(ctypes.cast(seg, ctypes.c_void_p), 1))
        out.append((Yf.download(np.float32, (P, Cout)), H.download(np.uint16, (P, Cout)),
                    db.download(np.float32, (Cout,))))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(ab.download(np.uint16, Ab.shape), nn.to_bf16(Ab))  # Ab is only read
    ov = out[0][0] * np.where(Ab.reshape(P, Cout) > 0, 1.0, 0.1).astype(np.float32)
    ref = ov.astype(np.float64).sum(axis=0)
    tol = 1e-5 * np.abs(ov).astype(np.float64).sum(axis=0) + 1e-6
    assert np.all(np.abs(out[1][2] - ref) <= tol), (out[1][2], ref)
    assert np.all(np.abs(out[0][2] - ref) <= tol)


@pytest.mark.gpu
@pytest.mark.parametrize("B,D", [(1, 32), (2, 64), (1, 128), (2, 16)])
def test_output_gradient_as_channel0(B, D):
    """The output layer's gradient kept as channel 0 alone (one bf16 per position):
    gp5_nn_mse_out0c writes channel 0 of gp5_nn_mse_out0's rows with the same loss
    and bias gradient (bit for bit); gp5_nn_conv3d_wgrad_out0c gives
    gp5_nn_conv3d_wgrad_out0's result bit for bit (and refuses D = 16);
    gp5_nn_conv3d_act_bwd_out0 (the 27 taps as the MFMA's K) agrees with
    gp5_nn_conv3d_act_bwd on the full rows: its bf16 data gradient and H to one bf16
    step, the bias gradient from its partials to float rounding, the data gradient
    with the oracle to bf16 rounding."""
    rng = np.random.default_rng(B * 5 + D)
    L = nn._lib()
    P = B * D ** 3
    y0 = rng.standard_normal(P).astype(np.float32)
    base = rng.standard_normal(P).astype(np.float32)
    tgt = (base + rng.standard_normal(P)).astype(np.float32)
    yb, bb, tb = _buf(y0), _buf(base), _buf(tgt)
    H, H0 = nn.DevBuf(P * 16 * 2), nn.DevBuf(P * 2)
    db = [nn.DevBuf(16 * 4), nn.DevBuf(16 * 4)]
    loss = [ctypes.c_double(), ctypes.c_double()]
    nn._ck(L.gp5_nn_mse_out0(yb.ptr, bb.ptr, tb.ptr, H.ptr, 16, db[0].ptr, P, ctypes.byref(loss[0])))
    nn._ck(L.gp5_nn_mse_out0c(yb.ptr, bb.ptr, tb.ptr, H0.ptr, 16, db[1].ptr, P, ctypes.byref(loss[1]), None))
    hf, h0 = H.download(np.uint16, (P, 16)), H0.download(np.uint16, (P,))
    assert np.array_equal(hf[:, 0], h0) and not np.any(hf[:, 1:])
    assert loss[0].value == loss[1].value
    assert np.array_equal(db[0].download(np.float32, (16,)), db[1].download(np.float32, (16,)))
    X = _rb(rng.standard_normal((B, D, D, D, 16)))
    xb = _buf(X, True)
    ws = nn.DevBuf(L.gp5_nn_wgrad_ws(B, D, 16, 16))
    dW = [nn.DevBuf(27 * 256 * 4), nn.DevBuf(27 * 256 * 4)]
    if D == 16:
        assert L.gp5_nn_conv3d_wgrad_out0c(xb.ptr, H0.ptr, dW[1].ptr, ws.ptr, B, D, 16, 16) != 0
    else:
        nn._ck(L.gp5_nn_conv3d_wgrad_out0(xb.ptr, H.ptr, dW[0].ptr, ws.ptr, B, D, 16, 16))
        nn._ck(L.gp5_nn_conv3d_wgrad_out0c(xb.ptr, H0.ptr, dW[1].ptr, ws.ptr, B, D, 16, 16))
        assert np.array_equal(dW[0].download(np.float32, (27, 16, 16)), dW[1].download(np.float32, (27, 16, 16)))
    W = _rb(rng.standard_normal((27, 16, 16)) * 0.1)  # [27][ci][co] of the output layer
    Ab = _rb(rng.standard_normal((B, D, D, D, 16)))
    wb, ab = _buf(W, True), _buf(Ab, True)
    out = []
    for f, g, gbytes in ((L.gp5_nn_conv3d_act_bwd, H, 4), (L.gp5_nn_conv3d_act_bwd_out0, H0, 2)):
        Yf, Hb = nn.DevBuf(P * 16 * gbytes), nn.DevBuf(P * 16 * 2)
        part, dbb = nn.DevBuf(L.gp5_nn_bias_part_floats(16) * 4), nn.DevBuf(16 * 4)
        nn._ck(f(g.ptr, wb.ptr, ab.ptr, Yf.ptr, Hb.ptr, part.ptr, B, D, 16, 16))
        seg = (nn.BiasSeg * 1)(nn.BiasSeg(part.ptr, dbb.ptr, 16))
        nn._ck(L.gp5_nn_bias_final_multi(ctypes.cast(seg, ctypes.c_void_p), 1))
        yf = (Yf.download(np.float32, (P, 16)) if gbytes == 4
              else nn.from_bf16(Yf.download(np.uint16, (P, 16))))
        out.append((yf, nn.from_bf16(Hb.download(np.uint16, (P, 16))), dbb.download(np.float32, (16,))))
    scale = np.max(np.abs(out[0][0]))
    step = 2.0 ** -7 * np.abs(out[0][0]) + 1e-6 * scale  # one bf16 step of the float gradient
    assert np.all(np.abs(out[1][0] - out[0][0]) <= step)
    assert np.all(np.abs(out[1][1] - out[0][1]) <= 2.0 ** -7 * np.abs(out[0][1]) + 1e-6 * scale)
    ov = out[0][0].astype(np.float64) * np.where(Ab.reshape(P, 16) > 0, 1.0, 0.1)
    tol = 1e-5 * np.abs(ov).sum(axis=0) + 1e-6
    assert np.all(np.abs(out[1][2] - ov.sum(axis=0)) <= tol)
    if D <= 64:
        G = np.zeros((B, D, D, D, 16))
        G[..., 0] = nn.from_bf16(h0).reshape(B, D, D, D)
        ref = R.conv_dgrad(G, W).reshape(P, 16)
        assert np.all(np.abs(out[1][0] - ref) <= 2.0 ** -8 * np.abs(ref) + 1e-4 * np.max(np.abs(ref)))


@pytest.mark.gpu
@pytest.mark.parametrize("B,D,C", [(2, 16, 128), (1, 8, 6), (3, 6, 20)])
def test_pool2_f32_adds_the_children_in_order(B, D, C):
    """gp5_nn_pool2_f32 (four channels per thread when C is a multiple of 4, one
    otherwise) gives bit for bit the float32 sum of the 8 children in the order
    k = 0..7 (z, y, x bits) times 0.125 * scale."""
    rng = np.random.default_rng(B * 100 + D + C)
    L = nn._lib()
    X = rng.standard_normal((B, D, D, D, C)).astype(np.float32)
    xb, yb = _buf(X), nn.DevBuf(B * D ** 3 * C // 2)
    nn._ck(L.gp5_nn_pool2_f32(xb.ptr, yb.ptr, B, D, C, 8.0))
    got = yb.download(np.float32, (B, D // 2, D // 2, D // 2, C))
    s = np.zeros((B, D // 2, D // 2, D // 2, C), np.float32)
    for k in range(8):
        s = s + X[:, (k >> 2)::2, ((k >> 1) & 1)::2, (k & 1)::2, :]
    assert np.array_equal(got, s * np.float32(0.125 * 8.0))


@pytest.mark.gpu
@pytest.mark.parametrize("B,D,C", [(1, 32, 32), (2, 16, 64)])
def test_channel0_data_gradient_over_channel_blocks(B, D, C):
    """gp5_nn_conv3d_act_bwd_out0 with C data-gradient channels (C / 16 blocks of 16 as
    grid rows, each with its own bias-partial columns) against gp5_nn_conv3d_act_bwd on
    the full 16-channel rows: the bf16 data gradient and H to one bf16 step, the bias
    gradient from the partials to float rounding."""
    rng = np.random.default_rng(B + D + C + 3)
    L = nn._lib()
    P = B * D ** 3
    g0 = _rb(rng.standard_normal(P))
    Gf = np.zeros((P, 16))
    Gf[:, 0] = g0
    W = _rb(rng.standard_normal((27, C, 16)) * 0.1)  # [27][ci][co]: ci = C data-gradient channels
    Ab = _rb(rng.standard_normal((P, C)))
    g0b, gfb, wb, ab = _buf(g0, True), _buf(Gf, True), _buf(W, True), _buf(Ab, True)
    out = []
    for f, g, gbytes in ((L.gp5_nn_conv3d_act_bwd, gfb, 4), (L.gp5_nn_conv3d_act_bwd_out0, g0b, 2)):
        Yf, Hb = nn.DevBuf(P * C * gbytes), nn.DevBuf(P * C * 2)
        part, dbb = nn.DevBuf(L.gp5_nn_bias_part_floats(C) * 4), nn.DevBuf(C * 4)
        nn._ck(f(g.ptr, wb.ptr, ab.ptr, Yf.ptr, Hb.ptr, part.ptr, B, D, 16, C))
        seg = (nn.BiasSeg * 1)(nn.BiasSeg(part.ptr, dbb.ptr, C))
        nn._ck(L.gp5_nn_bias_final_multi(ctypes.cast(seg, ctypes.c_void_p), 1))
        yf = Yf.download(np.float32, (P, C)) if gbytes == 4 else nn.from_bf16(Yf.download(np.uint16, (P, C)))
        out.append((yf, nn.from_bf16(Hb.download(np.uint16, (P, C))), dbb.download(np.float32, (C,))))
    scale = np.max(np.abs(out[0][0]))
    assert np.all(np.abs(out[1][0] - out[0][0]) <= 2.0 ** -7 * np.abs(out[0][0]) + 1e-6 * scale)
    assert np.all(np.abs(out[1][1] - out[0][1]) <= 2.0 ** -7 * np.abs(out[0][1]) + 1e-6 * scale)
    ov = out[0][0].astype(np.float64) * np.where(Ab > 0, 1.0, 0.1)
    assert np.all(np.abs(out[1][2] - ov.sum(axis=0)) <= 1e-5 * np.abs(ov).sum(axis=0) + 1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("B,D,C", [(1, 64, 16), (2, 32, 32)])
def test_act_bwd_part_with_bf16_gradient(B, D, C):
    """gp5_nn_act_bwd_part_bf16 gives, for a bf16 dA, what gp5_nn_act_bwd_part gives
    for the same values in float (H and the bias partials bit for bit); it needs G."""
    rng = np.random.default_rng(B + D + C)
    L = nn._lib()
    P = B * D ** 3
    dA = _rb(rng.standard_normal((P, C)))
    A = _rb(rng.standard_normal((P, C)))
    G = rng.standard_normal((B * (D // 2) ** 3, C)).astype(np.float32)
    ab, gb = _buf(A, True), _buf(G)
    out = []
    for f, d in ((L.gp5_nn_act_bwd_part, _buf(dA)), (L.gp5_nn_act_bwd_part_bf16, _buf(dA, True))):
        H, part = nn.DevBuf(P * C * 2), nn.DevBuf(L.gp5_nn_bias_part_floats(C) * 4)
        nn._ck(f(ab.ptr, d.ptr, gb.ptr, 0.125, H.ptr, part.ptr, B, D, C))
        out.append((H.download(np.uint16, (P, C)), part.download(np.float32, (L.gp5_nn_bias_part_floats(C),))))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    H = nn.DevBuf(P * C * 2)
    part = nn.DevBuf(L.gp5_nn_bias_part_floats(C) * 4)
    assert L.gp5_nn_act_bwd_part_bf16(ab.ptr, ab.ptr, None, 0.125, H.ptr, part.ptr, B, D, C) != 0


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B,skipped", [(64, 32, 16, 1, True), (32, 32, 32, 2, False),
                                                  (128, 32, 16, 1, True), (32, 32, 16, 1, True),
                                                  (32, 64, 32, 2, True), (64, 64, 16, 1, False),
                                                  (128, 64, 16, 1, True), (16, 128, 64, 2, True),
                                                  (32, 128, 16, 1, False), (64, 128, 16, 1, True)])
def test_upsampled_conv_by_output_parity(D, Cin, Cout, B, skipped, monkeypatch):
    """gp5_nn_conv3d_up with 32 or 64 input channels at D = 32..128, or 128 at
    D = 16..64, runs the convolution of the upsampled input by output parity
    (k_conv_up2p: 2x2x2 coarse taps with folded weights held as bf16 hi + lo; 64 and
    128 channels in 32-channel parts added through LDS; at D = 16 a tile covers two
    coarse rows).  Its float output agrees with the other kernels on the upsampled
    input (GP5_CONV_UP2P=0) to float rounding, its bf16 outputs (A, and Y with the
    skip input) to one bf16 step, and with the oracle to 1e-4."""
    rng = np.random.default_rng(D + Cin + Cout + B)
    L = nn._lib()
    Xh = _rb(rng.standard_normal((B, D // 2, D // 2, D // 2, Cin)))
    W = _rb(rng.standard_normal((27, Cin, Cout)) * 0.1)
    b = rng.standard_normal(Cout).astype(np.float32)
    S = _rb(rng.standard_normal((B, D, D, D, Cout)))
    xh, wf, bb, sk = _buf(Xh, True), _buf(np.ascontiguousarray(W.transpose(0, 2, 1)), True), _buf(b), _buf(S, True)
    P = B * D ** 3
    outs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("GP5_CONV_UP2P", flag)
        Yf = nn.DevBuf(P * Cout * 4)
        nn._ck(L.gp5_nn_conv3d_up(xh.ptr, wf.ptr, bb.ptr, None, None, None, Yf.ptr, B, D, Cin, Cout, 0, 1))
        A, Y = nn.DevBuf(P * Cout * 2), nn.DevBuf(P * Cout * 2)
        nn._ck(L.gp5_nn_conv3d_up(xh.ptr, wf.ptr, bb.ptr, A.ptr, sk.ptr if skipped else None, Y.ptr, None, B, D, Cin,
                                  Cout, 0, 1))
        outs.append((Yf.download(np.float32, (P, Cout)), nn.from_bf16(A.download(np.uint16, (P, Cout))),
                     nn.from_bf16(Y.download(np.uint16, (P, Cout)))))
    scale = np.max(np.abs(outs[1][0]))
    assert np.max(np.abs(outs[0][0] - outs[1][0])) <= 1e-5 * scale
    for k in (1, 2):
        assert np.max(np.abs(outs[0][k] - outs[1][k])) <= 2.0 ** -7 * (scale + (np.max(np.abs(S)) if k == 2 else 0))
    if D <= 64:
        U = Xh.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)
        ref = R.lrelu(R.conv(U, W, b.astype(np.float64))).reshape(P, Cout)
        assert np.max(np.abs(outs[0][0] - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("D,Cin,Cout,B", [(64, 16, 32, 2), (32, 32, 64, 4), (128, 16, 32, 1), (32, 16, 32, 1),
                                          (64, 32, 32, 1), (64, 16, 64, 1)])
def test_pooled_data_gradient_as_stride2_conv(D, Cin, Cout, B):
    """gp5_nn_conv3d_pool with 16 or 32 gradient channels runs the pooled data
    gradient as a stride-2 convolution of dY with folded weights (k_conv_pool2p):
    it agrees with gp5_nn_conv3d (mirrored) followed by gp5_nn_pool2_f32 to float
    rounding and with the oracle to 1e-4."""
    rng = np.random.default_rng(D + 2 * Cin + Cout + B)
    L = nn._lib()
    G = _rb(rng.standard_normal((B, D, D, D, Cin)))
    W = _rb(rng.standard_normal((27, Cout, Cin)) * 0.1)  # [27][ci][co] of the layer: its co is Cin here
    gb, wb = _buf(G, True), _buf(W, True)
    shape = (B, D // 2, D // 2, D // 2, Cout)
    h = nn.DevBuf(B * D ** 3 * Cout // 2)
    nn._ck(L.gp5_nn_conv3d_pool(gb.ptr, wb.ptr, h.ptr, B, D, Cin, Cout, 1))
    got = h.download(np.float32, shape)
    full, two = nn.DevBuf(B * D ** 3 * Cout * 4), nn.DevBuf(B * D ** 3 * Cout // 2)
    nn._ck(L.gp5_nn_conv3d(gb.ptr, wb.ptr, None, None, None, None, full.ptr, B, D, Cin, Cout, 1, 0))
    nn._ck(L.gp5_nn_pool2_f32(full.ptr, two.ptr, B, D, Cout, 8.0))
    ref2 = two.download(np.float32, shape)
    assert np.max(np.abs(got - ref2)) <= 1e-5 * np.max(np.abs(ref2))
    if D <= 64:
        ref = 8.0 * R.pool2(R.conv_dgrad(G, W))
        assert np.max(np.abs(got - ref)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("B,D,Cout", [(2, 64, 16), (1, 32, 16), (1, 128, 16), (2, 16, 16), (1, 64, 32)])
def test_output_layer_weight_gradient_from_channel0(B, D, Cout):
    """gp5_nn_conv3d_wgrad_out0 (dY nonzero in channel 0 only, as gp5_nn_mse_out0
    writes it; the taps as the MFMA's N dimension at D = 32..128, the full weight
    gradient at D = 16) agrees with gp5_nn_conv3d_wgrad to float rounding and with the
    oracle to 1e-4; the other output channels' gradients are exactly 0."""
    Cin = 16
    rng = np.random.default_rng(B + D + Cout + 9)
    L = nn._lib()
    X = _rb(rng.standard_normal((B, D, D, D, Cin)))
    G = np.zeros((B, D, D, D, Cout))
    G[..., 0] = _rb(rng.standard_normal((B, D, D, D)))
    xb, gb = _buf(X, True), _buf(G, True)
    ws = nn.DevBuf(L.gp5_nn_wgrad_ws(B, D, Cin, Cout))
    outs = []
    for f in (L.gp5_nn_conv3d_wgrad_out0, L.gp5_nn_conv3d_wgrad):
        dW = nn.DevBuf(27 * Cin * Cout * 4)
        nn._ck(f(xb.ptr, gb.ptr, dW.ptr, ws.ptr, B, D, Cin, Cout))
        outs.append(dW.download(np.float32, (27, Cin, Cout)))
    scale = np.max(np.abs(outs[1]))
    assert np.max(np.abs(outs[0] - outs[1])) <= 1e-5 * scale
    assert not np.any(outs[0][:, :, 1:])
    if D <= 64:
        ref = R.conv_wgrad(X, G)
        assert np.max(np.abs(outs[0] - ref)) <= 1e-4 * np.max(np.abs(ref))
