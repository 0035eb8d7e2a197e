"""The warm-start network on the GPU: a 4-level 3D convolutional encoder-decoder
(16 -> 32 -> 64 -> 128 channels, bfloat16 weights and activations, float
accumulation and gradients), run and trained entirely by the kernels of
include/gp5_nn.h (libgp5.so).  No counterpart in the reference.

Graph (channels-last cubes of side D, D divisible by 8):
    e1 = lrelu(c1(X))            D     16
    e2 = lrelu(c2(pool(e1)))     D/2   32
    e3 = lrelu(c3(pool(e2)))     D/4   64
    e4 = lrelu(c4(pool(e3)))     D/8   128
    d3 = lrelu(c5(up(e4))) + e3  D/4   64
    d2 = lrelu(c6(up(d3))) + e2  D/2   32
    d1 = lrelu(c7(up(d2))) + e1  D     16
    y  = base + c8(d1)[..., 0]   D     (float)
Loss: mean squared error of y against the target field.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

from . import capi

SHAPES = [("c1", 16, 16), ("c2", 16, 32), ("c3", 32, 64), ("c4", 64, 128),
          ("c5", 128, 64), ("c6", 64, 32), ("c7", 32, 16), ("c8", 16, 16)]
C0 = 16  # input feature channels


def _pool2p(d: int, co: int) -> bool:
    """Whether gp5_nn_conv3d_pool runs as the stride-2 convolution of the gradient
    (k_conv_pool2p: co = 16 gradient channels at side 32..128, or 32 at 32..64;
    GP5_CONV_POOL2P=0 turns it off)."""
    return os.environ.get("GP5_CONV_POOL2P", "1") != "0" and (
        (co == 16 and d in (32, 64, 128)) or (co == 32 and d in (32, 64)))


def _pool_fused(d: int, ci: int, co: int, B: int) -> bool:
    """Whether the data gradient of the layer of side d that reads the upsampled
    input (ci input channels of the layer, co output) comes out already summed over
    each cell's children (gp5_nn_conv3d_pool): the stride-2 kernel's shapes, or the
    z-sliding kernel's (co = 16 gradient channels in with ci a multiple of 32 out,
    or co = 32; side 32, 64 or 128; enough workgroups)."""
    if _pool2p(d, co):
        return True
    if d not in (32, 64, 128) or not ((co == 16 and ci % 32 == 0) or co == 32):
        return False
    nbz = 2 if co == 16 else 1
    per_z = B * (d // (256 // d)) * (ci // 16 // nbz)
    return per_z * (d // 2) >= 512


def _up_fused(d: int) -> bool:
    """Whether the convolutions of side d read their upsampled input in place
    (gp5_nn_conv3d_up / gp5_nn_conv3d_wgrad_up: a power of two in [8, 128])."""
    return 8 <= d <= 128 and d & (d - 1) == 0


def _out0c(d: int) -> bool:
    """Whether the output layer's gradient is kept as channel 0 alone (one bf16 per
    position; gp5_nn_mse_out0c, gp5_nn_conv3d_wgrad_out0c, gp5_nn_conv3d_act_bwd_out0:
    cube sides 32, 64 and 128); GP5_OUT0_COMPACT=0 keeps the 16-channel rows."""
    return d in (32, 64, 128) and os.environ.get("GP5_OUT0_COMPACT", "1") != "0"


def _lib():
    L = capi.load()
    if getattr(L, "_gp5_nn_sigs", False):
        return L
    P, I, F, S, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t, ctypes.c_longlong
    sig = {
        "gp5_nn_malloc": [ctypes.POINTER(ctypes.c_void_p), S], "gp5_nn_free": [P],
        "gp5_nn_upload": [P, P, S], "gp5_nn_download": [P, P, S], "gp5_nn_zero": [P, S], "gp5_nn_sync": [],
        "gp5_nn_copy": [P, P, S],
        "gp5_nn_conv3d": [P, P, P, P, P, P, P, I, I, I, I, I, I],
        "gp5_nn_conv3d_wgrad": [P, P, P, P, I, I, I, I],
        "gp5_nn_bias_grad": [P, P, LL, I],
        "gp5_nn_pool2_bf16": [P, P, I, I, I, F], "gp5_nn_up2_bf16": [P, P, I, I, I, F],
        "gp5_nn_pool2_f32": [P, P, I, I, I, F], "gp5_nn_up2_f32": [P, P, I, I, I, F],
        "gp5_nn_act_bwd": [P, P, P, LL], "gp5_nn_add_f32": [P, P, LL], "gp5_nn_cast_bf16": [P, P, LL],
        "gp5_nn_act_bwd_cast": [P, P, P, P, LL], "gp5_nn_up2_add_f32": [P, P, I, I, I, F],
        "gp5_nn_conv3d_up": [P, P, P, P, P, P, P, I, I, I, I, I, I],
        "gp5_nn_conv3d_wgrad_up": [P, P, P, P, I, I, I, I],
        "gp5_nn_mse_bf16": [P, I, P, P, P, P, P, LL, ctypes.POINTER(ctypes.c_double)],
        "gp5_nn_mse_out0": [P, P, P, P, I, P, LL, ctypes.POINTER(ctypes.c_double)],
        "gp5_nn_mse_out0_acc": [P, P, P, P, I, P, LL, P],
        "gp5_nn_conv3d_out0": [P, P, P, P, I, I, I, I],
        "gp5_nn_conv3d_pool": [P, P, P, I, I, I, I, I],
        "gp5_nn_conv3d_act_bwd": [P, P, P, P, P, P, I, I, I, I],
        "gp5_nn_conv3d_wgrad_out0": [P, P, P, P, I, I, I, I],
        "gp5_nn_conv3d_wgrad_out0c": [P, P, P, P, I, I, I, I],
        "gp5_nn_conv3d_act_bwd_out0": [P, P, P, P, P, P, I, I, I, I],
        "gp5_nn_mse_out0c": [P, P, P, P, I, P, LL, ctypes.POINTER(ctypes.c_double), P],
        "gp5_nn_act_bwd_part": [P, P, P, F, P, P, I, I, I],
        "gp5_nn_bias_final_multi": [P, I],
        "gp5_nn_act_bwd_cast_bias": [P, P, P, P, LL, I],
        "gp5_nn_up2_add_act_bwd_cast_bias": [P, P, P, F, P, P, I, I, I],
        "gp5_nn_act_bwd_part_bf16": [P, P, P, F, P, P, I, I, I],
        "gp5_nn_mse": [P, I, P, P, P, LL, ctypes.POINTER(ctypes.c_double)],
        "gp5_nn_adam": [P, P, P, P, LL, F, F, F, F, I],
        "gp5_nn_relayout": [P, P, P, I, I],
        "gp5_nn_adam_multi": [P, I, F, F, F, F, I],
        "gp5_nn_droplet_features": [P, P, I, ctypes.c_double, F, P, P],
        "gp5_nn_predict_out": [P, I, P, P, LL],
        "gp5_nn_stream_create": [ctypes.POINTER(ctypes.c_void_p)], "gp5_nn_stream_destroy": [P],
        "gp5_nn_set_stream": [P], "gp5_nn_event_create": [ctypes.POINTER(ctypes.c_void_p)],
        "gp5_nn_event_destroy": [P], "gp5_nn_event_record": [P, P], "gp5_nn_stream_wait": [P, P],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    L.gp5_nn_wgrad_ws.argtypes = [I, I, I, I]
    L.gp5_nn_wgrad_ws.restype = S
    L.gp5_nn_bias_part_floats.argtypes = [I]
    L.gp5_nn_bias_part_floats.restype = S
    L._gp5_nn_sigs = True
    return L


def _ck(rc):
    capi.check(rc)


class AdamSeg(ctypes.Structure):
    """gp5_nn_adam_seg (include/gp5_nn.h)."""
    _fields_ = [("p", ctypes.c_void_p), ("g", ctypes.c_void_p), ("m", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("wf", ctypes.c_void_p), ("wb", ctypes.c_void_p), ("n", ctypes.c_longlong), ("cin", ctypes.c_int),
                ("cout", ctypes.c_int)]


class BiasSeg(ctypes.Structure):
    """gp5_nn_bias_seg (include/gp5_nn.h)."""
    _fields_ = [("part", ctypes.c_void_p), ("db", ctypes.c_void_p), ("C", ctypes.c_int)]


class DevBuf:
    """Device allocation of the NN kernels (freed with the object)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _ck(_lib().gp5_nn_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p

    def zero(self):
        _ck(_lib().gp5_nn_zero(self.ptr, self.nbytes))

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.nbytes} B into {self.nbytes} B")
        _ck(_lib().gp5_nn_upload(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes != self.nbytes:
            raise ValueError("download size mismatch")
        _ck(_lib().gp5_nn_download(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes))
        return out

    def __del__(self):
        try:
            if self.ptr:
                _lib().gp5_nn_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def to_bf16(a: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    u = u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))
    return (u >> np.uint32(16)).astype(np.uint16)


def from_bf16(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << np.uint32(16)).view(np.float32)


def init_params(seed: int = 0) -> Dict[str, np.ndarray]:
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co in SHAPES:
        p[name + ".W"] = (rng.standard_normal((27, ci, co)) * np.sqrt(2.0 / (27 * ci))).astype(np.float32)
        p[name + ".b"] = np.zeros(co, dtype=np.float32)
    p["c8.W"] *= 0.1  # the network starts close to the baseline field
    return p


class UNet:
    """Weights (float masters, Adam moments, bf16 copies) and the activations of
    one batch shape (B, D)."""

    _main = None  # the stream of every pass but the side stream's weight gradients (None: the null stream)

    def __init__(self, B: int, D: int, params: Optional[Dict[str, np.ndarray]] = None, seed: int = 0):
        if D % 8:
            raise ValueError("D must be divisible by 8")
        self.B, self.D = int(B), int(D)
        params = params if params is not None else init_params(seed)
        self.w = {}
        for name, ci, co in SHAPES:
            W = np.ascontiguousarray(params[name + ".W"], dtype=np.float32)
            b = np.ascontiguousarray(params[name + ".b"], dtype=np.float32)
            ent = {"ci": ci, "co": co}
            for key, arr in (("W", W), ("b", b)):
                buf = DevBuf(arr.nbytes)
                buf.upload(arr)
                ent[key] = buf
                for k2 in ("g", "m", "v"):
                    t = DevBuf(arr.nbytes)
                    t.zero()
                    ent[key + k2] = t
            ent["Wf"] = DevBuf(W.size * 2)
            ent["Wb"] = DevBuf(W.size * 2)
            self.w[name] = ent
        self.step_count = 0
        self._relayout()
        self._alloc()

    # --- memory -----------------------------------------------------------
    def _alloc(self):
        B, D = self.B, self.D
        n = lambda d, c: B * d ** 3 * c  # noqa: E731
        sizes = {  # name: (elements, bytes per element)
            "X": (n(D, C0), 2), "e1": (n(D, 16), 2), "p1": (n(D // 2, 16), 2), "e2": (n(D // 2, 32), 2),
            "p2": (n(D // 4, 32), 2), "e3": (n(D // 4, 64), 2), "p3": (n(D // 8, 64), 2), "e4": (n(D // 8, 128), 2),
            "u4": (n(D // 4, 128), 2), "a5": (n(D // 4, 64), 2), "d3": (n(D // 4, 64), 2),
            "u3": (n(D // 2, 64), 2), "a6": (n(D // 2, 32), 2), "d2": (n(D // 2, 32), 2),
            "u2": (n(D, 32), 2), "a7": (n(D, 16), 2), "d1": (n(D, 16), 2),
            "out": (n(D, 1), 4), "base": (n(D, 1), 4), "target": (n(D, 1), 4),
            # gradients (float) and their bf16 copies
            "h_out": (n(D, 1 if _out0c(D) else 16), 2),  # channel 0 alone, or rows of 16
            "g16a": (n(D, 16), 2 if _out0c(D) else 4), "h16": (n(D, 16), 2),  # bf16 from the channel-0 form
            "g32a": (n(D // 2, 32), 4), "h32": (n(D // 2, 32), 2),
            "g64a": (n(D // 4, 64), 4), "h64": (n(D // 4, 64), 2),
            "g128": (n(D // 8, 128), 4), "h128": (n(D // 8, 128), 2),
            # the encoder layers' bf16 gradients (their own buffers: the decoder
            # layers' weight gradients may still read h64 / h32 / h16)
            "h64e": (n(D // 4, 64), 2), "h32e": (n(D // 2, 32), 2), "h16e": (n(D, 16), 2),
            "gu32": (n(D, 32), 4), "gu64": (n(D // 2, 64), 4), "gu128": (n(D // 4, 128), 4),
            "gp16": (n(D // 2, 16), 4), "gp32": (n(D // 4, 32), 4), "gp64": (n(D // 8, 64), 4),
        }
        for name, d in (("u4", D // 4), ("u3", D // 2), ("u2", D)):
            if _up_fused(d):  # read in place from the half-side tensor
                del sizes[name]
        if _pool_fused(D, 32, 16, B):  # c7's and c6's data gradients arrive summed over the children
            del sizes["gu32"]
        if _pool_fused(D // 2, 64, 32, B):
            del sizes["gu64"]
        self._o0c = _out0c(D)
        self.t = {k: DevBuf(e * b) for k, (e, b) in sizes.items()}
        self.t["loss_acc"] = DevBuf(8)  # sum of the losses of backward(read_loss=False) calls
        self.t["loss_acc"].zero()
        L = _lib()
        ws = max(L.gp5_nn_wgrad_ws(B, d, ci, co) for d, ci, co in
                 [(D, C0, 16), (D // 2, 16, 32), (D // 4, 32, 64), (D // 8, 64, 128), (D // 4, 128, 64),
                  (D // 2, 64, 32), (D, 32, 16), (D, 16, 16)])
        self.ws = DevBuf(ws)
        # c1's weight gradient runs last, on the main stream, beside the side stream's
        # tail: its own workspace
        self.ws1 = DevBuf(L.gp5_nn_wgrad_ws(B, D, C0, 16))
        # bias-gradient block partials of c1..c7, finished together at the end of backward
        self.bpart = {name: DevBuf(L.gp5_nn_bias_part_floats(co) * 4) for name, ci, co in SHAPES if name != "c8"}
        self._bsegs = None

    def _relayout(self):
        L = _lib()
        for name, e in self.w.items():
            _ck(L.gp5_nn_relayout(e["W"].ptr, e["Wf"].ptr, e["Wb"].ptr, e["ci"], e["co"]))

    def params(self) -> Dict[str, np.ndarray]:
        out = {}
        for name, e in self.w.items():
            out[name + ".W"] = e["W"].download(np.float32, (27, e["ci"], e["co"]))
            out[name + ".b"] = e["b"].download(np.float32, (e["co"],))
        return out

    def grads(self) -> Dict[str, np.ndarray]:
        out = {}
        for name, e in self.w.items():
            out[name + ".W"] = e["Wg"].download(np.float32, (27, e["ci"], e["co"]))
            out[name + ".b"] = e["bg"].download(np.float32, (e["co"],))
        return out

    # --- passes -----------------------------------------------------------
    def load_batch(self, Xs, bases, targets, idx, offset_bytes=(None, None, None)):
        """Copies samples idx of device-resident arrays (bf16 features [n, D^3*16],
        float base and target [n, D^3]) into this network's batch buffers."""
        L = _lib()
        P = self.D ** 3
        for j, i in enumerate(idx):
            _ck(L.gp5_nn_copy(ctypes.c_void_p(self.t["X"].ptr.value + j * P * C0 * 2),
                              ctypes.c_void_p(Xs.ptr.value + int(i) * P * C0 * 2), P * C0 * 2))
            _ck(L.gp5_nn_copy(ctypes.c_void_p(self.t["base"].ptr.value + j * P * 4),
                              ctypes.c_void_p(bases.ptr.value + int(i) * P * 4), P * 4))
            _ck(L.gp5_nn_copy(ctypes.c_void_p(self.t["target"].ptr.value + j * P * 4),
                              ctypes.c_void_p(targets.ptr.value + int(i) * P * 4), P * 4))

    def set_batch(self, X: np.ndarray, base: np.ndarray, target: Optional[np.ndarray] = None):
        """X: [B, D, D, D, 16] float features; base/target: [B, D, D, D] float.  The
        features go up as float and are rounded to bf16 on the device (the same
        round-to-nearest-even as to_bf16)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        if "Xf" not in self.t:
            self.t["Xf"] = DevBuf(self.t["X"].nbytes * 2)
        self.t["Xf"].upload(X)
        _ck(_lib().gp5_nn_cast_bf16(self.t["Xf"].ptr, self.t["X"].ptr, X.size))
        self.t["base"].upload(np.asarray(base, dtype=np.float32))
        if target is not None:
            self.t["target"].upload(np.asarray(target, dtype=np.float32))

    def _conv(self, name, x, A=None, skip=None, Y=None, Yf=None, D=None, act=1, up=False):
        e = self.w[name]
        P = lambda b: b.ptr if b is not None else None  # noqa: E731
        f = _lib().gp5_nn_conv3d_up if up else _lib().gp5_nn_conv3d
        _ck(f(x.ptr, e["Wf"].ptr, e["b"].ptr, P(A), P(skip), P(Y), P(Yf), self.B, D, e["ci"], e["co"], 0, act))

    def _conv_up(self, name, src, tmp, D, **kw):
        """The convolution of the nearest upsampling of src (side D/2) at side D:
        read in place where the kernels support the side, else formed in tmp."""
        if _up_fused(D):
            self._conv(name, self.t[src], D=D, up=True, **kw)
        else:
            e = self.w[name]
            _ck(_lib().gp5_nn_up2_bf16(self.t[src].ptr, self.t[tmp].ptr, self.B, D // 2, e["ci"], 1.0))
            self._conv(name, self.t[tmp], D=D, **kw)

    def forward(self) -> None:
        L, t, B, D = _lib(), self.t, self.B, self.D
        self._conv("c1", t["X"], A=t["e1"], D=D)
        _ck(L.gp5_nn_pool2_bf16(t["e1"].ptr, t["p1"].ptr, B, D, 16, 1.0))
        self._conv("c2", t["p1"], A=t["e2"], D=D // 2)
        _ck(L.gp5_nn_pool2_bf16(t["e2"].ptr, t["p2"].ptr, B, D // 2, 32, 1.0))
        self._conv("c3", t["p2"], A=t["e3"], D=D // 4)
        _ck(L.gp5_nn_pool2_bf16(t["e3"].ptr, t["p3"].ptr, B, D // 4, 64, 1.0))
        self._conv("c4", t["p3"], A=t["e4"], D=D // 8)
        self._conv_up("c5", "e4", "u4", D // 4, A=t["a5"], skip=t["e3"], Y=t["d3"])
        self._conv_up("c6", "d3", "u3", D // 2, A=t["a6"], skip=t["e2"], Y=t["d2"])
        self._conv_up("c7", "d2", "u2", D, A=t["a7"], skip=t["e1"], Y=t["d1"])
        e = self.w["c8"]  # output layer: channel 0 of its float output
        _ck(L.gp5_nn_conv3d_out0(t["d1"].ptr, e["Wf"].ptr, e["b"].ptr, t["out"].ptr, B, D, e["ci"], e["co"]))

    def set_droplet_features(self, s: float, a, consts, ub: float, index: int = 0):
        """Features of one droplet built on the device (gp5_nn_droplet_features) as
        sample `index` of the batch: the flat-top baseline with surface factor s, trap
        weights a, the five constant channels consts and the scaled half-box ub; the
        baseline also becomes that sample's `base`."""
        if not 0 <= index < self.B:
            raise ValueError(f"sample {index} of a batch of {self.B}")
        cube = self.D ** 3
        X = ctypes.c_void_p(self.t["X"].ptr.value + index * cube * C0 * 2)
        base = ctypes.c_void_p(self.t["base"].ptr.value + index * cube * 4)
        af = (ctypes.c_float * 3)(*[float(v) for v in a])
        cf = (ctypes.c_float * 5)(*[float(v) for v in consts])
        _ck(_lib().gp5_nn_droplet_features(X, base, self.D, float(ub), float(s), af, cf))

    def output0(self) -> np.ndarray:
        """The prediction base + out (channel 0 of c8) as [B, D, D, D] float, combined on the device."""
        B, D = self.B, self.D
        if "y" not in self.t:
            self.t["y"] = DevBuf(B * D ** 3 * 4)
        _ck(_lib().gp5_nn_predict_out(self.t["out"].ptr, 1, self.t["base"].ptr, self.t["y"].ptr, B * D ** 3))
        return self.t["y"].download(np.float32, (B, D, D, D))

    def output(self) -> np.ndarray:
        B, D = self.B, self.D
        out = self.t["out"].download(np.float32, (B, D, D, D))
        return self.t["base"].download(np.float32, (B, D, D, D)) + out

    def _wgrad(self, name, x, dz_bf, D, ws=None):
        e = self.w[name]
        _ck(_lib().gp5_nn_conv3d_wgrad(x.ptr, dz_bf.ptr, e["Wg"].ptr, (ws or self.ws).ptr, self.B, D, e["ci"],
                                       e["co"]))

    def _wgrad_up(self, name, src, tmp, dz_bf, D):
        """Weight gradient of _conv_up: the input is the upsampling of src."""
        if _up_fused(D):
            e = self.w[name]
            _ck(_lib().gp5_nn_conv3d_wgrad_up(self.t[src].ptr, dz_bf.ptr, e["Wg"].ptr, self.ws.ptr, self.B, D,
                                              e["ci"], e["co"]))
        else:
            self._wgrad(name, self.t[tmp], dz_bf, D)

    def _act_bwd(self, name, a, g, h, D):
        """Backward of layer name's leaky ReLU from its output t[a] and the incoming
        gradient t[g] (float): t[h] = bf16 gradient of the pre-activation, and the
        block partials of the layer's bias gradient (finished by _bias_final), in one
        pass."""
        _ck(_lib().gp5_nn_act_bwd_part(self.t[a].ptr, self.t[g].ptr, None, 0.0, self.t[h].ptr,
                                       self.bpart[name].ptr, self.B, D, self.w[name]["co"]))

    # --- the side stream of the weight gradients ----------------------------
    def _side_setup(self):
        """A non-blocking side stream and two events per layer, made on first use
        (GP5_NN_OVERLAP=0 keeps every pass on the null stream)."""
        if getattr(self, "_aux", None) is not None or os.environ.get("GP5_NN_OVERLAP", "1") == "0":
            return getattr(self, "_aux", None) is not None
        L = _lib()
        aux = ctypes.c_void_p()
        _ck(L.gp5_nn_stream_create(ctypes.byref(aux)))
        self._aux, self._ev = aux, {}
        for name, _, _ in SHAPES:
            for kind in ("dz", "w"):
                ev = ctypes.c_void_p()
                _ck(L.gp5_nn_event_create(ctypes.byref(ev)))
                self._ev[name, kind] = ev
        return True

    def _fork(self, name):
        """Marks the point of the main stream after which layer name's weight
        gradient may run (its inputs are queued): _side(name, ...) waits for this
        point, so the main stream's next pass can be queued before the side
        stream's calls, which cost the host more than a launch."""
        if self._overlapping:
            _ck(_lib().gp5_nn_event_record(self._ev[name, "dz"], self._main))
            self._forked.add(name)

    def _side(self, name, *fns):
        """Queues fns (weight gradients) on the side stream behind the work queued
        on the main stream up to _fork(name) (or so far, without one), which has
        produced their inputs; the data-gradient passes on the main stream run
        beside them."""
        if not self._overlapping:
            for fn in fns:
                fn()
            return
        L = _lib()
        if name not in self._forked:
            _ck(L.gp5_nn_event_record(self._ev[name, "dz"], self._main))
        self._forked.discard(name)
        _ck(L.gp5_nn_stream_wait(self._aux, self._ev[name, "dz"]))
        _ck(L.gp5_nn_set_stream(self._aux))
        try:
            for fn in fns:
                fn()
            _ck(L.gp5_nn_event_record(self._ev[name, "w"], self._aux))
        finally:
            _ck(L.gp5_nn_set_stream(self._main))

    def _after_side(self, name):
        """Later work on the main stream waits for layer name's weight gradient
        (and the side stream's earlier work)."""
        if self._overlapping:
            _ck(_lib().gp5_nn_stream_wait(self._main, self._ev[name, "w"]))

    def _side_free(self):
        aux = getattr(self, "_aux", None)
        if aux is None:
            return
        L = _lib()
        L.gp5_nn_sync()
        for ev in self._ev.values():
            L.gp5_nn_event_destroy(ev)
        L.gp5_nn_stream_destroy(aux)
        self._aux = None

    def __del__(self):
        try:
            self._side_free()
        except Exception:
            pass

    def _up_add_act_bwd(self, name, a, g, gh, h, D, g_bf16=False):
        """_act_bwd with the incoming gradient t[g] + t[gh] / 8 at the parent cell
        (t[gh] of side D/2: the adjoint of the pooling that fed the next level), in
        one pass that does not write t[g] back (t[g] in bf16 with g_bf16)."""
        f = _lib().gp5_nn_act_bwd_part_bf16 if g_bf16 else _lib().gp5_nn_act_bwd_part
        _ck(f(self.t[a].ptr, self.t[g].ptr, self.t[gh].ptr, 0.125, self.t[h].ptr, self.bpart[name].ptr, self.B, D,
              self.w[name]["co"]))

    def _bias_final(self):
        """The bias gradients of c1..c7 from their block partials, one launch."""
        if self._bsegs is None:
            segs = [BiasSeg(self.bpart[n].ptr, self.w[n]["bg"].ptr, self.w[n]["co"]) for n in self.bpart]
            self._bsegs = (BiasSeg * len(segs))(*segs)
        _ck(_lib().gp5_nn_bias_final_multi(ctypes.cast(self._bsegs, ctypes.c_void_p), len(self._bsegs)))

    def _dgrad(self, name, dz_bf, out_f, D):
        # data gradient: forward kernel on the [27][ci][co] copy with mirrored taps
        e = self.w[name]
        _ck(_lib().gp5_nn_conv3d(dz_bf.ptr, e["Wb"].ptr, None, None, None, None, out_f.ptr, self.B, D, e["co"],
                                 e["ci"], 1, 0))

    def backward(self, read_loss: bool = True) -> Optional[float]:
        """MSE loss against the target; fills every weight and bias gradient.
        Returns the loss (one device-to-host copy, which waits for the forward
        pass), or with read_loss=False adds it to a device sum that take_loss_sum()
        reads, and returns None without synchronising."""
        L, t, B, D = _lib(), self.t, self.B, self.D
        loss = ctypes.c_double()
        # loss, the bf16 output gradient and c8's bias gradient in one pass
        if self._o0c:
            _ck(L.gp5_nn_mse_out0c(t["out"].ptr, t["base"].ptr, t["target"].ptr, t["h_out"].ptr, 16,
                                   self.w["c8"]["bg"].ptr, B * D ** 3, ctypes.byref(loss) if read_loss else None,
                                   None if read_loss else t["loss_acc"].ptr))
        elif read_loss:
            _ck(L.gp5_nn_mse_out0(t["out"].ptr, t["base"].ptr, t["target"].ptr, t["h_out"].ptr, 16,
                                  self.w["c8"]["bg"].ptr, B * D ** 3, ctypes.byref(loss)))
        else:
            _ck(L.gp5_nn_mse_out0_acc(t["out"].ptr, t["base"].ptr, t["target"].ptr, t["h_out"].ptr, 16,
                                      self.w["c8"]["bg"].ptr, B * D ** 3, t["loss_acc"].ptr))
        # Weight gradients run on a side stream beside the data-gradient chain
        # (_side), each as soon as its dz is ready (grouping them, fewer forks,
        # measured slower); nothing on the main stream overwrites what they read
        # before the join at the end.
        self._overlapping = self._side_setup()
        self._forked = set()
        # Each weight gradient is queued on the side stream after the main stream's
        # next pass (_fork marks where its inputs are ready), so that the side
        # stream's host calls do not hold up the data-gradient chain.
        # c8, c7
        e8 = self.w["c8"]  # h_out is nonzero in channel 0 only (or holds channel 0 alone)
        wg8 = L.gp5_nn_conv3d_wgrad_out0c if self._o0c else L.gp5_nn_conv3d_wgrad_out0
        self._fork("c8")
        # c8's data gradient g16a = dd1 = de1 (skip part; bf16 in the channel-0 form) and c7's activation backward
        # dz7 = dd1 * act'(a7) (h16, c7's bias partials) in one pass where the shape allows
        e = self.w["c8"]
        act_bwd8 = L.gp5_nn_conv3d_act_bwd_out0 if self._o0c else L.gp5_nn_conv3d_act_bwd
        _ck(act_bwd8(t["h_out"].ptr, e["Wb"].ptr, t["a7"].ptr, t["g16a"].ptr, t["h16"].ptr,
                                    self.bpart["c7"].ptr, B, D, e["co"], e["ci"]))
        self._side("c8", lambda: _ck(wg8(t["d1"].ptr, t["h_out"].ptr, e8["Wg"].ptr, self.ws.ptr, B, D, e8["ci"],
                                         e8["co"])))
        self._fork("c7")
        if _pool_fused(D, 32, 16, B):  # gradient of u2 summed over each cell's children: dd2 = de2 (skip part)
            e = self.w["c7"]
            _ck(L.gp5_nn_conv3d_pool(t["h16"].ptr, e["Wb"].ptr, t["g32a"].ptr, B, D, e["co"], e["ci"], 1))
        else:
            self._dgrad("c7", t["h16"], t["gu32"], D)               # gradient of u2 (side D, 32 ch)
            _ck(L.gp5_nn_pool2_f32(t["gu32"].ptr, t["g32a"].ptr, B, D, 32, 8.0))  # dd2 = de2 (skip part)
        self._side("c7", lambda: self._wgrad_up("c7", "d2", "u2", t["h16"], D))
        # c6
        self._act_bwd("c6", "a6", "g32a", "h32", D // 2)
        self._fork("c6")
        if _pool_fused(D // 2, 64, 32, B):  # dd3 = de3 (skip part)
            e = self.w["c6"]
            _ck(L.gp5_nn_conv3d_pool(t["h32"].ptr, e["Wb"].ptr, t["g64a"].ptr, B, D // 2, e["co"], e["ci"], 1))
        else:
            self._dgrad("c6", t["h32"], t["gu64"], D // 2)
            _ck(L.gp5_nn_pool2_f32(t["gu64"].ptr, t["g64a"].ptr, B, D // 2, 64, 8.0))  # dd3 = de3 (skip part)
        self._side("c6", lambda: self._wgrad_up("c6", "d3", "u3", t["h32"], D // 2))
        # c5
        self._act_bwd("c5", "a5", "g64a", "h64", D // 4)
        self._fork("c5")
        self._dgrad("c5", t["h64"], t["gu128"], D // 4)
        _ck(L.gp5_nn_pool2_f32(t["gu128"].ptr, t["g128"].ptr, B, D // 4, 128, 8.0))  # de4
        self._side("c5", lambda: self._wgrad_up("c5", "e4", "u4", t["h64"], D // 4))
        # c4
        self._act_bwd("c4", "e4", "g128", "h128", D // 8)
        self._fork("c4")
        self._dgrad("c4", t["h128"], t["gp64"], D // 8)
        self._side("c4", lambda: self._wgrad("c4", t["p3"], t["h128"], D // 8))
        # c3 (de3 total = dd3 + the adjoint of the pooling into c4)
        self._up_add_act_bwd("c3", "e3", "g64a", "gp64", "h64e", D // 4)
        self._fork("c3")
        self._dgrad("c3", t["h64e"], t["gp32"], D // 4)
        self._side("c3", lambda: self._wgrad("c3", t["p2"], t["h64e"], D // 4))
        # c2 (de2 total)
        self._up_add_act_bwd("c2", "e2", "g32a", "gp32", "h32e", D // 2)
        self._fork("c2")
        self._dgrad("c2", t["h32e"], t["gp16"], D // 2)
        self._side("c2", lambda: self._wgrad("c2", t["p1"], t["h32e"], D // 2))
        # c1 (de1 total)
        self._up_add_act_bwd("c1", "e1", "g16a", "gp16", "h16e", D, g_bf16=self._o0c)
        # c1's weight gradient: on the main stream (nothing else is left to overlap it
        # with there, and a fork and join would only add their latency), in its own
        # workspace while the side stream may still run c2's; the bias gradients (every
        # partial is written once c1's activation backward is) beside it on the side stream
        self._fork("c1")
        self._wgrad("c1", t["X"], t["h16e"], D, ws=self.ws1)
        self._side("c1", self._bias_final)
        self._after_side("c1")  # every gradient is done before anything reads them
        return float(loss.value) if read_loss else None

    def take_loss_sum(self) -> float:
        """The sum of the losses of the backward(read_loss=False) calls since the
        last take (in call order, double), and a fresh zero sum."""
        v = float(self.t["loss_acc"].download(np.float64, (1,))[0])
        self.t["loss_acc"].zero()
        return v

    def _adam_segs(self):
        if getattr(self, "_segs", None) is None:
            segs = []
            for name, e in self.w.items():
                for key, n in (("W", 27 * e["ci"] * e["co"]), ("b", e["co"])):
                    w = key == "W"
                    segs.append(AdamSeg(e[key].ptr, e[key + "g"].ptr, e[key + "m"].ptr, e[key + "v"].ptr,
                                        e["Wf"].ptr if w else None, e["Wb"].ptr if w else None, n,
                                        e["ci"] if w else 0, e["co"] if w else 0))
            self._segs = (AdamSeg * len(segs))(*segs)
        return self._segs

    def step(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        """Adam update of the float masters and fresh bf16 copies, every tensor in one
        launch (gp5_nn_adam_multi: the arithmetic of gp5_nn_adam and gp5_nn_relayout)."""
        self.step_count += 1
        segs = self._adam_segs()
        _ck(_lib().gp5_nn_adam_multi(ctypes.cast(segs, ctypes.c_void_p), len(segs), lr, b1, b2, eps, self.step_count))

    def sync(self):
        _ck(_lib().gp5_nn_sync())
