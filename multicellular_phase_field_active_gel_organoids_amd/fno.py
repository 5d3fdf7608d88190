"""Python side of the FNO velocity surrogate (include/organoid_fno.h): the
network object over the gfx950 library, its initial weights, training on
(phi_n, rho_n) -> (vx, vy) pairs of the exact spectral solve, and the data
set those pairs come from.

Device buffers are torch tensors on the GPU (torch only allocates them and
orders the stream); every operation on them runs in the library's HIP
kernels.  There is no CPU path: without the built library these functions
raise.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import host

__all__ = ["FnoConfig", "FnoNet", "layout", "init_params", "fit_scales", "train", "TILES", "STAGES"]


# GEMM tile shapes and network stages (FNO_TILE_*, FNO_STAGE_* in include/organoid_fno.h)
TILES = ("64x64", "32x64", "64x32", "32x32")
STAGES = ("y", "x", "mix", "ix", "layer", "proj", "bwd_spectral", "bwd_pointwise", "wgrad", "api")
STAGE_COUNT = len(STAGES)


@dataclass
class FnoConfig:
    M: int = 128
    width: int = 32
    modes: int = 12
    layers: int = 4
    proj: int = 128
    batch_max: int = 64


class _Cfg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("M", "width", "modes", "layers", "proj", "batch_max")]


class GemmArgs(ctypes.Structure):
    """fno_gemm_args (include/organoid_fno.h)."""
    _fields_ = [("A", ctypes.c_void_p), ("sAm", ctypes.c_int64), ("sAk", ctypes.c_int64),
                ("sA1", ctypes.c_int64), ("sA2", ctypes.c_int64),
                ("B", ctypes.c_void_p), ("sBk", ctypes.c_int64), ("sBn", ctypes.c_int64),
                ("sB1", ctypes.c_int64), ("sB2", ctypes.c_int64),
                ("C", ctypes.c_void_p), ("sCm", ctypes.c_int64), ("sCn", ctypes.c_int64),
                ("sC1", ctypes.c_int64), ("sC2", ctypes.c_int64),
                ("C2", ctypes.c_void_p), ("bias", ctypes.c_void_p),
                ("M", ctypes.c_int32), ("N", ctypes.c_int32), ("K", ctypes.c_int32),
                ("nz1", ctypes.c_int32), ("nz2", ctypes.c_int32),
                ("a_f32", ctypes.c_int32), ("b_f32", ctypes.c_int32), ("c_mode", ctypes.c_int32),
                ("act", ctypes.c_int32), ("alpha", ctypes.c_float)]


def _lib():
    lib = host.load_library()
    if getattr(lib, "_fno_ready", False):
        return lib
    vp = ctypes.c_void_p
    lib.fno_create.argtypes = [ctypes.POINTER(_Cfg), ctypes.c_int, ctypes.POINTER(vp)]
    lib.fno_destroy.argtypes = [vp]
    lib.fno_param_count.argtypes = [vp]
    lib.fno_param_count.restype = ctypes.c_size_t
    lib.fno_set_params.argtypes = [vp, vp, ctypes.c_size_t]
    lib.fno_get_params.argtypes = [vp, vp, ctypes.c_size_t]
    lib.fno_set_scales.argtypes = [vp, vp, vp]
    lib.fno_set_stream.argtypes = [vp, vp]
    lib.fno_forward.argtypes = [vp, vp, ctypes.c_int, vp]
    lib.fno_loss_grad.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), vp]
    lib.fno_adam_step.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    lib.fno_gemm.argtypes = [vp, ctypes.POINTER(GemmArgs)]
    lib.fno_set_tiles.argtypes = [vp, vp]
    lib.fno_get_tiles.argtypes = [vp, vp]
    lib._fno_ready = True
    return lib


def layout(w: int, m: int, layers: int, proj: int) -> list:
    """(name, offset, shape) of the flat parameter array, header order."""
    nm = 2 * m * m
    out, off = [], 0
    shapes = [("P", (w, 4)), ("bP", (w,))]
    for l in range(layers):
        shapes += [(f"Rr{l}", (nm, w, w)), (f"Ri{l}", (nm, w, w)), (f"W{l}", (w, w)), (f"b{l}", (w,))]
    shapes += [("Q1", (proj, w)), ("c1", (proj,)), ("Q2", (2, proj)), ("c2", (2,))]
    for name, s in shapes:
        out.append((name, off, s))
        off += int(np.prod(s))
    return out


def init_params(cfg: FnoConfig, seed: int) -> np.ndarray:
    """Initial weights: linear layers U(-1/sqrt(fan_in), 1/sqrt(fan_in)),
    spectral weights U(0, 1/(w*w)) for real and imaginary parts (the FNO
    paper's initialisation)."""
    rng = np.random.default_rng(seed)
    lay = layout(cfg.width, cfg.modes, cfg.layers, cfg.proj)
    flat = np.zeros(lay[-1][1] + int(np.prod(lay[-1][2])), np.float32)
    fan = {"P": 4, "bP": 4, "Q1": cfg.width, "c1": cfg.width, "Q2": cfg.proj, "c2": cfg.proj}
    for name, o, s in lay:
        n = int(np.prod(s))
        if name[:2] in ("Rr", "Ri"):
            flat[o:o + n] = rng.uniform(0.0, 1.0 / (cfg.width * cfg.width), n)
        else:
            flat[o:o + n] = rng.uniform(-1.0, 1.0, n) / np.sqrt(fan.get(name, cfg.width))
    return flat


class FnoNet:
    """The surrogate network on the GPU (library-owned weights, Adam state
    and workspace sized for cfg.batch_max)."""

    def __init__(self, cfg: FnoConfig = FnoConfig(), device: int = 0):
        self.lib = _lib()
        self.cfg = cfg
        c = _Cfg(cfg.M, cfg.width, cfg.modes, cfg.layers, cfg.proj, cfg.batch_max)
        h = ctypes.c_void_p()
        host._chk(self.lib.fno_create(ctypes.byref(c), device, ctypes.byref(h)), "fno_create")
        self.h = h
        import torch
        host._chk(self.lib.fno_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                  "fno_set_stream")
        self.count = int(self.lib.fno_param_count(self.h))
        self.in_scale = np.ones(2, np.float32)
        self.out_scale = np.ones(2, np.float32)
        self.t = 0

    def close(self):
        if self.h:
            self.lib.fno_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat: np.ndarray) -> None:
        a = np.ascontiguousarray(flat, np.float32)
        host._chk(self.lib.fno_set_params(self.h, a.ctypes.data_as(ctypes.c_void_p), a.size), "fno_set_params")

    def get_params(self) -> np.ndarray:
        a = np.empty(self.count, np.float32)
        host._chk(self.lib.fno_get_params(self.h, a.ctypes.data_as(ctypes.c_void_p), a.size), "fno_get_params")
        return a

    def set_scales(self, in_scale, out_scale) -> None:
        self.in_scale = np.asarray(in_scale, np.float32).copy()
        self.out_scale = np.asarray(out_scale, np.float32).copy()
        host._chk(self.lib.fno_set_scales(self.h, self.in_scale.ctypes.data_as(ctypes.c_void_p),
                                          self.out_scale.ctypes.data_as(ctypes.c_void_p)), "fno_set_scales")

    def _check(self, t, B):
        import torch
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise ValueError("expected a contiguous float32 CUDA tensor")
        if tuple(t.shape) != (B, 2, self.cfg.M, self.cfg.M):
            raise ValueError(f"expected shape {(B, 2, self.cfg.M, self.cfg.M)}, got {tuple(t.shape)}")

    def forward(self, x, out=None):
        """(phi, rho) crops [B, 2, M, M] -> (vx, vy) [B, 2, M, M] (CUDA tensors)."""
        import torch
        B = x.shape[0]
        self._check(x, B)
        if out is None:
            out = torch.empty_like(x)
        self._check(out, B)
        host._chk(self.lib.fno_forward(self.h, ctypes.c_void_p(x.data_ptr()), B,
                                       ctypes.c_void_p(out.data_ptr())), "fno_forward")
        return out

    def loss_grad(self, x, target, want_grad: bool = False):
        B = x.shape[0]
        self._check(x, B)
        self._check(target, B)
        loss = ctypes.c_double()
        g = np.empty(self.count, np.float32) if want_grad else None
        host._chk(self.lib.fno_loss_grad(self.h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(target.data_ptr()),
                                         B, ctypes.byref(loss),
                                         g.ctypes.data_as(ctypes.c_void_p) if want_grad else None),
                  "fno_loss_grad")
        return (loss.value, g) if want_grad else loss.value

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.t += 1
        host._chk(self.lib.fno_adam_step(self.h, lr, beta1, beta2, eps, self.t), "fno_adam_step")

    def get_tiles(self) -> list:
        t = np.zeros(STAGE_COUNT, np.int32)
        host._chk(self.lib.fno_get_tiles(self.h, t.ctypes.data_as(ctypes.c_void_p)), "fno_get_tiles")
        return [int(v) for v in t]

    def set_tiles(self, tiles) -> None:
        t = np.ascontiguousarray(tiles, np.int32)
        if t.size != STAGE_COUNT:
            raise ValueError(f"{STAGE_COUNT} tile choices expected")
        host._chk(self.lib.fno_set_tiles(self.h, t.ctypes.data_as(ctypes.c_void_p)), "fno_set_tiles")

    def gemm(self, args: GemmArgs) -> None:
        host._chk(self.lib.fno_gemm(self.h, ctypes.byref(args)), "fno_gemm")


def fit_scales(x: np.ndarray, y: np.ndarray):
    """Input scales 1/std of (phi, rho) and output scales std of (vx, vy)
    over a data set ([S, 2, M, M] each)."""
    sx = np.array([1.0 / max(float(np.std(x[:, c])), 1e-12) for c in range(2)], np.float32)
    sy = np.array([max(float(np.std(y[:, c])), 1e-12) for c in range(2)], np.float32)
    return sx, sy


def train(net: FnoNet, x, y, steps: int, batch: int, lr: float = 1e-3, seed: int = 0, lr_final: float | None = None,
          log=None):
    """Adam on random mini-batches of the (CUDA tensor) data set x -> y
    ([S, 2, M, M] each); learning rate decays linearly to lr_final.  Returns
    the loss of every step."""
    import torch
    rng = np.random.default_rng(seed)
    S = x.shape[0]
    batch = min(batch, S, net.cfg.batch_max)
    lr_final = lr if lr_final is None else lr_final
    xb = torch.empty((batch,) + tuple(x.shape[1:]), dtype=torch.float32, device=x.device)
    yb = torch.empty_like(xb)
    losses = []
    for s in range(steps):
        idx = torch.as_tensor(rng.choice(S, batch, replace=False), device=x.device)
        torch.index_select(x, 0, idx, out=xb)
        torch.index_select(y, 0, idx, out=yb)
        losses.append(net.loss_grad(xb, yb))
        frac = s / max(steps - 1, 1)
        net.adam_step(lr + (lr_final - lr) * frac)
        if log is not None and (s % 100 == 0 or s == steps - 1):
            log(f"step {s}: loss {losses[-1]:.4e}")
    return losses
