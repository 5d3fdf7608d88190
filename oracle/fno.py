"""FP64 restatement of the Fourier-neural-operator velocity surrogate
(include/organoid_fno.h), forward and hand-derived backward.  TEST
INFRASTRUCTURE: only tests/ import it, as the checker of the GPU kernels.

The network is the one the header states (Li et al. 2020, FNO2d): lift
4 -> w, `layers` Fourier layers u = K(v) + W v + b (GELU between layers, none
after the last), projection w -> proj -> 2.  The spectral convolution K is
written from its definition with explicit DFT sums,

  Vh[kx, ky] = sum_{x,y} v[x, y] exp(-2 pi i (kx x + ky y) / M),
  Wh[o]      = sum_c Vh[c] R[mode][c][o]            (complex R = Rr + i Ri),
  w[x, y]    = 1/M^2 sum_{kx, ky} c_ky Re(Wh exp(+2 pi i (kx x + ky y) / M)),

kx in {0..m-1, M-m..M-1}, ky in {0..m-1}, c_0 = 1, c_ky = 2 otherwise, which
is numpy.fft.irfft2 of the truncated product spectrum (tests/test_fno_oracle.py
checks that identity).  The reference has no counterpart; the data the
network stands in for comes from the exact solve of rho_functions.cpp:134-190.

Pinned by tests/test_fno_oracle.py: the spectral convolution against
numpy.fft; plane waves inside / outside the retained modes; the whole
gradient against central finite differences; GELU against its integral
definition.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def layout(w: int, m: int, layers: int, proj: int) -> dict:
    """name -> (offset, shape) of the flat parameter array (header order)."""
    nm = 2 * m * m
    out, off = {}, 0

    def put(name, shape):
        nonlocal off
        out[name] = (off, shape)
        off += int(np.prod(shape))
    put("P", (w, 4)); put("bP", (w,))
    for l in range(layers):
        put(f"Rr{l}", (nm, w, w)); put(f"Ri{l}", (nm, w, w)); put(f"W{l}", (w, w)); put(f"b{l}", (w,))
    put("Q1", (proj, w)); put("c1", (proj,)); put("Q2", (2, proj)); put("c2", (2,))
    out["_count"] = off
    return out


def entries(lay: dict):
    """(name, offset, shape) of every tensor of a layout."""
    return [(k, v[0], v[1]) for k, v in lay.items() if not k.startswith("_")]


def unpack(flat: np.ndarray, lay: dict) -> dict:
    return {k: np.asarray(flat[o:o + int(np.prod(s))], np.float64).reshape(s) for k, o, s in entries(lay)}


def pack(d: dict, lay: dict) -> np.ndarray:
    out = np.zeros(lay["_count"])
    for k, o, s in entries(lay):
        out[o:o + int(np.prod(s))] = np.asarray(d[k], np.float64).ravel()
    return out


def kx_modes(M: int, m: int) -> np.ndarray:
    return np.concatenate([np.arange(m), np.arange(M - m, M)])


def _E(M: int, ks: np.ndarray) -> np.ndarray:
    """E[x, k] = exp(-2 pi i k x / M)."""
    x = np.arange(M)
    return np.exp(-2j * np.pi * np.outer(x, ks) / M)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


class Spectral:
    """Truncated spectral convolution on an M x M box with m modes per axis."""

    def __init__(self, M: int, m: int):
        self.M, self.m = M, m
        self.Ex = _E(M, kx_modes(M, m))        # [M, 2m]
        self.Ey = _E(M, np.arange(m))          # [M, m]
        self.c = np.where(np.arange(m) == 0, 1.0, 2.0)

    def fwd(self, v, R):
        """v [B, C, M, M] real, R [2m, m, C, O] complex -> (w [B, O, M, M], Vh)."""
        Vh = np.einsum("bcxy,xj,yk->bcjk", v, self.Ex, self.Ey, optimize=True)
        Wh = np.einsum("bcjk,jkco->bojk", Vh, R, optimize=True)
        w = np.einsum("bojk,xj,yk->boxy", Wh * self.c, np.conj(self.Ex), np.conj(self.Ey), optimize=True)
        return w.real / self.M ** 2, Vh

    def bwd(self, dw, Vh, R):
        """dw [B, O, M, M] -> (dv [B, C, M, M], gR [2m, m, C, O] complex with
        gR = dL/dRe R + i dL/dIm R)."""
        gW = np.einsum("boxy,xj,yk->bojk", dw, self.Ex, self.Ey, optimize=True) * self.c / self.M ** 2
        gR = np.einsum("bcjk,bojk->jkco", np.conj(Vh), gW, optimize=True)
        gV = np.einsum("bojk,jkco->bcjk", gW, np.conj(R), optimize=True)
        dv = np.einsum("bcjk,xj,yk->bcxy", gV, np.conj(self.Ex), np.conj(self.Ey), optimize=True).real
        return dv, gR


class FNO:
    """FP64 network with the header's parameter layout."""

    def __init__(self, M=128, width=32, modes=12, layers=4, proj=128):
        self.M, self.w, self.m, self.L, self.proj = M, width, modes, layers, proj
        self.lay = layout(width, modes, layers, proj)
        self.sp = Spectral(M, modes)
        self.in_scale = np.ones(2)
        self.out_scale = np.ones(2)

    @property
    def count(self) -> int:
        return self.lay["_count"]

    def _R(self, p, l):
        m, w = self.m, self.w
        return (p[f"Rr{l}"] + 1j * p[f"Ri{l}"]).reshape(2 * m, m, w, w)

    def lift_input(self, x):
        B, M = x.shape[0], self.M
        g = np.arange(M) / M
        xin = np.empty((B, 4, M, M))
        xin[:, 0] = x[:, 0] * self.in_scale[0]
        xin[:, 1] = x[:, 1] * self.in_scale[1]
        xin[:, 2] = g[:, None]
        xin[:, 3] = g[None, :]
        return xin

    def forward(self, flat, x, keep=False):
        """x [B, 2, M, M] (phi, rho) -> out [B, 2, M, M] before the output
        scale (and the cache for backward)."""
        p = unpack(flat, self.lay)
        xin = self.lift_input(np.asarray(x, np.float64))
        v = np.einsum("ci,bixy->bcxy", p["P"], xin) + p["bP"][None, :, None, None]
        cache = {"xin": xin, "v": [v], "u": [], "Vh": []}
        for l in range(self.L):
            s, Vh = self.sp.fwd(v, self._R(p, l))
            u = s + np.einsum("oc,bcxy->boxy", p[f"W{l}"], v) + p[f"b{l}"][None, :, None, None]
            v = gelu(u) if l < self.L - 1 else u
            cache["u"].append(u); cache["Vh"].append(Vh); cache["v"].append(v)
        qpre = np.einsum("jc,bcxy->bjxy", p["Q1"], v) + p["c1"][None, :, None, None]
        q = gelu(qpre)
        out = np.einsum("oj,bjxy->boxy", p["Q2"], q) + p["c2"][None, :, None, None]
        cache.update(qpre=qpre, q=q, p=p)
        return (out, cache) if keep else out

    def velocity(self, flat, x):
        return self.forward(flat, x) * self.out_scale[None, :, None, None]

    def loss_grad(self, flat, x, target):
        """Loss mean((out - target / out_scale)^2) and its gradient (flat)."""
        out, c = self.forward(flat, x, keep=True)
        p = c["p"]
        t = np.asarray(target, np.float64) / self.out_scale[None, :, None, None]
        d = out - t
        loss = float(np.mean(d * d))
        dout = 2.0 * d / d.size
        g = {}
        g["Q2"] = np.einsum("boxy,bjxy->oj", dout, c["q"]); g["c2"] = dout.sum(axis=(0, 2, 3))
        dq = np.einsum("oj,boxy->bjxy", p["Q2"], dout) * gelu_d(c["qpre"])
        g["Q1"] = np.einsum("bjxy,bcxy->jc", dq, c["v"][self.L]); g["c1"] = dq.sum(axis=(0, 2, 3))
        dv = np.einsum("jc,bjxy->bcxy", p["Q1"], dq)
        for l in range(self.L - 1, -1, -1):
            du = dv * gelu_d(c["u"][l]) if l < self.L - 1 else dv
            vl = c["v"][l]
            g[f"W{l}"] = np.einsum("boxy,bcxy->oc", du, vl); g[f"b{l}"] = du.sum(axis=(0, 2, 3))
            dvs, gR = self.sp.bwd(du, c["Vh"][l], self._R(p, l))
            gR = gR.reshape(2 * self.m * self.m, self.w, self.w)
            g[f"Rr{l}"] = gR.real; g[f"Ri{l}"] = gR.imag
            dv = np.einsum("oc,boxy->bcxy", p[f"W{l}"], du) + dvs
        g["P"] = np.einsum("bcxy,bixy->ci", dv, c["xin"]); g["bP"] = dv.sum(axis=(0, 2, 3))
        return loss, pack(g, self.lay)

