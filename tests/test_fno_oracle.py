"""Pinning of the FP64 FNO restatement (oracle/fno.py), CPU only: the
spectral convolution against numpy.fft, plane waves inside and outside the
retained modes, GELU against its defining integral, and the hand-derived
gradient against central finite differences of the loss."""
import numpy as np
from scipy import integrate

from oracle import fno as ofno


def test_spectral_conv_is_irfft2_of_truncated_product():
    rng = np.random.default_rng(1)
    M, m, C, O, B = 16, 3, 2, 3, 2
    sp = ofno.Spectral(M, m)
    v = rng.standard_normal((B, C, M, M))
    R = rng.standard_normal((2 * m, m, C, O)) + 1j * rng.standard_normal((2 * m, m, C, O))
    w, _ = sp.fwd(v, R)
    V = np.fft.rfft2(v)                           # [B, C, M, M//2+1]
    S = np.zeros((B, O, M, M // 2 + 1), complex)
    kx = ofno.kx_modes(M, m)
    for j, k in enumerate(kx):
        for ky in range(m):
            S[:, :, k, ky] = np.einsum("bc,co->bo", V[:, :, k, ky], R[j, ky])
    want = np.fft.irfft2(S, s=(M, M))
    np.testing.assert_allclose(w, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_plane_waves_pass_or_vanish_with_identity_weights():
    M, m = 32, 4
    sp = ofno.Spectral(M, m)
    R = np.zeros((2 * m, m, 1, 1), complex)
    R[..., 0, 0] = 1.0
    x, y = np.meshgrid(np.arange(M), np.arange(M), indexing="ij")
    inside = np.cos(2 * np.pi * (3 * x + 2 * y) / M) + 0.5 * np.sin(2 * np.pi * (-2 * x + 1 * y) / M) + 0.25
    outside = np.cos(2 * np.pi * (9 * x) / M) + np.sin(2 * np.pi * (1 * x + 7 * y) / M)
    w_in, _ = sp.fwd(inside[None, None], R)
    w_out, _ = sp.fwd(outside[None, None], R)
    np.testing.assert_allclose(w_in[0, 0], inside, atol=1e-12)
    np.testing.assert_allclose(w_out[0, 0], 0.0, atol=1e-12)


def test_gelu_matches_its_integral():
    for x in (-3.0, -0.7, 0.0, 0.4, 2.5):
        phi = integrate.quad(lambda t: np.exp(-t * t / 2) / np.sqrt(2 * np.pi), -np.inf, x)[0]
        assert abs(ofno.gelu(x) - x * phi) < 1e-12
        h = 1e-6
        assert abs(ofno.gelu_d(x) - (ofno.gelu(x + h) - ofno.gelu(x - h)) / (2 * h)) < 1e-8


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    net = ofno.FNO(M=16, width=4, modes=3, layers=2, proj=6)
    net.in_scale = np.array([1.0, 0.7]); net.out_scale = np.array([0.5, 2.0])
    flat = rng.standard_normal(net.count) * 0.3
    x = rng.standard_normal((2, 2, 16, 16))
    t = rng.standard_normal((2, 2, 16, 16))
    loss, g = net.loss_grad(flat, x, t)
    assert loss > 0
    lay = net.lay
    idx = []
    for name, o, s in ofno.entries(lay):      # a few entries of every tensor
        n = int(np.prod(s))
        idx += list(o + rng.choice(n, min(n, 3), replace=False))
    h = 1e-6
    for i in idx:
        e = np.zeros(net.count); e[i] = h
        lp, _ = net.loss_grad(flat + e, x, t)
        lm, _ = net.loss_grad(flat - e, x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_layout_matches_header_order():
    lay = ofno.layout(32, 12, 4, 128)
    nm, w = 2 * 12 * 12, 32
    assert lay["P"] == (0, (w, 4)) and lay["bP"][0] == 4 * w
    assert lay["Rr0"][0] == 5 * w and lay["Ri0"][0] == 5 * w + nm * w * w
    per_layer = 2 * nm * w * w + w * w + w
    assert lay["Q1"][0] == 5 * w + 4 * per_layer
    assert lay["_count"] == 5 * w + 4 * per_layer + 128 * w + 128 + 2 * 128 + 2
