"""The FNO surrogate's gfx950 kernels against numpy / the FP64 oracle
(oracle/fno.py): the BF16 MFMA GEMM over every operand layout and epilogue
the network uses, the forward pass, the full gradient, and Adam training.

Tolerances: GEMM relative 1e-5 against float64 products of the same BF16
inputs (products of BF16 values are exact in FP32; only the FP32 sums
round), BF16 outputs one BF16 rounding (2^-8).  Forward: relative L2 1e-2
against the FP64 network on the same weights (activations and weights pass
through BF16 at every matrix product, ~2^-9 each; measured 3.5e-3).
Gradient: relative L2 2.5e-2 per parameter tensor (the same roundings,
accumulated backwards; measured at most 1e-2).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from multicellular_phase_field_active_gel_organoids_amd import fno  # noqa: E402
from oracle import fno as ofno  # noqa: E402


def bf16_round(a: np.ndarray) -> np.ndarray:
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)


def rel_l2(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def small_net():
    net = fno.FnoNet(fno.FnoConfig(M=16, width=8, modes=3, layers=2, proj=12, batch_max=4))
    yield net
    net.close()


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


CASES = ["rowmajor", "transposed_f32", "bf16_gelu_out", "two_level_batch", "aligned_vector"]


@pytest.mark.parametrize("tile", range(len(fno.TILES)))
@pytest.mark.parametrize("case", CASES)
def test_mfma_gemm_matches_numpy(small_net, case, tile):
    """Ragged shapes take the element-wise staging, aligned ones the 16-byte
    vector loads (aligned_vector: BF16 k-contiguous A, n-contiguous B); every
    tile shape (64x64, 32x64, 64x32, 32x32)."""
    small_net.set_tiles([0] * (fno.STAGE_COUNT - 1) + [tile])
    rng = np.random.default_rng(CASES.index(case))
    M, N, K, nz1, nz2 = 100, 45, 70, 3, 1
    if case == "aligned_vector":
        M, N, K = 96, 72, 80
    if case == "two_level_batch":
        nz1, nz2 = 2, 3
    nz = nz1 * nz2
    A = rng.standard_normal((nz, M, K)).astype(np.float32)
    B = rng.standard_normal((nz, K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    C0 = rng.standard_normal((nz, M, N)).astype(np.float32)
    a = fno.GemmArgs()
    a.M, a.N, a.K, a.nz1, a.nz2, a.alpha = M, N, K, nz1, nz2, 0.5
    if case == "transposed_f32":
        # A stored [z][k][m] (m contiguous), B stored [z][n][k] (k contiguous), FP32 operands
        dA = _dev(A.transpose(0, 2, 1), torch.float32)
        dB = _dev(B.transpose(0, 2, 1), torch.float32)
        a.sAm, a.sAk, a.sBk, a.sBn = 1, M, 1, K
        a.a_f32 = a.b_f32 = 1
        Ar, Br = bf16_round(A), bf16_round(B)
    else:
        dA = _dev(A, torch.bfloat16)
        dB = _dev(B, torch.bfloat16)
        a.sAm, a.sAk, a.sBk, a.sBn = K, 1, N, 1
        Ar, Br = bf16_round(A), bf16_round(B)
    a.A, a.B = dA.data_ptr(), dB.data_ptr()
    if nz2 == 1:
        a.sA1, a.sB1, a.sC1 = M * K, K * N, M * N
    else:
        a.sA1, a.sA2 = nz2 * M * K, M * K
        a.sB1, a.sB2 = nz2 * K * N, K * N
        a.sC1, a.sC2 = nz2 * M * N, M * N
    a.sCm, a.sCn = N, 1
    dbias = _dev(bias, torch.float32)
    a.bias = dbias.data_ptr()
    want = 0.5 * np.einsum("zmk,zkn->zmn", Ar, Br) + bias[None, None, :]
    if case == "bf16_gelu_out":
        dC = torch.zeros((nz, M, N), dtype=torch.bfloat16, device="cuda")
        dC2 = torch.zeros((nz, M, N), dtype=torch.bfloat16, device="cuda")
        a.C, a.C2, a.c_mode, a.act = dC.data_ptr(), dC2.data_ptr(), 2, 1
        small_net.gemm(a)
        torch.cuda.synchronize()
        g = ofno.gelu(want)
        for got in (dC, dC2):
            got = got.float().cpu().numpy()
            assert np.max(np.abs(got - g) / (np.abs(g) + 1e-2)) < 2 ** -7
        return
    dC = _dev(C0, torch.float32)
    if case == "transposed_f32":
        a.c_mode = 1
        want = want + C0
    a.C = dC.data_ptr()
    small_net.gemm(a)
    torch.cuda.synchronize()
    got = dC.cpu().numpy()
    assert rel_l2(got, want) < 1e-5, rel_l2(got, want)
    assert np.max(np.abs(got - want)) < 1e-4 * np.max(np.abs(want))


def _inputs(B, M, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((B, 2, M, M), np.float32)
    ii, jj = np.meshgrid(np.arange(M), np.arange(M), indexing="ij")
    for b in range(B):
        c = M / 2 + rng.uniform(-4, 4, 2)
        r = np.hypot(ii - c[0], jj - c[1])
        R = rng.uniform(0.2, 0.3) * M
        phi = 0.5 * (1 - np.tanh((r - R) / 2.0))
        x[b, 0] = phi
        x[b, 1] = np.where((phi > 0.1) & (phi < 0.9), 1.0 + 0.1 * rng.standard_normal((M, M)), 0.0)
    return x


def _targets(x):
    """A smooth field of the inputs, stand-in for the velocity in the
    kernel-parity tests (the trained network's data come from the solve)."""
    g0 = np.gradient(x[:, 0], axis=1)
    g1 = np.gradient(x[:, 0], axis=2)
    return np.stack([g0 * (1 + x[:, 1]), g1 * (1 - 0.5 * x[:, 1])], axis=1).astype(np.float32) * 1e-2


@pytest.fixture(scope="module")
def full():
    cfg = fno.FnoConfig(M=128, width=32, modes=12, layers=4, proj=128, batch_max=4)
    net = fno.FnoNet(cfg)
    flat = fno.init_params(cfg, seed=3)
    net.set_params(flat)
    net.set_scales([1.0, 0.8], [1e-2, 2e-2])
    ref = ofno.FNO(M=128, width=32, modes=12, layers=4, proj=128)
    ref.in_scale = np.array([1.0, 0.8]); ref.out_scale = np.array([1e-2, 2e-2])
    yield net, flat, ref
    net.close()


def test_layout_and_init_agree_with_oracle(full):
    net, flat, ref = full
    assert net.count == ref.count == flat.size
    assert [(n, o, tuple(s)) for n, o, s in fno.layout(32, 12, 4, 128)] == \
        [(n, o, tuple(s)) for n, o, s in ofno.entries(ref.lay)]
    np.testing.assert_array_equal(net.get_params(), flat)


def test_forward_matches_oracle(full):
    net, flat, ref = full
    x = _inputs(2, 128, 11)
    got = net.forward(torch.as_tensor(x).cuda()).cpu().numpy()
    want = ref.velocity(flat.astype(np.float64), x)
    r = rel_l2(got, want)
    print(f"forward rel L2 {r:.3e}")
    assert r < 1e-2, r


def test_gradient_matches_oracle(full):
    net, flat, ref = full
    x = _inputs(2, 128, 12)
    t = _targets(x)
    loss, g = net.loss_grad(torch.as_tensor(x).cuda(), torch.as_tensor(t).cuda(), want_grad=True)
    loss_ref, g_ref = ref.loss_grad(flat.astype(np.float64), x, t)
    print(f"loss {loss:.6e} oracle {loss_ref:.6e}")
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    worst = 0.0
    for name, o, s in ofno.entries(ref.lay):
        n = int(np.prod(s))
        r = rel_l2(g[o:o + n], g_ref[o:o + n])
        worst = max(worst, r)
        print(f"  {name:4s} rel L2 {r:.3e}")
        assert r < 2.5e-2, (name, r)


def test_tile_shapes_give_identical_results(full):
    """Every tile shape accumulates the same products in the same K order:
    forward output and gradient are bit for bit those of the 64x64 tiles."""
    net, flat, ref = full
    x = torch.as_tensor(_inputs(2, 128, 14)).cuda()
    t = torch.as_tensor(_targets(_inputs(2, 128, 14))).cuda()
    net.set_tiles([0] * fno.STAGE_COUNT)
    out0 = net.forward(x).cpu().numpy()
    _, g0 = net.loss_grad(x, t, want_grad=True)
    for tile in range(1, len(fno.TILES)):
        net.set_tiles([tile] * fno.STAGE_COUNT)
        assert np.array_equal(net.forward(x).cpu().numpy(), out0), fno.TILES[tile]
        _, g = net.loss_grad(x, t, want_grad=True)
        assert np.array_equal(g, g0), fno.TILES[tile]
    with pytest.raises(Exception):
        net.set_tiles([4] * fno.STAGE_COUNT)
    net.set_tiles([0] * fno.STAGE_COUNT)


def test_tile_autotuner_picks_and_caches(full, tmp_path):
    from multicellular_phase_field_active_gel_organoids_amd import autotune
    net, flat, ref = full
    path = str(tmp_path / "tune.json")
    r = autotune.tuned_fno(net, 2, cache_path=path, reps=1, passes=1)
    assert r["source"] == "tuned" and r["forward_ms"] <= r["default_forward_ms"]
    assert net.get_tiles() == r["tile_ids"]
    print(f"tiles {r['tiles']}: forward {r['default_forward_ms']:.3f} -> {r['forward_ms']:.3f} ms, "
          f"train {r['default_train_ms']:.3f} -> {r['train_ms']:.3f} ms", flush=True)
    net.set_tiles([0] * fno.STAGE_COUNT)
    r2 = autotune.tuned_fno(net, 2, cache_path=path)
    assert r2["source"] == "cache" and net.get_tiles() == r["tile_ids"]
    net.set_tiles([0] * fno.STAGE_COUNT)


def test_training_reduces_loss():
    cfg = fno.FnoConfig(M=128, width=32, modes=12, layers=4, proj=128, batch_max=8)
    net = fno.FnoNet(cfg)
    try:
        net.set_params(fno.init_params(cfg, seed=5))
        x = _inputs(8, 128, 13)
        y = _targets(x)
        sx, sy = fno.fit_scales(x, y)
        net.set_scales(sx, sy)
        dx, dy = torch.as_tensor(x).cuda(), torch.as_tensor(y).cuda()
        losses = fno.train(net, dx, dy, steps=300, batch=8, lr=2e-3, lr_final=2e-4)
        print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
        assert np.isfinite(losses).all()
        assert losses[-1] < 0.1 * losses[0]
    finally:
        net.close()
