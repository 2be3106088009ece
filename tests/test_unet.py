"""The learned SI correction: the hand-written BF16 MFMA U-Net kernels against
the same network in PyTorch FP32, and the correction's plumbing in source
iteration (a zero network must leave SI bitwise unchanged; a useless network
must be rejected and still converge to the same answer)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

import doma_amd
from doma_amd import capi, unet, workloads


def test_param_layout_host():
    assert unet.param_count() == capi.net_param_count() == 144509
    torch = pytest.importorskip("torch")
    m = unet.make_model()
    m2 = unet.unflatten(unet.make_model(), unet.flatten(m))
    for a, b in zip(m.state_dict().values(), m2.state_dict().values()):
        assert torch.equal(a, b)


def _torch_forward(flat, feats):
    import torch
    m = unet.unflatten(unet.make_model(), flat)
    x = torch.from_numpy(feats).permute(3, 0, 1, 2)[None]
    with torch.no_grad():
        return m(x)[0, 0].numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(16, 20, 24), (32, 32, 32), (8, 12, 16), (48, 40, 64),
                                  (64, 64, 64)])
def test_net_forward_matches_torch(dims):
    """BF16 activations and weights, FP32 accumulation: relative L2 error of the
    output against FP32 PyTorch with the same (BF16-rounded) weights < 3e-2.
    The grids take every convolution tiling: tiles of one z-plane on the small
    grids, two at 48 x 40 x 64 and four at 64^3 (full resolution)."""
    torch = pytest.importorskip("torch")
    torch.manual_seed(1)
    flat = unet.bf16_round(unet.flatten(unet.make_model()))
    rng = np.random.default_rng(2)
    D, H, W = dims
    f = rng.standard_normal((D, H, W, unet.CIN)).astype(np.float32)
    print("net forward", dims, flush=True)
    got = capi.net_forward(flat, f)
    want = _torch_forward(flat, f)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < 3e-2, err
    # a layout mistake (axis swap, wrong tap) is an O(1) error, far above this
    assert np.abs(got - want).max() < 0.15 * np.abs(want).max()


@pytest.mark.gpu
def test_net_single_tap_exact():
    """One non-zero weight per layer path: with weights and inputs exactly
    representable in BF16 the output is exact, pinning tap and channel order."""
    flat = np.zeros(unet.param_count(), dtype=np.float32)
    torch = pytest.importorskip("torch")
    m = unet.make_model()
    with torch.no_grad():
        for conv in m.convs:
            conv.weight.zero_()
            conv.bias.zero_()
        # enc1a: out0 <- in channel 1 at offset (dz, dy, dx) = (+1, 0, -1)
        m.convs[0].weight[0, 1, 2, 1, 0] = 1.0
        for l in (1, 2, 3, 4, 5, 7, 9):
            m.convs[l].weight[0, 0, 1, 1, 1] = 1.0
        m.convs[6].weight[0, 0, 1, 1, 1] = 0.5       # skip path
        m.convs[6].weight[0, 24, 1, 1, 1] = 0.25     # upsampled path
        m.convs[8].weight[0, 0, 1, 1, 1] = 1.0
        m.convs[8].weight[0, 12, 0, 1, 2] = 0.5      # upsampled, offset (-1, 0, +1)
        m.convs[10].weight[0, 0] = 2.0
        m.convs[10].bias[0] = 0.125
    flat = unet.flatten(m)
    rng = np.random.default_rng(3)
    f = np.zeros((16, 16, 16, unet.CIN), dtype=np.float32)
    f[..., 1] = rng.integers(0, 8, size=(16, 16, 16)) / 8.0
    got = capi.net_forward(flat, f)
    want = _torch_forward(flat, f)
    assert np.allclose(got, want, rtol=1e-2, atol=1e-2)


def _slow_problem(n=16, seed=5, scheme=capi.SCHEME_TW_DD):
    spec = capi.MediaSpec(n=n, block=8, groups=1, sigt_log=1, sigt_lo=0.5, sigt_hi=10.0,
                          c_lo=0.95, c_hi=0.995, scatter_rule=capi.SCATTER_DOWN1,
                          source_rule=capi.SOURCE_BLOCKS, src_lo=0, src_hi=0,
                          source_fraction=0.2, cell_width=1.0, seed=seed)
    return workloads.block_media_problem(n, 1, 2, 2, scheme, spec, "slow16")


@pytest.mark.gpu
def test_zero_network_is_the_extrapolation():
    """A zero network leaves the bare geometric-series extrapolation g d_k: on a
    slowly converging diamond-difference problem it is accepted, cuts the
    sweeps, and the iteration ends at the plain fixed point (DD source
    iteration is linear)."""
    p = _slow_problem(scheme=capi.SCHEME_DD)
    with doma_amd.Context(p) as ctx:
        phi0, st0 = ctx.solve_fixed_source(1e-10, 5000, 1e-10, 10)
        ctx.set_learned_correction(np.zeros(unet.param_count(), np.float32), every=5)
        phi1, st1 = ctx.solve_fixed_source(1e-10, 5000, 1e-10, 10)
    print("zero network: sweeps", st0["sweeps"], "->", st1["sweeps"], "corrections",
          st1["corrections_accepted"], "+/", st1["corrections_rejected"], "-", flush=True)
    assert st0["success"] and st1["success"]
    assert st1["corrections_accepted"] > 0
    assert st1["sweeps"] < st0["sweeps"]
    assert rel_err(phi1, phi0).max() < 1e-7


@pytest.mark.gpu
def test_bad_network_is_rejected_and_converges():
    torch = pytest.importorskip("torch")
    torch.manual_seed(4)
    flat = unet.flatten(unet.make_model()) * 50.0  # large random corrections
    p = _slow_problem(seed=6)
    with doma_amd.Context(p) as ctx:
        phi0, st0 = ctx.solve_fixed_source(1e-9, 5000, 1e-9, 10)
        ctx.set_learned_correction(flat, every=5)
        phi1, st1 = ctx.solve_fixed_source(1e-9, 5000, 1e-9, 10)
    print("bad net", st1["corrections_accepted"], st1["corrections_rejected"], flush=True)
    assert st1["success"] and st1["corrections_rejected"] > 0
    if st1["corrections_accepted"] == 0:
        # every correction undone exactly: the plain iterates, bitwise
        assert np.array_equal(phi1, phi0)
    assert rel_err(phi1, phi0).max() < 1e-6


@pytest.mark.gpu
def test_correction_rejects_wrong_inputs():
    p = workloads.random_blocks(16, 1, 2, 2)
    with doma_amd.Context(p) as ctx:
        with pytest.raises(doma_amd.DomaError):
            ctx.set_learned_correction(np.zeros(10, np.float32), every=5)
    p2 = workloads.random_blocks(18, 1, 2, 2, block=6)  # 18 is not divisible by 4
    with doma_amd.Context(p2) as ctx:
        with pytest.raises(doma_amd.DomaError):
            ctx.set_learned_correction(np.zeros(unet.param_count(), np.float32), every=5)


@pytest.mark.gpu
def test_warm_start_same_fixed_point():
    """Warm-started group iterations (accelerated mode) reach the plain fixed
    point (diamond difference: linear SI) in fewer sweeps."""
    p = workloads.multigroup_full(16, 4, 2, 2, capi.SCHEME_DD)
    with doma_amd.Context(p) as ctx:
        phi0, st0 = ctx.solve_fixed_source(1e-11, 5000, 1e-11, 200)
        ctx.set_warm_start(True)
        phi1, st1 = ctx.solve_fixed_source(1e-11, 5000, 1e-11, 200)
    assert st0["success"] and st1["success"]
    assert st1["sweeps"] < st0["sweeps"]
    assert rel_err(phi1, phi0).max() < 1e-8


def _torch_grad(flat, feats, targets):
    """Loss sum (p - y)^2 / sum y^2 and its gradient by PyTorch autograd (FP32)."""
    import torch
    m = unet.unflatten(unet.make_model(), flat)
    x = torch.from_numpy(feats).permute(0, 4, 1, 2, 3).contiguous()
    y = torch.from_numpy(targets)[:, None]
    pred = m(x)
    loss = ((pred - y) ** 2).sum() / ((y ** 2).sum() + 1e-12)
    loss.backward()
    loss = loss.detach()
    g = []
    for conv in m.convs:
        g.append(conv.weight.grad.permute(0, 2, 3, 4, 1).reshape(-1).numpy())
        g.append(conv.bias.grad.numpy())
    return float(loss), np.concatenate(g)


@pytest.mark.gpu
def test_native_gradient_matches_autograd():
    """The hand-written backward pass (BF16 MFMA data gradients, FP32 weight
    gradients) against PyTorch FP32 autograd with the same BF16-rounded weights:
    per layer, relative L2 error of the weight gradient < 5e-2."""
    torch = pytest.importorskip("torch")
    torch.manual_seed(7)
    flat = unet.bf16_round(unet.flatten(unet.make_model()))
    rng = np.random.default_rng(8)
    B, Cr = 2, 16
    # inputs exactly representable in BF16, as the kernels store them
    f = unet.bf16_round(rng.standard_normal((B, Cr, Cr, Cr, unet.CIN))).astype(np.float32)
    y = rng.standard_normal((B, Cr, Cr, Cr)).astype(np.float32)
    print("native gradient", flush=True)
    with capi.NetTrainer(flat, crop=Cr, batch=B) as tr:
        loss, g = tr.gradient(f, y.reshape(B, -1))
    lt, gt = _torch_grad(flat, f, y)
    assert abs(loss - lt) <= 2e-2 * abs(lt)
    off, errs = 0, []
    for a, b, co, t in unet.SPEC:
        n = co * t * (a + b) + co
        ga, gb = g[off:off + n], gt[off:off + n]
        err = np.linalg.norm(ga - gb) / (np.linalg.norm(gb) + 1e-30)
        cos = float(ga @ gb / (np.linalg.norm(ga) * np.linalg.norm(gb) + 1e-30))
        errs.append((round(float(err), 4), round(cos, 5)))
        off += n
    print("per-layer relative error, cosine:", errs, flush=True)
    # BF16 activations and data gradients: the error grows with the depth of
    # the backward chain (layer 0 is last); a wrong index or sign is O(1).
    for err, cos in errs:
        assert err < 0.1 and cos > 0.995


@pytest.mark.gpu
def test_native_training_reduces_loss():
    """A few Adam steps on a fixed batch lower the loss."""
    torch = pytest.importorskip("torch")
    torch.manual_seed(9)
    flat = unet.flatten(unet.make_model())
    rng = np.random.default_rng(10)
    B, Cr = 2, 16
    f = rng.standard_normal((B, Cr, Cr, Cr, unet.CIN)).astype(np.float32)
    y = (0.5 * f[..., 0] + 0.1).astype(np.float32).reshape(B, -1)
    with capi.NetTrainer(flat, crop=Cr, batch=B) as tr:
        l0 = tr.step(f, y, 1e-3)
        for _ in range(80):
            l1 = tr.step(f, y, 1e-3)
        p = tr.params()
    print("loss", l0, "->", l1, flush=True)
    assert l1 < 0.7 * l0, (l0, l1)
    assert np.isfinite(p).all() and not np.array_equal(p, flat)


def test_extrapolation_ablation_is_the_zero_network():
    assert unet.extrapolation_params().shape == (unet.param_count(),)
    assert not unet.extrapolation_params().any()
    head = unet.init_params(seed=3, zero_head=True)[-(unet.C1 + 1):]
    assert not head.any()


def test_gain_and_features_host():
    rng = np.random.default_rng(14)
    d1 = rng.standard_normal((4, 4, 4))
    d0 = 2.0 * d1  # rms ratio 0.5 -> gain 1
    s, g = unet.gain(d1, d0)
    assert np.isclose(s, np.sqrt((d1 ** 2).mean())) and np.isclose(g, 1.0)
    f, s2, g2 = unet.features(d1, d0, np.full((4, 4, 4), 2.0), np.full((4, 4, 4), 0.9))
    assert f.shape == (4, 4, 4, unet.CIN) and s2 == s and g2 == g
    assert np.allclose(f[..., 1], 2.0 * f[..., 0], rtol=1e-6)
    assert np.allclose(f[..., 2], np.log(2.0)) and np.allclose(f[..., 3], np.log(1 / 0.1001))
    # growing increments clamp rho to 0.999
    assert np.isclose(unet.gain(d0, d1)[1], 0.999 / 0.001)
