"""GPU tests of the learned denoiser: forward activations and gradients of
the hand-written bf16 MFMA kernels against the CPU fp32 reference network
(oracle/unet_ref.py), the denoise entry point, and a short training run on
pairs rendered by the tracer that must beat its 1-spp input."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import unet_ref
from voxelpathtracer_amd.unet import UNet, bf16_bits, bf16_to_f32, make_dataset, tonemap_psnr, train

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _inputs(rng, n, h, w):
    x = np.zeros((n, h, w, 32), np.float32)
    x[..., :10] = rng.normal(0.3, 0.6, size=(n, h, w, 10))
    xb = bf16_bits(x)
    return xb, bf16_to_f32(xb).reshape(n, h, w, 32)


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def net(ctx):
    u = UNet(ctx, 8, 64, 64)
    yield u
    u.close()


@pytest.mark.parametrize("n,h,w", [(2, 32, 32), (1, 36, 20)])
def test_forward_matches_cpu(net, n, h, w):
    torch = _torch()
    rng = np.random.default_rng(5)
    xb, xf = _inputs(rng, n, h, w)
    net.init(11)
    params = net.get_params()
    out = torch.empty((n, h, w, 3), dtype=torch.float32, device="cuda")
    print(f"unet forward {n}x{h}x{w}", flush=True)
    net.forward(torch.from_numpy(xb.view(np.int16)).cuda(), n, h, w, out)
    torch.cuda.synchronize()
    with torch.no_grad():
        ref_out, acts = unet_ref.forward(unet_ref.unpack(params), torch.from_numpy(xf))
    # every stored activation: bf16 values of an fp32-accumulated product;
    # summation order differs, so a value may round to a neighbouring bf16
    for i, name in enumerate(["x", "e0a", "e0b", "p0", "e1a", "e1b", "p1", "b0", "b1", "d1a", "d1b", "d0a", "d0b"]):
        got = bf16_to_f32(net.activation(i, n, h, w))
        ref = acts[name].numpy()
        assert got.shape == ref.shape, name
        err = _rel(got, ref)
        assert err < 1e-2, (name, err)
        assert np.abs(got - ref).max() <= 0.02 * np.abs(ref).max() + 1e-6, name
    assert _rel(out.cpu().numpy(), ref_out.numpy()) < 1e-2


@pytest.mark.timeout(300)
def test_forward_matches_cpu_1080p(ctx):
    """The workload's resolution, 1920x1080: the persistent convolution
    kernels loop over many 16x32 tiles per workgroup and the pooled levels
    run at 960x540 and 480x270; every activation and the output equal the
    CPU fp32 reference to the same tolerance as the small cases."""
    torch = _torch()
    n, h, w = 1, 1080, 1920
    u = UNet(ctx, n, h, w)
    rng = np.random.default_rng(17)
    xb, xf = _inputs(rng, n, h, w)
    u.init(13)
    params = u.get_params()
    out = torch.empty((n, h, w, 3), dtype=torch.float32, device="cuda")
    print("unet forward 1x1080x1920", flush=True)
    u.forward(torch.from_numpy(xb.view(np.int16)).cuda(), n, h, w, out)
    torch.cuda.synchronize()
    with torch.no_grad():
        ref_out, acts = unet_ref.forward(unet_ref.unpack(params), torch.from_numpy(xf))
    for i, name in enumerate(["x", "e0a", "e0b", "p0", "e1a", "e1b", "p1", "b0", "b1", "d1a", "d1b", "d0a", "d0b"]):
        got = bf16_to_f32(u.activation(i, n, h, w))
        ref = acts[name].numpy()
        assert got.shape == ref.shape, name
        err = _rel(got, ref)
        assert err < 1e-2, (name, err)
        assert np.abs(got - ref).max() <= 0.02 * np.abs(ref).max() + 1e-6, name
    assert _rel(out.cpu().numpy(), ref_out.numpy()) < 1e-2
    u.close()


def test_gradients_match_cpu(net):
    torch = _torch()
    rng = np.random.default_rng(9)
    n, h, w = 2, 32, 48
    xb, xf = _inputs(rng, n, h, w)
    target = rng.normal(0.2, 0.5, size=(n, h, w, 3)).astype(np.float32)
    net.init(21)
    params = net.get_params()
    print("unet backward 2x32x48", flush=True)
    loss = net.backward(torch.from_numpy(xb.view(np.int16)).cuda(), torch.from_numpy(target).cuda(), n, h, w)
    g = net.get_grads()
    ref_loss, ref_g = unet_ref.loss_and_grads(params, xf, target)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    from voxelpathtracer_amd.unet import param_layout
    layout, total = param_layout()
    assert total == g.size == ref_g.size
    for name, off, shape in layout:
        size = int(np.prod(shape))
        a, b = g[off:off + size], ref_g[off:off + size]
        cos = float(np.dot(a, b) / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        # bf16 storage of activation gradients on the GPU: a few percent
        assert _rel(a, b) < 0.06 and cos > 0.998, (name, _rel(a, b), cos)


def _bf16_tensor(a):
    torch = _torch()
    return torch.from_numpy(bf16_to_f32(a.cpu().numpy().view(np.uint16))).cuda()


def test_denoise_entry_point(ctx, net):
    """vpt_unet_denoise (assemble + forward + exp) equals the separate calls."""
    torch = _torch()
    from voxelpathtracer_amd.unet import PairMaker
    net.init(4)
    pm = PairMaker(ctx, worlds=1, size=64, spp=2, seed=9)
    x1 = torch.empty((64, 64, 32), dtype=torch.int16, device="cuda")
    t1 = torch.empty((64, 64, 3), dtype=torch.float32, device="cuda")
    pm.pair(net, x1, t1)
    rgb = torch.empty((64, 64, 3), dtype=torch.float32, device="cuda")
    out = torch.empty((64, 64, 3), dtype=torch.float32, device="cuda")
    net.denoise(pm.gb, pm.tex, rgb)
    net.forward(x1, 1, 64, 64, out, mode=1)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(rgb.cpu().numpy(), out.cpu().numpy())
    pm.close()


def test_training_beats_input(ctx, net):
    """Tracer-rendered pairs (1 spp -> 64 spp at 64x64), 150 Adam steps: the
    loss falls and the network beats its own 1-spp input on held-out pairs."""
    torch = _torch()
    net.init(3)
    print("unet dataset", flush=True)
    x, t = make_dataset(ctx, net, 40, size=64, spp=64, seed=5, worlds=2)
    xtr, ttr, xte, tte = x[:32], t[:32], x[32:], t[32:]
    print("unet training", flush=True)
    losses = train(net, xtr, ttr, steps=150, batch=8, lr=2e-3)
    assert np.mean(losses[-10:]) < 0.7 * np.mean(losses[:5])
    out = torch.empty((8, 64, 64, 3), dtype=torch.float32, device="cuda")
    net.forward(xte, 8, 64, 64, out, mode=1)
    torch.cuda.synchronize()
    ref = torch.expm1(tte).cpu().numpy()
    noisy = torch.expm1(_bf16_tensor(xte)[..., :3]).cpu().numpy()
    p_net, p_noisy = tonemap_psnr(out.cpu().numpy(), ref), tonemap_psnr(noisy, ref)
    print(f"held-out PSNR: network {p_net:.2f} dB, 1-spp input {p_noisy:.2f} dB", flush=True)
    assert p_net > p_noisy + 3.0
