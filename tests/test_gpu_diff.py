"""GPU tests of the differentiable renderer (wpt_diff_*), which has no
reference counterpart: the forward image equals the oracle's FP32
restatement bit for bit; the path-replay adjoint matches the oracle's FP64
reverse-mode adjoint (relative error <= 1e-3) and central differences of the
device forward with common random numbers; the torch.autograd wrapper
returns the adjoint; Adam steps reduce the loss."""
import numpy as np
import pytest

import wpt

pytestmark = pytest.mark.gpu

TEXTURED_MATERIAL = 5


def setup(size, spp, texture=(32, 32), seed=1):
    import torch
    host = wpt.HostScene("diff-cornell")
    scene = wpt.Scene(host.desc(), device=True, owner=host)
    cam = host.camera(size, size)
    r = wpt.DiffRenderer(scene, cam, size, size, spp, 6, texture, TEXTURED_MATERIAL)
    lay = r.layout
    rng = np.random.default_rng(seed)
    params = np.concatenate([rng.uniform(0.2, 0.8, lay.texture_floats), rng.uniform(0.1, 0.9, lay.roughness_count),
                             rng.uniform(0.5, 1.5, lay.emission_count)]).astype(np.float32)
    return host, scene, cam, r, torch.tensor(params, device="cuda:0")


def oracle_for(oracle_mod, scene, r, texture):
    o = oracle_mod.Oracle(scene)
    return o, o.diff_config(r.width, r.height, r.spp, 6, texture, TEXTURED_MATERIAL)


def test_forward_matches_oracle_bit_exact(torch_cuda, oracle_mod):
    host, scene, cam, r, params = setup(48, 3)
    img = r.forward(params, sample_start=5).cpu().numpy()
    o, cfg = oracle_for(oracle_mod, scene, r, (32, 32))
    want = o.diff_render(cam, cfg, params.cpu().numpy(), sample_start=5)
    assert np.array_equal(img, want), f"max |diff| {np.abs(img - want).max()}"
    ref64 = o.diff_render(cam, cfg, params.cpu().numpy(), sample_start=5, precision=64)
    err = np.sqrt(np.mean((img.astype(np.float64) - ref64) ** 2)) / np.sqrt(np.mean(ref64.astype(np.float64) ** 2))
    print(f"diff forward: rel RMSE vs FP32 oracle 0, vs FP64 oracle {err:.2e}")


def test_adjoint_matches_fp64_host_adjoint(torch_cuda, oracle_mod):
    import torch
    host, scene, cam, r, params = setup(64, 4)
    img = r.forward(params)
    target = img * 0.8 + 0.01
    g_img = 2.0 * (img - target) / img.numel()
    grad = r.adjoint(params, g_img).cpu().numpy().astype(np.float64)
    o, cfg = oracle_for(oracle_mod, scene, r, (32, 32))
    want = o.diff_adjoint64(cam, cfg, params.cpu().numpy(), g_img.cpu().numpy())
    lay = r.layout
    t, rc = lay.texture_floats, lay.roughness_count
    for name, sl in (("texture", slice(0, t)), ("roughness", slice(t, t + rc)), ("emission", slice(t + rc, None))):
        rel = np.linalg.norm(grad[sl] - want[sl]) / np.linalg.norm(want[sl])
        print(f"adjoint {name}: relative error vs FP64 host adjoint {rel:.2e}")
        assert rel <= 1e-3, name
    st = r.stats()
    assert st.paths == 64 * 64 * 4 and st.forward_milliseconds > 0 and st.adjoint_milliseconds > 0


def test_adjoint_matches_central_differences(torch_cuda):
    # Central differences of the device forward (eps = 1e-3, the same random
    # numbers), loss summed in FP64, on 32 random parameters: emission
    # scales, roughness slots and texels of the visible back wall.
    import torch
    host, scene, cam, r, params = setup(96, 8, texture=(8, 8))
    lay = r.layout
    t, rc, ec = lay.texture_floats, lay.roughness_count, lay.emission_count
    target = (r.forward(params) * 0.7).double()

    def loss(p):
        return float(((r.forward(p).double() - target) ** 2).mean())

    img = r.forward(params)
    grad = r.adjoint(params, (2.0 * (img.double() - target) / img.numel()).float()).cpu().numpy()
    rng = np.random.default_rng(4)
    picks = list(rng.choice(t, 16, replace=False)) + list(t + rng.choice(rc, 10, replace=False)) + \
        list(t + rc + rng.choice(ec, 6, replace=False))
    eps = 1e-3
    fd, ad = [], []
    for i in picks:
        e = torch.zeros_like(params)
        e[int(i)] = eps
        fd.append((loss(params + e) - loss(params - e)) / (2 * eps))
        ad.append(float(grad[int(i)]))
    fd, ad = np.array(fd), np.array(ad)
    scale = np.abs(fd).max()
    print("FD vs adjoint max abs diff / max |grad|:", np.abs(fd - ad).max() / scale)
    assert np.abs(fd - ad).max() <= 2e-2 * scale
    big = np.abs(fd) > 0.1 * scale
    assert np.all(np.abs(fd[big] - ad[big]) <= 2e-2 * np.abs(fd[big]))


def test_roughness_gradient_matches_central_differences(torch_cuda):
    # Roughness gradients are small beside texture and emission gradients
    # (the F0 = 0.04 lobe), so they get a check on their own scale: central
    # differences (eps = 1e-2, the same random numbers, loss in FP64) on the
    # 8 roughness slots with the largest gradients, at 128x128, 16 spp, with
    # a target rendered from other samples and roughness shifted by 0.2.
    import torch
    host, scene, cam, r, params = setup(128, 16, texture=(8, 8), seed=3)
    lay = r.layout
    t, rc = lay.texture_floats, lay.roughness_count
    shifted = params.clone()
    shifted[t:t + rc] = (shifted[t:t + rc] + 0.2).clamp(0.05, 1.0)
    target = r.forward(shifted, sample_start=4096).double()

    def loss(p):
        return float(((r.forward(p).double() - target) ** 2).mean())

    img = r.forward(params)
    grad = r.adjoint(params, (2.0 * (img.double() - target) / img.numel()).float()).cpu().numpy()
    rough = grad[t:t + rc]
    picks = np.argsort(-np.abs(rough))[:8]
    assert np.abs(rough[picks]).min() > 0
    eps = 1e-2
    fd = []
    for k in picks:
        e = torch.zeros_like(params)
        e[t + int(k)] = eps
        fd.append((loss(params + e) - loss(params - e)) / (2 * eps))
    fd, ad = np.array(fd), rough[picks].astype(np.float64)
    scale = np.abs(fd).max()
    print("roughness FD:", fd, "adjoint:", ad)
    assert np.abs(fd - ad).max() <= 3e-2 * scale
    assert np.all(np.sign(fd) == np.sign(ad))


def test_torch_autograd_wrapper(torch_cuda):
    import torch
    host, scene, cam, r, params = setup(32, 2)
    fn = wpt.diff_render_function()
    p = params.clone().requires_grad_(True)
    img = fn.apply(p, r, 0)
    target = img.detach() * 0.5
    loss = ((img - target) ** 2).mean()
    loss.backward()
    direct = r.adjoint(p.detach(), (2.0 * (img.detach() - target) / img.numel()).contiguous())
    assert torch.allclose(p.grad, direct, rtol=1e-5, atol=1e-9)


def test_adam_steps_reduce_the_loss(torch_cuda):
    import torch
    host, scene, cam, r, params = setup(48, 4)
    truth = params.clone()
    target = r.forward(truth)
    gen = torch.Generator(device="cpu").manual_seed(3)
    p = (truth + 0.2 * torch.rand(truth.shape, generator=gen).to(truth.device) - 0.1).clamp(0.05, 1.0).contiguous()
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    losses = []
    for step in range(1, 21):
        img = r.forward(p)
        losses.append(float(((img - target) ** 2).mean()))
        g = r.adjoint(p, 2.0 * (img - target) / img.numel())
        r.adam(p, g, m, v, step, 0.02)
    assert losses[-1] < 0.5 * losses[0], losses
    lay = r.layout
    assert float(p.min()) >= 0.0 and float(p[lay.texture_floats:lay.texture_floats + lay.roughness_count].min()) >= 0.02


def test_adjoint_after_adam_needs_a_new_forward(torch_cuda):
    # wpt_diff_adam changes the parameters in place: the kept path radiance
    # then belongs to the old values, and an adjoint must not replay it.
    import torch
    host, scene, cam, r, params = setup(32, 2)
    img = r.forward(params)
    g = 2.0 * (img - 0.5 * img) / img.numel()
    grad = r.adjoint(params, g)
    m, v = torch.zeros_like(params), torch.zeros_like(params)
    r.adam(params, grad, m, v, 1, 0.01)
    with pytest.raises(wpt.WptError):
        r.adjoint(params, g)
    img = r.forward(params)
    r.adjoint(params, 2.0 * (img - 0.5 * img) / img.numel())
