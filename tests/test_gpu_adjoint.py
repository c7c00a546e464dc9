"""Path-replay adjoint (lmgpu_pt_backward) against central finite differences of the GPU
forward and of the FP64 oracle, with common sample streams.

With Russian roulette off (rr_depth = max_path_vertices) and common streams, every path
is a fixed function of the sample stream, and the image is a polynomial in the
reflectances and emissions: central differences with h = 1e-3 are then exact up to
rounding, so the adjoint must agree with them to well under 1 %."""
from __future__ import annotations

import numpy as np
import pytest

from lightmetrica_amd import scenes

pytestmark = pytest.mark.gpu

RENDERERS = ["pt", "pt.direct", "pt.mis"]
H_FD = 1e-3


def small_gradient_scene(num_boxes=150, w=48, h=48, spp=4, max_pv=6):
    return scenes.gradient_scene(w, h, spp, num_boxes=num_boxes, max_path_vertices=max_pv, rr_depth=max_pv)


def base_params(sc):
    return (np.array([m["R"] for m in sc.materials], np.float64), np.array([l["Le"] for l in sc.lights], np.float64))


def loss_and_grad(img, target):
    d = img - target
    return float((d * d).sum()), (2.0 * d).astype(np.float32)


def render(gpu, renderer, albedo, emission):
    gpu.set_params(albedo, emission)
    img, _ = gpu.render(renderer, film_precision="fp64", out_precision="fp64")
    return img


def pick_params(sc, rng, n_albedo=56, n_light=8):
    """(kind, index, channel) triples: box reflectances and light emissions."""
    boxes = np.asarray(sc.meta["box_materials"])
    out = [("albedo", int(rng.choice(boxes)), int(rng.integers(3))) for _ in range(n_albedo)]
    out += [("emission", int(rng.integers(len(sc.lights))), int(rng.integers(3))) for _ in range(n_light)]
    return out


def fd_gradient(render_fn, albedo, emission, target, params):
    g = []
    for kind, i, c in params:
        vals = []
        for s in (+1.0, -1.0):
            a, e = albedo.copy(), emission.copy()
            (a if kind == "albedo" else e)[i, c] += s * H_FD
            vals.append(loss_and_grad(render_fn(a, e), target)[0])
        g.append((vals[0] - vals[1]) / (2 * H_FD))
    return np.array(g)


def adjoint_at(ga, ge, params):
    return np.array([(ga if kind == "albedo" else ge)[i, c] for kind, i, c in params])


def compare(g_adj, g_fd, label, tol):
    err = np.abs(g_adj - g_fd)
    scale = np.abs(g_fd).max()
    rel = np.linalg.norm(g_adj - g_fd) / np.linalg.norm(g_fd)
    big = np.abs(g_fd) > 1e-2 * scale
    per = (err[big] / np.abs(g_fd[big])).max()
    print(f"{label}: |adj - fd| / |fd| = {rel:.2e}, worst per-parameter relative error {per:.2e} "
          f"over {big.sum()} parameters", flush=True)
    assert rel < tol
    assert per < 10 * tol


@pytest.mark.parametrize("renderer", RENDERERS)
def test_adjoint_matches_gpu_central_differences(gpu, renderer):
    sc = small_gradient_scene()
    gpu.upload(sc)
    gpu.build_bvh()
    albedo, emission = base_params(sc)
    boxes = np.asarray(sc.meta["box_materials"])
    hidden = albedo.copy()
    hidden[boxes] = scenes.hidden_box_albedo(len(boxes))
    target = render(gpu, renderer, hidden, emission)
    img = render(gpu, renderer, albedo, emission)
    _, dLdI = loss_and_grad(img, target)
    gpu.set_params(albedo, emission)
    ga, ge, info = gpu.backward(renderer, dLdI)
    print(f"{renderer}: replay residual {info['replay_residual']:.2e}", flush=True)
    assert info["replay_residual"] < 1e-9
    params = pick_params(sc, np.random.default_rng(7))
    g_fd = fd_gradient(lambda a, e: render(gpu, renderer, a, e), albedo, emission, target, params)
    compare(adjoint_at(ga, ge, params), g_fd, f"{renderer} vs GPU FD", 2e-3)


@pytest.mark.parametrize("renderer", RENDERERS)
def test_adjoint_matches_oracle_central_differences(gpu, oracle, renderer):
    """Central differences of the FP64 CPU restatement drawing the same streams."""
    sc = small_gradient_scene(num_boxes=60, w=32, h=32, spp=4, max_pv=5)
    gpu.upload(sc)
    gpu.build_bvh()
    albedo, emission = base_params(sc)
    boxes = np.asarray(sc.meta["box_materials"])
    hidden = albedo.copy()
    hidden[boxes] = scenes.hidden_box_albedo(len(boxes))

    def oracle_render(a, e):
        s = scenes.gradient_scene(32, 32, 4, num_boxes=60, box_albedo=a[boxes], max_path_vertices=5, rr_depth=5)
        s.materials["R"] = a.astype(np.float32)
        s.lights["Le"] = e.astype(np.float32)
        img, _ = oracle.Oracle(s, fp64=True, accel="bvh").render(renderer, threads=0)
        return img

    target = oracle_render(hidden, emission)
    img = render(gpu, renderer, albedo, emission)
    _, dLdI = loss_and_grad(img, target)
    gpu.set_params(albedo, emission)
    ga, ge, info = gpu.backward(renderer, dLdI)
    params = pick_params(sc, np.random.default_rng(11), n_albedo=12, n_light=4)
    # The oracle's loss gradient is taken at its own image, so its dL/dI matches the GPU's
    # up to the per-pixel differences of FP32 and FP64 paths.
    g_fd = fd_gradient(oracle_render, albedo, emission, target, params)
    compare(adjoint_at(ga, ge, params), g_fd, f"{renderer} vs oracle FD", 2e-2)


@pytest.mark.parametrize("case", [0, 3, 5, 6, 10])
def test_random_scenes_adjoint_matches_central_differences(gpu, case):
    """The random scenes of the same-stream parity test (mirror and diffuse-mirror boxes,
    several area lights, an environment in some), Russian roulette off: the adjoint of every
    renderer against central differences of the GPU forward over every non-zero reflectance
    channel and every emission channel."""
    from test_gpu_pt import random_scene

    sc = random_scene(case)
    sc.render.rr_depth = sc.render.max_path_vertices
    gpu.upload(sc)
    gpu.build_bvh()
    albedo, emission = base_params(sc)
    rng = np.random.default_rng(case)
    target_a = np.clip(albedo * rng.uniform(0.7, 1.3, albedo.shape), 0.0, 1.0)
    params = [("albedo", i, c) for i in range(len(albedo)) for c in range(3) if albedo[i, c] > 0]
    params += [("emission", i, c) for i in range(len(emission)) for c in range(3)]
    for renderer in RENDERERS:
        target = render(gpu, renderer, target_a, emission * 1.1)
        img = render(gpu, renderer, albedo, emission)
        _, dLdI = loss_and_grad(img, target)
        gpu.set_params(albedo, emission)
        ga, ge, info = gpu.backward(renderer, dLdI)
        assert info["replay_residual"] < 1e-9
        g_fd = fd_gradient(lambda a, e: render(gpu, renderer, a, e), albedo, emission, target, params)
        compare(adjoint_at(ga, ge, params), g_fd, f"random{case}/{renderer} vs GPU FD ({len(params)} parameters)",
                2e-3)


@pytest.mark.parametrize("renderer", RENDERERS)
def test_replay_is_exact_with_russian_roulette(gpu, renderer):
    """With Russian roulette on (its probabilities held constant) the replay still retraces
    every path: nothing is left of Q at the path ends."""
    sc = scenes.gradient_scene(48, 48, 8, num_boxes=150, max_path_vertices=9, rr_depth=2)
    gpu.upload(sc)
    gpu.build_bvh()
    img, _ = gpu.render(renderer, film_precision="fp64")
    rng = np.random.default_rng(3)
    dLdI = rng.standard_normal(img.shape).astype(np.float32)
    ga, ge, info = gpu.backward(renderer, dLdI)
    print(f"{renderer} with RR: replay residual {info['replay_residual']:.2e}", flush=True)
    assert info["replay_residual"] < 1e-9
    assert np.isfinite(ga).all() and np.isfinite(ge).all()
    assert np.abs(ga).max() > 0 and np.abs(ge).max() > 0


def test_linear_loss_gradient_is_the_emission_image(gpu):
    """For L = sum(w * I), dL/dLe[l] is sum(w * I_l) / Le_l where I_l is the image lit by
    light l alone: render each light alone and compare (pt.direct, RR on)."""
    sc = scenes.gradient_scene(32, 32, 16, num_boxes=100, max_path_vertices=6, rr_depth=3)
    gpu.upload(sc)
    gpu.build_bvh()
    albedo, emission = base_params(sc)
    rng = np.random.default_rng(5)
    w = rng.uniform(0.0, 1.0, (32, 32, 3)).astype(np.float32)
    gpu.set_params(albedo, emission)
    _, ge, _ = gpu.backward("pt.direct", w)
    for l in range(len(sc.lights)):
        e = np.zeros_like(emission)
        e[l] = emission[l]
        img = render(gpu, "pt.direct", albedo, e)
        want = (w * img).sum(axis=(0, 1)) / emission[l]
        np.testing.assert_allclose(ge[l], want, rtol=1e-4)


@pytest.mark.parametrize("renderer", ["pt.mis", "pt"])
def test_autograd_function(gpu, renderer):
    """render() as a torch.autograd.Function on CPU and GPU tensors (on the GPU the backward
    uses the forward's sample record) against the C ABI's backward."""
    import torch

    from lightmetrica_amd.autograd import render as drender

    sc = small_gradient_scene(num_boxes=80, w=32, h=32, spp=4, max_pv=5)
    gpu.upload(sc)
    gpu.build_bvh()
    albedo, emission = base_params(sc)
    boxes = np.asarray(sc.meta["box_materials"])
    hidden = albedo.copy()
    hidden[boxes] = scenes.hidden_box_albedo(len(boxes))
    for dev in ("cpu", "cuda"):
        target = drender(gpu, torch.tensor(hidden, device=dev), torch.tensor(emission, device=dev), renderer).detach()
        a = torch.tensor(albedo, device=dev, requires_grad=True)
        e = torch.tensor(emission, device=dev, requires_grad=True)
        img = drender(gpu, a, e, renderer)
        loss = ((img - target) ** 2).sum()
        loss.backward()
        gpu.set_params(albedo, emission)
        ga, ge, _ = gpu.backward(renderer, (2.0 * (img - target)).detach().cpu().numpy())
        np.testing.assert_allclose(a.grad.cpu().numpy(), ga, rtol=1e-5, atol=1e-9)
        np.testing.assert_allclose(e.grad.cpu().numpy(), ge, rtol=1e-5, atol=1e-9)
        # A few projected gradient steps on the box reflectances lower the loss.
        opt = torch.optim.Adam([a], lr=0.05)
        losses = []
        for _ in range(8):
            opt.zero_grad()
            l = ((drender(gpu, a, e.detach(), renderer) - target) ** 2).sum()
            l.backward()
            losses.append(l.item())
            opt.step()
            with torch.no_grad():
                a.clamp_(0.0, 1.0)
        print(f"{dev}: losses {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
        assert losses[-1] < 0.8 * losses[0]


@pytest.mark.parametrize("scene", ["boxes", "room"])
@pytest.mark.parametrize("renderer", RENDERERS)
def test_sample_record_backward_equals_two_replays(gpu, renderer, scene):
    """The forward's sample record (pt: radiance and pixel per sample; pt.direct / pt.mis: the
    sums at the path's pixel and at the camera connection's pixel): its per-pixel sums make the
    film, and a backward from it (one replay) gives the gradients of two replays.  "room" (36
    triangles) runs the kernels that trace without the BVH and shadow-test where they make a
    connection; "boxes" (7k triangles) the BVH and the shadow queue."""
    import torch

    from lightmetrica_amd.gpu import record_float4s

    if scene == "room":
        sc = scenes.room_scene(48, 48, 8)
    else:
        sc = scenes.gradient_scene(48, 48, 8, num_boxes=150, max_path_vertices=9, rr_depth=3)
    gpu.upload(sc)
    gpu.build_bvh()
    n = sc.render.num_samples
    rec = torch.empty((record_float4s(renderer, n), 4), dtype=torch.float32, device="cuda")
    img, _ = gpu.render(renderer, film_precision="fp64", sample_record_ptr=rec.data_ptr())
    r = rec.cpu().numpy()
    pix = r[:, 3].view(np.uint32)
    used = pix != 0xFFFFFFFF
    if renderer != "pt":
        # every sample's second slot is either unused or a pixel other than the first slot's
        pairs = pix.reshape(-1, 2)
        assert np.all((pairs[:, 1] == 0xFFFFFFFF) | (pairs[:, 1] != pairs[:, 0]))
    film = np.zeros((48 * 48, 3))
    np.add.at(film, pix[used], r[used, :3].astype(np.float64))
    np.testing.assert_allclose(film.reshape(48, 48, 3) * (48 * 48 / n), img, rtol=1e-5, atol=1e-9)
    dLdI = np.random.default_rng(9).standard_normal(img.shape).astype(np.float32)
    a2, e2, info2 = gpu.backward(renderer, dLdI)
    a1, e1, info1 = gpu.backward(renderer, dLdI, sample_record_ptr=rec.data_ptr())
    print(f"{renderer}: residuals: two replays {info2['replay_residual']:.1e}, with record "
          f"{info1['replay_residual']:.1e}; seconds {info2['seconds']:.4f} -> {info1['seconds']:.4f}", flush=True)
    assert info1["replay_residual"] < 1e-6
    np.testing.assert_allclose(a1, a2, rtol=1e-4, atol=1e-7 * np.abs(a2).max())
    np.testing.assert_allclose(e1, e2, rtol=1e-4, atol=1e-7 * np.abs(e2).max())


def test_tile_ordered_camera_rays_adjoint(gpu):
    """The full 2000-box gradient scene (a 7 MB BVH, so pt traces its camera rays in
    screen-tile order): the adjoint matches central differences of the same renders, and
    the backward from the sample record equals the backward with two replays."""
    sc = scenes.gradient_scene(48, 48, 8, max_path_vertices=6, rr_depth=6)
    gpu.upload(sc)
    info = gpu.build_bvh()
    assert info.device_bytes >= 4 << 20
    albedo, emission = base_params(sc)
    boxes = np.asarray(sc.meta["box_materials"])
    hidden = albedo.copy()
    hidden[boxes] = scenes.hidden_box_albedo(len(boxes))
    target = render(gpu, "pt", hidden, emission)
    img = render(gpu, "pt", albedo, emission)
    _, dLdI = loss_and_grad(img, target)
    gpu.set_params(albedo, emission)
    ga, ge, info2 = gpu.backward("pt", dLdI)
    assert info2["replay_residual"] < 1e-9
    params = pick_params(sc, np.random.default_rng(11), n_albedo=24, n_light=8)
    g_fd = fd_gradient(lambda a, e: render(gpu, "pt", a, e), albedo, emission, target, params)
    compare(adjoint_at(ga, ge, params), g_fd, "pt (tile-ordered camera rays) vs GPU FD", 2e-3)
