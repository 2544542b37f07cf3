"""GPU checks of the differentiable render: forward against the oracle's cast and shading,
backward against the FP64 CPU reference gradient and FP64 central finite differences, and
the libtorch autograd Function end to end."""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c5_small():
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    scene = W.C5_WORLD.build()
    world = V.World.from_scene(scene)
    cams = W.c5_cameras(n=4, size=256)
    nvox = world.info().num_bricks * 512  # a fresh world fills slots 0 .. num_bricks - 1
    g = torch.Generator(device="cuda").manual_seed(5)
    albedo = torch.rand((nvox, 3), generator=g, device="cuda", dtype=torch.float32)
    return world, scene, cams, albedo, nvox


def check_forward(V, W, oracle_mod, world, scene, cams, albedo):
    """Forward render against the oracle's cast and the Lambert weight, pixel by pixel."""
    image, sv, sw = V.render_saved(world, albedo, cams, W.MAX_ITERS)
    torch.cuda.synchronize()
    image, sv, sw = image.cpu().numpy(), sv.cpu().numpy().astype(np.int64), sw.cpu().numpy()
    ow = oracle_mod.OracleWorld.from_stream(scene.stream())
    L = np.array(V.sun_direction(), np.float32)
    alb = albedo.cpu().numpy()
    base = 0
    keys, ids = [], []
    for cam in cams:
        o, d = cam.rays_host()
        r = ow.cast_simd(o, d, cam.world_origin, W.MAX_ITERS)
        n = cam.n
        rec = r["rec"]
        hit = (rec[:, 3] & 0xFF) != 0
        s_v, s_w = sv[base:base + n], sw[base:base + n]
        assert np.array_equal(s_v >= 0, hit)
        info = rec[:, 3].astype(np.int64)
        N = np.stack([(((info >> (8 + 2 * a)) & 3) ^ 2) - 2 for a in range(3)], 1).astype(np.float32)
        ndl = N[:, 0] * L[0] + N[:, 1] * L[1] + N[:, 2] * L[2]
        wgt = np.maximum(ndl, np.float32(0)) * np.float32(5.0)
        assert np.array_equal(s_w[hit].view(np.uint32), wgt[hit].astype(np.float32).view(np.uint32))
        # the saved id is slot * 512 + voxel index in the brick, one id per voxel
        vx, vy, vz = rec[hit, 0].astype(np.int64), rec[hit, 1].astype(np.int64), rec[hit, 2].astype(np.int64)
        vi = (vx & 7) | (vz & 7) << 3 | (vy & 7) << 6
        assert np.array_equal(s_v[hit] & 511, vi)
        keys.append(vx | vy << 20 | vz << 40)
        ids.append(s_v[hit])
        exp = alb[s_v[hit]] * s_w[hit][:, None]
        assert np.array_equal(image[base:base + n][hit].view(np.uint32), exp.astype(np.float32).view(np.uint32))
        assert np.all(image[base:base + n][~hit] == 0)
        base += n
    # voxel <-> id is one to one: as many distinct (voxel, id) pairs as distinct voxels and ids
    keys, ids = np.concatenate(keys), np.concatenate(ids)
    nk, ni = len(np.unique(keys)), len(np.unique(ids))
    assert nk == ni == len(np.unique(np.stack([keys, ids], 1), axis=0))
    return base


# the camera anchor picks the traversal's aligned (sector), general (floor) or zero-origin form
@pytest.mark.parametrize("anchor", ["sector", "floor", "zero"])
def test_forward_matches_oracle(vrt, oracle_mod, cuda, c5_small, anchor):
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, albedo, nvox = c5_small
    if anchor != "sector":
        cams = W.c5_cameras(n=4, size=256, anchor=anchor)
    check_forward(V, W, oracle_mod, world, scene, cams, albedo)


def dense_gradient(V, world, grad_image, sv, sw, nvox, mode):
    """The albedo gradient of one backward form, as float64 [nvox, 3] on the host."""
    out = V.render_backward(world, grad_image, sv, sw, nvox, mode=mode)
    torch.cuda.synchronize()
    if mode != "sparse":
        return out.cpu().numpy().astype(np.float64)
    index, values = out
    dense = np.zeros((nvox, 3), np.float64)
    np.add.at(dense, index.cpu().numpy(), values.cpu().numpy().astype(np.float64))
    return dense


@pytest.mark.parametrize("mode", ["atomic", "bricks", "sparse"])
def test_backward_matches_reference_and_finite_differences(vrt, oracle_mod, cuda, c5_small, mode):
    """Tolerance 1e-4 relative to the sum of |terms| of each voxel's gradient: the scale of the
    rounding error of fp32 sums in any order (and of the fp32 image the loss is taken on)."""
    check_backward(oracle_mod, *c5_small, mode=mode)


def check_backward(oracle_mod, world, scene, cams, albedo, nvox, mode="bricks"):
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    g = torch.Generator(device="cuda").manual_seed(6)
    target_albedo = (albedo + 0.05 * torch.randn(albedo.shape, generator=g, device="cuda")).contiguous()
    target, _, _ = V.render_saved(world, target_albedo, cams, W.MAX_ITERS)
    image, sv, sw = V.render_saved(world, albedo, cams, W.MAX_ITERS)
    grad_image = 2.0 * (image - target)
    grad = dense_gradient(V, world, grad_image, sv, sw, nvox, mode)
    sv_h, sw_h = sv.cpu().numpy().view(np.uint32), sw.cpu().numpy()
    alb_h, tgt_h = albedo.cpu().numpy(), target.cpu().numpy()
    ref, abssum = oracle_mod.render_backward(sv_h, sw_h, grad_image.cpu().numpy(), nvox)
    assert np.abs(ref).max() > 0
    err = np.abs(grad - ref)
    assert np.all(err <= 1e-4 * abssum + 1e-12), float((err / (abssum + 1e-30)).max())
    # FP64 central differences of the FP64 loss on 256 seeded voxels that the views hit
    hit_ids = np.unique(sv_h[sv_h != 0xFFFFFFFF])
    rng = np.random.default_rng(7)
    picks = np.sort(rng.choice(hit_ids, size=min(256, len(hit_ids)), replace=False))
    # the loss restricted to the pixels that hit the picked voxels depends on their albedo only:
    # renumber them 0..255 so each finite difference perturbs a 256 x 3 array
    mask = np.isin(sv_h, picks)
    compact = np.searchsorted(picks, sv_h[mask]).astype(np.uint32)
    sw_sub, tgt_sub = sw_h[mask], tgt_h[mask]
    alb_c = alb_h[picks].astype(np.float64)
    # the loss the GPU differentiates is taken on the fp32 image: each pixel's residual carries
    # up to half an ulp of albedo * w, i.e. 2^-24 |a w| * 2w per gradient term
    img_err = np.zeros((len(picks), 3))
    np.add.at(img_err, compact, 2.0 ** -22 * np.abs(alb_c[compact] * sw_sub[:, None]) * sw_sub[:, None])
    h = 1e-3
    checked = 0
    for k, v in enumerate(picks):
        for c in range(3):
            a = alb_c.copy()
            a[k, c] += h
            lp = oracle_mod.render_loss(compact, sw_sub, a, tgt_sub)
            a[k, c] -= 2 * h
            lm = oracle_mod.render_loss(compact, sw_sub, a, tgt_sub)
            fd = (lp - lm) / (2 * h)
            tol = 1e-4 * abssum[v, c] + img_err[k, c] + 1e-9
            assert abs(grad[v, c] - fd) <= tol, (v, c, grad[v, c], fd, abssum[v, c])
            checked += 1
    assert checked >= 3 * min(256, len(hit_ids))


@pytest.fixture(scope="module")
def c5_full():
    """The differentiable configuration at full size: 512^3 terrain, 16 views of 512x512."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    scene = W.C5_WORLD.build()
    world = V.World.from_scene(scene)
    cams = W.c5_cameras(n=16, size=512)
    nvox = world.info().num_bricks * 512
    g = torch.Generator(device="cuda").manual_seed(5)
    albedo = torch.rand((nvox, 3), generator=g, device="cuda", dtype=torch.float32)
    return world, scene, cams, albedo, nvox


def test_full_configuration_forward_matches_oracle(vrt, oracle_mod, cuda, c5_full):
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, albedo, nvox = c5_full
    assert check_forward(V, W, oracle_mod, world, scene, cams, albedo) == 16 * 512 * 512


@pytest.mark.parametrize("mode", ["bricks", "sparse"])
def test_full_configuration_backward_matches_reference(vrt, oracle_mod, cuda, c5_full, mode):
    """16 x 512^2 pixels into 49.9 M albedo voxels: the gradient against the FP64 CPU reference
    gradient and FP64 central differences on 256 voxels, as for the small views."""
    check_backward(oracle_mod, *c5_full, mode=mode)


@pytest.mark.parametrize("mode", ["bricks", "atomic", "sparse"])
def test_autograd_function(vrt, cuda, c5_small, mode):
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, albedo, nvox = c5_small
    target, _, _ = V.render_saved(world, (albedo * 0.9).contiguous(), cams, W.MAX_ITERS)
    a = albedo.clone().requires_grad_(True)
    image = V.render(world, a, cams, W.MAX_ITERS, grad=mode)
    loss = ((image - target) ** 2).sum()
    loss.backward()
    image2, sv, sw = V.render_saved(world, albedo, cams, W.MAX_ITERS)
    assert torch.equal(image.detach(), image2)
    manual = V.render_backward(world, 2.0 * (image2 - target), sv, sw, nvox)
    torch.cuda.synchronize()
    assert a.grad.shape == albedo.shape
    assert a.grad.is_sparse == (mode == "sparse")
    got = a.grad.to_dense() if mode == "sparse" else a.grad
    err = (got - manual).abs().max().item()
    assert err <= 1e-4 * manual.abs().max().item()


def test_bricks_backward_writes_every_row(vrt, cuda, c5_small):
    """The written-whole gradient over an albedo longer than the slots (a partial last brick):
    rows no pixel hit are 0 whatever the buffer held, the rest equal the atomic form."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, albedo, nvox = c5_small
    n2 = nvox + 700
    alb = torch.cat([albedo, torch.rand((700, 3), device="cuda")]).contiguous()
    image, sv, sw = V.render_saved(world, alb, cams[:2], W.MAX_ITERS)
    gi = torch.randn(image.shape, device="cuda")
    ref = V.render_backward(world, gi, sv, sw, n2, mode="atomic")
    grad = torch.empty((n2, 3), device="cuda").fill_(float("nan"))
    nbytes = int(V.lib().vrt_render_bwd_workspace_bytes(sv.numel(), n2))
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    from voxelrt_amd import _lib as L

    L.check(V.lib().vrt_render_bwd_bricks(world.handle, sv.numel(), gi.data_ptr(), sv.data_ptr(), sw.data_ptr(), n2,
                                          grad.data_ptr(), ws.data_ptr(), nbytes, None))
    torch.cuda.synchronize()
    assert not bool(torch.isnan(grad).any())
    touched = torch.zeros(n2, dtype=torch.bool, device="cuda")
    touched[sv[(sv >= 0) & (sw != 0)].long()] = True
    assert bool((grad[~touched] == 0).all())
    assert float((grad - ref).abs().max()) <= 1e-4 * float(ref.abs().max())
    with pytest.raises(V.VrtError):  # shorter than the slots the world holds
        L.check(V.lib().vrt_render_bwd_bricks(world.handle, sv.numel(), gi.data_ptr(), sv.data_ptr(), sw.data_ptr(),
                                              nvox - 512, grad.data_ptr(), ws.data_ptr(), nbytes, None))


def test_forward_many_views_one_launch(vrt, cuda, c5_small):
    """18 views of uneven sizes (pixel counts that are not multiples of 64, more views than one
    launch takes): the same pixels, ids and weights as rendering each view on its own."""
    import math

    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, albedo, nvox = c5_small
    views = []
    for k in range(18):
        a = 2 * math.pi * k / 18
        pos = (256 + 180 * math.sin(a) + 0.13, 400.33, 256 + 180 * math.cos(a) + 0.41)
        w, h = 40 + 7 * k, 24 + 3 * (k % 5)
        views.append(V.CameraSpec(pos, math.degrees(a) + 180.0, -30.0, w, h, tile=(1, 1)))
    image, sv, sw = V.render_saved(world, albedo, views, W.MAX_ITERS)
    base = 0
    for cam in views:
        i1, v1, w1 = V.render_saved(world, albedo, [cam], W.MAX_ITERS)
        n = cam.n
        assert torch.equal(image[base:base + n], i1) and torch.equal(sv[base:base + n], v1)
        assert torch.equal(sw[base:base + n], w1)
        base += n
    assert base == image.shape[0] and bool((sv >= 0).any())
