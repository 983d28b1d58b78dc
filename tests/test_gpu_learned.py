"""GPU: the learned reconstruction library (libspimlearn.so) against a plain PyTorch float32
reference of the same network, and the device-side volume transforms it relies on.

The reference stores activations rounded to bf16 exactly where the kernels do (after each
32-channel layer, pooling and skip addition) and rounds the 32->32 weights to bf16, so the
remaining differences are float32 summation order and the bf16 rounding flips it causes:
forward outputs agree to 2% of their range (1e-3 mean); gradients (whose 32->32 input
gradients the kernels take in bf16) agree per parameter block to cosine >= 0.99 and 10%
relative L2 error."""
import numpy as np
import pytest

from spimagine_amd import _lib, learned, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert _lib.device_count() > 0
    return t


def ref_forward(torch, params, x):
    return ref_forward_t(torch, torch.from_numpy(np.asarray(params, np.float32)), x)


def random_params(seed):
    rng = np.random.default_rng(seed)
    flat = learned.init_params(seed)
    p = learned.split_params(flat)
    for k in p:
        if k.endswith(".b"):
            p[k][...] = rng.normal(0, .05, p[k].shape)
    p["head.w"][...] = rng.normal(0, .05, p["head.w"].shape)
    return flat


def smooth_inputs(rng, n, h, w):
    from scipy.ndimage import gaussian_filter
    x = rng.uniform(0, 1, (n, 2, h, w)).astype(np.float32)
    return np.stack([gaussian_filter(v, (0, 1.5, 1.5)) for v in x]).astype(np.float32)


@pytest.mark.parametrize("n,h,w", [(1, 64, 64), (2, 48, 80), (1, 36, 100)])
def test_forward_matches_torch(torch, n, h, w):
    rng = np.random.default_rng(n * 1000 + h + w)
    params = random_params(h + w)
    x = smooth_inputs(rng, n, h, w)
    net = learned.Net(0, params)
    dx = torch.from_numpy(x).cuda()
    dy = torch.full((n, 2 * h, 2 * w), -9., device="cuda")
    torch.cuda.synchronize()  # the fill runs on torch's stream, the network on its own
    net.forward(dx.data_ptr(), n, h, w, dy.data_ptr())
    torch.cuda.synchronize()
    got = dy.cpu().numpy()
    ref = ref_forward(torch, params, torch.from_numpy(x)).numpy()
    scale = np.abs(ref).max()
    d = np.abs(got - ref)
    assert d.max() <= 2e-2 * scale, (d.max(), scale)
    assert d.mean() <= 1e-3 * scale, (d.mean(), scale)
    np.testing.assert_array_equal(net.get_params(), params.astype(np.float32))
    net.close()


def test_gradients_match_torch_autograd(torch):
    n, h, w = 2, 32, 48
    rng = np.random.default_rng(5)
    params = random_params(11)
    x = smooth_inputs(rng, n, h, w)
    tgt = rng.uniform(0, 1, (n, 2 * h, 2 * w)).astype(np.float32)
    net = learned.Net(0, params)
    dx, dt = torch.from_numpy(x).cuda(), torch.from_numpy(tgt).cuda()
    g, sse = net.gradients(dx.data_ptr(), dt.data_ptr(), n, h, w)
    net.close()
    pt = torch.from_numpy(params.copy()).requires_grad_(True)
    # autograd through the reference (bf16 rounding passes the gradient straight through)
    y = ref_forward_t(torch, pt, torch.from_numpy(x))
    loss = ((y - torch.from_numpy(tgt)) ** 2).sum() / y.numel()
    loss.backward()
    gr = pt.grad.numpy()
    ref_sse = float(((y.detach() - torch.from_numpy(tgt)) ** 2).sum())
    assert abs(sse - ref_sse) <= 2e-2 * ref_sse
    k = 0
    for name, shape in learned.param_layout():
        m = int(np.prod(shape))
        a, b = g[k:k + m].astype(np.float64), gr[k:k + m].astype(np.float64)
        k += m
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        assert cos >= .99 and rel <= .1, (name, cos, rel)


def ref_forward_t(torch, flat, x):
    """ref_forward on a torch parameter vector (keeps the autograd graph)."""
    F = torch.nn.functional
    p, k = {}, 0
    for name, shape in learned.param_layout():
        m = int(np.prod(shape))
        p[name] = flat[k:k + m].reshape(shape)
        k += m

    def bf(t):
        return t.to(torch.bfloat16).to(torch.float32)

    def conv(t, w, b, round_w):
        cin, cout = w.shape[1], w.shape[2]
        wt = w.reshape(3, 3, cin, cout).permute(3, 2, 0, 1).contiguous()
        return F.conv2d(t, bf(wt) if round_w else wt, b, padding=1)

    a0 = bf(F.relu(conv(x, p["in.w"], p["in.b"], False)))
    e0 = bf(F.relu(conv(a0, p["e0.w"], p["e0.b"], True)))
    e0b = bf(F.relu(conv(e0, p["e0b.w"], p["e0b.b"], True)))
    p1 = bf(F.avg_pool2d(e0b, 2))
    e1 = bf(F.relu(conv(p1, p["e1.w"], p["e1.b"], True)))
    e1b = bf(F.relu(conv(e1, p["e1b.w"], p["e1b.b"], True)))
    p2 = bf(F.avg_pool2d(e1b, 2))
    e2a = bf(F.relu(conv(p2, p["e2a.w"], p["e2a.b"], True)))
    e2 = bf(F.relu(conv(e2a, p["e2.w"], p["e2.b"], True)))
    s1 = bf(F.interpolate(e2, scale_factor=2, mode="nearest") + e1b)
    u1 = bf(F.relu(conv(s1, p["u1.w"], p["u1.b"], True)))
    s0 = bf(F.interpolate(u1, scale_factor=2, mode="nearest") + e0b)
    u0 = bf(F.relu(conv(s0, p["u0.w"], p["u0.b"], True)))
    u0b = bf(F.relu(conv(u0, p["u0b.w"], p["u0b.b"], True)))
    r = conv(u0b, p["head.w"], p["head.b"], False)
    base = F.interpolate(x[:, :1], scale_factor=2, mode="bilinear", align_corners=False)
    return (F.pixel_shuffle(r, 2) + base)[:, 0]


def test_training_reduces_the_loss(torch):
    n, h, w = 4, 32, 32
    rng = np.random.default_rng(2)
    x = smooth_inputs(rng, n, h, w)
    # a target the network can represent: a fixed smooth function of the input planes
    up = np.repeat(np.repeat(x, 2, axis=2), 2, axis=3)
    tgt = np.clip(.7 * up[:, 0] + .3 * up[:, 1] ** 2, 0, 1).astype(np.float32)
    net = learned.Net(0, learned.init_params(3))
    dx, dt = torch.from_numpy(x).cuda(), torch.from_numpy(tgt).cuda()
    first = net.train_step(dx.data_ptr(), dt.data_ptr(), n, h, w, 1e-3)
    for _ in range(60):
        last = net.train_step(dx.data_ptr(), dt.data_ptr(), n, h, w, 1e-3)
    assert last < .5 * first, (first, last)
    net.close()


@pytest.mark.parametrize("dtype", [np.uint16, np.float32, np.uint8])
def test_dilated_volume_is_scipy_max_filter(torch, dtype):
    from scipy.ndimage import maximum_filter
    rng = np.random.default_rng(4)
    shape = (19, 23, 37)
    if dtype == np.float32:
        vol = rng.normal(0, 1, shape).astype(np.float32)
    else:
        vol = rng.integers(0, np.iinfo(dtype).max, shape).astype(dtype)
    ctx = _lib.Context(0)
    for layout in (_lib.LAYOUT_PADDED, _lib.LAYOUT_LINEAR):
        ctx.set_layout(layout)
        s = ctx.volume_add(vol)
        np.testing.assert_array_equal(ctx.volume_read(s), vol)
        for r in (1, 2):
            d = ctx.volume_add_dilated(s, r)
            np.testing.assert_array_equal(ctx.volume_read(d),
                                          maximum_filter(vol, size=2 * r + 1, mode="nearest"))
            ctx.volume_release(d)
        ctx.volume_release(s)


def test_learned_render_path_shapes_and_fallback(torch):
    vol = synth.blob_volume((48, 64, 64), n_blobs=12, sigma=(3., 6.), amp=(500., 3000.),
                            background=100., poisson=True, read_noise=5., dtype="uint16", seed=1)
    ctx = _lib.Context(0)
    slot = ctx.volume_add(vol)
    # threshold above any PSNR: validation must switch the learned path off
    lm = learned.LearnedMIP(ctx, params=learned.init_params(0), threshold=float("inf"))
    from spimagine_amd.utils import mat4_perspective
    from spimagine_amd.volumerender.volumerender import render_matrices
    P = mat4_perspective(45, 1., .1, 10)
    ps = []
    for M in synth.random_poses(2, seed=3):
        iP, iM = render_matrices(M, P, (64, 64, 48))
        ps.append(_lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), float(vol.min()), float(vol.max()),
                                   1., 0., .1, 21, 30, 128, _lib.FILTER_LINEAR))
    rep = lm.validate(slot, ps, 96, 96)
    assert len(rep) == 2 and all(np.isfinite(r["psnr"]) for r in rep)
    assert lm.enabled[slot] is False
    s = torch.cuda.Stream()
    o = torch.empty((96, 96), device="cuda")
    a = torch.empty_like(o)
    assert lm.render_or_exact(slot, ps[0], 96, 96, o.data_ptr(), a.data_ptr(), s.cuda_stream) is False
    s.synchronize()
    ref = torch.empty_like(o)
    ctx.render_slot(slot, ps[0], _lib.METHOD_MAX_PROJECT, 96, 96, ref.data_ptr(), a.data_ptr(),
                    stream=s.cuda_stream)
    s.synchronize()
    assert torch.equal(o, ref)


@pytest.mark.parametrize("W,H", [(128, 128), (256, 192)])
def test_sparse_judge_rays_are_the_full_frame_pixels(torch, W, H):
    """A (w/8) x (h/8) render with the same matrices traces the rays of pixels (8i, 8j) of
    the w x h frame, bit for bit (the judge shifts that grid to (8i+1, 8j+1), up to float
    rounding); the judge's estimate is the learned frame's PSNR on those pixels;
    render_judged falls back to the exact frame below its threshold and keeps the learned
    frame above it."""
    vol = synth.blob_volume((64, 80, 72), n_blobs=14, sigma=(3., 6.), amp=(500., 3000.),
                            background=100., poisson=True, read_noise=5., dtype="uint16", seed=2)
    ctx = _lib.Context(0)
    slot = ctx.volume_add(vol)
    from spimagine_amd.utils import mat4_perspective
    from spimagine_amd.volumerender.volumerender import render_matrices
    P = mat4_perspective(45, W / float(H), .1, 10)
    M = synth.random_poses(1, seed=9)[0]
    iP, iM = render_matrices(M, P, (72, 80, 64))
    p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), float(vol.min()), float(vol.max()), 1., 0.,
                         .1, 21, 30, 160, _lib.FILTER_LINEAR)
    s = torch.cuda.Stream()
    full = torch.empty((H, W), device="cuda")
    fa = torch.empty_like(full)
    sparse = torch.empty((H // 8, W // 8), device="cuda")
    sa = torch.empty_like(sparse)
    ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, W, H, full.data_ptr(), fa.data_ptr(),
                    stream=s.cuda_stream)
    ctx.render_slot(slot, learned.sparse_params(p, W, H, offset=0), _lib.METHOD_MAX_PROJECT, W // 8,
                    H // 8, sparse.data_ptr(), sa.data_ptr(), stream=s.cuda_stream)
    s.synchronize()
    assert torch.equal(sparse, full[::8, ::8]) and torch.equal(sa, fa[::8, ::8])
    out = torch.empty_like(full)
    alpha = torch.empty_like(full)
    strict = learned.LearnedMIP(ctx, params=learned.init_params(0), threshold=float("inf"))
    kept, est = strict.render_judged(slot, p, out, alpha, s)
    s.synchronize()
    assert kept is False and np.isfinite(est)
    assert torch.equal(out, full) and torch.equal(alpha, fa)
    lax = learned.LearnedMIP(ctx, params=learned.init_params(0), threshold=-float("inf"))
    kept, est2 = lax.render_judged(slot, p, out, alpha, s)
    ref = torch.empty_like(full)
    lax.render(slot, p, W, H, ref.data_ptr(), s.cuda_stream)
    s.synchronize()
    assert kept is True
    assert torch.equal(out, ref)
    # the judge's pixels: (S i + 1, S j + 1), traced by the shifted sparse render
    S = learned.JUDGE_STRIDE
    shifted = torch.empty((H // S, W // S), device="cuda")
    sa2 = torch.empty_like(shifted)
    ctx.render_slot(slot, learned.sparse_params(p, W, H), _lib.METHOD_MAX_PROJECT, W // S, H // S,
                    shifted.data_ptr(), sa2.data_ptr(), stream=s.cuda_stream)
    s.synchronize()
    mse = float(((ref[1::S, 1::S].double() - shifted.double()) ** 2).mean())
    assert abs(est2 - 10 * np.log10(1. / mse)) < 1e-3
    # and they are the full frame's rays at those pixels up to float rounding
    assert float((shifted - full[1::S, 1::S]).abs().max()) < 1e-3
    # the strict judge rejected the frame at the first stage and judged it again on all
    # four grids, i.e. on the pixels (S/2 i + 1, S/2 j + 1)
    h2 = S // 2
    mse4 = float(((ref[1::h2, 1::h2].double() - full[1::h2, 1::h2].double()) ** 2).mean())
    assert abs(est - 10 * np.log10(1. / mse4)) < .05
    # learned frames carry the coarse pass's alpha: exact at even pixels
    assert torch.equal(alpha[::2, ::2], fa[::2, ::2])


def test_judge_sse_patch_layout(torch):
    """spl_judge_sse on patches of 3 x 3 pixels at offset (2, 1) of 5 x 4 cells of a frame
    of arange values, against the same sum on the host."""
    net = learned.Net(0, learned.init_params(0))
    w, h, c, gx, gy, ox, oy = 40, 24, 3, 5, 4, 2, 1
    frame = torch.arange(w * h, device="cuda", dtype=torch.float32).reshape(h, w) / (w * h)
    pat = torch.rand((gy * gx, c, c), device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    sse = torch.full((2,), 5., device="cuda")
    torch.cuda.synchronize()
    net.judge_sse(frame.data_ptr(), w, h, pat.data_ptr(), c, gx, gy, ox, oy, sse.data_ptr())
    net.judge_sse(frame.data_ptr(), w, h, pat.data_ptr(), c, gx, gy, ox, oy, sse.data_ptr() + 4,
                  accumulate=True)
    torch.cuda.synchronize()
    F, Pt = frame.cpu().double().numpy(), pat.cpu().double().numpy()
    want = 0.
    for ty in range(gy):
        for tx in range(gx):
            y0, x0 = ty * (h // gy) + oy, tx * (w // gx) + ox
            want += ((F[y0:y0 + c, x0:x0 + c] - Pt[ty * gx + tx]) ** 2).sum()
    got = sse.cpu().double().numpy()
    assert abs(got[0] - want) < 1e-4 * want and abs(got[1] - 5. - want) < 1e-4 * want
    with pytest.raises(learned.SplError):
        net.judge_sse(frame.data_ptr(), w, h, pat.data_ptr(), c, gx, gy, 6, oy, sse.data_ptr())


def test_judge_sse_frames_layout(torch):
    """spl_judge_sse_frames on 21 frames (two launches of at most 16) against their sparse
    grids of 9 x 7 pixels at offset (3, 2), against the same sums on the host."""
    net = learned.Net(0, learned.init_params(0))
    n, w, h, gx, gy, ox, oy = 21, 90, 56, 9, 7, 3, 2
    g = torch.Generator("cuda").manual_seed(4)
    frames = [torch.rand((h, w), device="cuda", generator=g) for _ in range(n)]
    sparse = torch.rand((n, gy, gx), device="cuda", generator=g)
    sse = torch.full((n,), -1., device="cuda")
    torch.cuda.synchronize()
    sse4 = torch.full((n,), -1., device="cuda")
    net.judge_sse_frames([f.data_ptr() for f in frames], w, h, sparse.data_ptr(), gx, gy, ox, oy,
                         sse.data_ptr(), sse4.data_ptr())
    torch.cuda.synchronize()
    S = sparse.cpu().double().numpy()
    got, got4 = sse.cpu().double().numpy(), sse4.cpu().double().numpy()
    for k, f in enumerate(frames):
        F = f.cpu().double().numpy()
        d2 = (F[oy::h // gy, ox::w // gx][:gy, :gx] - S[k]) ** 2
        assert abs(got[k] - d2.sum()) < 1e-4 * d2.sum(), k
        assert abs(got4[k] - (d2 ** 2).sum()) < 1e-4 * (d2 ** 2).sum(), k
    # the bounded decision: keep iff m2 + z sqrt((m4 - m2^2) / npix) <= mse_max
    npix = gx * gy
    m2, m4 = got / npix, got4 / npix
    ucb = m2 + 2.5 * np.sqrt(np.maximum(m4 - m2 * m2, 0) / npix)
    flags = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    mse_max = float(np.median(ucb))
    net.judge_flags_bound(sse.data_ptr(), sse4.data_ptr(), n, npix, mse_max, 2.5, flags.data_ptr())
    torch.cuda.synchronize()
    want = (ucb > mse_max).astype(np.int32)
    close = np.abs(ucb - mse_max) < 1e-6 * mse_max  # float32 rounding at the boundary
    np.testing.assert_array_equal(flags.cpu().numpy()[~close], want[~close])
    assert 0 < want.sum() < n
    with pytest.raises(learned.SplError):
        net.judge_sse_frames([f.data_ptr() for f in frames], w, h, sparse.data_ptr(), gx, gy,
                             w // gx, oy, sse.data_ptr())


@pytest.mark.parametrize("dtype,shape", [(np.uint16, (64, 80, 72)), (np.float32, (48, 64, 64)),
                                         (np.uint8, (40, 52, 60))])
def test_coarse_pair_matches_exact_renders(torch, dtype, shape):
    """spr_render_coarse_pair (sample k at p0 + k d, rays split over four waves, both slots
    in one launch) against spr_render_slot on each slot: the sample positions differ only
    in their last bits, so the planes agree to 1e-4 of the normalised range (the
    north-star bar); alpha is identical."""
    amp = (.3, 1.) if dtype == np.float32 else (500., 3000.) if dtype == np.uint16 else (60., 200.)
    vol = synth.blob_volume(shape, n_blobs=14, sigma=(3., 6.), amp=amp, background=amp[0] / 10,
                            poisson=False, read_noise=0., dtype=dtype, seed=5)
    ctx = _lib.Context(0)
    a = ctx.volume_add(vol)
    b = ctx.volume_add_dilated(a, 2)
    from spimagine_amd.utils import mat4_perspective
    from spimagine_amd.volumerender.volumerender import render_matrices
    P = mat4_perspective(45, 1., .1, 10)
    s = torch.cuda.Stream()
    for M in synth.random_poses(3, seed=4):
        iP, iM = render_matrices(M, P, shape[::-1])
        for steps in (64, 128):
            p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), float(vol.min()), float(vol.max()),
                                 1., 0., .1, 21, 30, steps, _lib.FILTER_LINEAR)
            W, H = 100, 84
            pair = torch.full((2, H, W), -5., device="cuda")
            pa = torch.full((H, W), -5., device="cuda")
            torch.cuda.synchronize()
            ctx.render_coarse_pair(a, b, p, W, H, pair.data_ptr(), pa.data_ptr(), stream=s.cuda_stream)
            refs = []
            for slot in (a, b):
                o = torch.empty((H, W), device="cuda")
                al = torch.empty_like(o)
                ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, W, H, o.data_ptr(), al.data_ptr(),
                                stream=s.cuda_stream)
                refs.append((o, al))
            s.synchronize()
            for j in range(2):
                d = (pair[j] - refs[j][0]).abs().max().item()
                assert d <= 1e-4, (j, steps, d)
            assert torch.equal(pa, refs[0][1])
            assert (refs[0][0] > 0).sum().item() > 1000


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
def test_pooled_volumes(torch, dtype):
    """spr_volume_add_pooled: 2x2x2 mean (rounded to the voxel type) and max, odd sizes
    clamped at the far edge, from padded and linear slots."""
    rng = np.random.default_rng(8)
    shape = (17, 22, 35)
    if dtype == np.float32:
        vol = rng.normal(0, 1, shape).astype(np.float32)
    else:
        vol = rng.integers(0, np.iinfo(dtype).max, shape).astype(dtype)
    pad = np.pad(vol, ((0, 1), (0, 0), (0, 1)), mode="edge")  # odd z and x: clamp
    blocks = pad.reshape(9, 2, 11, 2, 18, 2).astype(np.float64)
    want_max = blocks.max(axis=(1, 3, 5)).astype(dtype)
    mean = blocks.sum(axis=(1, 3, 5)) / 8.
    ctx = _lib.Context(0)
    for layout in (_lib.LAYOUT_PADDED, _lib.LAYOUT_LINEAR):
        ctx.set_layout(layout)
        s = ctx.volume_add(vol)
        mx = ctx.volume_read(ctx.volume_add_pooled(s, 1))
        mn = ctx.volume_read(ctx.volume_add_pooled(s, 0))
        np.testing.assert_array_equal(mx, want_max)
        if dtype == np.float32:
            np.testing.assert_allclose(mn, mean, rtol=1e-6, atol=1e-6)
        else:
            np.testing.assert_array_equal(mn, np.rint(mean).astype(dtype))


def test_pooled_pair_volume_channels():
    """spr_volume_add_pooled_pair: channel 0 = the mean pooling, channel 1 = the max pooling,
    each equal to spr_volume_add_pooled's volume; the pair slot is refused by the
    one-channel entry points."""
    rng = np.random.default_rng(9)
    vol = rng.integers(0, 65535, (17, 22, 35)).astype(np.uint16)
    ctx = _lib.Context(0)
    for layout in (_lib.LAYOUT_PADDED, _lib.LAYOUT_LINEAR):
        ctx.set_layout(layout)
        s = ctx.volume_add(vol)
        pair = ctx.volume_add_pooled_pair(s)
        info = ctx.volume_info(pair)
        assert info["dtype"] == _lib.DTYPE_U16X2 and info["shape"] == (9, 11, 18)
        bits = ctx.volume_read(pair)
        np.testing.assert_array_equal((bits & 0xFFFF).astype(np.uint16),
                                      ctx.volume_read(ctx.volume_add_pooled(s, 0)))
        np.testing.assert_array_equal((bits >> 16).astype(np.uint16),
                                      ctx.volume_read(ctx.volume_add_pooled(s, 1)))
        eye = np.eye(4, dtype=np.float32)
        with pytest.raises(_lib.SprError):
            ctx.volume_add_pooled(pair, 0)
        with pytest.raises(_lib.SprError):
            ctx.render_slot(pair, _lib.make_params(eye, eye, max_steps=64), _lib.METHOD_MAX_PROJECT,
                            16, 16, 0, 0)
        ctx.volume_release(pair)
    with pytest.raises(_lib.SprError):
        ctx.volume_add_pooled_pair(ctx.volume_add(vol.astype(np.float32)))


@pytest.mark.parametrize("shape,steps,parts", [((64, 80, 72), 224, 1), ((128, 128, 128), 224, 1),
                                               ((40, 70, 50), 160, 2)])
def test_pair_render_planes_equal_single_renders(torch, shape, steps, parts):
    """spr_render_pair on a pooled pair: plane 0 is bit-identical to spr_render_slot on the
    mean-pooled slot, plane 1 to the max-pooled slot, alpha to both (multipass included);
    rendered with the staged kernel's chunk of 16, 8 and 4 samples."""
    vol = synth.blob_volume(shape, n_blobs=20, sigma=(2., 6.), amp=(500., 3000.), background=100.,
                            poisson=True, read_noise=5., dtype="uint16", seed=4)
    ctx = _lib.Context(0)
    s0 = ctx.volume_add(vol)
    pair = ctx.volume_add_pooled_pair(s0)
    mean, mx = ctx.volume_add_pooled(s0, 0), ctx.volume_add_pooled(s0, 1)
    from spimagine_amd.utils import mat4_perspective
    from spimagine_amd.volumerender.volumerender import render_matrices
    nz, ny, nx = ctx.volume_info(pair)["shape"]
    W, H = 150, 118
    P = mat4_perspective(45, W / float(H), .1, 10)
    s = torch.cuda.Stream()
    for k, M in enumerate(synth.random_poses(3, seed=21)):
        iP, iM = render_matrices(M, P, (nx, ny, nz))
        p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), 100., 3000., 1., 0., .1, 21, 30, steps,
                             _lib.FILTER_LINEAR)
        for chunk in (0, 16, 8, 4):
            ctx.set_tuning("chunk", chunk)
            planes = torch.empty((2, H, W), device="cuda")
            pa = torch.empty((H, W), device="cuda")
            refs = [(torch.empty((H, W), device="cuda"), torch.empty((H, W), device="cuda"))
                    for _ in range(2)]
            for part in range(parts):
                ctx.render_pair(pair, p, W, H, planes.data_ptr(), pa.data_ptr(), num_parts=parts,
                                current_part=part, stream=s.cuda_stream)
                for (o, a), slot in zip(refs, (mean, mx)):
                    ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, W, H, o.data_ptr(), a.data_ptr(),
                                    num_parts=parts, current_part=part, stream=s.cuda_stream)
            s.synchronize()
            assert torch.equal(planes[0], refs[0][0]), (k, chunk)
            assert torch.equal(planes[1], refs[1][0]), (k, chunk)
            assert torch.equal(pa, refs[0][1]) and torch.equal(pa, refs[1][1])
            assert (refs[1][0] > 0).sum().item() > 1000
    ctx.set_tuning("chunk", 0)
    with pytest.raises(_lib.SprError):  # plain max projections only
        ctx.render_pair(pair, _lib.make_params(iP, iM, alpha_pow=.5, max_steps=64), W, H,
                        planes.data_ptr(), pa.data_ptr(), stream=s.cuda_stream)
