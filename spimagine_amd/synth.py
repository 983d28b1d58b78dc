"""Seeded synthetic microscopy volumes (Gaussian blobs + camera noise) and camera poses.

These are the workloads the renderer is measured and checked on:
  parity   128^3 float32, 40 blobs, sigma U[2,6] per axis, amplitude U[0.3,1],
           +0.05 offset, + N(0, 0.02^2) noise -> 256^2, max_steps 224
  iso      512^3 uint16, 2000 blobs, sigma U[3,10], amplitude U[500,3000], background
           100, Poisson shot noise + N(0, 5^2) read noise, clipped -> 1024^2, max_steps 896
  headline 1024x1024x512 (x,y,z) uint16, 8000 blobs, same recipe -> 1024^2,
           max_steps 1536 (<= 1 voxel per step)
Blob parameters come from numpy.random.default_rng(seed); each blob is an axis-aligned
anisotropic Gaussian evaluated within +-3 sigma of its centre.  ``blob_volume`` builds the
volume with numpy on the host; ``blob_volume_device`` builds the same blobs with torch on
a GPU (the noise then comes from torch's generator, seeded the same way), which keeps the
1 GiB headline volume's generation to seconds.
"""
import math

import numpy as np

from .utils.transform_matrices import mat4_perspective, mat4_rotation, mat4_translate

CONFIGS = {
    "parity": dict(shape=(128, 128, 128), n_blobs=40, sigma=(2., 6.), amp=(.3, 1.),
                   background=.05, poisson=False, read_noise=.02, dtype="float32",
                   out=(256, 256), max_steps=224),
    "iso": dict(shape=(512, 512, 512), n_blobs=2000, sigma=(3., 10.), amp=(500., 3000.),
                background=100., poisson=True, read_noise=5., dtype="uint16",
                out=(1024, 1024), max_steps=896),
    "headline": dict(shape=(512, 1024, 1024), n_blobs=8000, sigma=(3., 10.), amp=(500., 3000.),
                     background=100., poisson=True, read_noise=5., dtype="uint16",
                     out=(1024, 1024), max_steps=1536),
}


# the serving workload's resident volumes: name, numpy shape (z, y, x), voxel type, blobs
SERVING_VOLUMES = [
    ("f32_256", (256, 256, 256), np.float32, 500),
    ("u8_256x512x512", (256, 512, 512), np.uint8, 1000),
    ("u16_512", (512, 512, 512), np.uint16, 2000),
    ("u16_512x1024x1024", (512, 1024, 1024), np.uint16, 8000),
    ("u16_512x2048x2048", (512, 2048, 2048), np.uint16, 16000),
]


def steps_for_shape(shape):
    """max_steps = 16 * ceil(|(Nx,Ny,Nz)| / 16): every ray steps at most one voxel."""
    return 16 * int(math.ceil(math.sqrt(sum(n * n for n in shape)) / 16.))


def blob_params(shape, n_blobs, sigma, amp, seed=0):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 1, (n_blobs, 3)) * np.array(shape, np.float64)  # (z, y, x)
    sigmas = rng.uniform(sigma[0], sigma[1], (n_blobs, 3))
    amps = rng.uniform(amp[0], amp[1], n_blobs)
    return rng, centres, sigmas, amps


def _axis_profile(c, s, n):
    lo = max(0, int(math.floor(c - 3 * s)))
    hi = min(n, int(math.ceil(c + 3 * s)) + 1)
    x = np.arange(lo, hi, dtype=np.float64)
    return lo, hi, np.exp(-.5 * ((x - c) / s) ** 2)


def _finish(vol, rng_noise, cfg, dtype):
    if cfg["poisson"]:
        vol = rng_noise.poisson(vol).astype(np.float32)
    if cfg["read_noise"]:
        vol = vol + rng_noise.normal(0, cfg["read_noise"], vol.shape).astype(np.float32)
    if np.dtype(dtype) == np.uint16:
        return np.clip(np.rint(vol), 0, 65535).astype(np.uint16)
    if np.dtype(dtype) == np.uint8:
        return np.clip(np.rint(vol), 0, 255).astype(np.uint8)
    return vol.astype(np.float32)


def blob_volume(shape, n_blobs, sigma, amp, background, poisson, read_noise, dtype, seed=0, **_):
    """Host (numpy) blob volume of numpy shape (Nz, Ny, Nx)."""
    rng, centres, sigmas, amps = blob_params(shape, n_blobs, sigma, amp, seed)
    vol = np.full(shape, background, np.float32)
    for (cz, cy, cx), (sz, sy, sx), a in zip(centres, sigmas, amps):
        z0, z1, gz = _axis_profile(cz, sz, shape[0])
        y0, y1, gy = _axis_profile(cy, sy, shape[1])
        x0, x1, gx = _axis_profile(cx, sx, shape[2])
        vol[z0:z1, y0:y1, x0:x1] += (a * gz[:, None, None] * gy[None, :, None]
                                     * gx[None, None, :]).astype(np.float32)
    cfg = dict(poisson=poisson, read_noise=read_noise)
    return _finish(vol, rng, cfg, dtype)


def blob_volume_device(shape, n_blobs, sigma, amp, background, poisson, read_noise, dtype,
                       seed=0, device="cuda", **_):
    """The same blobs built on a GPU with torch; returns a device tensor."""
    import torch
    _, centres, sigmas, amps = blob_params(shape, n_blobs, sigma, amp, seed)
    vol = torch.full(shape, float(background), dtype=torch.float32, device=device)
    for (cz, cy, cx), (sz, sy, sx), a in zip(centres, sigmas, amps):
        z0, z1, gz = _axis_profile(cz, sz, shape[0])
        y0, y1, gy = _axis_profile(cy, sy, shape[1])
        x0, x1, gx = _axis_profile(cx, sx, shape[2])
        prof = torch.from_numpy(np.concatenate([a * gz, gy, gx]).astype(np.float32)).to(device)
        tz, ty, tx = torch.split(prof, [len(gz), len(gy), len(gx)])
        vol[z0:z1, y0:y1, x0:x1] += tz[:, None, None] * ty[None, :, None] * tx[None, None, :]
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed) + 1)
    if poisson:
        vol = torch.poisson(vol, generator=gen)
    if read_noise:
        vol += float(read_noise) * torch.randn(vol.shape, generator=gen, device=device)
    dt = np.dtype(dtype)
    if dt == np.uint16:
        # torch has no general uint16 arithmetic: return the uint16 bit pattern as int16
        # (view the host copy with .view(np.uint16))
        v = torch.clamp(torch.round(vol), 0, 65535).to(torch.int32)
        return torch.where(v >= 32768, v - 65536, v).to(torch.int16)
    if dt == np.uint8:
        return torch.clamp(torch.round(vol), 0, 255).to(torch.uint8)
    return vol


def _splat_blobs_device(vol, centres, sigmas, amps, chunk=96):
    """Adds the blobs to a float32 device volume in chunks: each blob is evaluated on a
    64^3 window anchored at floor(c - 3 sigma) (wide enough for sigma <= 10), masked to
    [floor(c - 3 sigma), ceil(c + 3 sigma)] and the volume, and scattered with index_add_."""
    import torch
    dev = vol.device
    dims = torch.tensor(vol.shape, device=dev, dtype=torch.float32)
    L = 64
    if float(np.max(sigmas)) * 6 + 3 > L:
        raise ValueError("blob sigma too large for the splat window")
    ar = torch.arange(L, device=dev, dtype=torch.float32)
    c_all = torch.as_tensor(np.asarray(centres, np.float32), device=dev)
    s_all = torch.as_tensor(np.asarray(sigmas, np.float32), device=dev)
    a_all = torch.as_tensor(np.asarray(amps, np.float32), device=dev)
    ny, nx = vol.shape[1], vol.shape[2]
    flat = vol.view(-1)
    for k in range(0, len(amps), chunk):
        c, s, a = c_all[k:k + chunk], s_all[k:k + chunk], a_all[k:k + chunk]
        lo = torch.floor(c - 3 * s)
        hi = torch.ceil(c + 3 * s)
        idx = lo[:, :, None] + ar                                        # (m, 3, L)
        g = torch.exp(-.5 * ((idx - c[:, :, None]) / s[:, :, None]) ** 2)
        ok = (idx >= 0) & (idx < dims[None, :, None]) & (idx <= hi[:, :, None])
        g = torch.where(ok, g, torch.zeros_like(g))
        idx = torch.minimum(idx.clamp(min=0), dims[None, :, None] - 1).long()
        w = (a[:, None, None, None] * g[:, 0, :, None, None] * g[:, 1, None, :, None]
             * g[:, 2, None, None, :])
        lin = (idx[:, 0, :, None, None] * ny + idx[:, 1, None, :, None]) * nx \
            + idx[:, 2, None, None, :]
        flat.index_add_(0, lin.reshape(-1), w.reshape(-1))


def blob_series_device(shape, n_blobs, sigma, amp, background, poisson, read_noise, n_time,
                       drift=.5, seed=0, device="cuda"):
    """Yields n_time uint16 timepoints (int16 bit pattern, like blob_volume_device) of one
    blob population whose centres take seeded random walks of `drift` voxels per step
    (each step a uniformly random direction)."""
    import torch
    rng, centres, sigmas, amps = blob_params(shape, n_blobs, sigma, amp, seed)
    gen = torch.Generator(device=device)
    for t in range(n_time):
        if t:
            step = rng.normal(size=centres.shape)
            centres = centres + drift * step / np.linalg.norm(step, axis=1, keepdims=True)
        vol = torch.full(shape, float(background), dtype=torch.float32, device=device)
        _splat_blobs_device(vol, centres, sigmas, amps)
        gen.manual_seed(int(seed) * 1000 + t + 1)
        if poisson:
            vol = torch.poisson(vol, generator=gen)
        if read_noise:
            vol += float(read_noise) * torch.randn(vol.shape, generator=gen, device=device)
        v = torch.clamp(torch.round(vol), 0, 65535).to(torch.int32)
        del vol
        yield torch.where(v >= 32768, v - 65536, v).to(torch.int16)


def serving_volume(shape, dtype, n_blobs, seed):
    """Host (numpy) blob volume of the serving workload, built on the GPU: sigma U[3,10],
    amplitude U[0.3,1] (float32), U[60,200] (uint8) or U[500,3000] (uint16), background
    amplitude/10 (uint16: 100), no noise."""
    import torch
    dt = np.dtype(dtype)
    amp = (.3, 1.) if dt == np.float32 else (60., 200.) if dt == np.uint8 else (500., 3000.)
    bg = amp[0] / 10 if dt != np.uint16 else 100.
    _, c, s, a = blob_params(shape, n_blobs, (3., 10.), amp, seed)
    vol = torch.full(shape, bg, dtype=torch.float32, device="cuda")
    _splat_blobs_device(vol, c, s, a)
    if dt == np.float32:
        return vol.cpu().numpy()
    top = 255 if dt == np.uint8 else 65535
    v = torch.clamp(torch.round(vol), 0, top).to(torch.int32)
    del vol
    if dt == np.uint8:
        return v.to(torch.uint8).cpu().numpy()
    return torch.where(v >= 32768, v - 65536, v).to(torch.int16).cpu().numpy().view(np.uint16)


def device_volume_to_numpy(t, dtype):
    """Host copy of a volume from blob_volume_device, as the numpy dtype it encodes."""
    return t.cpu().numpy().view(np.dtype(dtype))


def config_volume(name, seed=0, device=None):
    cfg = CONFIGS[name]
    if device is None:
        return blob_volume(seed=seed, **cfg)
    return blob_volume_device(seed=seed, device=device, **cfg)


def headline_camera():
    """P = perspective(45, 1, 0.1, 10), M = translate(0,0,-4) . rotation(pi/6 about y)."""
    P = mat4_perspective(45, 1., .1, 10)
    M = np.dot(mat4_translate(0, 0, -4), mat4_rotation(np.pi / 6, 0, 1, 0))
    return M, P


def random_poses(n, seed=0, distance=4.):
    """n seeded poses: uniform axis on the sphere, angle U[0, 2pi), at the headline distance."""
    rng = np.random.default_rng(seed)
    poses = []
    for _ in range(n):
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        ang = rng.uniform(0, 2 * np.pi)
        poses.append(np.dot(mat4_translate(0, 0, -distance), mat4_rotation(ang, *axis)))
    return poses
