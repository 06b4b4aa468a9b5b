"""FP64 high-spp reference of the voxel path tracer -- test infrastructure only.

An independent double-precision Monte Carlo estimate of the image the
library's plain path tracer (ct_path_trace, and ct_adjoint_render for
Lambertian voxels) converges to: camera rays through jittered pixel
positions (rgen.glsl:18-42 with the jitter of the sample), the closest voxel
by a brute-force slab test over every AABB, emission collected at every
vertex, cosine-weighted scattering around the face normal with the
voxel's albedo as the weight, nothing from rays that leave the scene.  It
shares no code and no random numbers with the library or the C++ oracle
(numpy's generator, its own slab test and sampler), so agreement in
expectation checks the estimator, not the arithmetic.

Pinned by: a directly seen emitter gives exactly its emission at depth 0;
statistical agreement with the C++ oracle's float32 estimator on the same
scene; the per-pixel sample variance it reports sets the tests' bounds.
"""
import numpy as np


def camera_rays(inv_view, inv_proj, width, height, px, py, jx, jy):
    """World-space origins and unit directions for pixel (px, py) + jitter."""
    iv = np.asarray(inv_view, np.float64).reshape(4, 4).T   # column-major 4x4 -> row-major matrix
    ip = np.asarray(inv_proj, np.float64).reshape(4, 4).T
    dx = (px + jx) / width * 2.0 - 1.0
    dy = (py + jy) / height * 2.0 - 1.0
    t = np.stack([dx, dy, np.ones_like(dx), np.ones_like(dx)], 1) @ ip.T
    nt = t[:, :3] / np.linalg.norm(t[:, :3], axis=1, keepdims=True)
    d = np.concatenate([nt, np.zeros((nt.shape[0], 1))], 1) @ iv.T
    o = np.broadcast_to((iv @ np.array([0.0, 0.0, 0.0, 1.0]))[:3], d[:, :3].shape)
    d = d[:, :3]
    return np.array(o), d / np.linalg.norm(d, axis=1, keepdims=True)


def closest_hit(o, d, lo, hi, chunk=4096):
    """Closest box along each ray (t > 0): (index or -1, t, face normal)."""
    n = o.shape[0]
    idx = np.full(n, -1, np.int64)
    tbest = np.full(n, np.inf)
    nrm = np.zeros((n, 3))
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
    for s in range(0, n, chunk):
        oo, ii = o[s:s + chunk, None, :], inv[s:s + chunk, None, :]
        t0 = (lo[None] - oo) * ii
        t1 = (hi[None] - oo) * ii
        tmin_ax = np.where(np.isnan(t0), -np.inf, np.minimum(t0, t1))
        tmax_ax = np.where(np.isnan(t0), np.inf, np.maximum(t0, t1))
        tmin = tmin_ax.max(2)
        tmax = tmax_ax.min(2)
        ok = (tmax >= tmin) & (tmin > 0.0)
        tt = np.where(ok, tmin, np.inf)
        k = tt.argmin(1)
        tk = tt[np.arange(tt.shape[0]), k]
        hit = np.isfinite(tk)
        ax = tmin_ax[np.arange(tt.shape[0]), k].argmax(1)
        sl = slice(s, s + chunk)
        idx[sl] = np.where(hit, k, -1)
        tbest[sl] = tk
        nn = np.zeros((tt.shape[0], 3))
        nn[np.arange(tt.shape[0]), ax] = -np.sign(d[sl][np.arange(tt.shape[0]), ax])
        nrm[sl] = nn
    return idx, tbest, nrm


def cosine_directions(rng, n):
    """Cosine-weighted unit vectors around each normal n (N, 3)."""
    m = n.shape[0]
    u1, u2 = rng.random(m), rng.random(m)
    r, phi = np.sqrt(u1), 2.0 * np.pi * u2
    local = np.stack([r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - u1))], 1)
    a = np.where(np.abs(n[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(a, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    return local[:, :1] * t + local[:, 1:2] * b + local[:, 2:3] * n


def render(aabbs, albedo, emission, inv_view, inv_proj, width, height, spp, max_depth, seed=0):
    """Mean and per-pixel sample variance of the radiance, (H*W, 3) each.
    aabbs (N, 6) world boxes, albedo (N, 3), emission (N, 3)."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(aabbs, np.float64)[:, :3]
    hi = np.asarray(aabbs, np.float64)[:, 3:]
    alb = np.asarray(albedo, np.float64)
    emi = np.asarray(emission, np.float64)
    npx = width * height
    pix = np.repeat(np.arange(npx), spp)
    px, py = (pix % width).astype(np.float64), (pix // width).astype(np.float64)
    o, d = camera_rays(inv_view, inv_proj, width, height, px, py, rng.random(pix.size), rng.random(pix.size))
    beta = np.ones((pix.size, 3))
    L = np.zeros((pix.size, 3))
    alive = np.arange(pix.size)
    for k in range(max_depth + 1):
        if alive.size == 0:
            break
        idx, t, nrm = closest_hit(o[alive], d[alive], lo, hi)
        hit = idx >= 0
        alive, idx, t, nrm = alive[hit], idx[hit], t[hit], nrm[hit]
        L[alive] += beta[alive] * emi[idx]
        if k == max_depth:
            break
        pos = o[alive] + d[alive] * t[:, None]
        beta[alive] *= alb[idx]
        d[alive] = cosine_directions(rng, nrm)
        o[alive] = pos + nrm * 1e-7
    Ls = L.reshape(npx, spp, 3)
    return Ls.mean(1), Ls.var(1, ddof=1)
