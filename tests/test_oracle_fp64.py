"""Pins of the FP64 high-spp reference (oracle/fp64_pt.py): an emitter seen
directly gives its emission exactly; on a generated scene it agrees in
expectation with the C++ oracle's float32 path tracer."""
import numpy as np

import ctamd
import fp64_pt as R
import oracle as O
from ctamd import adjoint as AD


def test_direct_emitter_is_exact():
    iv, ip = ctamd.camera_matrices((0.0, 0.0, 5.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 0.2, 1.0)
    mean, var = R.render([[-1, -1, -1, 1, 1, 1]], [[0.5, 0.5, 0.5]], [[2.0, 3.0, 4.0]], iv.reshape(16), ip.reshape(16),
                         6, 6, 16, 0)
    assert np.array_equal(mean, np.tile([2.0, 3.0, 4.0], (36, 1))) and not var.any()
    # one bounce off a white wall lit by nothing stays exact too
    mean, _ = R.render([[-1, -1, -1, 1, 1, 1]], [[1.0, 1.0, 1.0]], [[0.0, 0.0, 0.0]], iv.reshape(16), ip.reshape(16),
                       4, 4, 8, 3)
    assert not mean.any()


def converged_scene():
    """8^3 voxels, half of them present and half of those emitting 1: most
    of the image is light that bounced, so the albedo shows in the mean.
    Depth 3, one 8x8 view at 0.5 rad."""
    w = AD.AdjointWorkload(grid=8, chunk=8, width=8, height=8, views=1, spp=1, max_depth=3)
    w.fill, w.emissive_frac = 0.5, 0.5
    sc, alb, emi = AD.build_scene(w)
    emi = (emi > 0).astype(np.float32)
    extent = w.grid * ctamd.CT_VOXEL_EXTENT
    views = AD.orbit_views(1, w.width, w.height, (0, 0, 0), 1.2 * extent, 0.3 * extent, fov_radians=0.5)
    return w, sc, alb, emi, views


REF_SPP = 2048


def fp64_reference(sc, alb, emi, view, depth):
    return R.render(sc.aabbs, alb, np.repeat(emi[:, None], 3, 1), view.inv_view[:], view.inv_proj[:], view.width,
                    view.height, REF_SPP, depth, seed=3)


def check_converged(img, spp, mean, var):
    """RMSE within 1.5x of the Monte Carlo error the reference's sample
    variance predicts for the two estimates, and the mean difference within
    4 standard errors.  (A 10 % albedo error gives 2.4x and 6 standard
    errors on this scene.)"""
    err = img - mean
    expected = np.sqrt((var / REF_SPP + var / spp).mean())
    rmse = np.sqrt((err ** 2).mean())
    assert mean.max() > 0 and rmse < 1.5 * expected, (rmse, expected)
    assert abs(err.mean()) < 4.0 * np.sqrt((var / REF_SPP + var / spp).mean() / err.size), err.mean()
    return rmse / expected


def test_agrees_with_the_float32_oracle_in_expectation():
    w, sc, alb, emi, views = converged_scene()
    mean, var = fp64_reference(sc, alb, emi, views[0], w.max_depth)
    o = O.OScene.from_scene(sc).adjoint_render(ctamd.AdjointDesc(1, 16384, w.max_depth, 11), views, alb, emi)
    check_converged(o, 16384, mean, var)
    # a wrong estimator is caught: albedo scaled by 0.9
    m2, v2 = R.render(sc.aabbs, alb * np.float32(0.9), np.repeat(emi[:, None], 3, 1), views[0].inv_view[:],
                      views[0].inv_proj[:], 8, 8, 512, w.max_depth, seed=4)
    err = o - m2
    assert abs(err.mean()) > 4.0 * np.sqrt((v2 / 512).mean() / err.size)
