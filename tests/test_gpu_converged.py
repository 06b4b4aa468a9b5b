"""Converged images on the GPU against the FP64 high-spp reference: the
adjoint's forward render and the plain path tracer (Lambertian materials)
at 16384 spp fall within the Monte Carlo error the reference's sample
variance predicts (RMSE within 1.5x, mean within 4 standard errors)."""
import numpy as np
import pytest

import ctamd
from ctamd import adjoint as AD
from test_oracle_fp64 import check_converged, converged_scene, fp64_reference

pytestmark = pytest.mark.gpu

SPP = 16384


def test_gpu_renders_converge_to_the_fp64_reference(gpu):
    import torch
    w, sc, alb, emi, views = converged_scene()
    acc = ctamd.Accel(max_instances=sc.info.instance_count + 8,
                      max_primitives=sc.info.primitive_count + 32 * sc.info.instance_count + 1024, max_cube_lights=8)
    AD._load_geometry(sc, acc)
    slots = AD.VoxelSlots(sc, acc)
    r = AD.DiffRenderer(acc, views, SPP, w.max_depth, 0x51)
    dev = torch.device("cuda")
    img = r.forward(torch.from_numpy(slots.scatter(alb)).to(dev), torch.from_numpy(slots.scatter(emi)).to(dev))
    mats = []
    for v in range(sc.info.primitive_count):  # voxel v has material v
        m = ctamd.Material()
        m.type = ctamd.CT_MATERIAL_LAMBERTIAN
        m.diffuse[:] = alb[v].tolist()
        m.emission[:] = [float(emi[v])] * 3
        mats.append(m)
    pt = AD.path_trace(acc, views, mats, spp=SPP, max_depth=w.max_depth, seed=0x52)
    print("fp64 reference", flush=True)
    mean, var = fp64_reference(sc, alb, emi, views[0], w.max_depth)
    print("adjoint forward rmse/expected", check_converged(img.cpu().numpy(), SPP, mean, var), flush=True)
    print("path trace rmse/expected", check_converged(pt.cpu().numpy(), SPP, mean, var), flush=True)
