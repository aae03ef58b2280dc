"""GPU test of svo_render_batch_noalloc, the batched render whose device
tables live in the library's fixed scratch array: called through the C ABI,
its outputs equal svo_render_batch_device's for the same views bit for bit,
on two calls in a row (the second with fewer and smaller views), in the grid
and the persistent form; more than 512 views and octrees not yet resident are
refused."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _views(svo, torch, ren, trees, specs):
    dev = torch.device("cuda", 0)
    outs, views = [], []
    for t, w, h, p, y, rs in specs:
        cam = t.orbit_camera(p, y, 1.0)
        o = {"hit": torch.full((w * h,), 9, dtype=torch.uint8, device=dev),
             "normal": torch.full((w * h,), -1, dtype=torch.int32, device=dev),
             "t": torch.full((w * h,), -1.0, dtype=torch.float32, device=dev),
             "leaf": torch.full((w * h,), -2, dtype=torch.int64, device=dev)}
        outs.append(o)
        views.append(ren.view(t, cam, w, h, o["hit"].data_ptr(), o["normal"].data_ptr(), o["t"].data_ptr(),
                              o["leaf"].data_ptr(), ray_scale=rs))
    return outs, views


def test_noalloc_batch_equals_batched_render(svo, golden):
    import torch
    dragon = svo.VoxelOctree(os.path.join(golden, "dragon.oct"))
    sdf = svo.VoxelOctree(os.path.join(golden, "build_sdf128.oct"))
    fresh = svo.VoxelOctree(os.path.join(golden, "build_sdf128.oct"))
    dragon.upload()
    sdf.upload()
    ren = svo.svo.Renderer(16)
    first = [(dragon, 320, 180, 20.0, 35.0, 0.0), (sdf, 257, 131, -30.0, 200.0, 0.0), (dragon, 7, 5, 10.0, 10.0, 0.0),
             (sdf, 640, 360, 45.0, 300.0, 0.0), (dragon, 200, 100, -60.0, 90.0, 0.02)]
    second = [(sdf, 96, 64, 5.0, 15.0, 0.0), (dragon, 33, 17, -20.0, 140.0, 0.0)]
    for form in (svo.svo.BATCH_GRID, svo.svo.BATCH_PERSISTENT):
        for k, specs in enumerate((first, second)):
            ref_out, ref_views = _views(svo, torch, ren, None, specs)
            ren.render_device(ref_views, None, 0, 0, form)
            outs, views = _views(svo, torch, ren, None, specs)
            print(f"[gpu] noalloc batch form {form} call {k}: {len(specs)} views", flush=True)
            svo.svo.render_batch_noalloc(views, None, 0, 0, form)
            torch.cuda.synchronize()
            for a, b in zip(outs, ref_out):
                for f in ("hit", "normal", "t", "leaf"):
                    assert torch.equal(a[f], b[f]), (form, k, f)
    # refused before any device work
    _, views = _views(svo, torch, ren, None, [(dragon, 8, 8, 0.0, 0.0, 0.0)])
    with pytest.raises(svo.SvoError) as e:
        svo.svo.render_batch_noalloc(views * 513)
    assert e.value.code == svo.svo.ERR_LIMIT
    _, views = _views(svo, torch, ren, None, [(fresh, 8, 8, 0.0, 0.0, 0.0)])
    with pytest.raises(svo.SvoError) as e:
        svo.svo.render_batch_noalloc(views)
    assert e.value.code == svo.svo.ERR_INVALID and "resident" in str(e.value)
