"""The Python drop-in (CudaRenderer.CudaRendererGpu, mirroring python/CudaRenderer.py) end to
end on the GPU: constructor runs the forward, getters and layouts, mask, target alias, and
torch.autograd gradients against the oracle with the reference's gradient routing."""
import numpy as np
import pytest
import torch

import CudaRenderer
from conftest import rel_err
from gvv_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _make(scene, device, albedo, requires_grad=True, **kw):
    t = lambda a, g=False: torch.tensor(np.ascontiguousarray(a), device=device, requires_grad=g)
    ins = dict(vpos=t(scene.vpos, requires_grad), vcol=t(scene.vcol, requires_grad), tex=t(scene.tex, requires_grad),
               sh=t(scene.sh, requires_grad), target=t(scene.target, requires_grad), extr=t(scene.extr, requires_grad),
               intr=t(scene.intr))
    r = CudaRenderer.CudaRendererGpu(
        faces_attr=scene.faces.reshape(-1).tolist(), texCoords_attr=scene.uv.reshape(-1).tolist(),
        numberOfVertices_attr=scene.N, numberOfCameras_attr=scene.C, renderResolutionU_attr=scene.width,
        renderResolutionV_attr=scene.height, albedoMode_attr=albedo, shadingMode_attr="shaded",
        vertexPos_input=ins["vpos"], vertexColor_input=ins["vcol"], texture_input=ins["tex"], shCoeff_input=ins["sh"],
        targetImage_input=ins["target"], extrinsics_input=ins["extr"], intrinsics_input=ins["intr"], **kw)
    return r, ins


@pytest.mark.parametrize("albedo", ["vertexColor", "textured"])
def test_dropin_forward_backward(device, albedo):
    s = scenes.random_soup(4, n_faces=90, n_cameras=2, width=40, height=32, batch=2)
    r, ins = _make(s, device, albedo)
    B, C, H, W = s.B, s.C, s.height, s.width
    assert r.getBaryCentricBufferTF().shape == (B, C, H, W, 2)
    assert r.getFaceBufferTF().shape == (B, C, H, W) and r.getFaceBufferTF().dtype == torch.int32
    assert r.getRenderBufferTF().shape == (B, C, H, W, 3)
    assert r.cudaRendererOperator[3].shape == (B, C, s.N, 3)
    tgt = r.getTargetBufferTF()
    assert tgt.data_ptr() == ins["target"].data_ptr()            # alias, no copy
    mask = r.getModelMaskTF()
    assert torch.equal(mask, (r.getFaceBufferTF() >= 0).float().unsqueeze(-1).expand(B, C, H, W, 3))
    assert r.getRenderBufferOpenCV(1, 0).shape == (H, W, 3)
    rng = np.random.Generator(np.random.PCG64(9))
    w1 = torch.tensor(rng.normal(size=(B, C, H, W, 3)).astype(np.float32), device=device)
    w2 = torch.tensor(rng.normal(size=(B, C, H, W, 3)).astype(np.float32), device=device)
    loss = (w1 * r.getRenderBufferTF()).sum() + (w2 * tgt).sum() + r.getBaryCentricBufferTF().sum()
    loss.backward()
    kw = s.renderer_kwargs()
    kw["albedo_mode"] = albedo
    orc = O.Oracle(**kw)
    ref = orc.forward(*s.inputs())
    g = orc.backward(w1.cpu().numpy(), w2.cpu().numpy(), s.target, s.vpos, s.vcol, s.tex, s.sh, ref["vnormal"],
                     ref["bary"], ref["face"], s.extr, s.intr)
    for name, key in (("vpos", "pos"), ("vcol", "col"), ("tex", "tex"), ("sh", "sh")):
        e = rel_err(ins[name].grad.cpu().numpy(), g[key])
        print(f"drop-in {albedo} grad {name} rel err {e:.3e}", flush=True)
        assert e <= 1e-4
    # target, extrinsics: zero gradient (CudaRenderer.py:215)
    assert torch.count_nonzero(ins["target"].grad) == 0 and torch.count_nonzero(ins["extr"].grad) == 0


@pytest.mark.parametrize("albedo", ["normal", "lighting"])
def test_dropin_zero_gradient_modes(device, albedo):
    s = scenes.random_soup(1, n_faces=40, n_cameras=1, width=24, height=24)
    r, ins = _make(s, device, albedo)
    r.getRenderBufferTF().sum().backward()
    for name in ("vpos", "vcol", "tex", "sh"):
        assert torch.count_nonzero(ins[name].grad) == 0


def test_dropin_normal_map_and_defaults(device):
    s = scenes.random_soup(3, n_faces=40, n_cameras=1, width=24, height=24, tex_size=(20, 12))
    r, _ = _make(s, device, "textured", requires_grad=False, compute_normal_map_attr=True)
    nm = r.getNormalMap()
    assert nm.shape == (1, 20, 12, 3)
    ref = O.Oracle(**s.renderer_kwargs(), compute_normal_map=True).forward(*s.inputs())
    assert rel_err(nm.cpu().numpy(), ref["normal_map"]) <= 1e-5
    r2, _ = _make(s, device, "textured", requires_grad=False)
    assert r2.getNormalMap() is None
    # python-wrapper defaults: filter sizes 1 (CudaRenderer.py:43-44)
    assert r2.image_filter_size_attr == 1 and r2.texture_filter_size_attr == 1
