"""GPU parity at the full workload sizes: the HIP renderer (through its C ABI) against the CPU
oracle on the benchmarked scenes themselves, not reductions of them.

  full cfg1   capped cylinder, 100,000 triangles, 8 views at 1024^2, 1024^2 texture:
              textured + shaded forward and backward (with and without the model-to-data
              term, with recomputed and with stored barycentrics) and the normal mode
  full cfg2   the 32-mesh uneven batch (2k-200k triangles, 1-8 views at 256/512/1024^2)
              through the multi-mesh runtime, every mesh against the single-mesh oracle
  full cfg3   the fit workload: 16 subjects x 4 views at 512^2 on the 70,000-triangle
              template at seeded ground-truth parameters, vertexColor + shaded

The accuracy policy is parity_util's: ids, barycentrics and vertex normals bit-exact; render
within 1e-5 and gradients within 1e-4 max-norm relative, plus elementwise bounds; every
output's errors are printed in the session summary."""
import numpy as np
import pytest
import torch

from gvv_amd import scenes
from oracle import oracle as O
from parity_util import check_exact, check_grad, check_render, run_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    import os
    O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))


@pytest.fixture(scope="module")
def cfg1():
    return scenes.config1()


def test_config1_textured_shaded(device, cfg1):
    run_parity(cfg1, device, "textured", "shaded", m2d=False, case="full cfg1 textured/shaded")


def test_config1_textured_shaded_model_to_data_stored_bary(device, cfg1):
    run_parity(cfg1, device, "textured", "shaded", m2d=True, use_bary=True, seed=1,
               case="full cfg1 textured/shaded m2d bary")


def test_config1_normal_mode(device, cfg1):
    """Forward of the normal albedo mode (its gradient is zero by the reference's routing)."""
    run_parity(cfg1, device, "normal", "shaded", m2d=False, check_grads=False, case="full cfg1 normal")


def test_config2_multi_mesh_batch(device):
    from gvv_amd.multi import MultiRenderer
    batch = scenes.config2()
    mr = MultiRenderer(batch, device=device)
    rng = np.random.Generator(np.random.PCG64(21))
    grads = []
    for b in mr.bufs:
        g = rng.normal(size=tuple(b["g_render"].shape)).astype(np.float32)
        b["g_render"].copy_(torch.from_numpy(g))
        grads.append(g)
    mr.run(backward=True)
    torch.cuda.synchronize()
    outs = {k: [] for k in ("render", "g_pos", "g_col", "g_sh")}
    refs = {k: [] for k in outs}
    for k, (s, b, g) in enumerate(zip(mr.scenes, mr.bufs, grads)):
        orc = O.Oracle(**s.renderer_kwargs())
        ref = orc.forward(*s.inputs())
        check_exact(f"cfg2 mesh {k} {s.meta['name']}", b["face"].cpu().numpy(), b["bary"].cpu().numpy(),
                    b["vnormal"].cpu().numpy(), ref)
        gref = orc.backward(g, np.zeros_like(g), s.target, s.vpos, s.vcol, s.tex, s.sh, ref["vnormal"], ref["bary"],
                            ref["face"], s.extr, s.intr)
        for name, r in (("render", ref["render"]), ("g_pos", gref["pos"]), ("g_col", gref["col"]), ("g_sh", gref["sh"])):
            outs[name].append(b[name].cpu().numpy().ravel())
            refs[name].append(np.asarray(r).ravel())
        # every mesh on its own scale as well: a small mesh's gradients are not hidden by a large one's
        for name, r in (("g_pos", gref["pos"]), ("g_col", gref["col"]), ("g_sh", gref["sh"])):
            check_grad(f"cfg2 mesh {k} {s.meta['name']}", name, b[name].cpu().numpy(), r)
    print(f"{len(batch)} meshes, {mr.pixels()} px, {mr.triangles()} triangle-views", flush=True)
    check_render("full cfg2 32 meshes", "render", np.concatenate(outs["render"]), np.concatenate(refs["render"]))
    for name in ("g_pos", "g_col", "g_sh"):
        check_grad("full cfg2 32 meshes", name, np.concatenate(outs[name]), np.concatenate(refs[name]))


def fit_scene(device):
    """The fit workload's batch as a Scene: 16 subjects at seeded ground-truth parameters."""
    from gvv_amd import fit
    model = fit.ShapeModel(fit.FitConfig(), device)
    cfg = model.cfg
    gt = fit.sample_params(np.random.Generator(np.random.PCG64(12345)), cfg.batch, cfg.n_basis, device)
    with torch.no_grad():
        vpos = model.vertices(gt).contiguous().cpu().numpy()
    B, C, R = cfg.batch, cfg.views, cfg.resolution
    sh = np.ascontiguousarray(np.broadcast_to(gt.sh.cpu().numpy()[:, None, :], (B, C, 27)), np.float32)
    return scenes.Scene(faces=model.faces.astype(np.int32), uv=model.uv.astype(np.float32), vpos=vpos.astype(np.float32),
                        vcol=np.ascontiguousarray(np.broadcast_to(model.colors.cpu().numpy(), (B, model.N, 3)), np.float32),
                        tex=np.full((B, 1, 1, 3), 0.5, np.float32), sh=sh,
                        extr=np.ascontiguousarray(np.broadcast_to(model.extr.cpu().numpy(), (B, C, 12)), np.float32),
                        intr=np.ascontiguousarray(np.broadcast_to(model.intr.cpu().numpy(), (B, C, 9)), np.float32),
                        target=np.zeros((B, C, R, R, 3), np.float32), width=R, height=R, albedo_mode="vertexColor",
                        shading_mode="shaded", meta=dict(name=f"fit{model.F}_{B}x{C}x{R}"))


def test_config3_fit_batch(device):
    s = fit_scene(device)
    run_parity(s, device, "vertexColor", "shaded", m2d=False, seed=3, case="full cfg3 vertexColor/shaded")
