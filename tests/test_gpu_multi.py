"""Multi-mesh runtime (include/gvv_multi.h): every mesh of an uneven batch against the CPU
oracle (a loop over the single-mesh oracle) and against the single-mesh entry points, with
hipGraph replay, plain launches and autotuned settings."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from gvv_amd import Rasterizer, scenes
from gvv_amd.multi import MultiRenderer
from oracle import oracle as O
from parity_util import GRAD_TOL, RENDER_TOL

pytestmark = pytest.mark.gpu


def small_batch(seed=0):
    return scenes.config2(seed=seed, n_meshes=6, tri_range=(300, 6000), resolutions=(48, 64, 96), max_views=3)


def fill_grads(mr, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for b in mr.bufs:
        g = rng.normal(size=tuple(b["g_render"].shape)).astype(np.float32)
        b["g_render"].copy_(torch.from_numpy(g))
        out.append(g)
    return out


def check_against_oracle(mr, grads):
    for k, (s, b, g) in enumerate(zip(mr.scenes, mr.bufs, grads)):
        orc = O.Oracle(**s.renderer_kwargs())
        ref = orc.forward(*s.inputs())
        face = b["face"].cpu().numpy()
        assert np.array_equal(face, ref["face"]), f"mesh {k}: face ids differ"
        assert np.array_equal(b["bary"].cpu().numpy().view(np.uint32), ref["bary"].view(np.uint32)), f"mesh {k}: bary"
        assert rel_err(b["render"].cpu().numpy(), ref["render"]) <= RENDER_TOL
        gref = orc.backward(g, np.zeros_like(g), s.target, s.vpos, s.vcol, s.tex, s.sh, ref["vnormal"], ref["bary"],
                            ref["face"], s.extr, s.intr)
        for name in ("pos", "col", "sh"):
            e = rel_err(b["g_" + name].cpu().numpy(), gref[name])
            print(f"mesh {k} {s.meta['name']} grad {name} rel err {e:.2e}", flush=True)
            assert e <= GRAD_TOL, (k, name, e)


@pytest.mark.parametrize("batched", [True, False])
def test_graph_run_matches_oracle_and_replays(device, batched):
    mr = MultiRenderer(small_batch(), device=device)
    mr.set_batched(batched)
    grads = fill_grads(mr, 1)
    mr.run(backward=True)
    torch.cuda.synchronize()
    first = [{k: v.clone() for k, v in b.items()} for b in mr.bufs]
    check_against_oracle(mr, grads)
    for _ in range(3):               # replays of the captured graph, eager work in between
        for b in mr.bufs:
            b["render"].zero_()
            b["face"].fill_(-7)
        torch.cuda.synchronize()
        mr.run(backward=True)
        torch.cuda.synchronize()
        for b, f in zip(mr.bufs, first):
            assert torch.equal(b["face"], f["face"]) and torch.equal(b["render"], f["render"])
            for name in ("g_pos", "g_col", "g_sh"):
                assert rel_err(b[name].cpu().numpy(), f[name].cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("batched", [True, False])
def test_plain_launches_equal_single_mesh_entry_points(device, batched):
    batch = small_batch(seed=3)
    mr = MultiRenderer(batch, device=device)
    mr.use_graph(False)
    mr.set_batched(batched)
    grads = fill_grads(mr, 2)
    mr.run(backward=True)
    torch.cuda.synchronize()
    for s, b, g in zip(batch, mr.bufs, grads):
        r = Rasterizer(**s.renderer_kwargs())
        ins = [torch.from_numpy(np.ascontiguousarray(a)).to(device) for a in s.inputs()]
        bary, face, render, vnormal, _ = r.forward(*ins)
        gp, gc, gt, gs = r.backward(torch.from_numpy(g).to(device), None, torch.from_numpy(s.target).to(device), ins[0],
                                    ins[1], ins[2], ins[3], vnormal, face, ins[4], ins[5])
        torch.cuda.synchronize()
        assert torch.equal(face, b["face"]) and torch.equal(render, b["render"]) and torch.equal(bary, b["bary"])
        assert torch.equal(vnormal, b["vnormal"])
        for x, name in ((gp, "g_pos"), (gc, "g_col"), (gs, "g_sh")):
            assert rel_err(x.cpu().numpy(), b[name].cpu().numpy()) <= 1e-6


def test_autotune_then_graph(device):
    mr = MultiRenderer(small_batch(seed=5), device=device)
    grads = fill_grads(mr, 3)
    mr.autotune(reps=1)
    fill_grads(mr, 3)
    settings = [mr.tuning(i) for i in range(len(mr.scenes))]
    print("tuning per mesh (tile, large_box, class):", settings, flush=True)
    for tile, box, cls in settings:
        assert tile in (8, 16, 32) and box in (64, 256, 1024) and 0 <= cls < 12
    mr.run(backward=True)
    torch.cuda.synchronize()
    check_against_oracle(mr, grads)


def test_batched_equals_per_mesh_launches_with_mixed_tiles(device):
    """The per-stage launches across meshes give the per-mesh launches' results, with meshes on
    different backward tile edges (one pixel-kernel launch per edge) and large-face thresholds."""
    batch = small_batch(seed=7)
    out = []
    for batched in (False, True):
        mr = MultiRenderer(batch, device=device)
        fill_grads(mr, 4)
        mr.set_batched(batched)
        for i in range(len(batch)):
            mr.set_tuning(i, (8, 16, 32)[i % 3], (64, 256, 1024)[(i // 3) % 3])
        mr.run(backward=True)
        torch.cuda.synchronize()
        out.append([{k: v.clone() for k, v in b.items()} for b in mr.bufs])
    for a, b in zip(*out):
        for k in ("face", "bary", "render", "vnormal"):
            assert torch.equal(a[k], b[k]), k
        for k in ("g_pos", "g_col", "g_sh"):
            assert rel_err(a[k].cpu().numpy(), b[k].cpu().numpy()) <= 1e-6, k


def test_autotune_cache_reused(device, tmp_path):
    """The autotuner's choices persist in a JSON file keyed by device and mode: a second
    runtime over meshes of the same size classes applies them without timing."""
    cache = str(tmp_path / "autotune.json")
    mr = MultiRenderer(small_batch(seed=5), device=device)
    fill_grads(mr, 3)
    assert mr.autotune(reps=1, cache=cache) == "timed"
    first = [mr.tuning(i) for i in range(len(mr.scenes))]
    mr2 = MultiRenderer(small_batch(seed=5), device=device)
    assert mr2.autotune(reps=1, cache=cache) == "cache"
    assert [mr2.tuning(i) for i in range(len(mr2.scenes))] == first
    grads = fill_grads(mr2, 6)
    mr2.run(backward=True)
    torch.cuda.synchronize()
    check_against_oracle(mr2, grads)
