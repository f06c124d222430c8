"""The single-mesh entry points under stream capture: a forward and backward captured into a
torch CUDA graph (a hipGraph) and replayed, with eager calls of the same handle between the
replays, give the eager results.  Under capture the forward keeps the large-face pass on the
caller's stream instead of forking it to the handle's second stream."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from gvv_amd import Rasterizer, scenes

pytestmark = pytest.mark.gpu


def test_forward_backward_graph_replays(device):
    s = scenes.config0(width=96, height=96, level=3, n_cameras=2)
    r = Rasterizer(**s.renderer_kwargs())
    ins = [torch.from_numpy(np.ascontiguousarray(a)).to(device) for a in s.inputs()]
    g = torch.from_numpy(np.random.default_rng(0).normal(size=(s.B, s.C, s.height, s.width, 3)).astype(np.float32)).to(device)

    def call():
        bary, face, render, vnormal, _ = r.forward(*ins)
        grads = r.backward(g, None, None, ins[0], ins[1], ins[2], ins[3], vnormal, face, ins[4], ins[5])
        return (bary, face, render, vnormal), grads

    ref_fwd, ref_grads = call()          # eager: sizes the handle's scratch, creates its streams
    ref_fwd = [x.clone() for x in ref_fwd]
    ref_grads = [x.clone() for x in ref_grads]
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        fwd, grads = call()
    for k in range(3):
        call()                           # eager work of the same handle between the replays
        for x in fwd:
            x.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for name, a, b in zip(("bary", "face", "render", "vnormal"), fwd, ref_fwd):
            assert torch.equal(a, b), (k, name)
        for name, a, b in zip(("pos", "col", "tex", "sh"), grads, ref_grads):
            e = rel_err(a.cpu().numpy(), b.cpu().numpy())
            assert e <= 1e-6, (k, name, e)
    print("3 replays equal the eager call", flush=True)
