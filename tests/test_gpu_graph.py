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


def test_forward_on_null_and_user_streams(device):
    """The forward's fork to the handle's second stream and join back work from the NULL
    (legacy default) stream as from a user stream: identical results, and the outputs are
    complete in the caller's stream order without a device-wide synchronisation."""
    s = scenes.config1(width=256, height=256, cols=80, rows=60, n_cameras=2, tex_size=128)
    ins = [torch.from_numpy(np.ascontiguousarray(a)).to(device) for a in s.inputs()]
    outs = []
    for stream in (torch.cuda.default_stream(device), torch.cuda.Stream(device)):
        r = Rasterizer(**s.renderer_kwargs())
        r.set_tuning(16, 64)                      # a low threshold: many faces take the large-face pass
        with torch.cuda.stream(stream):
            bary, face, render, _, _ = r.forward(*ins)
            # consumed on the same stream before any synchronisation
            checksum = (face.to(torch.int64).sum(), render.sum(), bary.sum())
        stream.synchronize()
        outs.append((face.clone(), render.clone(), bary.clone(), [float(c) for c in checksum]))
    (f0, r0, b0, c0), (f1, r1, b1, c1) = outs
    assert torch.equal(f0, f1) and torch.equal(r0, r1) and torch.equal(b0, b1)
    assert c0 == c1
    assert c0[0] == float(f0.to(torch.int64).sum())
