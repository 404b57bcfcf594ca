"""torch.autograd binding of the differentiable float SSSP (forward solve + hand-written backward kernel).

Both directions go through the C ABI of libadds.so (adds_sssp_multi and adds_sssp_backward_f32) on
torch's current stream; torch only provides device memory and streams.

    dist = sssp(w, graph, srcs)      # w: float32 CUDA tensor [m] (requires_grad), dist: [k, n]
    loss = (c * dist).sum(); loss.backward()   # w.grad = dL/dw through each source's tree

The tree of a source is rooted at the source: the parent edge of v is the smallest edge id (u -> v) with
d[u] < d[v] and fl(d[u] + w) == d[v]; vertices reached only through zero weights take a tight edge from an
equal-distance vertex resolved earlier (adds_sssp_backward_f32).  Unreached vertices (dist = +inf)
receive and pass no gradient.  The gradient is the exact sum of the fixed-point-scaled inputs rounded once
to float32: bit-identical run to run, whatever the order of the device's additions.
"""
import numpy as np
import torch

from .graph import Graph


def _stream():
    return torch.cuda.current_stream().cuda_stream


class SSSPFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, graph: Graph, srcs):
        if not w.is_cuda or w.dtype != torch.float32 or w.numel() != graph.m:
            raise ValueError("w must be a float32 CUDA tensor with one weight per directed edge")
        if graph.wtype != 1:
            raise ValueError("graph must have float weights")
        w = w.contiguous()
        srcs = np.ascontiguousarray(srcs, dtype=np.uint32)
        k = len(srcs)
        graph.set_weights_device(w.data_ptr(), stream=_stream())
        dist = torch.empty((k, graph.n), dtype=torch.float32, device=w.device)
        graph.sssp_multi_device(srcs, dist.data_ptr(), stream=_stream())
        ctx.graph = graph
        ctx.srcs = srcs
        ctx.save_for_backward(w, dist)
        ctx.mark_non_differentiable()
        return dist

    @staticmethod
    def backward(ctx, grad_dist):
        w, dist = ctx.saved_tensors
        graph = ctx.graph
        g = grad_dist.contiguous().to(torch.float32)
        # the handle's weights may have been replaced by a later forward: restore this call's weights
        graph.set_weights_device(w.data_ptr(), stream=_stream())
        gw = torch.empty(graph.m, dtype=torch.float32, device=w.device)
        graph.backward_device(ctx.srcs, dist.data_ptr(), g.data_ptr(), gw.data_ptr(), stream=_stream())
        return gw, None, None


def sssp(w, graph: Graph, srcs):
    """Differentiable multi-source SSSP: returns dist [k, n] (float32; +inf where unreached)."""
    return SSSPFunction.apply(w, graph, srcs)
