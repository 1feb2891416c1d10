"""torch.autograd binding of the weighted personalized PageRank operator.

``ppr(src, dst, w, s, alpha)`` returns r with r = (1-alpha) s + alpha P^T r,
P_e = w_e / (sum of the weights leaving src_e); its backward calls the
library's adjoint (cytom_ppr_graph_adjoint) on the operator the forward built,
for dL/dw and dL/ds.

Both directions run in libcytom_hip.so on the GPU; torch only does the
autograd bookkeeping on host tensors (the library owns its device buffers,
so no second HIP runtime has to share the device with torch's).
"""
import ctypes

import numpy as np
import torch

from . import CytomError, DeviceArray, lib

DEFAULT_TOL = 0.0
DEFAULT_MAX_ITER = 400


def _check(rc, what):
    if rc < 0:
        msg = lib().cytom_ppr_last_error()
        raise CytomError(rc, "%s: %s" % (what, msg.decode() if msg else ""))


class PPRGraph:
    """The operator on one edge list (cytom_ppr_graph): the CSR/CSC rows are
    built once, set_weights only recomputes the transition probabilities, and
    forward / adjoint calls reuse both."""

    def __init__(self, n_vertices, src, dst, w):
        self.V = int(n_vertices)
        self.E = len(src)
        src_d = DeviceArray.from_host(np.ascontiguousarray(src, np.uint32))
        dst_d = DeviceArray.from_host(np.ascontiguousarray(dst, np.uint32))
        w_d = DeviceArray.from_host(np.ascontiguousarray(w, np.float32))
        h = ctypes.c_void_p()
        _check(lib().cytom_ppr_graph_create(self.V, self.E, src_d.ptr, dst_d.ptr, w_d.ptr,
                                            ctypes.byref(h)), "ppr_graph_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().cytom_ppr_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the binding may already be gone
            pass

    def set_weights(self, w):
        w_d = DeviceArray.from_host(np.ascontiguousarray(w, np.float32))
        _check(lib().cytom_ppr_graph_set_weights(self._h, w_d.ptr), "ppr_graph_set_weights")

    def forward_device(self, s_dev, r_dev, alpha=0.85, tol=DEFAULT_TOL,
                       max_iter=DEFAULT_MAX_ITER):
        it = ctypes.c_uint32()
        _check(lib().cytom_ppr_graph_forward(self._h, s_dev.ptr, alpha, tol, max_iter, r_dev.ptr,
                                             ctypes.byref(it)), "ppr_graph_forward")
        return it.value

    def adjoint_device(self, r_dev, g_dev, dw_dev, ds_dev, alpha=0.85, tol=DEFAULT_TOL,
                       max_iter=DEFAULT_MAX_ITER):
        it = ctypes.c_uint32()
        _check(lib().cytom_ppr_graph_adjoint(self._h, r_dev.ptr, g_dev.ptr, alpha, tol, max_iter,
                                             dw_dev.ptr if dw_dev else None,
                                             ds_dev.ptr if ds_dev else None, ctypes.byref(it)),
               "ppr_graph_adjoint")
        return it.value

    def forward(self, s, alpha=0.85, tol=DEFAULT_TOL, max_iter=DEFAULT_MAX_ITER):
        s_dev = DeviceArray.from_host(np.ascontiguousarray(s, np.float32))
        r_dev = DeviceArray(self.V, np.float32)
        self.forward_device(s_dev, r_dev, alpha, tol, max_iter)
        return r_dev.to_host()

    def adjoint(self, r, g, alpha=0.85, tol=DEFAULT_TOL, max_iter=DEFAULT_MAX_ITER):
        r_dev = DeviceArray.from_host(np.ascontiguousarray(r, np.float32))
        g_dev = DeviceArray.from_host(np.ascontiguousarray(g, np.float32))
        dw_dev = DeviceArray(max(self.E, 1), np.float32)
        ds_dev = DeviceArray(self.V, np.float32)
        self.adjoint_device(r_dev, g_dev, dw_dev, ds_dev, alpha, tol, max_iter)
        return dw_dev.to_host()[:self.E], ds_dev.to_host()


class PPRFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, src, dst, w, s, alpha, tol, max_iter):
        G = PPRGraph(s.numel(), src.detach().cpu().numpy(), dst.detach().cpu().numpy(),
                     w.detach().cpu().numpy())
        r = G.forward(s.detach().cpu().numpy(), alpha, tol, max_iter)
        ctx.G = G
        ctx.cfg = (alpha, tol, max_iter)
        out = torch.from_numpy(r)
        ctx.save_for_backward(out)
        return out

    @staticmethod
    def backward(ctx, grad_r):
        (r,) = ctx.saved_tensors
        alpha, tol, max_iter = ctx.cfg
        dw, ds = ctx.G.adjoint(r.numpy(), grad_r.detach().cpu().float().numpy(), alpha, tol,
                               max_iter)
        return (None, None, torch.from_numpy(dw) if ctx.needs_input_grad[2] else None,
                torch.from_numpy(ds) if ctx.needs_input_grad[3] else None, None, None, None)


def ppr(src, dst, w, s, alpha=0.85, tol=DEFAULT_TOL, max_iter=DEFAULT_MAX_ITER):
    """Differentiable weighted personalized PageRank on host tensors (see
    module docstring): src/dst integer ids, w/s float32."""
    return PPRFunction.apply(src, dst, w, s, alpha, tol, max_iter)
