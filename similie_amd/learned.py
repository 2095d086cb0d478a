"""Python mirror of the learned-preconditioner and equal-size batch solver
entry points (sil_unet_*, sil_poisson2d_batch_*).

Tensors are torch CUDA tensors; every call runs asynchronously on torch's
current stream unless a stream is given.  There is no CPU fallback.
"""
import ctypes

import torch

from . import _lib as L

PARAM_COUNT = None


def param_count():
    global PARAM_COUNT
    if PARAM_COUNT is None:
        PARAM_COUNT = int(L.lib().sil_unet_param_count())
    return PARAM_COUNT


def _f32(t, name, shape=None):
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous float32 CUDA tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return L.ptr(t)


def _f64(t, name, shape=None):
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous float64 CUDA tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return L.ptr(t)


class UNet:
    """4-level, 32-channel U-Net with BF16 MFMA kernels (sil_unet)."""

    def __init__(self, max_batch, height, width, seed=0):
        h = ctypes.c_void_p()
        L.check(L.lib().sil_unet_create(ctypes.byref(h), max_batch, height, width, seed))
        self.handle = h
        self._destroy = L.lib().sil_unet_destroy
        self.max_batch, self.height, self.width = max_batch, height, width

    def __del__(self):
        # the destroy function was bound at construction: module globals may
        # already be gone when this runs at interpreter exit
        h = getattr(self, "handle", None)
        fn = getattr(self, "_destroy", None)
        if h is not None and h.value and fn is not None:
            fn(h)
            self.handle = None

    def get_params(self, stream=None):
        p = torch.empty(param_count(), dtype=torch.float32, device="cuda")
        L.check(L.lib().sil_unet_get_params(self.handle, L.ptr(p), L.stream_handle(stream)))
        return p

    def set_params(self, params, stream=None):
        params = params.to(device="cuda", dtype=torch.float32).contiguous()
        _f32(params, "params", (param_count(),))
        L.check(L.lib().sil_unet_set_params(self.handle, L.ptr(params), L.stream_handle(stream)))
        torch.cuda.current_stream().synchronize() if stream is None else stream.synchronize()

    def _shape(self, x):
        b = x.shape[0]
        _f32(x, "input", (b, self.height, self.width, 2))
        return b

    def forward(self, x, out=None, stream=None):
        b = self._shape(x)
        if out is None:
            out = torch.empty((b, self.height, self.width), dtype=torch.float32, device="cuda")
        _f32(out, "output", (b, self.height, self.width))
        L.check(L.lib().sil_unet_forward(self.handle, b, L.ptr(x), L.ptr(out), L.stream_handle(stream)))
        return out

    def infer(self, x, out=None, stream=None):
        """forward() without the activations a backward pass needs (the input
        layer and layer 1 fused); the same output bit for bit"""
        b = self._shape(x)
        if out is None:
            out = torch.empty((b, self.height, self.width), dtype=torch.float32, device="cuda")
        _f32(out, "output", (b, self.height, self.width))
        L.check(L.lib().sil_unet_infer(self.handle, b, L.ptr(x), L.ptr(out), L.stream_handle(stream)))
        return out

    ACTIVATIONS = ("a0", "e0", "a1", "e1", "a2", "e2", "a3", "b3", "c2", "f2", "c1", "f1", "c0")
    LEVELS = (0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0)

    def activation(self, name, batch, stream=None):
        """The stored activation `name` of the last forward pass (the output
        of conv_in or of one of the twelve layers), FP32 [batch][H_l][W_l][32]."""
        which = self.ACTIVATIONS.index(name)
        lv = self.LEVELS[which]
        out = torch.empty((batch, self.height >> lv, self.width >> lv, 32), dtype=torch.float32, device="cuda")
        L.check(L.lib().sil_unet_activation(self.handle, which, batch, L.ptr(out), L.stream_handle(stream)))
        return out

    def backward(self, x, dout, stream=None):
        """Parameter gradient of <dout, forward(x)> (call forward(x) first)."""
        b = self._shape(x)
        _f32(dout, "dout", (b, self.height, self.width))
        grad = torch.empty(param_count(), dtype=torch.float32, device="cuda")
        L.check(L.lib().sil_unet_backward(self.handle, b, L.ptr(x), L.ptr(dout), L.ptr(grad),
                                          L.stream_handle(stream)))
        return grad

    def loss(self, out, x, err, wx, wy, stream=None):
        b = self._shape(x)
        shp = (b, self.height, self.width)
        for t, name in ((out, "output"), (err, "err"), (wx, "wx"), (wy, "wy")):
            _f32(t, name, shp)
        dout = torch.empty(shp, dtype=torch.float32, device="cuda")
        loss = torch.empty(b, dtype=torch.float32, device="cuda")
        L.check(L.lib().sil_unet_loss(self.handle, b, L.ptr(out), L.ptr(x), L.ptr(err), L.ptr(wx), L.ptr(wy),
                                      L.ptr(dout), L.ptr(loss), L.stream_handle(stream)))
        return loss, dout

    def train_step(self, x, err, wx, wy, lr, loss=None, stream=None):
        """Forward, direction loss, backward and one Adam step; returns the
        per-image loss tensor (device, not synchronised)."""
        b = self._shape(x)
        shp = (b, self.height, self.width)
        for t, name in ((err, "err"), (wx, "wx"), (wy, "wy")):
            _f32(t, name, shp)
        if loss is None:
            loss = torch.empty(b, dtype=torch.float32, device="cuda")
        L.check(L.lib().sil_unet_train_step(self.handle, b, L.ptr(x), L.ptr(err), L.ptr(wx), L.ptr(wy),
                                            ctypes.c_float(lr), L.ptr(loss), L.stream_handle(stream)))
        return loss


METHODS = {
    "cg_plain": (L.PRECOND_IDENTITY, L.PRECISION_FP64),
    "cg": (L.PRECOND_JACOBI, L.PRECISION_FP64),
    "cg_mixed": (L.PRECOND_JACOBI, L.PRECISION_MIXED),
    "fcg_learned": (L.PRECOND_LEARNED, L.PRECISION_FP64),
    "fcg_learned_mixed": (L.PRECOND_LEARNED, L.PRECISION_MIXED),
}


class PoissonBatch:
    """B equal-size Dirichlet-zero 2D Poisson problems (sil_poisson2d_batch)."""

    def __init__(self, kappa):
        if kappa.dim() != 3 or kappa.shape[1] != kappa.shape[2]:
            raise ValueError("kappa must be [B][n][n]")
        self.kappa = kappa.to(device="cuda", dtype=torch.float64).contiguous()
        self.B, self.n = int(kappa.shape[0]), int(kappa.shape[1])
        h = ctypes.c_void_p()
        L.check(L.lib().sil_poisson2d_batch_create(ctypes.byref(h), self.B, self.n, L.ptr(self.kappa)))
        self.handle = h
        self._destroy = L.lib().sil_poisson2d_batch_destroy

    def __del__(self):
        # the destroy function was bound at construction: module globals may
        # already be gone when this runs at interpreter exit
        h = getattr(self, "handle", None)
        fn = getattr(self, "_destroy", None)
        if h is not None and h.value and fn is not None:
            fn(h)
            self.handle = None

    @property
    def shape(self):
        return (self.B, self.n, self.n)

    def apply(self, x, out=None, stream=None):
        _f64(x, "x", self.shape)
        if out is None:
            out = torch.empty(self.shape, dtype=torch.float64, device="cuda")
        L.check(L.lib().sil_poisson2d_batch_apply(self.handle, L.ptr(x), L.ptr(out), L.stream_handle(stream)))
        return out

    def weights(self, stream=None):
        """FP32 edge weights (wx, wy) and the normalised log κ channel."""
        wx, wy, lk = (torch.empty(self.shape, dtype=torch.float32, device="cuda") for _ in range(3))
        L.check(L.lib().sil_poisson2d_batch_weights(self.handle, L.ptr(wx), L.ptr(wy), L.ptr(lk),
                                                    L.stream_handle(stream)))
        return wx, wy, lk

    def solve(self, rhs, x=None, method="cg", tol=1e-10, max_iterations=20000, net=None, replace_interval=0,
              record_every=0, max_records=0, stream=None):
        """Solves every problem; returns (x, results list, seconds[, records]).

        method: "cg_plain" (unpreconditioned CG, FP64), "cg" (Jacobi CG,
        FP64), "cg_mixed", "fcg_learned" or "fcg_learned_mixed".  With record_every > 0 the FP32 residuals and
        iterates (FP64) before every record_every iterations are returned as
        (r_records, x_records) of shape [count][B][n][n]."""
        _f64(rhs, "rhs", self.shape)
        if x is None:
            x = torch.zeros(self.shape, dtype=torch.float64, device="cuda")
        _f64(x, "x", self.shape)
        precond, precision = METHODS[method]
        st = L.SolverSettings()
        L.lib().sil_default_solver_settings(ctypes.byref(st))
        st.max_iterations = max_iterations
        st.relative_tolerance = tol
        st.precision = precision
        st.preconditioner = precond
        st.replace_interval = replace_interval
        res = (L.BatchResult * self.B)()
        secs = ctypes.c_double()
        rec = None
        rr = xr = None
        count = ctypes.c_int32(0)
        if record_every:
            rr = torch.empty((max_records,) + self.shape, dtype=torch.float32, device="cuda")
            xr = torch.empty((max_records,) + self.shape, dtype=torch.float64, device="cuda")
            rec = L.SolveRecord(record_every, max_records, L.ptr(rr), L.ptr(xr), ctypes.pointer(count))
        L.check(L.lib().sil_poisson2d_batch_solve(self.handle, L.ptr(rhs), L.ptr(x), ctypes.byref(st),
                                                  net.handle if net is not None else None, res,
                                                  ctypes.byref(secs), ctypes.byref(rec) if rec else None,
                                                  L.stream_handle(stream)))
        results = [{"iterations": r.iterations, "converged": bool(r.converged),
                    "true_relative_residual": r.true_relative_residual} for r in res]
        if record_every:
            return x, results, secs.value, (rr[: count.value], xr[: count.value])
        return x, results, secs.value
