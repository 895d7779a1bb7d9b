"""torch.autograd binding of the H-matvec (no reference counterpart).

Y = H X is a torch.autograd.Function whose backward is computed by the
library: the input gradient is the transposed H-matvec H^T dY (U and V roles
swapped), the factor gradients are the per-tile outer products
dU_t = dY[rows] (V_t^T X[cols])^T and dV_t = X[cols] (U_t^T dY[rows])^T.
The factors are exposed as torch parameter tensors (one U and one V tensor
per level, in the library's layout); when a parameter is modified in place
(an optimizer step) the next forward copies it back into the library.
"""
import ctypes

import torch

from . import _check, _i32, _ptr, lib


class HMatrixOperator:
    """Torch-facing view of a built hmx.Pipeline (H-matrix on cuda)."""

    def __init__(self, pipe, requires_grad=True):
        self.pipe = pipe
        H = pipe.H
        self.levels = []
        self.U, self.V = {}, {}
        for l in range(1, H.num_levels):
            info = H.level_info(l)
            m = (1 << (H.num_levels - 1 - l)) * H.b
            if info["rankSum"] == 0:
                continue
            u = torch.empty(info["rankSum"] * m, dtype=torch.float64, device="cuda")
            v = torch.empty_like(u)
            _check(lib().hmx_hmatrix_level_get(pipe.ctx.handle, H.handle, l, _ptr(u), _ptr(v)), pipe.ctx)
            u.requires_grad_(requires_grad)
            v.requires_grad_(requires_grad)
            self.levels.append(l)
            self.U[l], self.V[l] = u, v
        self._seen = {l: (self.U[l]._version, self.V[l]._version) for l in self.levels}

    def parameters(self):
        out = []
        for l in self.levels:
            out += [self.U[l], self.V[l]]
        return out

    def sync(self):
        """Copy modified parameters back into the library's H-matrix."""
        for l in self.levels:
            ver = (self.U[l]._version, self.V[l]._version)
            if ver != self._seen[l]:
                _check(lib().hmx_hmatrix_set_level(self.pipe.ctx.handle, self.pipe.H.handle, l,
                                                   _ptr(self.U[l].detach()), _ptr(self.V[l].detach())), self.pipe.ctx)
                self._seen[l] = ver

    def __call__(self, X):
        return HMatVec.apply(X, self, *self.parameters())


class HMatVec(torch.autograd.Function):
    """Y = H X for X of shape (k, n) (an n x k column-major block) or (n,)."""

    @staticmethod
    def forward(ctx, X, op, *params):
        op.sync()
        Xd = X.detach().to(torch.float64).contiguous()
        Y = op.pipe.matvec(Xd)
        ctx.op = op
        ctx.in_dtype = X.dtype
        ctx.save_for_backward(Xd)
        return Y.to(X.dtype)

    @staticmethod
    def backward(ctx, dY):
        op = ctx.op
        (Xd,) = ctx.saved_tensors
        W = dY.detach().to(torch.float64).contiguous()
        dX = None
        if ctx.needs_input_grad[0]:
            dX = op.pipe.matvec(W, transpose=True).to(ctx.in_dtype)
        grads = [None] * (len(ctx.needs_input_grad) - 2)
        if any(ctx.needs_input_grad[2:]):
            H = op.pipe.H
            nl = H.num_levels - 1
            dU = {l: torch.zeros_like(op.U[l], requires_grad=False) for l in op.levels}
            dV = {l: torch.zeros_like(op.V[l], requires_grad=False) for l in op.levels}
            pu = (ctypes.c_void_p * nl)(*[dU[l].data_ptr() if l in dU else 0 for l in range(1, nl + 1)])
            pv = (ctypes.c_void_p * nl)(*[dV[l].data_ptr() if l in dV else 0 for l in range(1, nl + 1)])
            k = 1 if Xd.dim() == 1 else Xd.shape[0]
            _check(lib().hmx_hmatvec_grad_factors(op.pipe.ctx.handle, H.handle, k, _ptr(Xd), _ptr(W), pu, pv),
                   op.pipe.ctx)
            grads = []
            for l in op.levels:
                grads += [dU[l], dV[l]]
        return (dX, None, *grads)
