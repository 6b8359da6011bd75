"""Differentiable complex GEMM Y = alpha * A @ B + beta * C on the Cgemm3m kernels.

Forward and backward both call aclblasCgemm3mEx on torch's current stream; nothing here
computes a product itself. With row-major torch tensors (A: m x k, B: k x n, C: m x n):

  forward   Y^T = alpha * B^T A^T + beta * C^T          ops (N, N) on the column-major views
  dA        = conj(alpha) * dY @ B^H   -> (dA)^T = conj(alpha) * conj(B) dY^T   ops (C, N)
  dB        = conj(alpha) * A^H @ dY   -> (dB)^T = conj(alpha) * dY^T conj(A)   ops (N, C)
  dC        = conj(beta) * dY          (the library's beta-scale path: alpha = 0, k = 0)

(a row-major r x c tensor is the column-major c x r matrix with leading dimension c, and
conj(X) of a row-major X is op = C applied to that column-major view). These are torch's
conjugate-Wirtinger gradients for a real loss.
"""
import threading

import torch

from . import api

_tls = threading.local()


def _handle():
    h = getattr(_tls, "handle", None)
    if h is None:
        h = api.Handle()
        _tls.handle = h
    h.set_stream(torch.cuda.current_stream())
    return h


def _check(t, name):
    if t.dtype != torch.complex64 or not t.is_cuda:
        raise TypeError("%s must be a complex64 CUDA tensor" % name)
    return t.contiguous()


def _gemm(h, ta, tb, m, n, k, alpha, X, ldx, Y, ldy, beta, Z, ldz, cfg):
    api.cgemm3m(h, ta, tb, m, n, k, alpha, api.torch_ptr(X), ldx, api.torch_ptr(Y), ldy, beta,
                api.torch_ptr(Z), ldz, cfg)


class Cgemm3mFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B, C, alpha, beta, precision):
        A = _check(A, "A")
        B = _check(B, "B")
        m, k = A.shape
        k2, n = B.shape
        if k2 != k:
            raise ValueError("inner dimensions differ: %d vs %d" % (k, k2))
        cfg = api.config(precision)
        h = _handle()
        if C is None:
            Y = torch.empty((m, n), dtype=torch.complex64, device=A.device)
            beta_eff = 0.0
        else:
            C = _check(C, "C")
            if tuple(C.shape) != (m, n):
                raise ValueError("C must be %d x %d" % (m, n))
            Y = C.clone()
            beta_eff = beta
        _gemm(h, "N", "N", n, m, k, alpha, B, n, A, k, beta_eff, Y, n, cfg)
        ctx.save_for_backward(A, B)
        ctx.alpha, ctx.beta, ctx.precision = complex(alpha), complex(beta), precision
        ctx.has_c = C is not None
        return Y

    @staticmethod
    def backward(ctx, gY):
        A, B = ctx.saved_tensors
        gY = gY.contiguous()
        if gY.dtype != torch.complex64:
            gY = gY.to(torch.complex64)
        m, k = A.shape
        n = B.shape[1]
        h = _handle()
        cfg = api.config(ctx.precision)
        ca = ctx.alpha.conjugate()
        dA = dB = dC = None
        if ctx.needs_input_grad[0]:
            dA = torch.empty((m, k), dtype=torch.complex64, device=A.device)
            _gemm(h, "C", "N", k, m, n, ca, B, n, gY, n, 0.0, dA, k, cfg)
        if ctx.needs_input_grad[1]:
            dB = torch.empty((k, n), dtype=torch.complex64, device=A.device)
            _gemm(h, "N", "C", n, k, m, ca, gY, n, A, k, 0.0, dB, n, cfg)
        if ctx.has_c and ctx.needs_input_grad[2]:
            dC = gY.clone()
            # alpha = 0 takes the beta-scale kernel: dC = conj(beta) * dY
            api.cgemm3m(h, "N", "N", n, m, 0, 0.0, None, max(1, n), None, 1, ctx.beta.conjugate(),
                        api.torch_ptr(dC), n)
        return dA, dB, dC, None, None, None


def cgemm3m_mm(A, B, C=None, alpha=1.0, beta=0.0, precision="fp32"):
    """Y = alpha * A @ B + beta * C for complex64 CUDA tensors, differentiable in A, B, C."""
    return Cgemm3mFunction.apply(A, B, C, alpha, beta, precision)
