"""The torch.autograd.Function over the Cgemm3m kernels (forward and op=C backward).

Pinned two ways: (1) complex128 central finite differences of L = Re(sum conj(G) * Y) at
m = n = k = 32 (exact function evaluated with complex128 matmul on the CPU), against the
kernel's complex64 gradients; (2) torch.matmul autograd at the full 2048 x 1024 x 4096
size, in complex128, against both the FP32 and the bf16x3 backward.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rand(shape, seed):
    g = torch.Generator().manual_seed(seed)
    re = torch.rand(shape, generator=g, dtype=torch.float32) * 2 - 1
    im = torch.rand(shape, generator=g, dtype=torch.float32) * 2 - 1
    return torch.complex(re, im)


def _fd_grad(f, X, h=1e-6):
    """Conjugate-Wirtinger gradient dL/dRe + i dL/dIm by central differences (complex128)."""
    X = X.to(torch.complex128)
    g = torch.zeros_like(X)
    flat = X.reshape(-1)
    gf = g.reshape(-1)
    for idx in range(flat.numel()):
        for unit, part in ((1.0, "re"), (1j, "im")):
            Xp = flat.clone()
            Xm = flat.clone()
            Xp[idx] += h * unit
            Xm[idx] -= h * unit
            d = (f(Xp.reshape(X.shape)) - f(Xm.reshape(X.shape))) / (2 * h)
            gf[idx] += d if part == "re" else 1j * d
    return g


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_gradcheck_fd_complex128(gpu, precision):
    from cgemm3m.autograd import cgemm3m_mm

    m = n = k = 32
    alpha, beta = complex(0.8, 0.6), complex(0.5, 0.0)
    A, B, C, G = _rand((m, k), 1), _rand((k, n), 2), _rand((m, n), 3), _rand((m, n), 4)
    Ad, Bd, Cd = (x.cuda().requires_grad_(True) for x in (A, B, C))
    Y = cgemm3m_mm(Ad, Bd, Cd, alpha, beta, precision)
    Y.backward(G.cuda())
    A64, B64, C64, G64 = (x.to(torch.complex128) for x in (A, B, C, G))

    def loss(Ax, Bx, Cx):
        return float(torch.real((G64.conj() * (alpha * Ax @ Bx + beta * Cx)).sum()))

    tol = 2e-5 if precision == "fp32" else 2e-4
    y_ref = alpha * A64 @ B64 + beta * C64
    assert (Y.detach().cpu().to(torch.complex128) - y_ref).abs().max() <= tol * y_ref.abs().max()
    for name, got, fd in (
        ("A", Ad.grad, _fd_grad(lambda X: loss(X, B64, C64), A64)),
        ("B", Bd.grad, _fd_grad(lambda X: loss(A64, X, C64), B64)),
        ("C", Cd.grad, _fd_grad(lambda X: loss(A64, B64, X), C64)),
    ):
        err = (got.cpu().to(torch.complex128) - fd).abs().max() / fd.abs().max()
        print(precision, name, "fd rel err", float(err))
        assert err <= tol, (name, float(err))


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_full_size_vs_torch_matmul(gpu, precision):
    from cgemm3m.autograd import cgemm3m_mm

    m, k, n = 2048, 1024, 4096
    alpha, beta = complex(0.8, 0.6), complex(0.5, 0.0)
    A, B, C, G = (_rand(s, 10 + i).cuda() for i, s in enumerate(((m, k), (k, n), (m, n), (m, n))))
    Ad, Bd, Cd = (x.clone().requires_grad_(True) for x in (A, B, C))
    Y = cgemm3m_mm(Ad, Bd, Cd, alpha, beta, precision)
    Y.backward(G)
    Ar, Br, Cr = (x.to(torch.complex128).requires_grad_(True) for x in (A, B, C))
    Yr = alpha * torch.matmul(Ar, Br) + beta * Cr
    Yr.backward(G.to(torch.complex128))
    tol = 1.5e-5 if precision == "fp32" else 6e-5
    for name, got, ref in (("Y", Y.detach(), Yr.detach()), ("A", Ad.grad, Ar.grad), ("B", Bd.grad, Br.grad),
                           ("C", Cd.grad, Cr.grad)):
        err = float((got.to(torch.complex128) - ref).abs().max() / ref.abs().max())
        print(precision, name, "normwise err vs torch.matmul complex128", err)
        assert err <= tol, (name, err)
