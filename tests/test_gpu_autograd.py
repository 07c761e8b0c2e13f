"""The separable convolution as a torch.autograd.Function (test harness only).

The library exposes the forward and both adjoints through its C ABI
(ndl_sepconv_forward, ndl_sepconv_adjoint_input, ndl_sepconv_adjoint_taps);
this wraps them over torch tensors on the GPU so torch.autograd.gradcheck can
compare the hand-written adjoints with FP64 central differences of the
forward, for the input and for every tap of every axis, on every border.
torch is plumbing here (device memory, the current stream); the product
never imports it.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import ndl_amd
from ndl_amd import abi

pytestmark = pytest.mark.gpu

DT = {"float64": abi.F64, "float32": abi.F32}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _desc(t):
    """ndl descriptor of a torch tensor (ndl axis d = torch dim ndim-1-d)"""
    return abi.make_tensor(t.data_ptr(), DT[str(t.dtype).split(".")[1]], tuple(reversed(t.shape)),
                           tuple(reversed(t.stride())))


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def make_function():
    import torch

    class SepConv(torch.autograd.Function):
        """y = A x, A the separable operator of `taps` (one 1-D tensor per
        axis, ndl axis order) and `borders`, centres count // 2."""

        @staticmethod
        def forward(ctx, x, accum, borders, *taps):
            ks = abi.Kernels([t.detach().cpu().numpy() for t in taps], borders)
            axes = abi.Kernels.axes(list(range(len(taps))))
            xc = x.contiguous()
            y = torch.empty_like(xc)
            tx, ty = _desc(xc), _desc(y)
            ndl_amd.check(ndl_amd.lib().ndl_sepconv_forward(C.byref(tx), C.byref(ty), len(taps), axes, ks.arr,
                                                            accum, abi.ROUND_TRUNC, _stream()))
            ctx.save_for_backward(xc, *taps)
            ctx.accum, ctx.borders = accum, borders
            return y

        @staticmethod
        def backward(ctx, gy):
            x, *taps = ctx.saved_tensors
            ks = abi.Kernels([t.detach().cpu().numpy() for t in taps], ctx.borders)
            axes = abi.Kernels.axes(list(range(len(taps))))
            g = gy.contiguous()
            lib = ndl_amd.lib()
            gx = None
            if ctx.needs_input_grad[0]:
                gx = torch.empty_like(g)
                tg, tgx = _desc(g), _desc(gx)
                ndl_amd.check(lib.ndl_sepconv_adjoint_input(C.byref(tg), C.byref(tgx), len(taps), axes, ks.arr,
                                                            ctx.accum, _stream()))
            gt = [None] * len(taps)
            if any(ctx.needs_input_grad[3:]):
                out = torch.empty(sum(len(t) for t in taps), dtype=torch.float64, device=x.device)
                tx_, tg = _desc(x), _desc(g)
                ndl_amd.check(lib.ndl_sepconv_adjoint_taps(C.byref(tx_), C.byref(tg), len(taps), axes, ks.arr,
                                                           C.cast(out.data_ptr(), C.POINTER(C.c_double)),
                                                           ctx.accum, _stream()))
                flat = out.cpu()
                off = 0
                for i, t in enumerate(taps):
                    gt[i] = flat[off:off + len(t)].to(t.dtype)
                    off += len(t)
            return (gx, None, None, *gt)

    return SepConv


@pytest.mark.parametrize("border", ["clamp", "wrap", "reflect"])
def test_gradcheck_fp64_input_and_taps(border):
    # FP64 tensors, the reference arithmetic: the analytical input and tap
    # gradients against torch's central differences of the forward
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    SepConv = make_function()
    g = torch.Generator().manual_seed(5)
    x = torch.rand(6, 7, 9, dtype=torch.float64, generator=g).cuda().requires_grad_()
    # asymmetric taps, odd and even widths (even: centre count // 2)
    taps = [torch.randn(k, dtype=torch.float64, generator=g).requires_grad_() for k in (5, 4, 3)]
    say("gradcheck", border)
    ok = torch.autograd.gradcheck(lambda xx, a, b, c: SepConv.apply(xx, abi.ACCUM_REF, [border] * 3, a, b, c),
                                  (x, *taps), eps=1e-6, atol=1e-9, rtol=1e-7)
    assert ok


@pytest.mark.parametrize("border", ["clamp", "reflect"])
def test_fp32_backward_matches_fp64(border):
    # the FP32-accumulation forward / adjoints (fused kernels) against the
    # FP64 ones through the same autograd graph: loss = sum(w * A x)
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    SepConv = make_function()
    g = torch.Generator().manual_seed(9)
    x64 = torch.randn(40, 36, 44, dtype=torch.float64, generator=g)
    w64 = torch.randn(40, 36, 44, dtype=torch.float64, generator=g)
    t64 = [torch.randn(9, dtype=torch.float64, generator=g) for _ in range(3)]
    grads = {}
    for name, dt, accum in (("fp64", torch.float64, abi.ACCUM_REF), ("fp32", torch.float32, abi.ACCUM_F32)):
        x = x64.to(dt).cuda().requires_grad_()
        taps = [t.clone().to(dt).requires_grad_() for t in t64]
        y = SepConv.apply(x, accum, [border] * 3, *taps)
        (y * w64.to(dt).cuda()).sum().backward()
        grads[name] = (x.grad.double().cpu(), [t.grad.double() for t in taps])
    gx64, gt64 = grads["fp64"]
    gx32, gt32 = grads["fp32"]
    scale = float(w64.abs().max()) * float(np.prod([t.abs().sum() for t in t64]))
    assert float((gx32 - gx64).abs().max()) <= 2e-6 * scale
    # tap gradients: the error of FP32 U values is bounded by eps times the
    # abs-weighted sum sum_p |w_p| Ubar_p (Ubar = |x| through |taps|)
    import oracle
    bound = oracle.adjoint_taps(x64.abs().numpy(), w64.abs().numpy(), [0, 1, 2], [t.abs().numpy() for t in t64],
                                [border] * 3)
    err = np.concatenate([(a - b).abs().numpy() for a, b in zip(gt32, gt64)])
    assert np.all(err <= 4e-6 * bound), float(np.max(err / bound))
