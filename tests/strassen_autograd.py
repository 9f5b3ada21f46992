"""Test-only torch.autograd.Function over the library's contraction and its adjoints.

forward:  C[idx_C] = A[idx_A] . B[idx_B]     (Strassen contraction of the chosen level / impl)
backward: dA = dC . B,  dB = A . dC          (mistra_tensor_mult_backward: two Strassen contractions)
Device tensors only; every call goes through the C-ABI.
"""
import torch

import tblis_strassen as ts


class StrassenContract(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B, idx_A, idx_B, idx_C, shape_C, level, impl):
        C = torch.zeros(shape_C, dtype=A.dtype, device=A.device)
        stream = torch.cuda.current_stream()
        ts.tensor_mult(1.0, A, idx_A, B, idx_B, 0.0, C, idx_C, level=level, impl=impl, stream=stream)
        ctx.save_for_backward(A, B)
        ctx.idx = (idx_A, idx_B, idx_C)
        ctx.level, ctx.impl = level, impl
        return C

    @staticmethod
    def backward(ctx, dC):
        A, B = ctx.saved_tensors
        idx_A, idx_B, idx_C = ctx.idx
        need_a, need_b = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        dA = torch.empty_like(A) if need_a else None
        dB = torch.empty_like(B) if need_b else None
        ts.tensor_mult_backward(1.0, A, idx_A, B, idx_B, dC.contiguous(), idx_C, dA=dA, dB=dB, beta=0.0,
                                level=ctx.level, impl=ctx.impl, stream=torch.cuda.current_stream())
        return dA, dB, None, None, None, None, None, None


def contract(A, B, idx_A, idx_B, idx_C, shape_C, level=1, impl="abc"):
    return StrassenContract.apply(A, B, idx_A, idx_B, idx_C, tuple(shape_C), level, impl)
