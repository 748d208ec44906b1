"""torch.autograd wrappers of the device NUFFT (no reference counterpart;
apollo-nufft has no autograd).

Type-2 ``c = T2(f, x)``, c_j = sum_k f_k exp(+2 pi i k.x_j / L):
  grad_f = T1(g) at the same points (the opposite-type NUFFT),
  grad_x[j, a] = Re(conj(g_j) d c_j / d x_j^(a)), evaluated on the oversampled
  grid with the kernel derivative (anufft_execute_grad_points); when both
  are needed on a plan with the MFMA kernels, one call computes both from a
  single binning (anufft_execute_type2_backward); likewise for Type-1
  (anufft_execute_type1_backward: one FFT of G, one binning).
Type-1 ``F = T1(c, x)``, F_k = sum_j c_j exp(-2 pi i k.x_j / L):
  grad_c = T2(G) at the same points,
  grad_x[j, a] = Re(conj(c_j) d/dx^(a) sum_k G_k exp(+2 pi i k.x_j / L)).
Gradients follow torch's convention for complex tensors (the conjugate
Wirtinger derivative times 2 for real losses).  Plans must use the
direct-sum scaling (compat_reference_type2_3d_scaling off).
"""
import torch

from .plan import NufftPlan1D, NufftPlan3D  # noqa: F401


def _check(plan):
    if plan.meta["compat_reference_type2_3d_scaling"]:
        raise ValueError("autograd needs a plan with the direct-sum Type-2 scaling")


class _Type2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, modes, positions, plan):
        _check(plan)
        out = plan.type2(positions.contiguous(), modes.contiguous())
        ctx.plan = plan
        ctx.save_for_backward(modes, positions)
        return out

    @staticmethod
    def backward(ctx, g):
        modes, positions = ctx.saved_tensors
        plan = ctx.plan
        g = g.contiguous()
        gm = gx = None
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1] and plan.fused_backward:
            # both gradients from one binning of the points
            gm, gx = plan.type2_backward(positions.contiguous(), g, modes.contiguous())
            return gm, gx.reshape(positions.shape).to(positions.dtype), None
        if ctx.needs_input_grad[0]:
            gm = plan.type1(positions.contiguous(), g)
        if ctx.needs_input_grad[1]:
            gx = plan.grad_points(positions.contiguous(), g, modes.contiguous())
            gx = gx.reshape(positions.shape).to(positions.dtype)
        return gm, gx, None


class _Type1(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, positions, plan):
        _check(plan)
        out = plan.type1(positions.contiguous(), values.contiguous())
        ctx.plan = plan
        ctx.save_for_backward(values, positions)
        return out

    @staticmethod
    def backward(ctx, G):
        values, positions = ctx.saved_tensors
        plan = ctx.plan
        G = G.contiguous()
        gv = gx = None
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1] and plan.fused_backward:
            # both gradients from one FFT of G and one binning of the points
            gv, gx = plan.type1_backward(positions.contiguous(), values.contiguous(), G)
            return gv, gx.reshape(positions.shape).to(positions.dtype), None
        if ctx.needs_input_grad[0]:
            gv = plan.type2(positions.contiguous(), G)
        if ctx.needs_input_grad[1]:
            gx = plan.grad_points(positions.contiguous(), values.contiguous(), G)
            gx = gx.reshape(positions.shape).to(positions.dtype)
        return gv, gx, None


def nufft_type2(modes, positions, plan):
    """Differentiable Type-2: modes (plan mode shape, complex64), positions
    (M, 3) or (M,) float32 on the plan's GPU."""
    return _Type2.apply(modes, positions, plan)


def nufft_type1(values, positions, plan):
    """Differentiable Type-1: values (M,) complex64, positions (M, 3) or (M,)."""
    return _Type1.apply(values, positions, plan)
