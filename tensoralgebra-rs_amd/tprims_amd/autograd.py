"""torch.autograd integration: a contraction whose forward runs through the C
ABI on DLPack operands and whose backward runs the hand-written adjoint
contractions (tprims_product_vjp): dA and dB, with diagonal scatter and
broadcast over reduced indices."""
import torch

from . import DType, Executor, Labels, LayoutSpec, Op, OperandSpec, Plan, PlanConfig, Problem, CSpec

_EXECUTORS = {}


def _executor(device):
    """An asynchronous executor on torch's current stream of `device`, so the
    library's kernels are ordered with the torch work around them (one per
    stream, kept for the process)."""
    s = torch.cuda.current_stream(device)
    if s.cuda_stream == 0:
        return None  # the null stream: the library's default executor runs there
    key = (device.index, s.cuda_stream)
    ex = _EXECUTORS.get(key)
    if ex is None:
        ex = _EXECUTORS[key] = Executor(device.index if device.index is not None else 0, s.cuda_stream,
                                        asynchronous=True)
    return ex


class _Contract(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, alpha, a, b, out_shape, out_stride):
        d = torch.empty_strided(out_shape, out_stride, dtype=a.dtype, device=a.device)
        plan.execute_dlpack(alpha, a, b, 0.0, None, d, _executor(a.device))
        ctx.plan, ctx.alpha = plan, alpha
        ctx.save_for_backward(a, b)
        ctx.out_stride = out_stride
        return d

    @staticmethod
    def backward(ctx, g):
        a, b = ctx.saved_tensors
        plan = ctx.plan
        if tuple(g.stride()) != tuple(ctx.out_stride):
            gg = torch.empty_strided(g.shape, ctx.out_stride, dtype=g.dtype, device=g.device)
            gg.copy_(g)
            g = gg
        da = db = None
        need_a, need_b = ctx.needs_input_grad[2], ctx.needs_input_grad[3]
        if need_a:
            da = torch.empty_strided(a.shape, a.stride(), dtype=a.dtype, device=a.device)
        if need_b:
            db = torch.empty_strided(b.shape, b.stride(), dtype=b.dtype, device=b.device)
        if need_a or need_b:
            plan.vjp(ctx.alpha, a, b, g, da, db, executor=_executor(g.device))
        return None, None, da, db, None, None


def _dtype(t):
    return DType.of(t)


def contract(labels_a, a, labels_b, b, labels_d, alpha=1.0, conj_a=False, conj_b=False, config=None):
    """D[labels_d] = alpha * sum op_A(A) op_B(B), differentiable in A and B.

    A and B are torch tensors on the GPU (any strides the plan accepts; a
    repeated label is a diagonal, a label of one input only is a reduction).
    D is a new compact row-major tensor.
    """
    extent = {}
    for ls, t in ((labels_a, a), (labels_b, b)):
        for l, e in zip(ls, t.shape):
            extent[l] = e
    out_shape = tuple(extent[l] for l in labels_d)
    stride, acc = [0] * len(out_shape), 1
    for k in range(len(out_shape) - 1, -1, -1):
        stride[k] = acc
        acc *= max(out_shape[k], 1)
    out_stride = tuple(stride)
    dt = _dtype(a)
    A = OperandSpec(LayoutSpec.of(a), Op.CONJUGATE if conj_a else Op.IDENTITY)
    B = OperandSpec(LayoutSpec.of(b), Op.CONJUGATE if conj_b else Op.IDENTITY)
    D = OperandSpec(LayoutSpec(list(out_shape), list(out_stride)))
    la = [ord(x) if isinstance(x, str) else int(x) for x in labels_a]
    lb = [ord(x) if isinstance(x, str) else int(x) for x in labels_b]
    ld = [ord(x) if isinstance(x, str) else int(x) for x in labels_d]
    plan = Plan(Problem.from_labels(dt, A, B, CSpec.separate(D), D, Labels(la, lb, ld, ld)), config or PlanConfig())
    return _Contract.apply(plan, alpha, a, b, out_shape, out_stride)
