"""Differentiable slosh loads: a torch.autograd.Function over the solver's
hand-written adjoint (CryoSPH.Solver.adjoint_slosh, cryo_solver_adjoint_slosh).

    loss = SloshLoss.apply(segments, make_solver, steps, checkpoint_every)
    loss = SloshLoss.apply(segments, make_solver, steps, checkpoint_every, lee)

`segments` is an (S, 3) tensor of frame accelerations, held piecewise constant
over S equal spans of `steps` steps (step k uses segment k * S // steps).
`make_solver()` returns a fresh, loaded CryoSPH.Solver (the run advances it);
with a scalar tensor `lee` it is called as make_solver(lee) and must set the
evaporation and condensation coefficients to that value, and d loss / d lee is
the sum of the two coefficients' derivatives.
The loss is sum_k dt |load_k,xy|^2 with load_k the wall load (the negated
slosh force) after step k.  When a gradient is wanted the forward call runs the
reverse sweep too (checkpoints every `checkpoint_every` steps, segments re-run
forward on the GPU), so backward() only scales the stored d loss / d segments.
The reference has no adjoint or autodiff: this module has no counterpart there.
"""
import numpy as np
import torch


def step_segments(steps: int, count: int) -> np.ndarray:
    """Segment index of every step: k * count // steps."""
    return (np.arange(steps, dtype=np.int64) * count) // steps


class SloshLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, segments, make_solver, steps, checkpoint_every=20, lee=None):
        if segments.dim() != 2 or segments.shape[1] != 3:
            raise ValueError("segments must have shape (S, 3)")
        seg = segments.detach().to("cpu", torch.float64).numpy()
        idx = step_segments(int(steps), seg.shape[0])
        acc = np.ascontiguousarray(seg[idx])
        want_lee = lee is not None and bool(ctx.needs_input_grad[4])
        want = bool(ctx.needs_input_grad[0]) or want_lee
        solver = make_solver() if lee is None else make_solver(float(lee))
        loss, grad, _, _, dlee = solver.adjoint_slosh(acc, int(checkpoint_every), want)
        gseg = np.zeros_like(seg)
        if want:
            np.add.at(gseg, idx, grad)
        ctx.save_for_backward(torch.from_numpy(gseg).to(segments.device, segments.dtype))
        ctx.dlee = (dlee[0] + dlee[1]) if want_lee else 0.0
        return torch.tensor(loss, dtype=segments.dtype, device=segments.device)

    @staticmethod
    def backward(ctx, grad_out):
        (gseg,) = ctx.saved_tensors
        if len(ctx.needs_input_grad) == 4:  # called without lee
            return grad_out * gseg, None, None, None
        glee = grad_out * ctx.dlee if ctx.needs_input_grad[4] else None
        return grad_out * gseg, None, None, None, glee
