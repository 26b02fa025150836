"""torch.autograd harness over the C-ABI forward/adjoint pair.

``tbl_adjoint_window`` (include/tbl.h) does two things for a window of steps:

* it evaluates the downstream wall-shear objective J;
* it computes the gradient of J with respect to the wall-normal wall velocity,
  by the hand-written discrete adjoint.

This module wraps that pair in a ``torch.autograd.Function``, so that a
gradient check (central differences along seeded directions) can be written
against ``J.backward()``. It is a thin harness: all compute runs in
libtbl_hip.so.

The convection's Peclet switch is discontinuous in the velocity, so a finite
difference that crosses a switch measures a jump, not a derivative. With
``freeze=True`` the behaviour is:

* a gradient evaluation records the switch pattern of its trajectory;
* a gradient-free evaluation replays that pattern.

Central differences then see the same piecewise-smooth map that the adjoint
differentiates.
"""
import numpy as np
import torch

from . import tbl


class WallShearWindow(torch.autograd.Function):
    """J(wall_w): wall_w is an (nX, nY) float64 tensor.  `solver` is a
    tbl.Solver that `reset(solver)` puts back at the window's first state."""

    @staticmethod
    def forward(ctx, wall_w, solver, reset, nsteps, i0, i1, nsnap, freeze):
        need = ctx.needs_input_grad[0]
        mode = (1 if need else 2) if freeze else 0
        reset(solver)
        J, g, _, _ = solver.adjoint_window(nsteps, i0, i1, wall_w=wall_w.detach().cpu().double().numpy(),
                                           nsnap=nsnap, forward_only=not need, branch_mode=mode)
        if need:
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(g)).to(wall_w.device))
        return wall_w.new_tensor(J)

    @staticmethod
    def backward(ctx, gout):
        (g,) = ctx.saved_tensors
        return gout * g, None, None, None, None, None, None, None


def wall_shear_objective(wall_w, solver, reset, nsteps, i0, i1, nsnap=10, freeze=False):
    """J of a window of `nsteps` steps: the time- and span-mean wall shear over
    1-based stations i0..i1 (differentiable in wall_w)"""
    return WallShearWindow.apply(wall_w, solver, reset, nsteps, i0, i1, nsnap, freeze)
