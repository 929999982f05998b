"""torch.autograd binding of the DGLM solver: u(T) = S(u0) with a discrete adjoint.

    solver = Solver(problem, "euler", cfl=0.4)
    uT = dglm_evolve(u0, solver, nsteps=300)       # u0: cuda float64 [nvar, K*Np]
    J = objective(uT); J.backward()                 # u0.grad = (dS/du0)^T dJ/duT

The forward runs the HIP step loop and keeps checkpoints every `every` steps
(default ~sqrt(nsteps)), the last segments whole where device memory allows
(Solver.set_adjoint_keep); backward recomputes every other segment from its
checkpoint and runs the FP64 adjoint kernels (including the tau(u) and dt(u)
sensitivities for Euler).  A solver created with precision="fp32" gives an
FP32-MFMA forward with the FP64 adjoint along the same trajectory.

A solver holds one recorded trajectory: a second dglm_evolve on the same
solver replaces the first one's recording, and backward through the first
then raises instead of returning the gradient of the second trajectory (use
one solver per differentiable evolve in a graph).
"""
import torch


class DGLMEvolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, solver, nsteps, every, t0):
        if not u0.is_cuda or u0.dtype != torch.float64:
            raise TypeError("u0 must be a float64 CUDA tensor")
        u0c = u0.detach().contiguous()
        if u0c.numel() != solver.nvar * solver.K * solver.Np:
            raise ValueError("u0 has the wrong number of entries")
        torch.cuda.current_stream().synchronize()  # the solver runs on its own stream
        solver.set_state_device(u0c.data_ptr())
        solver.set_time(t0, 0)
        solver.adjoint_record(nsteps, every)
        out = torch.empty_like(u0c)
        solver.get_state_device(out.data_ptr())  # synchronises the solver stream
        ctx.solver = solver
        ctx.recording = solver.recording
        return out

    @staticmethod
    def backward(ctx, grad_out):
        if ctx.solver.recording != ctx.recording:
            raise RuntimeError("dglm_evolve: the solver's tape was re-recorded by a later dglm_evolve before this "
                               "backward; one solver holds one recorded trajectory (use a solver per evolve)")
        g = grad_out.contiguous()
        gin = torch.empty_like(g)
        torch.cuda.current_stream().synchronize()
        ctx.solver.adjoint_backward_device(g.data_ptr(), gin.data_ptr())
        return gin, None, None, None, None


def dglm_evolve(u0, solver, nsteps, every=None, t0=0.0):
    """Differentiable nsteps of the solver starting from u0 at time t0."""
    if every is None:
        every = max(1, int(round(nsteps ** 0.5)))
    return DGLMEvolve.apply(u0, solver, nsteps, every, t0)
