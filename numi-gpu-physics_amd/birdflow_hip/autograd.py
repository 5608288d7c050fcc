"""torch.autograd bridge for the adjoint: a load functional of mesh
kinematic parameters as a differentiable torch function.

    J(theta) = sum_s w[s] . L_s(theta)

where L_s are the force / torque of step s of a simulation whose mesh moves
with theta: frames generated from theta (LoadFunctional), or an articulated
mesh whose Fourier coefficients are theta (KinematicsFunctional).  forward()
runs the solver and the reverse sweep (bf_adjoint_gradient) in one go;
backward() scales the stored dJ/dtheta.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from .simulation import ArticulatedMesh, BirdFlowSimulation, SolverConfig


@dataclass
class LoadFunctional:
    """A load functional of mesh parameters.

    config: the solver configuration; frames(theta) -> (MeshSequence,
    tangents [K, frames, vertices, 3]); load_weights [steps, 6]; the solver
    starts from equilibrium on the first frame's geometry, optionally runs
    `warmup` untracked steps, then the `steps` of the objective."""

    config: SolverConfig
    frames: Callable
    load_weights: np.ndarray
    warmup: int = 0
    checkpoint_interval: int = 100
    last: dict = field(default_factory=dict)

    @property
    def steps(self) -> int:
        return int(np.asarray(self.load_weights).shape[0])

    def evaluate(self, theta: np.ndarray):
        seq, tangents = self.frames(np.asarray(theta, np.float64))
        with BirdFlowSimulation(self.config) as sim:
            sim.upload_mesh(seq.points, seq.times, seq.triangles, seq.parts, seq.half_thickness)
            sim.adjoint_set_tangents(np.asarray(tangents, np.float32))
            sim.init_equilibrium()
            if self.warmup:
                sim.advance(self.warmup)
            J, grad, timings = sim.adjoint_gradient(self.steps, self.load_weights, self.checkpoint_interval)
        self.last = {"objective": J, "gradient": grad, **timings}
        return J, grad


@dataclass
class KinematicsFunctional:
    """A load functional of an articulated mesh's Fourier coefficients.

    mesh: the articulated mesh; parameters: KinematicParameters, theta_k the
    offset of parameter k's coefficient in every joint it names (theta = 0:
    the mesh as given, ArticulatedMesh.perturbed); the solver starts
    from equilibrium at time_start, optionally runs `warmup` untracked
    steps, then the `steps` of the objective.  Optional `body` state places
    the mesh (with move_mesh body parameters)."""

    config: SolverConfig
    mesh: ArticulatedMesh
    parameters: Sequence
    load_weights: np.ndarray
    time_start: float = 0.0
    warmup: int = 0
    checkpoint_interval: int = 100
    body: Optional[tuple] = None
    last: dict = field(default_factory=dict)

    @property
    def steps(self) -> int:
        return int(np.asarray(self.load_weights).shape[0])

    def evaluate(self, theta: np.ndarray):
        mesh = self.mesh.perturbed(self.parameters, np.asarray(theta, np.float64))
        with BirdFlowSimulation(self.config) as sim:
            if self.body is not None:
                sim.set_body(*self.body)
            sim.upload_articulated(mesh, self.time_start)
            sim.adjoint_set_kinematic_parameters(self.parameters)
            sim.init_equilibrium()
            if self.warmup:
                sim.advance(self.warmup)
            J, grad, timings = sim.adjoint_gradient(self.steps, self.load_weights, self.checkpoint_interval)
        self.last = {"objective": J, "gradient": grad, **timings}
        return J, grad


class _LoadFunctionalFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, functional):
        J, grad = functional.evaluate(theta.detach().cpu().numpy())
        ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype, device=theta.device))
        return theta.new_tensor(J)

    @staticmethod
    def backward(ctx, grad_output):
        (grad,) = ctx.saved_tensors
        return grad_output * grad, None


def load_functional(theta: torch.Tensor, functional) -> torch.Tensor:
    """J(theta) as a differentiable scalar (gradient by the hand-written adjoint)."""
    return _LoadFunctionalFn.apply(theta, functional)
