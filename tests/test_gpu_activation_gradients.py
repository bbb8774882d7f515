"""The reference's tests/test_activation_gradients.py on cuda:0 through torch
autograd: gradients of a three-substep loss of the differentiable HIP step
against central differences of the same HIP forward (as the reference checks
its own path), on the squeezed 27-node cube:

- :81-120 per-step tet activations, free and on the floor (ground penalty,
  damping and friction), along a random direction and the largest
  component, with every perturbed trajectory on the same contact branch;
- :122-145 the weights of the reference bear's phase network (eight phases,
  one tanh layer) driving the activations on the floor;
- :169-181 the default stretch clamp bounds (0.05, 20) give a finite state
  derivative that central differences confirm in the clamp's smooth region.
"""

import math

import numpy as np
import pytest
import torch

from helpers import cube
from newton_solver_fem_amd.model import Control

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DT = 0.002


def make_solver(model, **options):
    from newton_solver_fem_amd.solver import SolverFEM

    args = dict(dt=DT, differentiable=True, linear_solver_maxiter=128, solver_type="cg",
                preconditioner_type="block_diag", self_contact_ke=0.0, min_stretch=-1.0)
    args.update(options)
    return SolverFEM(model, **args)


def scene(*, contact=False, clamp=False):
    model = cube("tet", gravity=(0, 0, -9.81) if contact else (0, 0, 0), device=DEV)
    options = {"enable_tet_activations": True}
    if contact:
        options.update(ground_plane=(0, 0, 1, 0), ground_ke=10.0, ground_kd=0.02, ground_kf=0.01, ground_mu=0.5)
    if clamp:
        options.update(min_stretch=0.05, max_stretch=20.0)
    solver = make_solver(model, **options)
    q = model.host["particle_q"].astype(np.float64)
    origin = q.min(axis=0)
    q = origin + (q - origin) * [1.03, 0.97, 1.01]
    if contact:
        q[:, 2] -= 0.298
    v0 = (0.4, -0.3, -0.3) if contact else (0.04, -0.03, -0.04)
    velocity = np.tile(v0, (len(q), 1)).astype(np.float32)
    return model, solver, q.astype(np.float32), velocity


def objective(q, v):
    """sum_i q_i . ((i + 1) / P, 0.13, -0.41) + 0.03 |v|^2, in fp64."""
    n = q.shape[0]
    w = torch.stack([(torch.arange(n, device=q.device, dtype=torch.float64) + 1.0) / n,
                     torch.full((n,), 0.13, device=q.device, dtype=torch.float64),
                     torch.full((n,), -0.41, device=q.device, dtype=torch.float64)], 1)
    return (q.double() * w).sum() + 0.03 * (v.double() ** 2).sum()


def evaluate(model, solver, q, v, activations):
    """Three differentiable substeps from (q, v) with per-step activations;
    returns the loss and the trajectory's states."""
    from newton_solver_fem_amd.model import State

    st = State(torch.tensor(q, device=DEV), torch.tensor(v, device=DEV), None)
    states = [st]
    for k in range(len(activations)):
        out = State(torch.empty_like(st.particle_q), torch.empty_like(st.particle_q), None)
        solver.step(st, out, Control(activations[k]), None, DT)
        st = out
        states.append(st)
    return objective(st.particle_q, st.particle_qd), states


def contact_signature(model, states, baseline, reference=None):
    """assert_contact_branch (reference :64-79): the same nodes below their
    radius in every state, sliding down, and the same normal / friction
    regime per node."""
    radius = model.host["particle_radius"]
    mass = model.host["particle_mass"]
    signatures = []
    for k, st in enumerate(states):
        q = st.particle_q.detach().cpu().numpy()
        v = st.particle_qd.detach().cpu().numpy()
        np.testing.assert_array_equal(q[:, 2] < radius, baseline)
        assert np.all(v[baseline, 2] < -0.01)
        assert np.all(np.linalg.norm(v[baseline, :2], axis=1) > 0.01)
        normal = 10.0 * (radius - q[:, 2]) - 0.02 * v[:, 2]
        weight = mass * 9.81
        speed = np.linalg.norm(v[:, :2], axis=1)
        sig = np.stack((normal > weight, 0.01 * speed > 0.5 * np.minimum(normal, weight)))[:, baseline]
        signatures.append(sig)
        if reference is not None:
            np.testing.assert_array_equal(sig, reference[k])
    return signatures


@pytest.mark.parametrize("contact", [False, True])
def test_tet_activation_direction_and_component_derivatives(contact):
    print(f"case activation gradients contact={contact}", flush=True)
    model, solver, q, velocity = scene(contact=contact)
    activation = np.random.default_rng(11).uniform(-0.02, 0.02, (3, model.tet_count)).astype(np.float32)
    act = [torch.tensor(a, device=DEV, requires_grad=True) for a in activation]
    loss, states = evaluate(model, solver, q, velocity, act)
    loss.backward()
    derivative = np.array([a.grad.cpu().numpy() for a in act])
    assert np.isfinite(derivative).all()
    assert np.linalg.norm(derivative) > 1e-5
    direction = np.random.default_rng(12).normal(size=activation.shape).astype(np.float32)
    direction /= np.linalg.norm(direction)
    component = np.zeros_like(direction)
    component.flat[np.argmax(np.abs(derivative))] = 1.0
    branch = q[:, 2] < model.host["particle_radius"]
    branches = contact_signature(model, states, branch) if contact else None
    for delta in (direction, component):
        exact = float(np.sum(derivative * delta))
        for eps in (0.004, 0.002):
            losses = []
            for sign in (-1, 1):
                pert = [torch.tensor(a, device=DEV) for a in activation + sign * eps * delta]
                with torch.no_grad():
                    value, traj = evaluate(model, solver, q, velocity, pert)
                losses.append(float(value))
                if contact:
                    contact_signature(model, traj, branch, branches)
            np.testing.assert_allclose(exact, (losses[1] - losses[0]) / (2 * eps), rtol=0.015, atol=2e-5)


def phase_network(weights, time):
    """The reference bear's eight phases and one tanh layer (test :11-18)."""
    phases = torch.tensor([math.sin(5.0 * time + p * (2.0 * math.pi / 8.0)) for p in range(8)], device=DEV)
    return 0.3 * torch.tanh(weights @ phases)


def network_rollout(model, solver, q, velocity, weights):
    act = [phase_network(weights, i * DT) for i in range(3)]
    return evaluate(model, solver, q, velocity, act)


def test_network_weights_multistep_gradient_with_floor_contact():
    print("case activation network gradients", flush=True)
    model, solver, q, velocity = scene(contact=True)
    values = np.random.default_rng(42).uniform(-math.sqrt(1 / 8), math.sqrt(1 / 8),
                                               (model.tet_count, 8)).astype(np.float32)
    w = torch.tensor(values, device=DEV, requires_grad=True)
    loss, states = network_rollout(model, solver, q, velocity, w)
    loss.backward()
    gradient = w.grad.cpu().numpy()
    assert np.isfinite(gradient).all()
    assert np.linalg.norm(gradient) > 1e-5
    branch = q[:, 2] < model.host["particle_radius"]
    branches = contact_signature(model, states, branch)
    direction = np.random.default_rng(6).normal(size=values.shape).astype(np.float32)
    direction /= np.linalg.norm(direction)
    component = np.zeros_like(direction)
    component.flat[np.argmax(np.abs(gradient))] = 1.0
    for delta in (direction, component):
        exact = float(np.sum(gradient * delta))
        for eps in (0.008, 0.004):
            losses = []
            for sign in (-1, 1):
                with torch.no_grad():
                    value, traj = network_rollout(model, solver, q, velocity,
                                                  torch.tensor(values + sign * eps * delta, device=DEV))
                losses.append(float(value))
                contact_signature(model, traj, branch, branches)
            np.testing.assert_allclose(exact, (losses[1] - losses[0]) / (2 * eps), rtol=0.02, atol=2e-5)


def test_default_stretch_clamp_has_finite_state_derivative_in_smooth_region():
    print("case activation clamp derivative", flush=True)
    model, solver, q, velocity = scene(clamp=True)
    zero = [torch.zeros(model.tet_count, device=DEV)] * 2
    q0 = torch.tensor(q, device=DEV, requires_grad=True)
    from newton_solver_fem_amd.model import State

    def run(qq):
        st = State(qq, torch.tensor(velocity, device=DEV), None)
        for a in zero:
            out = State(torch.empty_like(st.particle_q), torch.empty_like(st.particle_q), None)
            solver.step(st, out, Control(a), None, DT)
            st = out
        return objective(st.particle_q, st.particle_qd)

    run(q0).backward()
    gradient = q0.grad.cpu().numpy()
    assert np.isfinite(gradient).all()
    direction = np.random.default_rng(5).normal(size=q.shape).astype(np.float32)
    direction /= np.linalg.norm(direction)
    exact = float(np.sum(gradient * direction))
    for eps in (2e-4, 1e-4):
        with torch.no_grad():
            plus = float(run(torch.tensor(q + eps * direction, device=DEV)))
            minus = float(run(torch.tensor(q - eps * direction, device=DEV)))
        np.testing.assert_allclose(exact, (plus - minus) / (2 * eps), rtol=0.015, atol=1e-4)
