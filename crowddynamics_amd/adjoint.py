"""Differentiable rollout: trajectories of the fused step as a
``torch.autograd.Function`` whose backward pass is the hand-written adjoint
(cd_adjoint_step), with timestep checkpointing.

Scope: circular agents in fp64 with every node of the standard tree
(crowddynamics/examples/simulations.py:123-136): Reset, Integrator at a
fixed or adaptive dt, Adjusting, Navigation (the direction looked up again
at each step's positions), AgentAgentInteractions,
AgentObstacleInteractions and the (parameter-free) Fluctuation.
Differentiable inputs: tau_0, mu, kappa, damping (one value for all agents,
or one per agent) and the per-agent target_velocity.

The forward pass stores the state (position, velocity, force_prev) every
``checkpoint`` steps; the backward pass re-runs each segment from its
checkpoint (bit-identical to the first run), keeps the segment's states and
sweeps the adjoint back through it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .device import Workspace, stream_handle
from .exceptions import InvalidValue
from .stepper import Stepper

ROLLOUT_NODES = _lib.CD_USE_AGENT_AGENT | _lib.CD_USE_ADJUST | _lib.CD_USE_RESET


class Rollout:
    """Fixed-dt rollout of `steps` steps from the records `array`."""

    def __init__(self, array, steps, dt=0.005, cell_size=3.6, checkpoint=25, seed=1,
                 fluctuation=False, cell_capacity=None, bounds=None, device=None, lanes=0,
                 order_every=0, segments=None, nav=None, dt_min=None, dt_max=None):
        """dt: a fixed step; dt_min < dt_max instead select the adaptive
        step (core/integrator.py:29-97).  segments ([W, 4]) add the wall
        contacts, nav (as Stepper) the navigation lookup."""
        if 'position_ls' in array.dtype.names:
            raise InvalidValue('the adjoint covers circular agents')
        if steps < 1 or checkpoint < 1:
            raise InvalidValue('steps and checkpoint must be positive')
        flags = ROLLOUT_NODES | (_lib.CD_USE_FLUCT if fluctuation else 0)
        if segments is not None and len(segments):
            flags |= _lib.CD_USE_OBSTACLE
        if nav is not None:
            flags |= _lib.CD_USE_NAV
        self.array0 = array.copy()
        self.steps = int(steps)
        self.checkpoint = int(checkpoint)
        self.dt_min = float(dt if dt_min is None else dt_min)
        self.dt_max = float(dt if dt_max is None else dt_max)
        if not 0.0 < self.dt_min <= self.dt_max:
            raise InvalidValue('need 0 < dt_min <= dt_max')
        self.dt = self.dt_max
        self.adaptive = self.dt_min < self.dt_max
        self.st = Stepper(array.copy(), segments=segments, nav=nav, cell_size=cell_size,
                          dt_min=self.dt_min, dt_max=self.dt_max, seed=seed,
                          flags=flags, precision='f64', cell_capacity=cell_capacity,
                          bounds=bounds, device=device, lanes=lanes, order_every=order_every)
        self.agents = self.st.agents
        self.n = self.agents.n
        dev = self.agents.device
        f64 = torch.float64
        n_ck = (self.steps + self.checkpoint - 1) // self.checkpoint
        self.ck = {f: torch.zeros((n_ck, self.n, 2), dtype=f64, device=dev)
                   for f in ('position', 'velocity', 'force_prev')}
        L = self.checkpoint + 1
        self.tape_x = torch.zeros((L, self.n, 2), dtype=f64, device=dev)
        self.tape_v = torch.zeros((L, self.n, 2), dtype=f64, device=dev)
        self.adj = {k: torch.zeros((self.n, 2), dtype=f64, device=dev)
                    for k in ('lam_x', 'lam_v', 'lam_g', 'lam_f', 'lam_vt')}
        self.grad = torch.zeros((self.n, 5), dtype=f64, device=dev)
        a = _lib.CDAdjoint()
        for k, t in self.adj.items():
            setattr(a, k, C.c_void_p(t.data_ptr()))
        a.grad = C.c_void_p(self.grad.data_ptr())
        self.adj_struct = a

    # -- parameters and state -------------------------------------------
    def set_parameters(self, tau_0=None, mu=None, kappa=None, damping=None, v0=None):
        f = self.agents.fields
        for name, val in (('tau_0', tau_0), ('mu', mu), ('kappa', kappa),
                          ('damping', damping), ('target_velocity', v0)):
            if val is None:
                continue
            t = torch.as_tensor(val, dtype=torch.float64, device=self.agents.device).detach()
            if t.dim() == 0:
                f[name].fill_(t)
            elif t.shape == (self.n,):
                f[name].copy_(t)
            else:
                raise InvalidValue('%s must be a scalar or one value per agent' % name)

    def _restore_initial(self):
        self.agents.upload(self.array0, fields=('position', 'velocity', 'force', 'force_prev',
                                                'target_direction'))

    # -- passes ------------------------------------------------------------
    def forward(self):
        """Run the rollout from the initial records; returns positions
        [steps + 1, n, 2] (fp64, on the device) and stores checkpoints."""
        self._restore_initial()
        self.st.prepare(0)
        X = torch.empty((self.steps + 1, self.n, 2), dtype=torch.float64,
                        device=self.agents.device)
        ck = self.ck
        _lib.check(_lib.lib().cd_rollout_record(
            self.agents.ref(), C.byref(self.st.params), self.st.ws.ref(), self.steps,
            self.checkpoint, C.c_void_p(X.data_ptr()), C.c_void_p(ck['position'].data_ptr()),
            C.c_void_p(ck['velocity'].data_ptr()), C.c_void_p(ck['force_prev'].data_ptr()),
            stream_handle()), 'rollout_record')
        self.st.check()
        return X

    def backward(self, grad_X):
        """Adjoint sweep for dL/dX = grad_X [steps + 1, n, 2]; returns the
        per-agent parameter gradients [n, 5] (tau_0, mu, kappa, damping,
        target_velocity) and dL/d(position_0, velocity_0)."""
        grad_X = grad_X.to(torch.float64).contiguous()
        for t in self.adj.values():
            t.zero_()
        self.grad.zero_()
        self.adj['lam_x'].copy_(grad_X[self.steps])
        L = _lib.lib()
        s_h = stream_handle()
        n_seg = (self.steps + self.checkpoint - 1) // self.checkpoint
        ck = self.ck
        for s in reversed(range(n_seg)):
            k0 = s * self.checkpoint
            k1 = min(k0 + self.checkpoint, self.steps)
            _lib.check(L.cd_adjoint_segment(
                self.agents.ref(), C.byref(self.st.params), self.st.ws.ref(), k0, k1,
                C.c_void_p(ck['position'][s].data_ptr()), C.c_void_p(ck['velocity'][s].data_ptr()),
                C.c_void_p(ck['force_prev'][s].data_ptr()), C.c_void_p(self.tape_x.data_ptr()),
                C.c_void_p(self.tape_v.data_ptr()), C.c_void_p(grad_X.data_ptr()),
                C.byref(self.adj_struct), s_h), 'adjoint_segment')
        self.st.check()
        return self.grad.clone(), self.adj['lam_x'].clone(), self.adj['lam_v'].clone()

    def tangent(self, dp, dx0=None, dv0=None):
        """Forward-mode derivative of the trajectory: dX [steps + 1, n, 2],
        the derivative of the positions along the parameter tangent dp
        ([n, 5]: tau_0, mu, kappa, damping, target_velocity of each agent)
        and the initial tangents dx0, dv0 ([n, 2], default zero) of
        position and velocity.  Runs the rollout from the initial records,
        one cd_tangent_step (from state k, binned in a workspace of its
        own) before each fused step k -> k + 1."""
        dev = self.agents.device
        f64 = torch.float64
        n = self.n
        T = {k: torch.zeros((n, 2), dtype=f64, device=dev) for k in ('dx', 'dv', 'dg', 'df')}
        if dx0 is not None:
            T['dx'].copy_(torch.as_tensor(dx0, dtype=f64, device=dev))
        if dv0 is not None:
            T['dv'].copy_(torch.as_tensor(dv0, dtype=f64, device=dev))
        dp = torch.as_tensor(dp, dtype=f64, device=dev).contiguous()
        if dp.shape != (n, 5):
            raise InvalidValue('dp must be [n, 5]')
        tan = _lib.CDTangent()
        for k, t in T.items():
            setattr(tan, k, C.c_void_p(t.data_ptr()))
        tan.dp = C.c_void_p(dp.data_ptr())
        if getattr(self, 'tan_ws', None) is None:
            w = self.st.ws
            self.tan_ws = Workspace(n, w.cell_capacity, 0, cell_slots=w.cell_slots,
                                    device=dev, lanes=w.desc.agent_lanes)
        dX = torch.empty((self.steps + 1, n, 2), dtype=f64, device=dev)
        dX[0] = T['dx']
        self._restore_initial()
        self.st.prepare(0)
        L = _lib.lib()
        s_h = stream_handle()
        par = C.byref(self.st.params)
        # the tangent step of k -> k + 1 runs after the fused step, from a
        # copy of state k, with the agents' new state as state k + 1
        xk = torch.empty((n, 2), dtype=f64, device=dev)
        vk = torch.empty((n, 2), dtype=f64, device=dev)
        view = _lib.CDAgents.from_buffer_copy(self.agents.struct)
        view.position = C.c_void_p(xk.data_ptr())
        view.velocity = C.c_void_p(vk.data_ptr())
        tan.x_next = self.agents.struct.position
        tan.v_next = self.agents.struct.velocity
        f = self.agents.fields
        for k in range(self.steps):
            xk.copy_(f['position'])
            vk.copy_(f['velocity'])
            _lib.check(L.cd_step(self.agents.ref(), par, self.st.ws.ref(), s_h), 'step')
            _lib.check(L.cd_tangent_step(C.byref(view), par, C.byref(tan),
                                         self.tan_ws.ref(), s_h), 'tangent_step')
            dX[k + 1] = T['dx']
        self.st.check()
        if self.tan_ws.state().status & ~(_lib.CD_STATUS_SLOTS_FULL | _lib.CD_STATUS_NAV_KEPT):
            raise InvalidValue('tangent binning: status %d' % self.tan_ws.state().status)
        return dX

    def adjoint_step(self, grad_x=None, x_next=None, v_next=None):
        """One cd_adjoint_step from the agents' current position and
        velocity, updating the adjoint buffers ``adj`` and ``grad`` in place
        (x_next, v_next: the state after the step, for adaptive dt)."""
        gx = C.c_void_p(grad_x.data_ptr()) if grad_x is not None else C.c_void_p(0)
        a = _lib.CDAdjoint.from_buffer_copy(self.adj_struct)
        a.x_next = C.c_void_p(x_next.data_ptr()) if x_next is not None else None
        a.v_next = C.c_void_p(v_next.data_ptr()) if v_next is not None else None
        _lib.check(_lib.lib().cd_adjoint_step(self.agents.ref(), C.byref(self.st.params), gx,
                                              C.byref(a), self.st.ws.ref(),
                                              stream_handle()), 'adjoint_step')


class RolloutFunction(torch.autograd.Function):
    """X = RolloutFunction.apply(rollout, tau_0, mu, kappa, damping, v0).

    Each parameter is a 0-d tensor (one value for all agents) or a tensor of
    one value per agent; v0 is per agent.  X holds the positions of steps
    0..steps, [steps + 1, n, 2]."""

    @staticmethod
    def forward(ctx, rollout, tau_0, mu, kappa, damping, v0):
        rollout.set_parameters(tau_0, mu, kappa, damping, v0)
        X = rollout.forward()
        ctx.rollout = rollout
        ctx.shapes = [t.shape for t in (tau_0, mu, kappa, damping, v0)]
        ctx.save_for_backward(tau_0, mu, kappa, damping, v0)
        return X

    @staticmethod
    def backward(ctx, grad_X):
        r = ctx.rollout
        r.set_parameters(*ctx.saved_tensors)
        g, _, _ = r.backward(grad_X)
        out = []
        for k, shape in enumerate(ctx.shapes):
            gk = g[:, k]
            out.append(gk.sum().reshape(shape) if len(shape) == 0 else gk.reshape(shape))
        return (None, *out)


def rollout_positions(rollout, tau_0, mu, kappa, damping, v0):
    """Positions [steps + 1, n, 2] of the rollout, differentiable in its
    parameters."""
    return RolloutFunction.apply(rollout, tau_0, mu, kappa, damping, v0)
