"""Force-matching fit of LJ (epsilon, sigma) per type and per-atom charges
through the library's adjoint (mi355nb_compute_force_matching_grad).

A batch of frames of one molecule is evaluated as one batched compute object
(one pair-kernel launch and one adjoint launch for all frames).  The loss is

    L(theta) = sum_frames sum_atoms |F(theta) - F_target|^2,

F = fast + slow nonbonded force, theta = (sigma_t, epsilon_t, q_i).  The
kernel returns dL/dA_ab, dL/dB_ab and dL/dq; the LJ mixing rule of
LJTable::compute_vdw_params (src/LJTable.C:118-135; sigma_ab = (s_a+s_b)/2,
eps_ab = sqrt(e_a e_b), B = 4 eps sigma^6, A = B sigma^6) is differentiated
here in FP64.  ForceMatchingLoss wraps it as a torch.autograd.Function.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import systems
from .compute import ComputeNonbonded


def lj_tables(sigma, eps):
    """A, B (ntypes x ntypes) by arithmetic sigma / geometric epsilon mixing, FP64."""
    s = np.asarray(sigma, np.float64)
    e = np.asarray(eps, np.float64)
    sab = 0.5 * (s[:, None] + s[None, :])
    eab = np.sqrt(e[:, None] * e[None, :])
    B = 4.0 * eab * sab ** 6
    A = B * sab ** 6
    return A, B


def lj_param_grad(sigma, eps, gA, gB):
    """dL/dsigma, dL/deps from ordered-bin dL/dA, dL/dB (bin [a][b] is the
    gradient of the table entry used for pairs with types (a, b))."""
    s = np.asarray(sigma, np.float64)
    e = np.asarray(eps, np.float64)
    A, B = lj_tables(s, e)
    sab = 0.5 * (s[:, None] + s[None, :])
    WA = np.asarray(gA) * A
    WB = np.asarray(gB) * B
    # d eps_ab / d eps_a = eps_ab / (2 eps_a); A, B are linear in eps_ab
    we = WA + WB
    deps = (we.sum(1) + we.sum(0)) / (2.0 * e)
    # d sigma_ab / d sigma_a = 1/2; dA/dsigma_ab = 12 A / sigma_ab, dB/dsigma_ab = 6 B / sigma_ab
    ws = (6.0 * WA + 3.0 * WB) / sab
    dsig = ws.sum(1) + ws.sum(0)
    return dsig, deps


def make_frames(base: systems.System, n_frames: int = 16, seed: int = 0, jitter: float = 0.05):
    """n_frames copies of base with seeded Gaussian displacements (sd `jitter` A)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_frames):
        pos = base.pos + rng.normal(0.0, jitter, base.pos.shape)
        out.append(dataclasses.replace(base, name=f"{base.name}_frame{k}", pos=np.mod(pos, base.box)))
    return out


def perturbed_parameters(base: systems.System, seed: int = 1, rel: float = 0.05):
    """theta * (1 + rel * N(0, 1)), seeded: (sigma, eps, q)."""
    rng = np.random.default_rng(seed)
    sig = base.lj_sigma * (1.0 + rel * rng.standard_normal(base.lj_sigma.shape))
    eps = base.lj_eps * (1.0 + rel * rng.standard_normal(base.lj_eps.shape))
    q = base.charge * (1.0 + rel * rng.standard_normal(base.charge.shape))
    return sig, eps, q


class ForceMatching:
    """Frames of one molecule on one GPU; theta -> (loss, gradient)."""

    def __init__(self, frames, stream=None, device=0):
        self.frames = frames
        self.nmol = frames[0].natoms
        self.nframes = len(frames)
        self.compute = ComputeNonbonded.from_systems(frames, device=device, stream=stream)
        self.compute.launch_work(do_energy=False, do_virial=False, do_slow=True, do_pairlist=True)
        self.compute.finish(forces=False)
        self.target = None

    def set_parameters(self, sigma, eps, q):
        A, B = lj_tables(sigma, eps)
        self.compute.update_parameters(np.tile(np.asarray(q, np.float64), self.nframes), A, B)

    def forces(self, sigma, eps, q):
        """Total (fast + slow) force per frame, (nframes, natoms, 3)."""
        self.set_parameters(sigma, eps, q)
        self.compute.launch_work(do_energy=False, do_virial=False, do_slow=True, do_pairlist=False)
        _, _, f, fs = self.compute.finish_batch()
        return (f + fs).reshape(self.nframes, self.nmol, 3)

    def set_target(self, forces):
        self.target = np.ascontiguousarray(forces, np.float64).reshape(-1, 3)

    def loss_and_grad(self, sigma, eps, q):
        """(loss, dL/dsigma, dL/deps, dL/dq) in FP64."""
        if self.target is None:
            raise RuntimeError("set_target first")
        self.set_parameters(sigma, eps, q)
        loss, gA, gB, gq = self.compute.force_matching_grad(self.target)
        dsig, deps = lj_param_grad(sigma, eps, gA, gB)
        dq = gq.reshape(self.nframes, self.nmol).sum(0)
        return loss, dsig, deps, dq

    def close(self):
        self.compute.close()


def force_matching_loss():
    """torch.autograd.Function class: loss = apply(sigma, eps, q, fm) with FP64
    CPU tensors for the parameters; backward returns the adjoint gradient."""
    import torch

    class ForceMatchingLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, sigma, eps, q, fm):
            loss, ds, de, dq = fm.loss_and_grad(sigma.detach().cpu().numpy(), eps.detach().cpu().numpy(),
                                                q.detach().cpu().numpy())
            ctx.save_for_backward(torch.from_numpy(ds).to(sigma), torch.from_numpy(de).to(eps),
                                  torch.from_numpy(dq).to(q))
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            ds, de, dq = ctx.saved_tensors
            return gout * ds, gout * de, gout * dq, None

    return ForceMatchingLoss
