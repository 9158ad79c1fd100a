"""Differentiable blob trajectories: a torch.autograd.Function over the C ABI.

`BlobTrajectory.apply(G, E, A, B, blob, pos0, nsteps, interval)` runs `nsteps` mass-less
steps of `blob` (a `GpuBlob` with calc_noise = 0 and pinned nodes) from positions `pos0`
with per-element moduli G, E and viscosities A, B (meso units, float64 CPU tensors of
shape [num_elements]) and returns the final node positions ([num_nodes, 3]). Its
backward is the library's discrete adjoint (ffea_gpu_blob_adjoint): one viscosity-
matrix CG solve and one element reverse pass per step, segments of `interval` steps
recomputed from position checkpoints. The step reversed is the reference's
Blob::update_internal_forces + Blob::update_positions (Blob.cpp:400-508).

Tensors stay on the host: the GPU work is the library's, on its own HIP stream (torch's
own HIP runtime is never initialised here).
"""
from __future__ import annotations

import numpy as np
import torch


class BlobTrajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, G, E, A, B, blob, pos0, nsteps: int, interval: int = 20):
        for t in (G, E, A, B):
            if t.device.type != "cpu" or t.dtype != torch.float64:
                raise TypeError("G, E, A, B must be float64 CPU tensors")
        blob.set_material(A.detach().numpy(), B.detach().numpy(), G.detach().numpy(), E.detach().numpy())
        blob.set_positions(np.asarray(pos0, np.float64))
        blob.run_checkpointed(int(nsteps), int(interval))
        ctx.blob = blob
        ctx.run_generation = blob.run_generation
        return torch.from_numpy(blob.positions())

    @staticmethod
    def backward(ctx, grad_pos):
        # the library reverses the run the blob keeps, with its current material: refuse
        # when another trajectory or a material change has replaced this forward's run
        if ctx.blob.run_generation != ctx.run_generation:
            raise RuntimeError("BlobTrajectory.backward: the blob has run another trajectory or changed its "
                               "material since this forward; run each differentiated trajectory on its own blob "
                               "or call backward before the next one")
        dA, dB, dG, dE, _ = ctx.blob.adjoint(grad_pos.detach().contiguous().numpy())
        t = torch.from_numpy
        return t(dG), t(dE), t(dA), t(dB), None, None, None, None


def trajectory(blob, G, E, A, B, pos0, nsteps, interval=20):
    """final positions of `blob` after nsteps steps, differentiable in G, E, A, B"""
    return BlobTrajectory.apply(G, E, A, B, blob, pos0, nsteps, interval)
