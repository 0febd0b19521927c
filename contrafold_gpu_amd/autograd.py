"""Differentiable logZ: the outside pass as the hand-written adjoint.

``logz(theta, seqs)`` returns per-sequence FP32 logZ as a torch tensor that
is differentiable with respect to ``theta``, the raw CONTRAfold weight vector
(708 entries in the order of the parameter file).  Forward runs the inside
pass on the GPU; backward reuses the resident inside tables, runs the outside
pass and the expected-count kernel with the upstream gradient as per-sequence
weights, and maps the table gradient to the raw weights through the (linear)
raw -> table map: d logZ / d theta = J^T E[feature counts].

PyTorch is only the autograd plumbing here; all arithmetic is in the HIP
library.  Import this module (or torch) before the engine first runs: torch
brings its own HIP runtime, and loading it after the engine's would put two
runtimes in one process (``_lib.check_single_runtime`` raises then).
"""
import numpy as np
import torch

from . import params as _params
from ._lib import CF_GRAD, CF_LOGZ, check_single_runtime

check_single_runtime("contrafold_gpu_amd.autograd")


class LogZFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, seqs, masks, engine, pmap):
        raw = theta.detach().to("cpu", torch.float64).numpy()
        tables = pmap.tables(raw)
        engine.set_params(tables)
        batch = engine.batch(seqs, masks)
        batch.run(CF_LOGZ)
        z = batch.logz()
        ctx.batch = batch
        ctx.tables = tables
        ctx.engine = engine
        ctx.pmap = pmap
        return torch.from_numpy(z.astype(np.float64)).to(theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        w = grad_out.detach().to("cpu", torch.float32).numpy()
        ctx.engine.set_params(ctx.tables)  # no-op unless another call changed them
        ctx.batch.run(CF_GRAD, weights=w)
        tg = ctx.batch.grad()
        ctx.batch.close()
        rg = ctx.pmap.raw_grad(tg)
        return torch.from_numpy(rg).to(grad_out.dtype), None, None, None, None


def logz(theta, seqs, masks=None, engine=None, params_path=None):
    """Per-sequence logZ (float64 tensor on the CPU), differentiable in theta."""
    if engine is None:
        from . import default_engine
        engine = default_engine()
    pmap = _params.param_map(params_path)
    if theta.numel() != pmap.n_raw:
        raise ValueError(f"theta has {theta.numel()} entries, the parameter file has {pmap.n_raw}")
    return LogZFunction.apply(theta, list(seqs), masks, engine, pmap)


def raw_theta(params_path=None, dtype=torch.float64):
    """The bundled raw weights as a leaf tensor that requires grad."""
    _names, values = _params.read_raw(params_path)
    return torch.tensor(values, dtype=dtype, requires_grad=True)
