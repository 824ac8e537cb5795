"""Pair energies with hand-written backward kernels behind torch.autograd.Function.

``pair_energy(positions, params, potential, system_length, charges=None, molecule=None,
scope="all")`` returns the total pair energy as a 0-d FP64 tensor.  Its backward runs
ecmc_energy_backward, which gives

  * dU/dx: for unit i along axis d, the sum over partners j of the potential's own
    directional derivative (the quantity ECMC consumes, potential/potential.py:153-181);
  * dU/dtheta for the potential's parameters (params, laid out like the C ABI's par[4]);
  * dU/dc_i for charged terms,

so that torch.autograd delivers dU/dtheta for parameter-sensitivity estimators
<O dU/dtheta>.  ``sensitivity`` turns samples of O and dU/dtheta into d<O>/dtheta.

Potentials (names as in chain.potential_code): inverse_power (power, prefactor),
inverse_power_truncated (power, prefactor, cutoff), lennard_jones (prefactor,
characteristic_length), displaced_even_power (equilibrium_separation, power, prefactor),
merged_image_coulomb (alpha, fourier_cutoff, position_cutoff, prefactor; only the prefactor
gets a gradient).
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from .chain import potential_code

_SCOPES = {"all": L.SCOPE_ALL, "inter": L.SCOPE_INTER, "intra": L.SCOPE_INTRA}
_CHARGED = {"inverse_power", "inverse_power_truncated", "merged_image_coulomb"}


def _term(potential, params, scope, use_charge):
    t = L.EnergyTerm()
    t.potential = potential_code(potential)
    t.scope = _SCOPES[scope]
    t.use_charge = 1 if use_charge else 0
    vals = [float(v) for v in params]
    if len(vals) > 4:
        raise ValueError("at most four parameters")
    for i, v in enumerate(vals):
        t.par[i] = v
    return t


def _workspace(term, n, device):
    b = ctypes.c_int64()
    L.check(L.lib().ecmc_energy_workspace_bytes(ctypes.byref(term), n, ctypes.byref(b)))
    return torch.empty(max(1, (b.value + 7) // 8), dtype=torch.float64, device=device)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class _PairEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, params, charges, molecule, potential, system_length, scope):
        L.require_device()
        ctx.n_params = params.numel()
        ctx.par_device = params.device
        if positions.device.type != "cuda" or positions.dtype != torch.float64 or positions.dim() != 2 \
                or positions.shape[1] != 3:
            raise ValueError("positions must be an (n, 3) float64 device tensor")
        use_charge = charges is not None
        if use_charge and potential not in _CHARGED:
            raise ValueError("%s takes no charges" % potential)
        pos = positions.detach().contiguous()
        chg = charges.detach().to(torch.float64).contiguous() if use_charge else None
        mol = molecule.to(torch.int32).contiguous() if molecule is not None else None
        term = _term(potential, params.detach().cpu().tolist(), scope, use_charge)
        n = pos.shape[0]
        work = _workspace(term, n, pos.device)
        out = torch.empty((), dtype=torch.float64, device=pos.device)
        L.check(L.lib().ecmc_energy_forward(ctypes.byref(term), n, L.dptr(pos), L.dptr(chg), L.dptr(mol),
                                            float(system_length), L.dptr(out), L.dptr(work), _stream()))
        ctx.save_for_backward(pos, chg if chg is not None else torch.empty(0), mol if mol is not None else
                              torch.empty(0, dtype=torch.int32))
        ctx.term, ctx.n, ctx.L, ctx.use_charge, ctx.has_mol = term, n, float(system_length), use_charge, \
            mol is not None
        return out

    @staticmethod
    def backward(ctx, grad_out):
        pos, chg, mol = ctx.saved_tensors
        chg = chg if ctx.use_charge else None
        mol = mol if ctx.has_mol else None
        n = ctx.n
        work = _workspace(ctx.term, n, pos.device)
        gpos = torch.empty((n, 3), dtype=torch.float64, device=pos.device)
        gpar = torch.empty(4, dtype=torch.float64, device=pos.device)
        gchg = torch.empty(n, dtype=torch.float64, device=pos.device) if ctx.use_charge else None
        L.check(L.lib().ecmc_energy_backward(ctypes.byref(ctx.term), n, L.dptr(pos), L.dptr(chg), L.dptr(mol), ctx.L,
                                             L.dptr(gpos), L.dptr(gpar), L.dptr(gchg), L.dptr(work), _stream()))
        g = grad_out.to(torch.float64)
        npar = ctx.needs_input_grad[1]
        return (g * gpos if ctx.needs_input_grad[0] else None,
                (g * gpar)[: ctx.n_params].to(ctx.par_device) if npar else None,
                g * gchg if (gchg is not None and ctx.needs_input_grad[2]) else None,
                None, None, None, None)


def pair_energy(positions, params, potential, system_length, charges=None, molecule=None, scope="all"):
    """Total pair energy (0-d FP64 device tensor), differentiable in positions, params and
    charges through the hand-written backward kernel."""
    if not torch.is_tensor(params):
        params = torch.tensor(params, dtype=torch.float64, device=positions.device)
    if params.numel() > 4:
        raise ValueError("at most four parameters")
    return _PairEnergy.apply(positions, params, charges, molecule, potential, system_length, scope)


def sensitivity(observable, dU_dtheta, beta):
    """d<O>/dtheta = -beta (<O dU/dtheta> - <O><dU/dtheta>) from equilibrium samples."""
    o = np.asarray(observable, dtype=np.float64)
    g = np.asarray(dU_dtheta, dtype=np.float64)
    if g.ndim == 1:
        g = g[:, None]
    return -beta * (np.mean(o[:, None] * g, axis=0) - o.mean() * g.mean(axis=0))
