"""Evaluate the reference's potential methods (by class name and constructor arguments, with
the reference's argument conventions) on the oracle or on the device pair kernels."""
import math

import numpy as np

POT_CODE = {"HardSpherePotential": 0, "InversePowerPotential": 1, "LennardJonesPotential": 2,
            "MergedImageCoulombPotential": 4, "InversePowerCoulombBoundingPotential": 5,
            "DisplacedEvenPowerPotential": 6, "HardDipolePotential": 8}


def ctor_params(pot):
    """Constructor arguments -> the C ABI's par[4] (in the order the header documents)."""
    cls, a, k = pot["cls"], pot["args"], pot["kwargs"]

    def get(name, pos, default=None):
        if name in k:
            return k[name]
        if len(a) > pos:
            return a[pos]
        return default
    if cls == "HardSpherePotential":
        return [get("radius", 0), 0, 0, 0]
    if cls == "HardDipolePotential":
        return [get("minimum_separation", 0), get("maximum_separation", 1), 0, 0]
    if cls == "InversePowerPotential":
        return [get("power", 0), get("prefactor", 1), 0, 0]
    if cls == "LennardJonesPotential":
        return [get("prefactor", 0, 1.0), get("characteristic_length", 1, 1.0), 0, 0]
    if cls == "DisplacedEvenPowerPotential":
        return [get("equilibrium_separation", 0), get("power", 1), get("prefactor", 2, 1.0), 0]
    if cls == "InversePowerCoulombBoundingPotential":
        return [get("prefactor", 0, 1.5837), 0, 0, 0]
    if cls == "MergedImageCoulombPotential":
        return [get("alpha", 0, 3.45), get("fourier_cutoff", 1, 6), get("position_cutoff", 2, 2), get("prefactor", 3, 1.0)]
    raise ValueError(cls)


def axis_velocity(v):
    """(direction, speed) of a velocity with one positive component, else None."""
    nz = [i for i, c in enumerate(v) if c != 0.0]
    if len(nz) != 1 or v[nz[0]] <= 0.0:
        return None
    return nz[0], float(v[nz[0]])


def split_args(vec):
    """(velocity, separation, c1, c2, potential_change) from the call arguments."""
    args = list(vec["args"])
    kw = vec.get("kwargs", {})
    velocity, separation = args[0], args[1]
    rest = args[2:]
    cls = vec["potential"]["cls"]
    charged = cls in ("InversePowerPotential", "InversePowerCoulombBoundingPotential", "MergedImageCoulombPotential")
    c1 = c2 = 1.0
    if charged:
        c1, c2 = rest[0], rest[1]
        rest = rest[2:]
    dU = kw.get("potential_change", rest[0] if rest else 0.0)
    return velocity, separation, c1, c2, dU


def expected_value(vec):
    e = vec["expected"]
    if e == "inf":
        return math.inf
    if e == "-inf":
        return -math.inf
    return float(e)


def matches(got, vec):
    exp = expected_value(vec)
    if vec["exact"] or math.isinf(exp):
        return got == exp
    places = vec["places"] if vec["places"] is not None else 7
    return round(abs(got - exp), places) == 0


def oracle_eval(O, vec, math_mode=0):
    """The reference method evaluated with the oracle's restated functions."""
    cls, meth = vec["potential"]["cls"], vec["method"]
    par = ctor_params(vec["potential"])
    v, s, c1, c2, dU = split_args(vec)
    L = O.lib()
    sv, vv = O.vec(s), O.vec(v)
    dim = len(s)
    pd = O._dp
    if cls == "HardSpherePotential":
        return L.ora_hs_displacement(dim, pd(vv), pd(sv), par[0])
    if cls == "HardDipolePotential":
        return L.ora_hd_displacement(dim, pd(vv), pd(sv), par[0], par[1])
    ax = axis_velocity(v)
    d, speed = ax
    if cls == "InversePowerPotential":
        if meth == "derivative":
            return L.ora_ip_derivative(dim, d, speed, pd(sv), c1, c2, par[0], par[1], math_mode)
        return L.ora_ip_displacement(dim, d, speed, pd(sv), c1, c2, par[0], par[1], dU, math_mode)
    if cls == "LennardJonesPotential":
        if meth == "derivative":
            return L.ora_lj_derivative(dim, d, speed, pd(sv), par[0], par[1], math_mode)
        return L.ora_lj_displacement(dim, d, speed, pd(sv), par[0], par[1], dU, math_mode)
    if cls == "DisplacedEvenPowerPotential":
        if meth == "derivative":
            return L.ora_dep_derivative(dim, d, speed, pd(sv), par[0], int(par[1]), par[2], math_mode)
        return L.ora_dep_displacement(dim, d, speed, pd(sv), par[0], int(par[1]), par[2], dU, math_mode)
    p = [s[(d + i) % 3] for i in range(3)]   # vectors.permutation_3d
    if cls == "InversePowerCoulombBoundingPotential":
        pp = par[0] * c1 * c2
        if meth == "derivative":
            return L.ora_ipcb_derivative(pp, *p) * speed
        return L.ora_ipcb_displacement(pp, p[0], p[1], p[2], dU, vec["system_length"]) / speed
    if cls == "MergedImageCoulombPotential":
        h = L.ora_ewald_create(int(par[1]), int(par[2]), par[0], vec["system_length"])
        try:
            return par[3] * c1 * c2 * L.ora_ewald_derivative(h, *p) * speed
        finally:
            L.ora_ewald_destroy(h)
    raise ValueError(cls)


def device_eval(vectors):
    """Evaluate a list of vectors (all with axis velocities) with ecmc_pair_* on the GPU,
    one launch per (potential, constructor, method, dimension, system length) group."""
    import torch
    from jellyfysh_amd import _lib as DL
    out = [None] * len(vectors)
    groups = {}
    for i, vec in enumerate(vectors):
        key = (vec["potential"]["cls"], tuple(ctor_params(vec["potential"])), vec["method"], len(vec["args"][1]),
               vec["system_length"] or 1.0)
        groups.setdefault(key, []).append(i)
    for (cls, par, meth, dim, L), idx in groups.items():
        rows = [split_args(vectors[i]) for i in idx]
        dirs = torch.tensor([axis_velocity(r[0])[0] for r in rows], dtype=torch.int32, device="cuda")
        speed = torch.tensor([axis_velocity(r[0])[1] for r in rows], dtype=torch.float64, device="cuda")
        sep = torch.tensor([list(r[1]) for r in rows], dtype=torch.float64, device="cuda").contiguous()
        c1 = torch.tensor([r[2] for r in rows], dtype=torch.float64, device="cuda")
        c2 = torch.tensor([r[3] for r in rows], dtype=torch.float64, device="cuda")
        dU = torch.tensor([r[4] for r in rows], dtype=torch.float64, device="cuda")
        res = torch.empty(len(rows), dtype=torch.float64, device="cuda")
        p = (np.ctypeslib.as_ctypes(np.array(par, dtype=np.float64)))
        import ctypes
        pptr = ctypes.cast(p, ctypes.POINTER(ctypes.c_double))
        if meth == "derivative":
            DL.check(DL.lib().ecmc_pair_derivative(POT_CODE[cls], pptr, dim, len(rows), DL.dptr(dirs), DL.dptr(speed),
                                                   DL.dptr(sep), DL.dptr(c1), DL.dptr(c2), L, DL.dptr(res), None))
        else:
            DL.check(DL.lib().ecmc_pair_displacement(POT_CODE[cls], pptr, dim, len(rows), DL.dptr(dirs),
                                                     DL.dptr(speed), DL.dptr(sep), DL.dptr(c1), DL.dptr(c2),
                                                     DL.dptr(dU), L, DL.dptr(res), None))
        torch.cuda.synchronize()
        vals = res.cpu().numpy()
        for j, i in enumerate(idx):
            out[i] = float(vals[j])
    return out
