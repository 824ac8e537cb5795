"""Potentials with the reference's plugin interface, evaluated by the device pair kernels.

Each class keeps the name, constructor arguments and call signature of its counterpart in
the reference's jellyfysh/potential package, so an event handler written against that
interface can hold one of these instead:

    derivative(velocity, separation[, charge_one, charge_two]) -> float
    displacement(velocity, separation[, charge_one, charge_two][, potential_change]) -> float

(potential/potential.py:153-253).  ``potential_change_required``,
``number_separation_arguments`` and ``number_charge_arguments`` have the values the
reference infers from those signatures (potential/potential.py:141-151, 293-301).

Every call runs on the GPU through ecmc_pair_displacement / ecmc_pair_derivative of the C
ABI; the array forms ``displacements`` / ``derivatives`` evaluate many rows in one launch
(device tensors in, device tensor out).  Velocities must point along +x, +y or +z (the only
velocities straight-chain ECMC uses; the reference's standard-velocity potentials assert
the same, potential/abstracts.py:115-141); any other velocity raises AssertionError here,
also for the hard sphere, which the reference evaluates for any velocity.
"""
import ctypes

from . import _lib as L


def _axis(velocity):
    """(direction, speed) of a velocity with exactly one nonzero, positive component."""
    nz = [i for i, c in enumerate(velocity) if c != 0.0]
    if len(nz) != 1 or not velocity[nz[0]] > 0.0:
        raise AssertionError("the velocity must point in the positive direction of one cartesian axis")
    return nz[0], float(velocity[nz[0]])


class DevicePotential:
    """Base: parameter block, argument counts and the two device entry points."""
    code = None
    charged = False
    takes_potential_change = True

    def __init__(self, params, system_length=1.0):
        self._par = (ctypes.c_double * 4)(*([float(p) for p in params] + [0.0] * (4 - len(params))))
        self.system_length = float(system_length)

    @property
    def potential_change_required(self):
        return self.takes_potential_change

    @property
    def number_separation_arguments(self):
        return 1

    @property
    def number_charge_arguments(self):
        return 2 if self.charged else 0

    # ---- array forms -------------------------------------------------------------------
    def _launch(self, derivative, direction, speed, separation, charge_one, charge_two, potential_change):
        import torch
        dev = separation.device
        if dev.type != "cuda":
            raise ValueError("device tensors are expected")
        n, dim = separation.shape
        f64 = dict(dtype=torch.float64, device=dev)
        direction = torch.as_tensor(direction, dtype=torch.int32, device=dev).expand(n).contiguous()
        speed = torch.as_tensor(speed, **f64).expand(n).contiguous()
        sep = separation.to(torch.float64).contiguous()
        c1 = None if charge_one is None else torch.as_tensor(charge_one, **f64).expand(n).contiguous()
        c2 = None if charge_two is None else torch.as_tensor(charge_two, **f64).expand(n).contiguous()
        out = torch.empty(n, **f64)
        lib = L.lib()
        if derivative:
            L.check(lib.ecmc_pair_derivative(self.code, self._par, dim, n, L.dptr(direction), L.dptr(speed),
                                             L.dptr(sep), L.dptr(c1), L.dptr(c2), self.system_length, L.dptr(out),
                                             None))
        else:
            dU = None
            if self.takes_potential_change:
                dU = torch.as_tensor(potential_change, **f64).expand(n).contiguous()
            L.check(lib.ecmc_pair_displacement(self.code, self._par, dim, n, L.dptr(direction), L.dptr(speed),
                                               L.dptr(sep), L.dptr(c1), L.dptr(c2), L.dptr(dU), self.system_length,
                                               L.dptr(out), None))
        return out

    def displacements(self, direction, speed, separation, charge_one=None, charge_two=None, potential_change=None):
        """Rows of (direction, speed, separation[, charges][, potential change]) -> displacements
        (time until the event; +inf for none)."""
        return self._launch(False, direction, speed, separation, charge_one, charge_two, potential_change)

    def derivatives(self, direction, speed, separation, charge_one=None, charge_two=None):
        """Rows -> directional time derivatives."""
        return self._launch(True, direction, speed, separation, charge_one, charge_two, None)

    # ---- reference-style scalar calls ---------------------------------------------------
    def _scalar(self, derivative, velocity, separation, c1, c2, dU):
        import torch
        d, speed = _axis(velocity)
        sep = torch.tensor([list(map(float, separation))], dtype=torch.float64, device="cuda")
        out = self._launch(derivative, d, speed, sep, c1, c2, dU)
        return float(out.cpu()[0])


class HardSpherePotential(DevicePotential):
    """potential/hard_sphere_potential.py:33-119: U = inf for |s| < 2 radius, else 0."""
    code = L.POT_HARD_SPHERE
    takes_potential_change = False

    def __init__(self, radius):
        super().__init__([radius])

    def displacement(self, velocity, separation):
        return self._scalar(False, velocity, separation, None, None, None)

    def derivative(self, velocity, separation):
        # hard_sphere_potential.py:101-119: not defined for ECMC
        raise NotImplementedError


class HardDipolePotential(DevicePotential):
    """potential/hard_dipole_potential.py:32-125: U = 0 for minimum <= |s| <= maximum, else inf."""
    code = L.POT_HARD_DIPOLE
    takes_potential_change = False

    def __init__(self, minimum_separation, maximum_separation):
        if not minimum_separation > 0.0 or not maximum_separation > 0.0 or \
                not minimum_separation < maximum_separation:
            raise ValueError("0 < minimum_separation < maximum_separation is required")
        super().__init__([minimum_separation, maximum_separation])

    def displacement(self, velocity, separation):
        return self._scalar(False, velocity, separation, None, None, None)

    def derivative(self, velocity, separation):
        # hard_dipole_potential.py:112-125: not defined for ECMC
        raise NotImplementedError


class InversePowerPotential(DevicePotential):
    """potential/inverse_power_potential.py:32-179: U = c1 c2 k / |s|^p."""
    code = L.POT_INVERSE_POWER
    charged = True

    def __init__(self, power, prefactor):
        if not power > 0.0:
            raise ValueError("power must be positive")
        super().__init__([power, prefactor])
        self.power, self.prefactor = float(power), float(prefactor)

    def derivative(self, velocity, separation, charge_one, charge_two):
        return self._scalar(True, velocity, separation, charge_one, charge_two, None)

    def displacement(self, velocity, separation, charge_one, charge_two, potential_change):
        return self._scalar(False, velocity, separation, charge_one, charge_two, potential_change)

    def potential(self, charge_product, separation):
        """inverse_power_potential.py:127-143 (host arithmetic: a single value)."""
        return charge_product * self.prefactor / (sum(c * c for c in separation) ** (self.power / 2.0))


class TruncatedInversePowerPotential(DevicePotential):
    """Inverse power truncated and shifted at a cutoff: U = k/|s|^p - k/rc^p inside rc, 0
    outside.  No reference counterpart (the reference's inverse power has no cutoff); the
    batched Lennard-Jones workload uses it, and for rc = inf it equals InversePowerPotential."""
    code = L.POT_INVERSE_POWER_TRUNC

    def __init__(self, power, prefactor, cutoff):
        super().__init__([power, prefactor, cutoff])

    def derivative(self, velocity, separation):
        return self._scalar(True, velocity, separation, None, None, None)

    def displacement(self, velocity, separation, potential_change):
        return self._scalar(False, velocity, separation, None, None, potential_change)


class LennardJonesPotential(DevicePotential):
    """potential/lennard_jones_potential.py:31-135 (Mexican-hat inversion around 2^(1/6) sigma)."""
    code = L.POT_LENNARD_JONES

    def __init__(self, prefactor=1.0, characteristic_length=1.0):
        super().__init__([prefactor, characteristic_length])

    def derivative(self, velocity, separation):
        return self._scalar(True, velocity, separation, None, None, None)

    def displacement(self, velocity, separation, potential_change):
        return self._scalar(False, velocity, separation, None, None, potential_change)


class DisplacedEvenPowerPotential(DevicePotential):
    """potential/displaced_even_power_potential.py:32-142: U = k (|s| - r0)^p, p even."""
    code = L.POT_DISPLACED_EVEN_POWER

    def __init__(self, equilibrium_separation, power, prefactor=1.0):
        if int(power) != power or int(power) % 2 or power <= 0:
            raise ValueError("the power must be a positive even integer")
        super().__init__([equilibrium_separation, power, prefactor])

    def derivative(self, velocity, separation):
        return self._scalar(True, velocity, separation, None, None, None)

    def displacement(self, velocity, separation, potential_change):
        return self._scalar(False, velocity, separation, None, None, potential_change)


class InversePowerCoulombBoundingPotential(DevicePotential):
    """inverse_power_coulomb_bounding_potential.py:39-140 and its C file: k c1 c2 / |s| with
    image-aware inversion over windings of the box (needs the system length)."""
    code = L.POT_INVERSE_POWER_COULOMB_BOUND
    charged = True

    def __init__(self, prefactor=1.5837, system_length=1.0):
        super().__init__([prefactor], system_length)

    def derivative(self, velocity, separation, charge_one, charge_two):
        return self._scalar(True, velocity, separation, charge_one, charge_two, None)

    def displacement(self, velocity, separation, charge_one, charge_two, potential_change):
        return self._scalar(False, velocity, separation, charge_one, charge_two, potential_change)


class MergedImageCoulombPotential(DevicePotential):
    """merged_image_coulomb_potential.py:38-248 and its C file: the Ewald-summed directional
    derivative of the Coulomb interaction with all images (alpha in units of 1/L).

    The instance owns an ecmc_ewald handle (its Fourier table on the device), as the reference's
    owns the C struct: destroyed when the instance is collected (ffi.gc, :124-125), copied by
    deepcopy through ecmc_ewald_copy (:176-209), dropped by pickling and rebuilt on unpickling
    (:211-248).  ``estimated_size`` gives the handle's bytes."""
    code = L.POT_MERGED_IMAGE_COULOMB
    charged = True
    takes_potential_change = False

    def __init__(self, alpha=3.45, fourier_cutoff=6, position_cutoff=2, prefactor=1.0, system_length=1.0):
        super().__init__([alpha, fourier_cutoff, position_cutoff, prefactor], system_length)
        self._args = (float(alpha), int(fourier_cutoff), int(position_cutoff), float(prefactor), float(system_length))
        self._construct()

    def _construct(self, source=None):
        import weakref
        h = ctypes.c_void_p()
        if source is None:
            alpha, fc, pc, _, length = self._args
            L.check(L.lib().ecmc_ewald_create(fc, pc, alpha, length, ctypes.byref(h)))
        else:
            L.check(L.lib().ecmc_ewald_copy(source, ctypes.byref(h)))
        if not h.value:
            raise MemoryError("the merged-image Coulomb handle could not be built")
        self._handle = h
        self._finalizer = weakref.finalize(self, L.lib().ecmc_ewald_destroy, h)

    @property
    def estimated_size(self):
        b = ctypes.c_int64()
        L.check(L.lib().ecmc_ewald_size(self._handle, ctypes.byref(b)))
        return b.value

    def __deepcopy__(self, memo):
        out = MergedImageCoulombPotential.__new__(MergedImageCoulombPotential)
        out.__dict__.update({k: v for k, v in self.__dict__.items() if k not in ("_handle", "_finalizer")})
        out._construct(self._handle)
        memo[id(self)] = out
        return out

    def __getstate__(self):
        return {k: v for k, v in self.__dict__.items() if k not in ("_handle", "_finalizer", "_par")}

    def __setstate__(self, state):
        self.__dict__.update(state)
        alpha, fc, pc, prefactor, _ = self._args
        self._par = (ctypes.c_double * 4)(alpha, float(fc), float(pc), prefactor)
        self._construct()

    def ewald_derivatives(self, separation_along_motion):
        """The C file's raw derivative for rows (n x 3 device tensor, x along the motion)."""
        import torch
        sep = separation_along_motion.to(torch.float64).contiguous()
        out = torch.empty(sep.shape[0], dtype=torch.float64, device=sep.device)
        L.check(L.lib().ecmc_ewald_derivative(self._handle, sep.shape[0], L.dptr(sep), L.dptr(out),
                                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return out

    def derivative(self, velocity, separation, charge_one, charge_two):
        """prefactor * c1 * c2 * derivative(permuted separation) * speed (:150-160)."""
        import torch
        d, speed = _axis(velocity)
        p = [float(separation[(d + i) % 3]) for i in range(3)]
        raw = self.ewald_derivatives(torch.tensor([p], dtype=torch.float64, device="cuda"))
        return self._args[3] * charge_one * charge_two * float(raw.cpu()[0]) * speed

    def displacement(self, *args, **kwargs):
        # merged_image_coulomb_potential.py: a StandardVelocityPotential, not invertible
        raise NotImplementedError


__all__ = ["HardSpherePotential", "HardDipolePotential", "InversePowerPotential", "TruncatedInversePowerPotential", "LennardJonesPotential",
           "DisplacedEvenPowerPotential", "InversePowerCoulombBoundingPotential", "MergedImageCoulombPotential",
           "DevicePotential"]
