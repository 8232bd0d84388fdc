"""Python mirror of Tarang's basis / field / solver interface over the C ABI.

The classes keep the reference's names and argument meaning so that a
harness written against Tarang reads the same:

  FFF_Basis               ~ Universal / FFF_PENCIL (universal.h:48-773) with
                            fftk's transform plan (FFF_pencil.cc:91)
  FluidVF                 ~ FluidVF (FluidVF.h:73-180): cvf V1..V3, nlin1..3,
                            dissipation_coefficient, forcing switch
  Nlin_incompress         ~ Nlin_incompress::Compute_nlin (Nlin.h:64-112)
  Time_advance_incompress ~ Time_advance_incompress::Time_advance_step RK4
                            (Time_advance.h:69-187)
  EnergyTr                ~ EnergyTr::Compute_flux / Compute_shell_tr
                            (EnergyTr.h:99-349)

Device memory is held in torch tensors (complex128 on the GPU); torch is
only the allocator and stream provider, every computation runs in
libtarang_amd.so.  Errors raise TarangGpuError.
"""
import ctypes

import numpy as np
import torch

from . import _lib, h5io
from ._lib import (DEALIAS_NONE, DEALIAS_REF_CORNER, DEALIAS_TWOTHIRDS, PREC_FP32, PREC_FP64, STEP_CONTINUE,
                   TgConfig, TgFluid, check, ptr_array, BASIS_CHFF, BASIS_SFF, SC_CHEB, SC_COS, SC_SIN, TgRbc)

__all__ = ["FFF_Basis", "SFF_Basis", "FluidVF", "FluidSF", "Nlin_incompress", "Time_advance_incompress",
           "Time_advance_incompress_MHD", "Time_advance_incompress_RBC", "Adjoint_incompress", "Correlation",
           "EnergyTr", "SC_SIN", "SC_COS",
           "DEALIAS_NONE", "DEALIAS_REF_CORNER", "DEALIAS_TWOTHIRDS", "PREC_FP64", "PREC_FP32"]


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class FFF_Basis:
    """Periodic Fourier basis on an (Nx, Ny, Nz) grid, one plan per instance."""

    basis_code = 0
    fourier_x = True   # x is a Fourier axis (HDF5 reduced-plan layout)

    @property
    def N(self):
        return (self.Nx, self.Ny, self.Nz)

    def __init__(self, N, kfactor=(1.0, 1.0, 1.0), precision=PREC_FP64, dealias=DEALIAS_REF_CORNER,
                 strong_reality=True, device=0):
        if isinstance(N, int):
            N = (N, N, N)
        self.Nx, self.Ny, self.Nz = (int(n) for n in N)
        self.kfactor = tuple(float(k) for k in kfactor)
        self.precision = precision
        self.dealias_mode = dealias
        self.device = torch.device("cuda", device)
        self.L = _lib.load()
        cfg = TgConfig(self.basis_code, self.Nx, self.Ny, self.Nz, (ctypes.c_double * 3)(*self.kfactor), precision,
                       dealias, int(strong_reality), device)
        h = ctypes.c_void_p()
        check("tg_plan_create", self.L.tg_plan_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.plan = h
        self.bind_stream()

    def bind_stream(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        check("tg_set_stream", self.L.tg_set_stream(self.plan, ctypes.c_void_p(s.cuda_stream)))

    def close(self):
        if getattr(self, "plan", None):
            self.L.tg_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # shapes (fftk Get_FA_shape / Get_RA_shape)
    @property
    def shape_complex_array(self):
        return (self.Nx, self.Ny, self.Nz // 2 + 1)

    @property
    def shape_real_array(self):
        return (self.Nx, self.Ny, self.Nz + 2)

    def complex_array(self):
        return torch.zeros(self.shape_complex_array, dtype=torch.complex128, device=self.device)

    def real_array(self):
        return torch.zeros(self.shape_real_array, dtype=torch.float64, device=self.device)

    def device_bytes(self):
        return int(self.L.tg_plan_device_bytes(self.plan))

    def synchronize(self):
        check("tg_synchronize", self.L.tg_synchronize(self.plan))

    # transforms (FFF_pencil_tr.cc:43-66)
    def Forward_transform(self, Ar, A):
        check("tg_forward", self.L.tg_forward(self.plan, 0, _vp(Ar), _vp(A)))

    def Inverse_transform(self, A, Ar):
        check("tg_inverse", self.L.tg_inverse(self.plan, 0, _vp(A), _vp(Ar)))

    def Dealias(self, *arrays):
        check("tg_dealias", self.L.tg_dealias(self.plan, ptr_array([a.data_ptr() for a in arrays]), len(arrays)))

    def Satisfy_strong_reality_condition_in_Array(self, *arrays):
        check("tg_strong_reality",
              self.L.tg_strong_reality(self.plan, ptr_array([a.data_ptr() for a in arrays]), len(arrays)))

    # reductions (universal_energy.cc)
    def Get_total_energy(self, A):
        out = ctypes.c_double()
        check("tg_energy", self.L.tg_energy(self.plan, _vp(A), ctypes.byref(out)))
        return out.value

    def Get_total_Sn(self, A, n):
        out = ctypes.c_double()
        check("tg_Sn", self.L.tg_Sn(self.plan, _vp(A), int(n), ctypes.byref(out)))
        return out.value

    def Min_radius_outside(self):
        a, b, c = ((self.Nx // 2) * self.kfactor[0], (self.Ny // 2) * self.kfactor[1],
                   (self.Nz // 2) * self.kfactor[2])
        return int(np.ceil(np.sqrt(a * a + b * b + c * c)))

    def Max_radius_inside(self):
        return int(min((self.Nx // 2) * self.kfactor[0], (self.Ny // 2) * self.kfactor[1],
                       (self.Nz // 2) * self.kfactor[2]))

    def Compute_shell_spectrum(self, A, n=0):
        nb = self.Min_radius_outside() + 1
        out = (ctypes.c_double * nb)()
        check("tg_shell_spectrum", self.L.tg_shell_spectrum(self.plan, _vp(A), int(n), out, nb))
        return np.array(out[:])

    def default_radii(self, n):
        """Log-spaced radii (Global_process_advanced_vars.cc:137-175), n+1 entries."""
        mri = self.Max_radius_inside()
        r = np.zeros(n + 1)
        r[0], r[min(1, n)], r[min(2, n)], r[min(3, n)] = 0.0, 2.0, 4.0, 8.0
        r[n - 2] = mri / 2.0
        r[n - 1] = mri
        r[n] = np.inf
        if n > 6:
            s = np.log2(mri / 16.0) / (n - 5)
            for i in range(4, n - 2):
                r[i] = 8 * 2 ** (s * (i - 3))
        return r


class SFF_Basis(FFF_Basis):
    """sin/cos along x (walls at x = 0 and L_x), Fourier along y and z
    (SFF_pencil.cc).  Default kfactor is the reference's RBC box
    (pi, pi/sqrt 2, pi/sqrt 2) (SFF_pencil.cc:67-72); reality 0 none, 1 strong,
    2 weak (the reference's RBC example applies the weak condition)."""

    basis_code = BASIS_SFF
    fourier_x = False

    def __init__(self, N, kfactor=(np.pi, np.pi / np.sqrt(2.0), np.pi / np.sqrt(2.0)), precision=PREC_FP64,
                 dealias=DEALIAS_REF_CORNER, reality=2, device=0):
        super().__init__(N, kfactor=kfactor, precision=precision, dealias=dealias, strong_reality=reality,
                         device=device)

    # SFF_pencil_tr.cc:43-67; sincos SC_SIN ('S') or SC_COS ('C')
    def Forward_transform(self, Ar, A, sincos=SC_SIN):
        check("tg_forward", self.L.tg_forward(self.plan, sincos, _vp(Ar), _vp(A)))

    def Inverse_transform(self, A, Ar, sincos=SC_SIN):
        check("tg_inverse", self.L.tg_inverse(self.plan, sincos, _vp(A), _vp(Ar)))

    def Satisfy_weak_reality_condition_in_Array(self, *arrays):
        check("tg_weak_reality",
              self.L.tg_weak_reality(self.plan, ptr_array([a.data_ptr() for a in arrays]), len(arrays)))


class ChFF_Basis(FFF_Basis):
    """Chebyshev along x (walls at x = +-1 of the Chebyshev variable, kfactor[0]
    = 2 / L_x), Fourier along y and z: the basis of the reference's Chebyshev
    solvers (time_advance/chebyshev/).  Chebyshev-Gauss sample points; FP64."""

    basis_code = BASIS_CHFF
    fourier_x = False

    def __init__(self, N, kfactor=(1.0, 1.0, 1.0), device=0):
        super().__init__(N, kfactor=kfactor, precision=PREC_FP64, dealias=DEALIAS_NONE, strong_reality=False,
                         device=device)

    def Forward_transform(self, Ar, A):
        check("tg_forward", self.L.tg_forward(self.plan, SC_CHEB, _vp(Ar), _vp(A)))

    def Inverse_transform(self, A, Ar):
        check("tg_inverse", self.L.tg_inverse(self.plan, SC_CHEB, _vp(A), _vp(Ar)))

    def Xderiv(self, A, B):
        """d/dx in Chebyshev space (B distinct from A)."""
        check("tg_cheb_xderiv", self.L.tg_cheb_xderiv(self.plan, _vp(A), _vp(B)))

    def Helmholtz(self, R, U, lambda_supplement, value_plus1=0j, value_minus1=0j):
        """Helmholtz_complex_full_array: u'' - (Kperp^2 + lambda_supplement) u = -R
        per (ky, kz) line, u(+1) = value_plus1, u(-1) = value_minus1."""
        vp = (ctypes.c_double * 2)(complex(value_plus1).real, complex(value_plus1).imag)
        vm = (ctypes.c_double * 2)(complex(value_minus1).real, complex(value_minus1).imag)
        check("tg_helmholtz", self.L.tg_helmholtz(self.plan, _vp(R), _vp(U), float(lambda_supplement), vp, vm))


def _fluid(U, dt, force_kind=0, force_eps=0.0, force_r_in=0.0, force_r_out=0.0, force_helicity=0.0,
           force_tau=0.0, force_seed=0):
    """tg_fluid of a FluidVF: dissipation, dt, forcing, hyperviscosity."""
    return TgFluid(U.dissipation_coefficient, float(dt), int(force_kind), float(force_eps), float(force_r_in),
                   float(force_r_out), getattr(U, "hyper_dissipation_coefficient", 0.0),
                   getattr(U, "hyper_dissipation_exponent", 0), float(force_helicity), float(force_tau),
                   int(force_seed))


class Time_advance_incompress_BDF1:
    """Time_advance_incompress::BDF1 on a ChFF basis (BDF1.cc:63-121): U.cvf
    advanced in place, the pressure in self.P."""

    def __init__(self, U, dt):
        self.U = U
        self.fluid = _fluid(U, dt)
        self.P = U.basis.complex_array()

    def Compute_nlin(self):
        b = self.U.basis
        check("tg_compute_nlin_chff", b.L.tg_compute_nlin_chff(b.plan, self.U.ptrs(), self.U.ptrs("nlin")))

    def Time_advance_step(self, nsteps=1):
        b = self.U.basis
        check("tg_bdf1_step", b.L.tg_bdf1_step(b.plan, self.U.ptrs(), _vp(self.P), ctypes.byref(self.fluid),
                                               int(nsteps)))


class FluidSF:
    """Scalar field (temperature): csf.F, nlin, diffusion coefficient."""

    def __init__(self, basis, diffusion_coefficient=0.0, name="T"):
        self.basis = basis
        self.name = name
        self.diffusion_coefficient = float(diffusion_coefficient)
        self.F = basis.complex_array()
        self.nlin = basis.complex_array()

    def set_field(self, F):
        self.F.copy_(torch.as_tensor(F, dtype=torch.complex128))

    def Compute_total_energy(self):
        self.total_energy = self.basis.Get_total_energy(self.F)
        return self.total_energy

    def Write_complex_field(self, out_dir, t, reduced=None):
        """CSF::Write_complex_field: <out_dir>/time_<t>/<name>.F.h5 (reduced:
        <out_dir>/reduced_<t>/..., the low modes of N_out_reduced)."""
        b = self.basis
        if reduced is None:
            h5io.write(self.F, b.N, h5io.FULL, out_dir, h5io.time_folder(t), self.name + ".F", b.fourier_x)
        else:
            h5io.write(self.F, b.N, h5io.REDUCED, out_dir, "reduced_" + h5io.time_folder(t)[5:], self.name + ".F",
                       b.fourier_x, reduced)

    def Read_complex_field(self, in_dir, reduced=None):
        """CSF::Read_complex_field (or the reduced read: N_in_reduced modes)."""
        b = self.basis
        plan = h5io.FULL if reduced is None else h5io.REDUCED
        h5io.read(self.F, b.N, plan, in_dir, self.name + ".F", b.fourier_x, reduced)


class FluidVF:
    """Solver-level vector field: cvf (V1, V2, V3), nlin1..3, dissipation."""

    def __init__(self, basis, dissipation_coefficient=0.0, force_switch=False, name="U",
                 hyper_dissipation_coefficient=0.0, hyper_dissipation_exponent=0):
        """hyper_dissipation_*: FluidVF.cc:72-84 (exponent 0: none); taken by
        the RK4 step and the adjoint."""
        self.basis = basis
        self.name = name
        self.dissipation_coefficient = float(dissipation_coefficient)
        self.hyper_dissipation_coefficient = float(hyper_dissipation_coefficient)
        self.hyper_dissipation_exponent = int(hyper_dissipation_exponent)
        self.force_switch = force_switch
        self.cvf = [basis.complex_array() for _ in range(3)]
        self.nlin = [basis.complex_array() for _ in range(3)]

    @property
    def V1(self):
        return self.cvf[0]

    @property
    def V2(self):
        return self.cvf[1]

    @property
    def V3(self):
        return self.cvf[2]

    def ptrs(self, which="cvf"):
        arr = self.cvf if which == "cvf" else self.nlin
        return ptr_array([a.data_ptr() for a in arr])

    def set_field(self, V):
        for dst, src in zip(self.cvf, V):
            dst.copy_(torch.as_tensor(src, dtype=torch.complex128))

    def Get_dt(self, courant, dt_fixed, dt_option=1):
        """FluidVF::Get_dt (FluidVF.cc:546-572): dt_fixed (dt_option 0), or the
        CFL step min(Courant / sum_i max|u_i| / Delta_x_i, dt_fixed) of the
        dealiased field (dt_option 1; dealiases cvf in place).  Returns dt."""
        if dt_option == 0:
            return float(dt_fixed)
        b = self.basis
        dt = ctypes.c_double()
        um = (ctypes.c_double * 3)()
        check("tg_get_dt", b.L.tg_get_dt(b.plan, self.ptrs(), float(courant), float(dt_fixed), ctypes.byref(dt), um))
        self.umax = tuple(um)
        return dt.value

    def Init_cond_spectrum(self, kind=0, k0=4.0, kmin=2.0, kmax=170.0, energy=0.5, seed=1):
        """Seeded solenoidal field (kind 0: K^4 exp(-2(K/k0)^2); kind 1: K^-5/3)."""
        b = self.basis
        check("tg_init_field", b.L.tg_init_field(b.plan, kind, k0, kmin, kmax, energy, seed, self.ptrs()))

    def Write_complex_field(self, out_dir, t, reduced=None, vx_vy=False):
        """CVF::Write_complex_field (Cvf.cc:494-506): <out_dir>/time_<t>/<name>.V1.h5,
        .V2, .V3 (vx_vy: V3's kz = 0 plane only, as <name>.V3kz0); with
        reduced = N_out_reduced, CVF::Write_reduced_complex_field into
        reduced_<t>/."""
        b = self.basis
        folder = h5io.time_folder(t) if reduced is None else "reduced_" + h5io.time_folder(t)[5:]
        full, kz0 = (h5io.FULL, h5io.KZ0_FULL) if reduced is None else (h5io.REDUCED, h5io.KZ0_REDUCED)
        for c, a in enumerate(self.cvf):
            if c == 2 and vx_vy:
                h5io.write(a, b.N, kz0, out_dir, folder, self.name + ".V3kz0", b.fourier_x, reduced)
            else:
                h5io.write(a, b.N, full, out_dir, folder, f"{self.name}.V{c + 1}", b.fourier_x, reduced)

    def Read_complex_field(self, in_dir, reduced=None, vx_vy=False):
        """CVF::Read_complex_field (Cvf.cc:528-539) / Read_reduced_complex_field
        (:541-556); with vx_vy, V3 is rebuilt for kz > 0 from incompressibility
        (FluidIO::Init_cond_complex_field, init_cond_field.cc:49-56)."""
        b = self.basis
        full, kz0 = (h5io.FULL, h5io.KZ0_FULL) if reduced is None else (h5io.REDUCED, h5io.KZ0_REDUCED)
        for c, a in enumerate(self.cvf):
            if c == 2 and vx_vy:
                if reduced is not None:
                    a.zero_()
                h5io.read(a, b.N, kz0, in_dir, self.name + ".V3kz0", b.fourier_x, reduced)
            else:
                h5io.read(a, b.N, full, in_dir, f"{self.name}.V{c + 1}", b.fourier_x, reduced)
        if vx_vy:
            h5io.fill_vz(self.cvf[0], self.cvf[1], self.cvf[2], b.N, b.kfactor)

    def Compute_total_energy(self):
        e = [self.basis.Get_total_energy(a) for a in self.cvf]
        self.total_E1, self.total_E2, self.total_E3 = e
        self.total_energy = sum(e)
        return self.total_energy

    def Compute_total_k2energy(self):
        self.total_k2energy = 2 * sum(self.basis.Get_total_Sn(a, 2) for a in self.cvf)
        return self.total_k2energy

    def dissipation(self):
        return self.dissipation_coefficient * self.Compute_total_k2energy()


class Nlin_incompress:
    @staticmethod
    def Compute_nlin(U, W=None, mhd=False):
        """Compute_nlin(U); Compute_nlin(U, Giver) (giver form); or, with
        mhd=True, Compute_nlin(U, W) of MHD filling U.nlin and W.nlin."""
        b = U.basis
        if W is None:
            check("tg_compute_nlin", b.L.tg_compute_nlin(b.plan, U.ptrs(), U.ptrs("nlin")))
        elif mhd:
            check("tg_compute_nlin_mhd", b.L.tg_compute_nlin_mhd(b.plan, U.ptrs(), W.ptrs(), U.ptrs("nlin"),
                                                                  W.ptrs("nlin")))
        else:
            check("tg_compute_nlin_giver", b.L.tg_compute_nlin_giver(b.plan, U.ptrs(), W.ptrs(), U.ptrs("nlin")))


class Correlation:
    @staticmethod
    def Compute_ring_spectrum(U, n=0, sector_angles=None, dirn=3):
        """Universal::Compute_ring_spectrum of U.cvf (or of a FluidSF's F):
        [component, shell, sector]; Correlation.cc:301-311 multiplies the n = 2
        form by 2 for the dissipation spectrum."""
        b = U.basis
        arrs = U.cvf if hasattr(U, "cvf") else [U.F]
        a = np.ascontiguousarray(sector_angles if sector_angles is not None else sector_angles_equi(4), np.float64)
        out = np.zeros((len(arrs), max_radius_inside(b) + 1, len(a)))
        dp = ctypes.POINTER(ctypes.c_double)
        check("tg_ring_spectrum", b.L.tg_ring_spectrum(b.plan, ptr_array([x.data_ptr() for x in arrs]), len(arrs),
                                                       int(n), int(dirn), a.ctypes.data_as(dp), len(a) - 1,
                                                       out.ctypes.data_as(dp)))
        return out

    @staticmethod
    def Compute_cylindrical_ring_spectrum(U, n=0, kpll=None, dirn=3):
        """Universal::Compute_cylindrical_ring_spectrum: (S1, S2) [Kperp shell, slab]."""
        b = U.basis
        h = np.ascontiguousarray(kpll if kpll is not None else kpll_edges_equi(b, 4, dirn), np.float64)
        km = aniso_max_krho_inside(b, dirn)
        S1, S2 = np.zeros((km + 1, len(h))), np.zeros((km + 1, len(h)))
        dp = ctypes.POINTER(ctypes.c_double)
        check("tg_cylindrical_ring_spectrum",
              b.L.tg_cylindrical_ring_spectrum(b.plan, U.ptrs(), int(n), int(dirn), h.ctypes.data_as(dp), len(h) - 1,
                                               S1.ctypes.data_as(dp), S2.ctypes.data_as(dp)))
        return S1, S2

    @staticmethod
    def Get_Nusselt_no(U, T, Ra, Pr):
        """Correlation.cc:144-160 (PRLARGE / ULARGE)."""
        out = ctypes.c_double()
        b = U.basis
        prm = TgRbc(float(Ra), float(Pr), 1.0)
        check("tg_nusselt", b.L.tg_nusselt(b.plan, _vp(U.cvf[0]), _vp(T.F), ctypes.byref(prm), ctypes.byref(out)))
        return out.value

    @staticmethod
    def Get_cross_helicity(U, W):
        out = ctypes.c_double()
        b = U.basis
        check("tg_cross_helicity", b.L.tg_cross_helicity(b.plan, U.ptrs(), W.ptrs(), ctypes.byref(out)))
        return out.value


class Time_advance_incompress:
    """RK4 time advance of an incompressible fluid on one GPU."""

    def __init__(self, U, dt, force_kind=0, force_eps=0.0, force_r_in=0.0, force_r_out=0.0, force_helicity=0.0,
                 force_tau=0.0, force_seed=0):
        """force_kind 1: constant energy supply (rate force_eps on the band
        (r_in, r_out]); 2: random band forcing (compute_force_random.cc:
        65-130) with energy and helicity supply rates force_eps,
        force_helicity, random interval force_tau: the phases are drawn anew
        (seed force_seed + n in the n-th interval) once the time has passed
        the end of the current interval."""
        self.U = U
        self.fluid = _fluid(U, dt, force_kind, force_eps, force_r_in, force_r_out, force_helicity, force_tau,
                            force_seed)
        self._continue = False
        self._versions = None
        self.time = 0.0
        self._seed0, self._interval, self._random_next = int(force_seed), 0, float(force_tau)

    def _state_versions(self):
        # torch bumps a tensor's version on every in-place write (set_field,
        # copy_, zero_, ...); the library's own kernels write through raw
        # pointers and leave it unchanged
        return tuple((id(t), t._version) for t in self.U.cvf)

    def Compute_force_TO_rhs(self):
        b = self.U.basis
        check("tg_compute_rhs", b.L.tg_compute_rhs(b.plan, self.U.ptrs(), ctypes.byref(self.fluid),
                                                   self.U.ptrs("nlin")))
        self._continue = False

    def Time_advance_step(self, nsteps=1, chain=True):
        """nsteps RK4 steps.  chain: reuse the stage-1 transform left by the
        previous call when the state tensors were not written in place since
        (raises TarangGpuError with code TG_ERR_NAN = -5 when the state turns
        non-finite)."""
        b = self.U.basis
        if self.fluid.force_kind == 2 and nsteps > 1:
            for _ in range(nsteps):
                self.Time_advance_step(1, chain)
            return
        if self.fluid.force_kind == 2 and self.time >= self._random_next:
            # Compute_force_using_random_noise_assign (compute_force_random.cc:143-167)
            self._interval += 1
            self._random_next += self.fluid.force_tau
            self.fluid.force_seed = self._seed0 + self._interval
        cont = chain and self._continue and self._versions == self._state_versions()
        flags = STEP_CONTINUE if cont else 0
        self._continue = False
        check("tg_rk4_step", b.L.tg_rk4_step(b.plan, self.U.ptrs(), ctypes.byref(self.fluid), int(nsteps), flags))
        self._continue = True
        self._versions = self._state_versions()
        self.time += self.fluid.dt * nsteps

    def invalidate(self):
        self._continue = False


class Time_advance_incompress_MHD:
    """RK4 of velocity U (nu) and magnetic field W (eta), RK_VF_MHD.cc:126-190."""

    def __init__(self, U, W, dt):
        self.U, self.W = U, W
        self.fluid = _fluid(U, dt)

    def Time_advance_step(self, nsteps=1):
        b = self.U.basis
        check("tg_rk4_step_mhd", b.L.tg_rk4_step_mhd(b.plan, self.U.ptrs(), self.W.ptrs(), ctypes.byref(self.fluid),
                                                     self.W.dissipation_coefficient, int(nsteps)))


class Time_advance_incompress_RBC:
    """Rayleigh-Benard convection on an SFF basis, PRLARGE / ULARGE
    (nu = sqrt(Pr/Ra), kappa = 1/sqrt(Pr Ra)): RK4 of U and T (RK_SF.cc:97-215)."""

    def __init__(self, U, T, Ra, Pr, dt):
        self.U, self.T = U, T
        self.prm = TgRbc(float(Ra), float(Pr), float(dt))
        U.dissipation_coefficient = float(np.sqrt(Pr / Ra))
        T.diffusion_coefficient = float(1.0 / np.sqrt(Pr * Ra))

    def Compute_nlin(self):
        """Nlin_incompress::Compute_nlin(U, T) into U.nlin, T.nlin (dealiases U, T)."""
        b = self.U.basis
        check("tg_compute_nlin_rbc", b.L.tg_compute_nlin_rbc(b.plan, self.U.ptrs(), _vp(self.T.F),
                                                             self.U.ptrs("nlin"), _vp(self.T.nlin)))

    def Compute_force_TO_rhs(self):
        b = self.U.basis
        check("tg_compute_rhs_rbc", b.L.tg_compute_rhs_rbc(b.plan, self.U.ptrs(), _vp(self.T.F),
                                                           ctypes.byref(self.prm), self.U.ptrs("nlin"),
                                                           _vp(self.T.nlin)))

    def Time_advance_step(self, nsteps=1):
        b = self.U.basis
        check("tg_rk4_step_rbc", b.L.tg_rk4_step_rbc(b.plan, self.U.ptrs(), _vp(self.T.F), ctypes.byref(self.prm),
                                                     int(nsteps)))


class Adjoint_incompress:
    """Adjoint of the unforced Lawson RK4 map of U (no counterpart in the
    reference).  Forward() advances U.cvf in place, keeping a checkpoint every
    checkpoint_every steps and the real-space stage velocities of the last
    segment; Backward(lam) maps dL/dV_n to dL/dV_0 in place, recomputing the
    earlier segments.  Gradients of complex coefficients are dL/dRe + i dL/dIm
    (the convention of torch.autograd for complex tensors).  The plan must have
    strong_reality off."""

    def __init__(self, U, dt, checkpoint_every=20):
        self.U = U
        self.fluid = _fluid(U, dt)
        self.checkpoint_every = int(checkpoint_every)

    def Forward(self, nsteps):
        b = self.U.basis
        check("tg_adjoint_forward", b.L.tg_adjoint_forward(b.plan, self.U.ptrs(), ctypes.byref(self.fluid),
                                                           int(nsteps), self.checkpoint_every))

    def Backward(self, lam):
        b = self.U.basis
        check("tg_adjoint_backward", b.L.tg_adjoint_backward(b.plan, ptr_array([x.data_ptr() for x in lam])))
        return lam

    def Rhs_vjp(self, lam):
        """f'(V)^T lam for f(V) = -P NL(D V) (the unforced Compute_rhs) at V = U.cvf."""
        b = self.U.basis
        out = [torch.empty_like(self.U.cvf[0]) for _ in range(3)]
        check("tg_rhs_vjp", b.L.tg_rhs_vjp(b.plan, self.U.ptrs(), ptr_array([x.data_ptr() for x in lam]),
                                           ptr_array([x.data_ptr() for x in out])))
        return out

    def Free(self):
        b = self.U.basis
        check("tg_adjoint_free", b.L.tg_adjoint_free(b.plan))


class EnergyTr:
    def __init__(self, basis, flux_radii=None, shell_radii=None):
        self.basis = basis
        self.flux_radii = flux_radii
        self.shell_radii = shell_radii

    def Compute_flux(self, U):
        r = np.ascontiguousarray(self.flux_radii, dtype=np.float64)
        n = len(r) - 1
        out = np.zeros(n + 1)
        b = self.basis
        check("tg_flux", b.L.tg_flux(b.plan, U.ptrs(), r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        self.flux_self = out
        return out

    def Compute_flux_MHD(self, U, W):
        """u< -> b> and b< -> b> fluxes (flux_VF_Uin_Wout, flux_VF_Win_Wout)."""
        r = np.ascontiguousarray(self.flux_radii, dtype=np.float64)
        n = len(r) - 1
        u2b, b2b = np.zeros(n + 1), np.zeros(n + 1)
        b = self.basis
        dp = ctypes.POINTER(ctypes.c_double)
        check("tg_flux_mhd", b.L.tg_flux_mhd(b.plan, U.ptrs(), W.ptrs(), r.ctypes.data_as(dp), n,
                                             u2b.ctypes.data_as(dp), b2b.ctypes.data_as(dp)))
        return u2b, b2b

    def Compute_flux_and_shell_tr(self, U):
        """Compute_flux and Compute_shell_tr together, for flux spheres equal to
        the shell radii (one giver evaluation per shell; tg_energy_transfer)."""
        if not np.array_equal(np.asarray(self.flux_radii), np.asarray(self.shell_radii)):
            raise ValueError("Compute_flux_and_shell_tr needs flux_radii == shell_radii")
        r = np.ascontiguousarray(self.shell_radii, dtype=np.float64)
        n = len(r) - 1
        flux = np.zeros(n + 1)
        S = np.zeros((n + 1, n + 1))
        b = self.basis
        dp = ctypes.POINTER(ctypes.c_double)
        check("tg_energy_transfer", b.L.tg_energy_transfer(b.plan, U.ptrs(), r.ctypes.data_as(dp), n,
                                                           flux.ctypes.data_as(dp), S.ctypes.data_as(dp)))
        self.flux_self, self.shelltoshell_self = flux, S
        return flux, S

    def Compute_shell_tr(self, U):
        r = np.ascontiguousarray(self.shell_radii, dtype=np.float64)
        n = len(r) - 1
        out = np.zeros((n + 1, n + 1))
        b = self.basis
        check("tg_shell_to_shell",
              b.L.tg_shell_to_shell(b.plan, U.ptrs(), r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        self.shelltoshell_self = out
        return out

    def Compute_ring_tr(self, U, radii=None, sector_angles=None, dirn=3):
        """EnergyTr::Compute_ring_tr (compute_ring_tr.cc:62-77):
        ring_to_ring_self[giver shell, giver sector, receiver shell, receiver sector]."""
        b = self.basis
        r = np.ascontiguousarray(radii if radii is not None else ring_radii(b, 8), dtype=np.float64)
        a = np.ascontiguousarray(sector_angles if sector_angles is not None else sector_angles_equi(4), np.float64)
        out = np.zeros((len(r), len(a), len(r), len(a)))
        dp = ctypes.POINTER(ctypes.c_double)
        check("tg_ring_to_ring", b.L.tg_ring_to_ring(b.plan, U.ptrs(), int(dirn), r.ctypes.data_as(dp), len(r) - 1,
                                                     a.ctypes.data_as(dp), len(a) - 1, out.ctypes.data_as(dp)))
        self.ring_to_ring_self = out
        return out

    def Compute_cylindrical_ring_tr(self, U, radii=None, kpll=None, dirn=3):
        """EnergyTr::Compute_cylindrical_ring_tr (compute_cylin_ring_tr.cc:51-64):
        [giver shell, giver slab, receiver shell, receiver slab]."""
        b = self.basis
        r = np.ascontiguousarray(radii if radii is not None else cylindrical_ring_radii(b, 8, dirn), np.float64)
        h = np.ascontiguousarray(kpll if kpll is not None else kpll_edges_equi(b, 4, dirn), np.float64)
        out = np.zeros((len(r), len(h), len(r), len(h)))
        dp = ctypes.POINTER(ctypes.c_double)
        check("tg_cylindrical_ring_to_ring",
              b.L.tg_cylindrical_ring_to_ring(b.plan, U.ptrs(), int(dirn), r.ctypes.data_as(dp), len(r) - 1,
                                              h.ctypes.data_as(dp), len(h) - 1, out.ctypes.data_as(dp)))
        self.cylindrical_ring_to_ring_self = out
        return out


def sector_angles_equi(nsectors, max_theta=np.pi / 2):
    """EQUISPACED sector angles i max_theta / n, the last one max_theta (pi/2 on
    FFF, Get_max_polar_angle) (Global_process_advanced_vars.cc:100-107)."""
    d = max_theta / nsectors
    return np.array([i * d for i in range(nsectors)] + [max_theta])


def _aniso_max_kpll(basis, dirn):
    # Anis_max_Kpll (FFF_pencil_inline.h:767-781)
    n = (basis.Nx, basis.Ny, basis.Nz)[dirn - 1]
    return (n // 2) * basis.kfactor[dirn - 1]


def aniso_max_krho_inside(basis, dirn):
    """Anis_max_Krho_radius_inside (FFF_pencil_inline.h:784-798)."""
    h = [(basis.Nx // 2) * basis.kfactor[0], (basis.Ny // 2) * basis.kfactor[1], (basis.Nz // 2) * basis.kfactor[2]]
    del h[dirn - 1]
    return int(min(h))


def max_radius_inside(basis):
    return int(min((basis.Nx // 2) * basis.kfactor[0], (basis.Ny // 2) * basis.kfactor[1],
                   (basis.Nz // 2) * basis.kfactor[2]))


def kpll_edges_equi(basis, nslabs, dirn=3):
    """EQUISPACED slab edges over [0, Anis_max_Kpll] (Global_process_advanced_vars.cc:112-123)."""
    d = _aniso_max_kpll(basis, dirn) / nslabs
    return np.array([i * d for i in range(nslabs + 1)])


def ring_radii(basis, nshells):
    """Default ring-to-ring radii (Global_process_advanced_vars.cc:180-195):
    0, 4, 8, log-spaced, Kmax/2, Kmax, inf with Kmax = Max_radius_inside."""
    kmax = max_radius_inside(basis)
    r = np.zeros(nshells + 1)
    r[1], r[2] = 4.0, 8.0
    r[nshells - 2], r[nshells - 1], r[nshells] = kmax / 2.0, kmax, np.inf
    if nshells > 5:
        s = np.log2(kmax / 16.0) / (nshells - 4)
        for i in range(3, nshells - 2):
            r[i] = 8 * 2 ** (s * (i - 2))
    return r


def cylindrical_ring_radii(basis, nshells, dirn=3):
    """Default cylindrical-ring radii (Global_process_advanced_vars.cc:224-238):
    as ring_radii, with the log spacing over Anis_max_Krho_radius_inside."""
    kmax = max_radius_inside(basis)
    kc = aniso_max_krho_inside(basis, dirn)
    r = np.zeros(nshells + 1)
    r[1], r[2] = 4.0, 8.0
    r[nshells - 2], r[nshells - 1], r[nshells] = kmax / 2.0, kmax, np.inf
    if nshells > 5:
        s = np.log2(kc / 16.0) / (nshells - 4)
        for i in range(3, nshells - 2):
            r[i] = 8 * 2 ** (s * (i - 2))
    return r
