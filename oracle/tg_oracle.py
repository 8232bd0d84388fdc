"""ctypes binding of the FP64 CPU oracle (oracle/libtg_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (tarang_amd) never
imports it.

Arrays: complex spectral fields are numpy complex128 of shape (Nx, Ny, Nz/2+1);
real fields are float64 of shape (Nx, Ny, Nz+2).  All arrays must be
C-contiguous.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtg_oracle.so")

DEALIAS_NONE, DEALIAS_REF_CORNER, DEALIAS_TWOTHIRDS = 0, 1, 2


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("kfactor", ctypes.c_double * 3), ("dealias", ctypes.c_int),
                ("strong_reality", ctypes.c_int)]


class Fluid(ctypes.Structure):
    _fields_ = [("nu", ctypes.c_double), ("dt", ctypes.c_double), ("force_kind", ctypes.c_int),
                ("force_eps", ctypes.c_double), ("force_r_in", ctypes.c_double),
                ("force_r_out", ctypes.c_double), ("nu_hyper", ctypes.c_double), ("hyper_exponent", ctypes.c_int),
                ("force_helicity", ctypes.c_double), ("force_tau", ctypes.c_double),
                ("force_seed", ctypes.c_ulonglong)]


class Rbc(ctypes.Structure):
    """PRLARGE / ULARGE Rayleigh-Benard parameters; reality 0 none, 1 strong, 2 weak."""
    _fields_ = [("Ra", ctypes.c_double), ("Pr", ctypes.c_double), ("dt", ctypes.c_double),
                ("reality", ctypes.c_int)]


SIN, COS = 0, 1


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        d = ctypes.c_double
        p = ctypes.c_void_p
        i = ctypes.c_int
        gp = ctypes.POINTER(Grid)
        fp = ctypes.POINTER(Fluid)
        pp = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "or_fft1d": (i, [p, i, i]),
            "or_forward": (None, [gp, p, p]),
            "or_inverse": (None, [gp, p, p]),
            "or_dealias": (None, [gp, p]),
            "or_strong_reality": (None, [gp, p]),
            "or_weak_reality": (None, [gp, p]),
            "or_multiplicity": (d, [gp, i, i, i]),
            "or_kmag": (d, [gp, i, i, i]),
            "or_min_radius_outside": (i, [gp]),
            "or_max_radius_inside": (i, [gp]),
            "or_compute_nlin": (None, [gp, pp, pp]),
            "or_compute_nlin_giver": (None, [gp, pp, pp, pp]),
            "or_compute_force": (None, [gp, fp, pp, pp]),
            "or_compute_rhs": (None, [gp, pp, pp]),
            "or_rk4_step": (None, [gp, fp, pp]),
            "or_euler_step": (None, [gp, fp, pp]),
            "or_energy": (d, [gp, p]),
            "or_Sn": (d, [gp, p, i]),
            "or_shell_spectrum": (i, [gp, p, i, p, i]),
            "or_fill_array_shell": (None, [gp, p, p, d, d]),
            "or_shell_mult_single": (d, [gp, p, p, d, d]),
            "or_shell_mult_all": (None, [gp, pp, pp, p, i, p]),
            "or_flux": (None, [gp, pp, p, i, p]),
            "or_shell_to_shell": (None, [gp, pp, p, i, p]),
            "or_default_radii": (None, [gp, i, p]),
            "or_init_field": (None, [gp, i, d, d, d, d, ctypes.c_uint64, pp]),
            "or_compute_nlin_mhd": (None, [gp, pp, pp, pp, pp]),
            "or_rk4_step_mhd": (None, [gp, fp, d, pp, pp]),
            "or_cross_energy": (d, [gp, p, p]),
            "or_flux_mhd": (None, [gp, pp, pp, p, i, p, p]),
            "or_rhs_jvp": (None, [gp, pp, pp, pp]),
            "or_rhs_vjp": (None, [gp, pp, pp, pp]),
            "or_rk4_step_vjp": (None, [gp, fp, pp, pp]),
            "or_adjoint_shell_energy": (d, [gp, fp, pp, i, d, d, pp]),
            "or_sff_forward": (None, [gp, i, p, p]),
            "or_sff_inverse": (None, [gp, i, p, p]),
            "or_sff_dealias": (None, [gp, p]),
            "or_get_dt": (d, [gp, i, pp, d, d, p]),
            "or_modes_in_shell": (None, [gp, i, p]),
            "or_azimuthal_angle": (d, [gp, i, i, i, i]),
            "or_sff_reality": (None, [gp, p, i]),
            "or_sff_energy": (d, [gp, p, p]),
            "or_sff_kmag": (d, [gp, i, i, i]),
            "or_compute_nlin_rbc": (None, [gp, pp, p, pp, p]),
            "or_compute_rhs_rbc": (None, [gp, p, p, pp, p]),
            "or_rk4_step_rbc": (None, [gp, ctypes.POINTER(Rbc), pp, p]),
            "or_nusselt": (d, [gp, ctypes.POINTER(Rbc), p, p]),
            "or_chff_forward": (None, [gp, p, p]),
            "or_chff_inverse": (None, [gp, p, p]),
            "or_cheb_xderiv": (None, [gp, p, p]),
            "or_helmholtz_line": (None, [i, p, p, d, d, d]),
            "or_helmholtz": (None, [gp, p, p, d, p, p]),
            "or_chff_influence": (None, [gp, d, d, p, p, p, p, p]),
            "or_chff_correct": (None, [gp, p, p, p, p, p, p, p]),
            "or_compute_nlin_chff": (None, [gp, pp, pp]),
            "or_bdf1_chff": (None, [gp, d, d, pp, p]),
            "or_aniso_kpll": (d, [gp, i, i, i, i]),
            "or_aniso_kperp": (d, [gp, i, i, i, i]),
            "or_polar_angle": (d, [gp, i, i, i, i]),
            "or_aniso_max_krho_inside": (i, [gp, i]),
            "or_ring_spectrum": (None, [gp, pp, i, i, i, p, i, p]),
            "or_cyl_ring_spectrum": (None, [gp, pp, i, i, p, i, p, p]),
            "or_fill_array_ring": (None, [gp, p, p, d, d, d, d, i]),
            "or_fill_array_cyl_ring": (None, [gp, p, p, d, d, d, d, i]),
            "or_ring_mult_all": (None, [gp, pp, pp, i, p, i, p, i, p]),
            "or_cyl_ring_mult_all": (None, [gp, pp, pp, i, p, i, p, i, p]),
            "or_ring_to_ring": (None, [gp, pp, i, p, i, p, i, p]),
            "or_cyl_ring_to_ring": (None, [gp, pp, i, p, i, p, i, p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _cptr(a):
    """Pointer to a complex128 spectral array (rejects any other dtype)."""
    assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags["C_CONTIGUOUS"])
    return a.ctypes.data_as(ctypes.c_void_p)


def _pp(arrs):
    for a in arrs:
        assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class Oracle:
    """FP64 CPU restatement of the reference FFF path on one grid."""

    def __init__(self, n, kfactor=(1.0, 1.0, 1.0), dealias=DEALIAS_REF_CORNER, strong_reality=True):
        if isinstance(n, int):
            n = (n, n, n)
        self.nx, self.ny, self.nz = n
        self.kfactor = tuple(float(k) for k in kfactor)
        self.g = Grid(self.nx, self.ny, self.nz, (ctypes.c_double * 3)(*kfactor), dealias,
                      1 if strong_reality else 0)
        self.L = lib()

    @property
    def cshape(self):
        return (self.nx, self.ny, self.nz // 2 + 1)

    @property
    def rshape(self):
        return (self.nx, self.ny, self.nz + 2)

    def czeros(self):
        return np.zeros(self.cshape, np.complex128)

    def forward(self, real):
        real = np.ascontiguousarray(real, np.float64)
        out = self.czeros()
        self.L.or_forward(ctypes.byref(self.g), _ptr(real), _ptr(out))
        return out

    def inverse(self, spec):
        spec = np.ascontiguousarray(spec, np.complex128)
        out = np.zeros(self.rshape, np.float64)
        self.L.or_inverse(ctypes.byref(self.g), _ptr(spec), _ptr(out))
        return out

    def dealias(self, spec):
        self.L.or_dealias(ctypes.byref(self.g), _cptr(spec))

    def strong_reality(self, spec):
        self.L.or_strong_reality(ctypes.byref(self.g), _cptr(spec))

    def compute_nlin(self, V):
        N = [self.czeros() for _ in range(3)]
        self.L.or_compute_nlin(ctypes.byref(self.g), _pp(V), _pp(N))
        return N

    def compute_nlin_giver(self, V, W):
        N = [self.czeros() for _ in range(3)]
        self.L.or_compute_nlin_giver(ctypes.byref(self.g), _pp(V), _pp(W), _pp(N))
        return N

    def compute_force(self, fluid, V):
        F = [self.czeros() for _ in range(3)]
        self.L.or_compute_force(ctypes.byref(self.g), ctypes.byref(fluid), _pp(V), _pp(F))
        return F

    def compute_rhs(self, N, F=None):
        self.L.or_compute_rhs(ctypes.byref(self.g), _pp(N), _pp(F) if F is not None else None)

    def rk4_step(self, fluid, V):
        self.L.or_rk4_step(ctypes.byref(self.g), ctypes.byref(fluid), _pp(V))

    def euler_step(self, fluid, V):
        self.L.or_euler_step(ctypes.byref(self.g), ctypes.byref(fluid), _pp(V))

    def energy(self, A):
        return self.L.or_energy(ctypes.byref(self.g), _cptr(A))

    def Sn(self, A, n):
        return self.L.or_Sn(ctypes.byref(self.g), _cptr(A), n)

    def multiplicity(self, lx, ly, lz):
        return self.L.or_multiplicity(ctypes.byref(self.g), lx, ly, lz)

    def kmag(self, lx, ly, lz):
        return self.L.or_kmag(ctypes.byref(self.g), lx, ly, lz)

    def min_radius_outside(self):
        return self.L.or_min_radius_outside(ctypes.byref(self.g))

    def max_radius_inside(self):
        return self.L.or_max_radius_inside(ctypes.byref(self.g))

    def shell_spectrum(self, A, n=0):
        nb = self.min_radius_outside() + 1
        Sk = np.zeros(nb)
        self.L.or_shell_spectrum(ctypes.byref(self.g), _cptr(A), n, _ptr(Sk), nb)
        return Sk

    def fill_array_shell(self, A, r_in, r_out):
        B = self.czeros()
        self.L.or_fill_array_shell(ctypes.byref(self.g), _cptr(A), _cptr(B), r_in, r_out)
        return B

    def shell_mult_single(self, A, B, r_in, r_out):
        return self.L.or_shell_mult_single(ctypes.byref(self.g), _cptr(A), _cptr(B), r_in, r_out)

    def shell_mult_all(self, A, B, radii):
        radii = np.ascontiguousarray(radii, np.float64)
        out = np.zeros(len(radii))
        self.L.or_shell_mult_all(ctypes.byref(self.g), _pp(A), _pp(B), _ptr(radii), len(radii), _ptr(out))
        return out

    def default_radii(self, n):
        r = np.zeros(n + 1)
        self.L.or_default_radii(ctypes.byref(self.g), n, _ptr(r))
        return r

    def flux(self, V, radii):
        radii = np.ascontiguousarray(radii, np.float64)
        out = np.zeros(len(radii))
        self.L.or_flux(ctypes.byref(self.g), _pp(V), _ptr(radii), len(radii) - 1, _ptr(out))
        return out

    def shell_to_shell(self, V, radii):
        radii = np.ascontiguousarray(radii, np.float64)
        n = len(radii) - 1
        out = np.zeros((n + 1, n + 1))
        self.L.or_shell_to_shell(ctypes.byref(self.g), _pp(V), _ptr(radii), n, _ptr(out))
        return out

    def compute_nlin_mhd(self, V, W):
        NU = [self.czeros() for _ in range(3)]
        NW = [self.czeros() for _ in range(3)]
        self.L.or_compute_nlin_mhd(ctypes.byref(self.g), _pp(V), _pp(W), _pp(NU), _pp(NW))
        return NU, NW

    def rk4_step_mhd(self, fluid, eta, V, W):
        self.L.or_rk4_step_mhd(ctypes.byref(self.g), ctypes.byref(fluid), eta, _pp(V), _pp(W))

    def cross_energy(self, A, B):
        return self.L.or_cross_energy(ctypes.byref(self.g), _cptr(A), _cptr(B))

    def flux_mhd(self, V, W, radii):
        radii = np.ascontiguousarray(radii, np.float64)
        u2b = np.zeros(len(radii))
        b2b = np.zeros(len(radii))
        self.L.or_flux_mhd(ctypes.byref(self.g), _pp(V), _pp(W), _ptr(radii), len(radii) - 1, _ptr(u2b), _ptr(b2b))
        return u2b, b2b

    def rhs_jvp(self, V, dV):
        out = [self.czeros() for _ in range(3)]
        self.L.or_rhs_jvp(ctypes.byref(self.g), _pp(V), _pp(dV), _pp(out))
        return out

    def rhs_vjp(self, V, Lam):
        out = [self.czeros() for _ in range(3)]
        self.L.or_rhs_vjp(ctypes.byref(self.g), _pp(V), _pp(Lam), _pp(out))
        return out

    def rk4_step_vjp(self, fluid, V, lam):
        lam = [l.copy() for l in lam]
        self.L.or_rk4_step_vjp(ctypes.byref(self.g), ctypes.byref(fluid), _pp(V), _pp(lam))
        return lam

    def adjoint_shell_energy(self, fluid, V0, nsteps, r_lo, r_hi):
        grad = [self.czeros() for _ in range(3)]
        loss = self.L.or_adjoint_shell_energy(ctypes.byref(self.g), ctypes.byref(fluid), _pp(V0), nsteps, r_lo, r_hi,
                                              _pp(grad))
        return loss, grad

    # ---- Chebyshev (ChFF): Chebyshev-Gauss points along x ----
    def chff_forward(self, real):
        real = np.ascontiguousarray(real, np.float64)
        out = self.czeros()
        self.L.or_chff_forward(ctypes.byref(self.g), _ptr(real), _ptr(out))
        return out

    def chff_inverse(self, spec):
        out = np.zeros(self.rshape)
        self.L.or_chff_inverse(ctypes.byref(self.g), _cptr(spec), _ptr(out))
        return out

    def cheb_xderiv(self, A):
        B = self.czeros()
        self.L.or_cheb_xderiv(ctypes.byref(self.g), _cptr(A), _cptr(B))
        return B

    def helmholtz_line(self, f, lam, vp, vm):
        f = np.ascontiguousarray(f, np.float64)
        x = np.zeros_like(f)
        self.L.or_helmholtz_line(len(f), _ptr(f), _ptr(x), lam, vp, vm)
        return x

    def helmholtz(self, R, lam_sup, vp=0j, vm=0j):
        U = self.czeros()
        a = np.array([complex(vp).real, complex(vp).imag])
        b = np.array([complex(vm).real, complex(vm).imag])
        self.L.or_helmholtz(ctypes.byref(self.g), _cptr(R), _cptr(U), lam_sup, _ptr(a), _ptr(b))
        return U

    def chff_influence(self, nu, dt):
        shp = self.cshape
        pp, pm, vxp, vxm = (np.zeros(shp) for _ in range(4))
        M = np.zeros((shp[1], shp[2], 4))
        self.L.or_chff_influence(ctypes.byref(self.g), nu, dt, _ptr(pp), _ptr(pm), _ptr(vxp), _ptr(vxm), _ptr(M))
        return pp, pm, vxp, vxm, M

    def compute_nlin_chff(self, V):
        N = [self.czeros() for _ in range(3)]
        self.L.or_compute_nlin_chff(ctypes.byref(self.g), _pp(V), _pp(N))
        return N

    def bdf1_chff(self, nu, dt, V):
        P = self.czeros()
        self.L.or_bdf1_chff(ctypes.byref(self.g), nu, dt, _pp(V), _cptr(P))
        return P

    # ---- anisotropic diagnostics (dirn = anisotropy_dirn 1..3) ----
    def polar_angle(self, lx, ly, lz, dirn=3):
        return self.L.or_polar_angle(ctypes.byref(self.g), lx, ly, lz, dirn)

    def aniso_max_krho_inside(self, dirn=3):
        return self.L.or_aniso_max_krho_inside(ctypes.byref(self.g), dirn)

    def ring_spectrum(self, A, n, angles, dirn=3):
        """[ncomp, Max_radius_inside+1, nsectors+1]"""
        angles = np.ascontiguousarray(angles, np.float64)
        nsec = len(angles) - 1
        kmax = self.max_radius_inside()
        out = np.zeros((len(A), kmax + 1, nsec + 1))
        self.L.or_ring_spectrum(ctypes.byref(self.g), _pp(A), len(A), n, dirn, _ptr(angles), nsec, _ptr(out))
        return out

    def cyl_ring_spectrum(self, V, n, kpll, dirn=3):
        kpll = np.ascontiguousarray(kpll, np.float64)
        ns = len(kpll) - 1
        km = self.aniso_max_krho_inside(dirn)
        S1, S2 = np.zeros((km + 1, ns + 1)), np.zeros((km + 1, ns + 1))
        self.L.or_cyl_ring_spectrum(ctypes.byref(self.g), _pp(V), n, dirn, _ptr(kpll), ns, _ptr(S1), _ptr(S2))
        return S1, S2

    def fill_array_ring(self, A, r_in, r_out, a_lo, a_hi, dirn=3):
        B = self.czeros()
        self.L.or_fill_array_ring(ctypes.byref(self.g), _cptr(A), _cptr(B), r_in, r_out, a_lo, a_hi, dirn)
        return B

    def fill_array_cyl_ring(self, A, r_in, r_out, h_lo, h_hi, dirn=3):
        B = self.czeros()
        self.L.or_fill_array_cyl_ring(ctypes.byref(self.g), _cptr(A), _cptr(B), r_in, r_out, h_lo, h_hi, dirn)
        return B

    def ring_mult_all(self, A, B, radii, angles, dirn=3):
        radii = np.ascontiguousarray(radii, np.float64)
        angles = np.ascontiguousarray(angles, np.float64)
        out = np.zeros((len(radii), len(angles)))
        self.L.or_ring_mult_all(ctypes.byref(self.g), _pp(A), _pp(B), dirn, _ptr(radii), len(radii) - 1,
                                _ptr(angles), len(angles) - 1, _ptr(out))
        return out

    def cyl_ring_mult_all(self, A, B, radii, kpll, dirn=3):
        radii = np.ascontiguousarray(radii, np.float64)
        kpll = np.ascontiguousarray(kpll, np.float64)
        out = np.zeros((len(radii), len(kpll)))
        self.L.or_cyl_ring_mult_all(ctypes.byref(self.g), _pp(A), _pp(B), dirn, _ptr(radii), len(radii) - 1,
                                    _ptr(kpll), len(kpll) - 1, _ptr(out))
        return out

    def ring_to_ring(self, V, radii, angles, dirn=3):
        """[giver shell, giver sector, receiver shell, receiver sector]"""
        radii = np.ascontiguousarray(radii, np.float64)
        angles = np.ascontiguousarray(angles, np.float64)
        ns, na = len(radii), len(angles)
        out = np.zeros((ns, na, ns, na))
        self.L.or_ring_to_ring(ctypes.byref(self.g), _pp(V), dirn, _ptr(radii), ns - 1, _ptr(angles), na - 1,
                               _ptr(out))
        return out

    def cyl_ring_to_ring(self, V, radii, kpll, dirn=3):
        radii = np.ascontiguousarray(radii, np.float64)
        kpll = np.ascontiguousarray(kpll, np.float64)
        ns, nb = len(radii), len(kpll)
        out = np.zeros((ns, nb, ns, nb))
        self.L.or_cyl_ring_to_ring(ctypes.byref(self.g), _pp(V), dirn, _ptr(radii), ns - 1, _ptr(kpll), nb - 1,
                                   _ptr(out))
        return out

    # ---- SFF basis / Rayleigh-Benard ----
    def sff_forward(self, parity, real):
        real = np.ascontiguousarray(real, np.float64)
        out = self.czeros()
        self.L.or_sff_forward(ctypes.byref(self.g), parity, _ptr(real), _ptr(out))
        return out

    def sff_inverse(self, parity, spec):
        spec = np.ascontiguousarray(spec, np.complex128)
        out = np.zeros(self.rshape, np.float64)
        self.L.or_sff_inverse(ctypes.byref(self.g), parity, _cptr(spec), _ptr(out))
        return out

    def modes_in_shell(self, imax):
        c = np.zeros(imax + 1)
        self.L.or_modes_in_shell(ctypes.byref(self.g), imax, _ptr(c))
        return c

    def azimuthal_angle(self, lx, ly, lz, dirn=3):
        return self.L.or_azimuthal_angle(ctypes.byref(self.g), lx, ly, lz, dirn)

    def get_dt(self, V, courant, dt_fixed, sff=False):
        """FluidVF::Get_dt with dt_option 1 (dealiases V in place): (dt, umax[3])."""
        um = np.zeros(3)
        dt = self.L.or_get_dt(ctypes.byref(self.g), 1 if sff else 0, _pp(V), courant, dt_fixed, _ptr(um))
        return dt, um

    def sff_dealias(self, A):
        self.L.or_sff_dealias(ctypes.byref(self.g), _cptr(A))

    def sff_reality(self, A, mode):
        self.L.or_sff_reality(ctypes.byref(self.g), _cptr(A), mode)

    def sff_energy(self, A, B=None):
        return self.L.or_sff_energy(ctypes.byref(self.g), _cptr(A), _cptr(A if B is None else B))

    def sff_kmag(self, lx, ly, lz):
        return self.L.or_sff_kmag(ctypes.byref(self.g), lx, ly, lz)

    def compute_nlin_rbc(self, V, T):
        N = [self.czeros() for _ in range(3)]
        NT = self.czeros()
        self.L.or_compute_nlin_rbc(ctypes.byref(self.g), _pp(V), _cptr(T), _pp(N), _cptr(NT))
        return N, NT

    def compute_rhs_rbc(self, Fu1, FT, N, NT):
        self.L.or_compute_rhs_rbc(ctypes.byref(self.g), _cptr(Fu1), _cptr(FT), _pp(N), _cptr(NT))

    def rk4_step_rbc(self, prm, V, T):
        self.L.or_rk4_step_rbc(ctypes.byref(self.g), ctypes.byref(prm), _pp(V), _cptr(T))

    def nusselt(self, prm, V1, T):
        return self.L.or_nusselt(ctypes.byref(self.g), ctypes.byref(prm), _cptr(V1), _cptr(T))

    def init_field(self, kind=0, k0=4.0, kmin=2.0, kmax=170.0, energy=0.5, seed=1):
        V = [self.czeros() for _ in range(3)]
        self.L.or_init_field(ctypes.byref(self.g), kind, k0, kmin, kmax, energy, seed, _pp(V))
        return V


def fft1d(x, sign):
    x = np.array(x, dtype=np.complex128, copy=True)
    lib().or_fft1d(_ptr(x), x.size, sign)
    return x
