"""ORACLE — test infrastructure only (ctypes front end of liboracle.so).

The CPU restatement of the K17 path lives in oracle.cpp / oracle_k17.h; this
module loads it and, where the reference tree was present at build time, the
reference's own header-only arithmetic compiled into oracle/_ref/.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg use it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None

# BC kinds of orc_bc_post (oracle.cpp)
BC_NOSLIP_BOUNCEBACK = 0
BC_VELOCITY_BOUNCEBACK = 1
BC_NOSLIP_INTERPOLATED = 2
BC_VELOCITY_INTERPOLATED = 3
BC_VELOCITY_PRESSURE_INTERPOLATED = 4


def build(with_ref: bool | None = None) -> None:
    """Compile liboracle.so (and oracle/_ref/ when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/src/lbm")
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _ctype(dtype):
    return ctypes.c_double if np.dtype(dtype) == np.float64 else ctypes.c_float


def _sfx(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def k17_nodes(fin, omega, limiters=(0.01, 0.01, 0.01), force=(0.0, 0.0, 0.0), dtype=np.float64):
    """K17 collision of n independent nodes, fin shape (n, 27). Returns (fout, macro[n,4])."""
    fin = np.ascontiguousarray(fin, dtype=dtype)
    n = fin.shape[0]
    fout = np.empty_like(fin)
    macro = np.empty((n, 4), dtype=dtype)
    lim = np.asarray(limiters, dtype=dtype)
    frc = np.asarray(force, dtype=dtype)
    getattr(lib(), "orc_k17_nodes_" + _sfx(dtype))(
        _p(fin), _p(fout), ctypes.c_longlong(n), _ctype(dtype)(omega), _p(lim), _p(frc), _p(macro))
    return fout, macro


def chimera(kind, abc, u, u2, kinv=0.0, k=0.0):
    a = np.ascontiguousarray(abc, dtype=np.float64).copy()
    lib().orc_chimera_f64(ctypes.c_int(kind), _p(a), ctypes.c_double(u), ctypes.c_double(u2),
                          ctypes.c_double(kinv), ctypes.c_double(k))
    return a


def moments(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty(8)
    lib().orc_moments_f64(_p(f), _p(out))
    return out


def eso_slot(d, even):
    return lib().orc_eso_slot(ctypes.c_int(d), ctypes.c_int(1 if even else 0))


class Lattice:
    """One population buffer of 27*N values (SoA) plus +x/+y/+z neighbour arrays."""

    def __init__(self, nbx, nby, nbz, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.nbx = np.ascontiguousarray(nbx, dtype=np.uint32)
        self.nby = np.ascontiguousarray(nby, dtype=np.uint32)
        self.nbz = np.ascontiguousarray(nbz, dtype=np.uint32)
        self.N = int(self.nbx.shape[0])
        self.f = np.zeros(27 * self.N, dtype=self.dtype)
        self.even = True
        self._fn = lambda name: getattr(lib(), name + "_" + _sfx(self.dtype))
        self._T = _ctype(self.dtype)

    def _nb(self):
        return (_p(self.f), ctypes.c_ulonglong(self.N), _p(self.nbx), _p(self.nby), _p(self.nbz))

    def collide(self, fluid, omega, limiters=(0.01, 0.01, 0.01), force=(0.0, 0.0, 0.0), macro=None, porosity=None):
        fluid = np.ascontiguousarray(fluid, dtype=np.uint32)
        lim = np.asarray(limiters, dtype=self.dtype)
        frc = np.asarray(force, dtype=self.dtype)
        if porosity is not None:
            assert macro is None
            por = np.ascontiguousarray(porosity, dtype=self.dtype)
            self._fn("orc_collide_porous")(*self._nb(), _p(fluid), ctypes.c_ulonglong(fluid.size), self._T(omega),
                                           _p(lim), _p(frc), _p(por), ctypes.c_int(int(self.even)))
            return
        r = vx = vy = vz = None
        if macro is not None:
            r, vx, vy, vz = macro
        self._fn("orc_collide_indexed")(*self._nb(), _p(fluid), ctypes.c_ulonglong(fluid.size), self._T(omega),
                                        _p(lim), _p(frc), ctypes.c_int(int(self.even)), _p(r), _p(vx), _p(vy),
                                        _p(vz))

    def collide_body_force(self, fluid, omega, forces, force_factor, body_force, limiters=(0.01, 0.01, 0.01),
                           macro=None):
        """runCollision<..., applyBodyForce=true> (RunCollision.cuh:87-99): per node the half-step
        shift (forces + body_force[k]) * 1/2 * force_factor; body_force (three arrays of N values,
        modified in place) is reset to zero at every collided node."""
        fluid = np.ascontiguousarray(fluid, dtype=np.uint32)
        lim = np.asarray(limiters, dtype=self.dtype)
        frc = np.asarray(forces, dtype=self.dtype)
        for b in body_force:
            assert b.dtype == self.dtype and b.flags.c_contiguous and b.size == self.N
        r = vx = vy = vz = None
        if macro is not None:
            r, vx, vy, vz = macro
        self._fn("orc_collide_bodyforce")(*self._nb(), _p(fluid), ctypes.c_ulonglong(fluid.size), self._T(omega),
                                          _p(lim), _p(frc), self._T(force_factor), *(_p(b) for b in body_force),
                                          ctypes.c_int(int(self.even)), _p(r), _p(vx), _p(vy), _p(vz))

    def bc_post(self, kind, k, q, omega=0.0, velocity=None):
        k = np.ascontiguousarray(k, dtype=np.int32)
        q = np.ascontiguousarray(q, dtype=self.dtype)
        nbc = k.size
        assert q.size == 27 * nbc
        vx = vy = vz = None
        if velocity is not None:
            vx, vy, vz = (np.ascontiguousarray(v, dtype=self.dtype) for v in velocity)
        self._fn("orc_bc_post")(ctypes.c_int(kind), *self._nb(), _p(k), _p(q), ctypes.c_ulonglong(nbc), _p(vx),
                                _p(vy), _p(vz), self._T(omega), ctypes.c_int(int(self.even)))

    def bc_pressure(self, k, kN, rho_bc, direction):
        k = np.ascontiguousarray(k, dtype=np.int32)
        kN = np.ascontiguousarray(kN, dtype=np.int32)
        rho_bc = np.ascontiguousarray(rho_bc, dtype=self.dtype)
        self._fn("orc_bc_pressure_neq")(*self._nb(), _p(k), _p(kN), _p(rho_bc), ctypes.c_ulonglong(k.size),
                                        ctypes.c_int(int(self.even)), ctypes.c_int(direction))

    def init_eq(self, rho, ux, uy, uz, geo=None):
        arrs = [np.ascontiguousarray(a, dtype=self.dtype) for a in (rho, ux, uy, uz)]
        g = None if geo is None else np.ascontiguousarray(geo, dtype=np.uint32)
        self._fn("orc_init_eq")(*self._nb(), _p(g), *(_p(a) for a in arrs), ctypes.c_int(int(self.even)))

    def init_neq(self, rho, ux, uy, uz, nbinv, omega, geo=None):
        """Equilibrium plus the velocity-gradient non-equilibrium part (needs the (-1,-1,-1)
        neighbour array nbinv)."""
        arrs = [np.ascontiguousarray(a, dtype=self.dtype) for a in (rho, ux, uy, uz)]
        g = None if geo is None else np.ascontiguousarray(geo, dtype=np.uint32)
        inv = np.ascontiguousarray(nbinv, dtype=np.uint32)
        self._fn("orc_init_neq")(*self._nb(), _p(inv), _p(g), *(_p(a) for a in arrs), self._T(omega),
                                 ctypes.c_int(int(self.even)))

    def macroscopic(self, geo=None):
        g = None if geo is None else np.ascontiguousarray(geo, dtype=np.uint32)
        out = [np.empty(self.N, dtype=self.dtype) for _ in range(5)]
        self._fn("orc_macroscopic")(*self._nb(), _p(g), ctypes.c_int(int(self.even)), *(_p(a) for a in out))
        return out  # drho, vx, vy, vz, p

    def swap(self):
        self.even = not self.even


def drag_lift(lattice, bc_k, q27, post):
    """Per-node momentum-exchange halves (DragLiftPost27 with post=True after the collision,
    DragLiftPre27 after the BC kernels and the swap), at the lattice's current parity:
    an array (3, nbc) of doubles."""
    k = np.ascontiguousarray(bc_k, dtype=np.int32)
    q = np.ascontiguousarray(q27, dtype=lattice.dtype)
    nbc = k.size
    out = np.zeros((3, nbc))
    lattice._fn("orc_drag_lift")(*lattice._nb(), _p(k), _p(q), ctypes.c_ulonglong(nbc), ctypes.c_int(int(lattice.even)),
                                 ctypes.c_int(int(post)), _p(out[0]), _p(out[1]), _p(out[2]))
    return out


def kinetic_energy(nbx, nby, nbz, nbinv, geo, vx, vy, vz, rho):
    """KineticEnergyAnalyzer's sum and node count over fluid nodes with six fluid face
    neighbours, from macroscopic fields (double)."""
    arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (nbx, nby, nbz, nbinv, geo)]
    vals = [np.ascontiguousarray(a, dtype=np.float64) for a in (vx, vy, vz, rho)]
    out = np.zeros(2)
    lib().orc_kinetic_energy_f64(ctypes.c_ulonglong(arrs[0].size), *(_p(a) for a in arrs), *(_p(a) for a in vals),
                                 _p(out))
    return out


def collide_box(f, nx, ny, nz, omega, even, limiters=(0.01, 0.01, 0.01), mask=None):
    """Dense periodic box collision (node k = x + nx*(y + ny*z)), in place on f (27*N)."""
    lim = np.asarray(limiters, dtype=f.dtype)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    getattr(lib(), "orc_collide_box_" + _sfx(f.dtype))(
        _p(f), ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(nz), _p(m), _ctype(f.dtype)(omega), _p(lim),
        None, ctypes.c_int(1 if even else 0))


# ---- the reference's own arithmetic, compiled from its headers (present only where built) ----
_refs = {}


def ref(precision="f64"):
    """ctypes handle of oracle/_ref/libref_<precision>.so, or None where it was not built."""
    if precision not in _refs:
        path = os.path.join(REF_DIR, f"libref_{precision}.so")
        _refs[precision] = ctypes.CDLL(path) if os.path.exists(path) else None
    return _refs[precision]


def ref_k17_nodes(fin, omega, limiters=(0.01, 0.01, 0.01), force=(0.0, 0.0, 0.0), precision="f64"):
    r = ref(precision)
    dtype = np.float64 if precision == "f64" else np.float32
    T = _ctype(dtype)
    fin = np.ascontiguousarray(fin, dtype=dtype)
    n = fin.shape[0]
    fout = np.empty_like(fin)
    macro = np.empty((n, 4), dtype=dtype)
    lim = np.asarray(limiters, dtype=dtype)
    frc = np.asarray(force, dtype=dtype)
    r.ref_k17_nodes(_p(fin), _p(fout), ctypes.c_longlong(n), T(omega), _p(lim), _p(frc), _p(macro))
    return fout, macro


def ref_chimera(kind, abc, u, u2, kinv=0.0, k=0.0, precision="f64"):
    r = ref(precision)
    dtype = np.float64 if precision == "f64" else np.float32
    T = _ctype(dtype)
    a = np.ascontiguousarray(abc, dtype=dtype).copy()
    r.ref_chimera(ctypes.c_int(kind), _p(a), T(u), T(u2), T(kinv), T(k))
    return a


def ref_moments(f, precision="f64"):
    r = ref(precision)
    dtype = np.float64 if precision == "f64" else np.float32
    f = np.ascontiguousarray(f, dtype=dtype)
    out = np.empty(9, dtype=dtype)
    r.ref_moments(_p(f), _p(out))
    return out


# ---- grid refinement --------------------------------------------------------------------------
COEFF_NAMES = ("a000 a100 a010 a001 a200 a020 a002 a110 a101 a011 "
               "b000 b100 b010 b001 b200 b020 b002 b110 b101 b011 "
               "c000 c100 c010 c001 c200 c020 c002 c110 c101 c011 "
               "d000 d100 d010 d001 d110 d101 d011 a111 b111 c111 d111 LaplaceRho").split()


def interp_cell(f8, omega_src, offset=(0.0, 0.0, 0.0), mode=0, omega_dst=1.0, eps=0.5, pos=(0.0, 0.0, 0.0),
                lib_handle=None, fn="orc_interp_cell_f64"):
    """One refinement cell: f8 node-major (8, 27), n = ix + 2 iy + 4 iz. Returns (coeffs[42], f[27])."""
    h = lib_handle or lib()
    f8 = np.ascontiguousarray(f8, dtype=np.float64)
    c = np.empty(42)
    out = np.empty(27)
    D = ctypes.c_double
    getattr(h, fn)(_p(f8), D(omega_src), D(offset[0]), D(offset[1]), D(offset[2]), ctypes.c_int(mode),
                   D(omega_dst), D(eps), D(pos[0]), D(pos[1]), D(pos[2]), _p(c), _p(out))
    return c, out


def interp_cell_vjp(f8, omega_src, g, offset=(0.0, 0.0, 0.0), mode=0, omega_dst=1.0):
    """J^T g of one refinement cell by dual numbers: g (8, 27) of the CF children (mode 0) or
    (27,) of the FC node (mode 1); returns (8, 27)."""
    f8 = np.ascontiguousarray(f8, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.empty((8, 27))
    D = ctypes.c_double
    lib().orc_interp_cell_vjp_f64(_p(f8), D(omega_src), D(offset[0]), D(offset[1]), D(offset[2]), ctypes.c_int(mode),
                                  D(omega_dst), _p(g), _p(out))
    return out


def scale_adjoint(fine_to_coarse, coarse, fine, even_coarse, cells_coarse, cells_fine, offsets, omega_c, omega_f,
                  lam_coarse, lam_fine):
    """Reference adjoint of `scale` (dual numbers per cell), in place on the FP64 adjoint arrays
    lam_coarse / lam_fine (population layout of each level)."""
    cc = np.ascontiguousarray(cells_coarse, dtype=np.uint32)
    cf = np.ascontiguousarray(cells_fine, dtype=np.uint32)
    n = cc.size
    ox = oy = oz = None
    if offsets is not None:
        ox, oy, oz = (np.ascontiguousarray(o, dtype=np.float64) for o in offsets)
    for a in (coarse.f, fine.f, lam_coarse, lam_fine):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    lib().orc_scale_adjoint_f64(ctypes.c_int(int(fine_to_coarse)), *coarse._nb(), *fine._nb(),
                                ctypes.c_int(int(even_coarse)), _p(cc), _p(cf), ctypes.c_ulonglong(n), _p(ox), _p(oy),
                                _p(oz), ctypes.c_double(omega_c), ctypes.c_double(omega_f), _p(lam_coarse),
                                _p(lam_fine))


def ref_interp_cell(f8, omega_src, offset=(0.0, 0.0, 0.0), mode=0, omega_dst=1.0, eps=0.5, pos=(0.0, 0.0, 0.0)):
    return interp_cell(f8, omega_src, offset, mode, omega_dst, eps, pos, lib_handle=ref("f64"), fn="ref_interp_cell")


def scale(fine_to_coarse, coarse, fine, even_coarse, cells_coarse, cells_fine, offsets, omega_c, omega_f):
    """orc_scale on two oracle Lattices (same dtype); offsets = (ox, oy, oz) arrays per cell."""
    assert coarse.dtype == fine.dtype
    T = coarse._T
    cc = np.ascontiguousarray(cells_coarse, dtype=np.uint32)
    cf = np.ascontiguousarray(cells_fine, dtype=np.uint32)
    ox, oy, oz = (np.ascontiguousarray(o, dtype=coarse.dtype) for o in offsets)
    getattr(lib(), "orc_scale_" + _sfx(coarse.dtype))(
        ctypes.c_int(1 if fine_to_coarse else 0), *coarse._nb(), *fine._nb(), ctypes.c_int(1 if even_coarse else 0),
        _p(cc), _p(cf), ctypes.c_ulonglong(cc.size), _p(ox), _p(oy), _p(oz), T(omega_c), T(omega_f))


# ---- adjoint (forward-mode dual numbers through the oracle's collision) ------------------------
def k17_vjp(fin, g, omega, limiters=(0.01, 0.01, 0.01), force=None):
    """(d post / d pre)^T g per node (FP64), plus the adjoint of the velocity shift (n, 3)."""
    fin = np.ascontiguousarray(fin, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    n = fin.shape[0]
    fb = np.empty_like(fin)
    ub = np.empty((n, 3))
    lim = np.asarray(limiters, dtype=np.float64)
    frc = None if force is None else np.asarray(force, dtype=np.float64)
    lib().orc_k17_vjp_f64(_p(fin), _p(g), ctypes.c_longlong(n), ctypes.c_double(omega), _p(lim), _p(frc), _p(fb),
                          _p(ub))
    return fb, ub


def adjoint_step(lattice, lam, fluid, omega, limiters=(0.01, 0.01, 0.01), force=None, shift_adjoint=None,
                 porosity=None, porosity_adjoint=None):
    """Reverse of one Lattice.collide(fluid, ...) step: lattice.f holds the state before the
    step and lattice.even its parity; lam (27*N, FP64) is updated in place. shift_adjoint
    (optional (3, N) array) accumulates the adjoint of each node's velocity shift;
    porosity_adjoint (N, with porosity given) that of the Brinkman porosity."""
    assert lattice.dtype == np.float64 and lam.dtype == np.float64 and lam.flags.c_contiguous
    fluid = np.ascontiguousarray(fluid, dtype=np.uint32)
    lim = np.asarray(limiters, dtype=np.float64)
    frc = None if force is None else np.asarray(force, dtype=np.float64)
    ub = None
    if shift_adjoint is not None:
        ub = np.zeros((lattice.N, 3))
    por = None if porosity is None else np.ascontiguousarray(porosity, dtype=np.float64)
    if porosity_adjoint is not None:
        assert porosity_adjoint.dtype == np.float64 and porosity_adjoint.flags.c_contiguous
    lib().orc_adjoint_step_f64(_p(lattice.f), _p(lam), ctypes.c_ulonglong(lattice.N), _p(lattice.nbx), _p(lattice.nby),
                               _p(lattice.nbz), _p(fluid), ctypes.c_ulonglong(fluid.size), ctypes.c_double(omega),
                               _p(lim), _p(frc), ctypes.c_int(int(lattice.even)), _p(ub), _p(por),
                               _p(porosity_adjoint))
    if shift_adjoint is not None:
        shift_adjoint += ub.T


def bc_post_adjoint(kind, post_lattice, lam, bc_k, q27, omega, velocity=None, ubar=None):
    """Reverse of Lattice.bc_post: post_lattice.f holds the post-collision state (its parity the
    one the BC used); lam (27*N, FP64) is updated in place; ubar ((n, 3), optional) accumulates the
    wall-velocity adjoint."""
    L = post_lattice
    assert L.dtype == np.float64 and lam.dtype == np.float64
    k = np.ascontiguousarray(bc_k, dtype=np.int32)
    q = np.ascontiguousarray(q27, dtype=np.float64)
    vx = vy = vz = None
    if velocity is not None:
        vx, vy, vz = (np.ascontiguousarray(v, dtype=np.float64) for v in velocity)
    lib().orc_bc_post_adjoint_f64(ctypes.c_int(kind), _p(L.f), _p(lam), ctypes.c_ulonglong(L.N), _p(L.nbx), _p(L.nby),
                                  _p(L.nbz), _p(k), _p(q), ctypes.c_ulonglong(k.size), _p(vx), _p(vy), _p(vz),
                                  ctypes.c_double(omega), ctypes.c_int(int(L.even)), _p(ubar))


def bc_pressure_adjoint(lattice, lam, bc_k, kN, rho_bc, direction, rhobar=None):
    """Reverse of Lattice.bc_pressure: lattice.f holds the state after the step (parity the BC
    used); lam updated in place; rhobar (optional) accumulates the RhoBC adjoint."""
    L = lattice
    k = np.ascontiguousarray(bc_k, dtype=np.int32)
    kn = np.ascontiguousarray(kN, dtype=np.int32)
    rb = np.ascontiguousarray(rho_bc, dtype=np.float64)
    lib().orc_bc_pressure_adjoint_f64(_p(L.f), _p(lam), ctypes.c_ulonglong(L.N), _p(L.nbx), _p(L.nby), _p(L.nbz),
                                      _p(k), _p(kn), _p(rb), ctypes.c_ulonglong(k.size), ctypes.c_int(int(L.even)),
                                      ctypes.c_int(direction), _p(rhobar))
