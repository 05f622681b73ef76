"""FP64 reference of a refinement hierarchy and of its adjoint, built from the oracle's pieces.

The forward follows UpdateGrid27::updateGrid level by level (src/gpu/core/Calculation/
UpdateGrid27.cpp:54-96; a level lets its finer level take two steps, collides, applies its BCs,
swaps, applies the pressure BC, then refreshes the interface: fine -> coarse, coarse -> fine). The
reverse walks the recorded steps backwards with the oracle's FP64 reference adjoints (forward-mode
dual numbers per kernel): the interface (coarse -> fine, then fine -> coarse), the pressure BC, the
post-collision BCs in reverse order, the collision, then the finer level's two steps. Test
infrastructure: only tests import it.

The case: a channel (velocity inflow, pressure outlet, interpolated no-slip walls) on the coarsest
level, boxes nested level in level, porosity fields on any level, started from a uniform flow.
Objective: the kinetic energy of a coarse plane averaged over the last `avg` of `steps` coarse
steps.
"""
import numpy as np

from helpers import ORACLE_POST_BC, W, equilibrium, pack_pre
from virtualfluids_amd.adjoint import pre_collision_index
from virtualfluids_amd.grid import CX, CY, CZ, channel, gaussian_random_field
from virtualfluids_amd.refinement import nested_box

NU = 0.01
_C = np.stack([np.asarray(CX, float), np.asarray(CY, float), np.asarray(CZ, float)])
VELOCITY_BC = ORACLE_POST_BC["VelocityInterpolatedCompressible"]
NOSLIP_BC = ORACLE_POST_BC["NoSlipInterpolatedCompressible"]


def omega(level):
    """omega of a level whose lattice spacing and time step are 2^-level of the coarse level's."""
    return 1.0 / (3 * NU * 2 ** level + 0.5)


def three_levels():
    """Grids (coarse channel, fine box, finer box), interfaces, the observed coarse plane and a
    porosity field of the finest level."""
    coarse = channel((28, 18, 18), inflow=0.03)
    nb1 = nested_box(None, (6, 4, 4), (16, 14, 14), coarse=coarse)
    nb2 = nested_box(None, (6, 6, 6), (14, 14, 14), coarse=nb1.fine)
    grids = [nb1.coarse, nb2.coarse, nb2.fine]
    interfaces = [(nb1.cf, nb1.fc), (nb2.cf, nb2.fc)]
    x, _, _ = grids[0].coords()
    plane = np.nonzero((x == 18) & (grids[0].geo == 19))[0]  # downstream of the boxes, within reach of the inflow
    g2 = grids[2]
    phi = gaussian_random_field((g2.nx, g2.ny, g2.nz), 2.0, seed=1600, low=0.0, high=0.05)
    return grids, interfaces, plane, phi.astype(np.float32).astype(np.float64)


def initial_states(grids, seed):
    """Per level: the uniform flow (u_x = 0.03) with seeded perturbations of density, velocity and
    the non-equilibrium populations, stored FP32-exact. An exactly uniform equilibrium would start
    every limited third-order cumulant at zero, where the limiter x N / (|x| + L) has a jump in its
    second derivative, so differences of J across that point do not converge."""
    rng = np.random.default_rng(seed)
    out = []
    for g in grids:
        u = np.zeros((g.N, 3))
        u[:, 0] = 0.03
        u += rng.normal(0.0, 3e-3, (g.N, 3))
        f = equilibrium(rng.normal(0.0, 1e-3, g.N), u) + W * rng.normal(0.0, 1e-3, (g.N, 27))
        out.append(pack_pre(f, g, True).astype(np.float32).astype(np.float64))
    return out


def _ke_grad(f):
    """0.5 |j|^2 / (1 + drho) summed over the columns of f (27, M), and its gradient."""
    drho = f.sum(0)
    j = _C @ f
    ir = 1.0 / (1.0 + drho)
    ke = (0.5 * (j * j).sum(0) * ir).sum()
    g = (_C.T @ (j * ir)) - 0.5 * ((j * j).sum(0) * ir * ir)[None, :]
    return ke, g


def run(orc, grids, interfaces, plane, f0s, vx, porosity, steps, avg, record=False, limiters=(0.01, 0.01, 0.01),
        final_states=False):
    """The FP64 forward; returns J (and, with record, what the reverse reads; with final_states,
    every level's buffer after the last step). limiters: the K17 quadric limiters of every level."""
    levels = [orc.Lattice(g.nbx, g.nby, g.nbz) for g in grids]
    for P, f in zip(levels, f0s):
        P.f[:] = f
    g0 = grids[0]
    vel, ns, pr = g0.bcs["velocity"], g0.bcs["noslip"], g0.bcs["pressure"]
    velocity = (vx, vel.velocity[1], vel.velocity[2])
    idx = {e: pre_collision_index(g0, e) for e in (True, False)}
    last = len(grids) - 1

    def update(level):
        subs = [update(level + 1), update(level + 1)] if level < last else []
        P = levels[level]
        rec = {"even": P.even, "before": P.f.copy() if record else None, "subs": subs}
        P.collide(grids[level].fluid, omega(level), limiters=limiters, porosity=porosity[level])
        if level == 0:
            rec["post_velocity"] = P.f.copy() if record else None
            P.bc_post(VELOCITY_BC, vel.k, vel.q27, omega(0), velocity)
            rec["post_noslip"] = P.f.copy() if record else None
            P.bc_post(NOSLIP_BC, ns.k, ns.q27, omega(0))
        P.swap()
        if level == 0:
            P.bc_pressure(pr.k, pr.kN, pr.rho_bc, pr.direction)
        rec["after"] = P.f.copy() if record else None
        if level < last:
            cf, fc = interfaces[level]
            F = levels[level + 1]
            rec["fine"] = F.f.copy() if record else None
            orc.scale(True, P, F, P.even, fc.coarse, fc.fine, [np.zeros(fc.n)] * 3, omega(level), omega(level + 1))
            rec["coarse_after_fc"] = P.f.copy() if record else None
            orc.scale(False, P, F, P.even, cf.coarse, cf.fine, [np.zeros(cf.n)] * 3, omega(level), omega(level + 1))
        return rec

    J, records = 0.0, []
    C = levels[0]
    for step in range(steps):
        records.append(update(0))
        if step >= steps - avg:
            J += _ke_grad(C.f[idx[C.even][:, plane]])[0] / avg
    if final_states:
        return J, [P.f.copy() for P in levels]
    return (J, records) if record else J


def gradient(orc, grids, interfaces, plane, f0s, vx, porosity, steps, avg, limiters=(0.01, 0.01, 0.01)):
    """J and its FP64 reference gradients: (J, [d J / d f0 per level], d J / d (inflow wall
    velocities) as (n, 3), [d J / d porosity per level, None where a level has none])."""
    J, records = run(orc, grids, interfaces, plane, f0s, vx, porosity, steps, avg, record=True, limiters=limiters)
    last = len(grids) - 1
    lams = [np.zeros_like(np.asarray(f, dtype=np.float64)) for f in f0s]
    padj = [None if p is None else np.zeros(grids[i].N) for i, p in enumerate(porosity)]
    g0 = grids[0]
    vel, ns, pr = g0.bcs["velocity"], g0.bcs["noslip"], g0.bcs["pressure"]
    velocity = (vx, vel.velocity[1], vel.velocity[2])
    ubar = np.zeros((vel.n, 3))
    idx = {e: pre_collision_index(g0, e) for e in (True, False)}

    def lattice(level, f, even):
        g = grids[level]
        P = orc.Lattice(g.nbx, g.nby, g.nbz)
        P.f[:] = f
        P.even = even
        return P

    def reverse(level, rec):
        even = rec["even"]
        if level < last:
            cf, fc = interfaces[level]
            C = lattice(level, rec["coarse_after_fc"], not even)
            F = lattice(level + 1, rec["fine"], True)
            orc.scale_adjoint(False, C, F, not even, cf.coarse, cf.fine, None, omega(level), omega(level + 1),
                              lams[level], lams[level + 1])
            orc.scale_adjoint(True, C, F, not even, fc.coarse, fc.fine, None, omega(level), omega(level + 1),
                              lams[level], lams[level + 1])
        if level == 0:
            orc.bc_pressure_adjoint(lattice(0, rec["after"], not even), lams[0], pr.k, pr.kN, pr.rho_bc, pr.direction)
            orc.bc_post_adjoint(NOSLIP_BC, lattice(0, rec["post_noslip"], even), lams[0], ns.k, ns.q27, omega(0))
            orc.bc_post_adjoint(VELOCITY_BC, lattice(0, rec["post_velocity"], even), lams[0], vel.k, vel.q27, omega(0),
                                velocity=velocity, ubar=ubar)
        orc.adjoint_step(lattice(level, rec["before"], even), lams[level], grids[level].fluid, omega(level),
                         limiters=limiters, porosity=porosity[level], porosity_adjoint=padj[level])
        for sub in reversed(rec["subs"]):
            reverse(level + 1, sub)

    for step in reversed(range(steps)):
        rec = records[step]
        if step >= steps - avg:
            cells = idx[not rec["even"]][:, plane]
            _, g = _ke_grad(rec["coarse_after_fc"].reshape(-1)[cells])
            np.add.at(lams[0], cells.reshape(-1), (g / avg).reshape(-1))
        reverse(0, rec)
    return J, lams, ubar, padj
