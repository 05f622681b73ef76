"""The FP64 reference adjoint of a three-level hierarchy (multilevel_reference.gradient: the oracle's
per-kernel reference adjoints walked in the reverse order of UpdateGrid27's recursion) against FP64
central differences (fourth order) of the oracle's forward run. This pins the order and the linearisation points
of the level recursion that the GPU's k17_multilevel_steps follows (CPU only)."""
import numpy as np

import multilevel_reference as ref

STEPS, AVG = 20, 4


def test_three_level_reference_adjoint_matches_fp64_differences(orc):
    grids, interfaces, plane, phi2 = ref.three_levels()
    rng = np.random.default_rng(1620)
    phi1 = rng.uniform(0, 0.05, grids[1].N)
    por = [None, phi1, phi2]
    f0s = ref.initial_states(grids, 1621)
    vel = grids[0].bcs["velocity"]
    vx = np.asarray(vel.velocity[0], dtype=np.float64)
    J, lams, ubar, padj = ref.gradient(orc, grids, interfaces, plane, f0s, vx, por, STEPS, AVG)
    assert J == ref.run(orc, grids, interfaces, plane, f0s, vx, por, STEPS, AVG)

    def J_of(f0s_=f0s, vx_=vx, por_=por):
        return ref.run(orc, grids, interfaces, plane, f0s_, vx_, por_, STEPS, AVG)

    def unit(x):
        return x / np.linalg.norm(x)

    def fd4(J_at, h):
        """d/dt J_at(t) at 0, fourth-order central differences (truncation O(h^4)): the derivatives
        here are small (the inflow's is 7e-7) and strongly curved (the porosity derivatives move
        by 4e-3 relative over a step of 6e-4), which second-order differences cannot resolve to
        1e-6 with any step. Agreement to 1e-5 pins the recursion: a wrong order or linearisation
        point of any level's kernel moves these derivatives by far more."""
        return (-J_at(2 * h) + 8 * J_at(h) - 8 * J_at(-h) + J_at(-2 * h)) / (12 * h)

    checks = []
    v = unit(rng.normal(size=vx.size))
    checks.append(("inflow u_x", ubar[:, 0] @ v, fd4(lambda t: J_of(vx_=vx + t * v), 1e-4)))
    for lev, p in ((1, phi1), (2, phi2)):
        w = unit(rng.normal(size=p.size))

        def J_at(t, lev=lev, p=p, w=w):
            q = list(por)
            q[lev] = p + t * w
            return J_of(por_=q)

        checks.append((f"porosity level {lev}", padj[lev] @ w, fd4(J_at, 5e-6)))
    for lev in range(3):
        w = unit(rng.normal(size=f0s[lev].size) * (np.abs(lams[lev]) > 0))

        def J_at(t, lev=lev, w=w):
            q = list(f0s)
            q[lev] = f0s[lev] + t * w
            return J_of(f0s_=q)

        # the finest level's cells that feed the fine -> coarse interpolation make J stiff in its
        # states (the derivative along one such cell moves from 0.1 to 9.9 between steps of 1e-4
        # and 1e-8): a step of 1e-9 resolves it; the coarser levels take 1e-7, where FP64
        # round-off stays below 1e-7 of their smaller differences
        checks.append((f"initial state level {lev}", lams[lev] @ w, fd4(J_at, 1e-9 if lev == 2 else 1e-7)))
    for name, ad, fd in checks:
        print(f"{name}: reference adjoint {ad:.10e} FD {fd:.10e} rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) < 1e-5 * abs(fd), name
