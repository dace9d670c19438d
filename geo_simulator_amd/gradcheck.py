"""Central finite differences of an inversion loss, without cancellation.

New surface with no reference counterpart: the check of the adjoint
gradients (geo_simulator_amd.adjoint) against central differences of
L(theta) = 0.5 ||y(theta)[obs] - y_obs||^2 in one element parameter theta_e
(log k_e or log E_e) at a time.

Formed directly, a 1e-6 relative change of k_e shares its digits with the
rest of each row of K, and L(theta + h) - L(theta - h) loses the difference
in L's sixteenth digit.  Both are avoided here:
  - each perturbed solution is y0 + dy with K(theta) dy = f - K(theta) y0,
    whose right-hand side is r0 - (K(theta) - K(theta0)) y0, the second term
    taken from the residual difference of an O(1) change of the same elements
    scaled back (K is linear in the element parameters exp(theta_e), the
    loads do not depend on them);
  - the loss difference is 0.5 sum (dy+ - dy-)(2 y0 + dy+ + dy- - 2 y_obs).
"""
from __future__ import annotations

import numpy as np


def corrected_solver(assemble, theta0, y0, cg):
    """theta -> y(theta) - y0 for theta near theta0.

    assemble(theta) -> (K_bc, f_bc) with K linear in exp(theta_e); y0 the
    solution at theta0; cg a ConjugateGradient (its tolerance applies to
    ||f - K(theta) y0||).  With w_e = expm1(theta_e - theta0_e),
        (K(theta) - K(theta0)) y0 = (r0 - r_S) / S,  r0 = f - K(theta0) y0,
    where r_S is the residual of y0 at theta0 + log1p(S w), S scaling the
    largest |S w_e| to 1/4."""
    import torch
    K0, f0 = assemble(theta0)
    r0 = f0 - K0.apply(y0)

    def solve(theta):
        w = torch.expm1(theta - theta0)
        S = 0.25 / float(w.abs().max())
        KS, fS = assemble(theta0 + torch.log1p(S * w))
        d = r0 - (fS - KS.apply(y0))
        K, _ = assemble(theta)
        dy, _ = cg.solve(K, r0 - d / S)
        return dy
    return solve


def fd_central(solve_obs, theta, elems, h, y_obs):
    """Central differences of L in theta_e for e in elems.  solve_obs(theta)
    returns (y(theta) - y0, y0) at the observations."""
    import torch
    out = []
    for e in elems:
        tp = theta.clone()
        tm = theta.clone()
        tp[e] += h
        tm[e] -= h
        dp, y0 = solve_obs(tp)
        dm, _ = solve_obs(tm)
        dl = 0.5 * torch.sum((dp - dm) * (2.0 * y0 + dp + dm - 2.0 * y_obs)).item()
        out.append(dl / (tp[e] - tm[e]).item())  # 2h as represented
    return np.array(out)


def fd_direction(solve_obs, theta, v, h, y_obs):
    """Central difference of L along the direction v (a device tensor)."""
    import torch
    dp, y0 = solve_obs(theta + h * v)
    dm, _ = solve_obs(theta - h * v)
    return 0.5 * torch.sum((dp - dm) * (2.0 * y0 + dp + dm - 2.0 * y_obs)).item() / (2.0 * h)
