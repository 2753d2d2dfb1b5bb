"""Tetrahedron quadrature rules used to build the reference tables.

The reference extracts MFEM's ``IntRules.Get(TETRAHEDRON, 2p-2)`` and
``IntRules.Get(TETRAHEDRON, 2p)`` (solver/asm_data.inc:108-116) and hands the
points and weights to the kernels as run-time tables.  MFEM is not available
here, so this module produces rules of the same polynomial order with the
same point counts MFEM's symmetric (Keast-type) rules have for orders 0-8
(1/4/11/24/45 points).  The orbit parameters are refined by Newton iteration
on the moment equations until every monomial of total degree <= order
integrates to machine precision against the closed form
i! j! k! / (i+j+k+3)!.

The degree-8 rule (the order-4 mass rule) is Keast's 45-point rule: the
centroid with a negative weight, two S31, two S22 and two S211 orbits -- 15
unknowns for the 15 S4-invariant moment conditions of degree <= 8.  Its
starting values below came from a multi-start Newton search over the orbit
parameters of that structure (the S31 and the first S22 parameters seeded at
Keast's values); the solution's weights agree with Keast's table.

Orders above 8 use a collapsed Gauss-Jacobi (Stroud conical product) rule.

The kernels never assume a particular rule: points and weights are inputs,
exactly as in the reference (solver/asm_gpu.inc:70-77).
"""

from __future__ import annotations

import itertools
from functools import lru_cache
from math import factorial

import numpy as np

# Reference tetrahedron: (0,0,0), (1,0,0), (0,1,0), (0,0,1); volume 1/6.


def monomial_exponents(order: int) -> list[tuple[int, int, int]]:
    return [(i, j, k) for s in range(order + 1)
            for i in range(s, -1, -1) for j in range(s - i, -1, -1)
            for k in [s - i - j]]


def exact_moment(i: int, j: int, k: int) -> float:
    """Integral of x^i y^j z^k over the reference tetrahedron."""
    return factorial(i) * factorial(j) * factorial(k) / factorial(i + j + k + 3)


def _orbit_points(kind: str, par: np.ndarray) -> np.ndarray:
    """Barycentric points (n x 4) of one S4 orbit."""
    if kind == "s4":
        base = [(0.25, 0.25, 0.25, 0.25)]
    elif kind == "s31":
        a = par[0]
        base = [(a, a, a, 1.0 - 3.0 * a)]
    elif kind == "s22":
        a = par[0]
        base = [(a, a, 0.5 - a, 0.5 - a)]
    elif kind == "s211":
        a, b = par[0], par[1]
        base = [(a, a, b, 1.0 - 2.0 * a - b)]
    elif kind == "s1111":
        a, b, c = par[0], par[1], par[2]
        base = [(a, b, c, 1.0 - a - b - c)]
    else:
        raise ValueError(kind)
    pts = set()
    out = []
    for p in base:
        for perm in itertools.permutations(range(4)):
            q = tuple(p[perm[m]] for m in range(4))
            key = tuple(round(v, 12) for v in q)
            if key not in pts:
                pts.add(key)
                out.append(q)
    return np.array(out, dtype=np.float64)


_NPAR = {"s4": 0, "s31": 1, "s22": 1, "s211": 2, "s1111": 3}


def _assemble(orbits, x):
    """orbits: list of kind; x: flat [params..., weight] per orbit."""
    pts, wts = [], []
    o = 0
    for kind in orbits:
        npar = _NPAR[kind]
        par = x[o:o + npar]
        w = x[o + npar]
        o += npar + 1
        b = _orbit_points(kind, par)
        pts.append(b)
        wts.append(np.full(len(b), w))
    bary = np.vstack(pts)
    # barycentric (l0,l1,l2,l3) for vertices V0..V3 -> x = l1, y = l2, z = l3
    return bary[:, 1:4].copy(), np.concatenate(wts)


def _residual(orbits, x, order):
    pts, w = _assemble(orbits, x)
    res = []
    for (i, j, k) in monomial_exponents(order):
        res.append(np.sum(w * pts[:, 0] ** i * pts[:, 1] ** j * pts[:, 2] ** k)
                   - exact_moment(i, j, k))
    return np.array(res)


def _refine(orbits, x0, order, iters=60):
    x = np.array(x0, dtype=np.float64)
    for _ in range(iters):
        r = _residual(orbits, x, order)
        if np.max(np.abs(r)) < 1e-17:
            break
        jac = np.empty((len(r), len(x)))
        for m in range(len(x)):
            h = 1e-7 * max(1.0, abs(x[m]))
            xp = x.copy()
            xp[m] += h
            xm = x.copy()
            xm[m] -= h
            jac[:, m] = (_residual(orbits, xp, order) - _residual(orbits, xm, order)) / (2 * h)
        dx, *_ = np.linalg.lstsq(jac, -r, rcond=None)
        x = x + dx
    return x


# Starting values (published Keast orbit parameters, weights for volume 1/6);
# refined below to full double precision on the moment equations.
_KEAST = {
    2: (["s31"], [0.1381966011250105, 1.0 / 24.0]),
    4: (["s4", "s31", "s22"],
        [-0.01315555555555556, 0.0714285714285714, 0.007622222222222222,
         0.399403576166799, 0.02488888888888889]),
    6: (["s31", "s31", "s31", "s211"],
        [0.214602871259151684, 0.00665379170969464506,
         0.0406739585346113397, 0.00167953517588677620,
         0.322337890142275646, 0.00922619692394239843,
         0.0636610018750175299, 0.269672331458315867, 0.00803571428571428248]),
    8: (["s4", "s31", "s31", "s22", "s22", "s211", "s211"],
        [-0.039327006641137995,
         0.12747093656672795, 0.0040813160593389216,
         0.032078830392628654, 0.0006580867733034576,
         0.049777095643258565, 0.004384258825121977,
         0.31626955260150313, 0.013830063842480502,
         0.037970048471853264, 0.7303134278075847, 0.00223873973961593,
         0.2319010893971481, 0.5132800333608629, 0.004240437424685757]),
}


def _conical_rule(order: int):
    """Collapsed Gauss-Jacobi product rule exact for degree `order`."""
    from scipy.special import roots_jacobi
    n = (order + 2) // 2 + 0
    n = max(1, (order + 1 + 1) // 2)
    # x1 with weight (1-t)^2, x2 with (1-t)^1, x3 with (1-t)^0 on [0,1]
    t1, w1 = roots_jacobi(n, 2.0, 0.0)
    t2, w2 = roots_jacobi(n, 1.0, 0.0)
    t3, w3 = roots_jacobi(n, 0.0, 0.0)
    # map [-1,1] -> [0,1]: (1-s)^a ds -> factor 2^{-(a+1)}
    s1, c1 = (t1 + 1) / 2, w1 / 8.0
    s2, c2 = (t2 + 1) / 2, w2 / 4.0
    s3, c3 = (t3 + 1) / 2, w3 / 2.0
    pts, wts = [], []
    for a in range(n):
        for b in range(n):
            for c in range(n):
                x = s1[a]
                y = s2[b] * (1 - x)
                z = s3[c] * (1 - x) * (1 - s2[b])
                pts.append((x, y, z))
                wts.append(c1[a] * c2[b] * c3[c])
    return np.array(pts), np.array(wts)


@lru_cache(maxsize=None)
def tet_rule(order: int) -> tuple[np.ndarray, np.ndarray]:
    """(points nq x 3, weights nq) exact for total degree <= order."""
    order = max(order, 0)
    if order <= 1:
        return np.array([[0.25, 0.25, 0.25]]), np.array([1.0 / 6.0])
    key = order + (order % 2)  # odd orders use the next even rule
    if key in _KEAST:
        orbits, x0 = _KEAST[key]
        x = _refine(orbits, x0, key)
        pts, w = _assemble(orbits, x)
        err = np.max(np.abs(_residual(orbits, x, key)))
        if err > 1e-15:
            raise RuntimeError(f"tet rule order {key} did not converge ({err})")
        return pts, w
    return _conical_rule(order)


def rule_error(order: int, test_order: int | None = None) -> float:
    pts, w = tet_rule(order)
    test_order = order if test_order is None else test_order
    err = 0.0
    for (i, j, k) in monomial_exponents(test_order):
        v = np.sum(w * pts[:, 0] ** i * pts[:, 1] ** j * pts[:, 2] ** k)
        err = max(err, abs(v - exact_moment(i, j, k)) / exact_moment(i, j, k))
    return err
