"""All-FP64 restatement of one velocity-based IP step (TEST AND MEASUREMENT INFRASTRUCTURE ONLY).

The reference computes in float32 (Warp `float`).  This module evaluates the same step in
float64 -- kernels, density, Aii, the IP outer loop with its CG, explicit or Weiler18
viscosity, advection and the domain clamp -- so that the float32 GPU path's per-step error
can be stated against a higher-precision evaluation of the same algorithm.  It takes the
neighbour SETS from the float32 oracle's build of the same positions (sets are exact), and
vectorises every row pass over the (row, neighbour) pair list with numpy, on a thread pool over
chunks of whole rows.  Summation order is numpy's (each row's pairs in order, whatever the
thread count), not the reference's: this is a precision yardstick, not a bit-level restatement.

Followed files (reference root): sph_kernel.py:8-61; pressure_base.py:12-203, 394-447;
ip.py:19-404; viscosity.py:9-108; weiler18.py:21-401; framework.py:26-41, 117-136;
boundary.py:27-91.  Only tests/ and bench.py load this module.
"""

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

PI_LIT = 3.14159265359

# Row sums over the pair list run on threads, each over whole rows (numpy releases the GIL in
# the gathers, the arithmetic and bincount); every row sums its pairs in the same order as one
# bincount over all pairs would, so the results do not depend on the thread count.
_POOL = None


def _pool():
    global _POOL
    if _POOL is None:
        _POOL = ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1)))
    return _POOL


def cubic_kernel(r, h):
    """sph_kernel.py:8-26 in float64."""
    k = 8.0 / PI_LIT / (h * h * h)
    q = r / h
    res = np.zeros_like(q)
    a = q <= 0.5
    b = (q > 0.5) & (q <= 1.0)
    res[a] = k * (6.0 * q[a] ** 3 - 6.0 * q[a] ** 2 + 1.0)
    res[b] = k * 2.0 * (1.0 - q[b]) ** 3
    return res


def spiky_grad(r, h):
    """sph_kernel.py:52-61 in float64; r is (P, 3)."""
    k = 45.0 / (PI_LIT * h ** 6)
    rn = np.maximum(np.linalg.norm(r, axis=1), 1e-6)
    t = h - rn
    return (-k * t * t / rn)[:, None] * r


class Pairs:
    """(row, neighbour) pairs of the dynamic rows, from the reference-layout cache."""

    def __init__(self, count, ids, C, is_dynamic):
        n = count.shape[0]
        cnt = np.minimum(count, C).astype(np.int64)
        rows = np.nonzero(is_dynamic == 1)[0]
        ids2 = ids.reshape(n, C)
        mask = np.arange(C)[None, :] < cnt[rows, None]
        self.i = np.repeat(rows, cnt[rows])
        self.j = ids2[rows][mask].astype(np.int64)
        self.n = n
        # pair ranges of whole rows (the pairs are in row order), about 4 per thread
        P = self.i.shape[0]
        K = max(1, min(64, P // 65536))
        cuts = [0]
        for k in range(1, K):
            c = int(np.searchsorted(self.i, self.i[min(P - 1, k * P // K)], side="left"))
            if c > cuts[-1]:
                cuts.append(c)
        cuts.append(P)
        self.chunks = [(a, b) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]

    def row_sums(self, weights, width=0):
        """sum over each row's pairs of weights(sl) (per-pair values of the pair slice sl, shape
        (len,) or (len, width)); rows without pairs get 0."""
        out = np.zeros((self.n, width) if width else self.n)

        def task(ch):
            a, b = ch
            r0, r1 = int(self.i[a]), int(self.i[b - 1]) + 1
            w = weights(slice(a, b))
            loc = self.i[a:b] - r0
            if width:
                for c in range(width):
                    out[r0:r1, c] = np.bincount(loc, weights=w[:, c], minlength=r1 - r0)
            else:
                out[r0:r1] = np.bincount(loc, weights=w, minlength=r1 - r0)

        list(_pool().map(task, self.chunks))
        return out


def _sum_rows(n, i, w):
    if w.ndim == 1:
        return np.bincount(i, weights=w, minlength=n)
    return np.stack([np.bincount(i, weights=w[:, a], minlength=n) for a in range(w.shape[1])], axis=1)


class Step64:
    """One step of framework.py:423-544 (velocity-based branch) in float64."""

    def __init__(self, x, m, m_inv, density0, is_dynamic, material, h, pairs, material_fluid=1):
        self.x = np.asarray(x, np.float64)
        self.m = np.asarray(m, np.float64)
        self.m_inv = np.asarray(m_inv, np.float64)
        self.rho0 = np.asarray(density0, np.float64)
        self.dyn = np.asarray(is_dynamic) == 1
        self.fluid = np.asarray(material) == material_fluid
        self.h = float(h)
        self.P = pairs
        i, j = pairs.i, pairs.j
        P = i.shape[0]
        # per-pair arrays, filled chunk by chunk on the pool (disjoint slices)
        self.xij = np.empty((P, 3))
        self.rn = np.empty(P)
        self.g = np.empty((P, 3))
        self.dynj = np.empty(P)
        self.mj = np.empty(P)
        self.mi = np.empty(P)

        def fill(ch):
            sl = slice(*ch)
            self.xij[sl] = self.x[i[sl]] - self.x[j[sl]]
            self.rn[sl] = np.linalg.norm(self.xij[sl], axis=1)
            self.g[sl] = spiky_grad(self.xij[sl], self.h)
            self.dynj[sl] = self.dyn[j[sl]]
            self.mj[sl] = self.m[j[sl]]
            self.mi[sl] = self.m[i[sl]]

        list(_pool().map(fill, pairs.chunks))
        n = pairs.n
        h, mj, dynj, g, rn = self.h, self.mj, self.dynj, self.g, self.rn
        # compute_density_kernel, pressure_base.py:12-46
        rho = self.m * cubic_kernel(np.zeros(n), h) + pairs.row_sums(lambda sl: mj[sl] * cubic_kernel(rn[sl], h))
        self.density = np.where(self.dyn, rho, self.rho0)
        # compute_Aii_kernel, pressure_base.py:49-87
        minvj = lambda sl: self.m_inv[j[sl]]

        def a1w(sl):
            Jij = mj[sl][:, None] * g[sl]
            return (Jij * Jij).sum(1) * minvj(sl) * dynj[sl]

        a1 = pairs.row_sums(a1w)
        Jii = -pairs.row_sums(lambda sl: mj[sl][:, None] * g[sl], 3)
        aii = a1 + (Jii * Jii).sum(1) * self.m_inv
        self.k = np.where(self.dyn, 1.0 / (aii + 1e-3), 0.0)

    # -- IP (ip.py) -------------------------------------------------------------------------
    def J(self, d):
        """compute_J_x_kernel form: sum_j m_j (d_i - [dyn_j] d_j) . gradW_ij"""
        i, j, mj, dynj, g = self.P.i, self.P.j, self.mj, self.dynj, self.g

        def w(sl):
            dl = d[i[sl]] - dynj[sl, None] * d[j[sl]]
            return mj[sl] * (dl * g[sl]).sum(1)

        return np.where(self.dyn, self.P.row_sums(w), 0.0)

    def JT(self, y):
        """sum_j (m_j y_i + [dyn_j] m_i y_j) gradW_ij"""
        i, j, mj, mi, dynj, g = self.P.i, self.P.j, self.mj, self.mi, self.dynj, self.g

        def w(sl):
            s = mj[sl] * y[i[sl]] + dynj[sl] * mi[sl] * y[j[sl]]
            return s[:, None] * g[sl]

        out = self.P.row_sums(w, 3)
        out[~self.dyn] = 0.0
        return out

    def ip_solve(self, v, dt, omega=10.0, tol_opt=1e-3, max_iteration_opt=1000, use_fixed_iters=False,
                 tol_pcg=1e-6, max_iteration_pcg=1000):
        dyn = self.dyn
        c = np.where(dyn, self.density - self.rho0, 0.0)
        d = np.where(dyn[:, None], dt * v, 0.0)
        nd = max(int(dyn.sum()), 1)
        errs, pcg = [], []
        opt_iter = 0
        while True:
            t = self.J(d) + c
            f = np.maximum(t, 0.0)
            dfdt = (t >= 1e-6).astype(np.float64)
            err = float(np.maximum(f[dyn] / np.maximum(self.rho0[dyn], 1e-6), 0.0).sum() / nd)
            errs.append(err)
            if (err < tol_opt and not use_fixed_iters and opt_iter >= 1) or opt_iter >= max_iteration_opt:
                break
            rhs = omega * self.JT(f * self.k)
            s = dfdt * self.k

            def A(p):
                return np.where(dyn[:, None], self.m[:, None] * p + omega * self.JT(s * self.J(p)), 0.0)

            sol = np.zeros_like(d)
            r = rhs.copy()
            rr_old = float((r[dyn] * r[dyn]).sum())
            it = 0
            if rr_old > 1e-30:
                p = r.copy()
                while it < max_iteration_pcg:
                    Ap = A(p)
                    pAp = float((p[dyn] * Ap[dyn]).sum())
                    if pAp < 0.0:
                        raise RuntimeError("[IP] non-positive definite matrix (FP64 restatement)")
                    alpha = rr_old / max(pAp, 1e-30)
                    sol += alpha * p
                    r -= alpha * Ap
                    rr_new = float((r[dyn] * r[dyn]).sum())
                    it += 1
                    if (rr_new < tol_pcg and it >= 2) or it >= max_iteration_pcg:
                        break
                    p = r + (rr_new / max(rr_old, 1e-30)) * p
                    rr_old = rr_new
            pcg.append(it)
            d = d - sol
            opt_iter += 1
        v_new = np.where(dyn[:, None], d / dt, v)
        return v_new, dict(outer_iters=opt_iter, pcg_iters=pcg, density_errors=errs, final_density_error=errs[-1])

    # -- viscosity ---------------------------------------------------------------------------
    def viscosity_explicit(self, v, coeff, dt):
        """viscosity.py:9-108, every row reading the pre-step velocities."""
        i, j = self.P.i, self.P.j
        ok = self.fluid[j] & self.fluid[i]
        eps2 = (0.01 * self.h) ** 2
        vij = v[i] - v[j]
        vxy = (vij * self.xij).sum(1)
        K = 2.0 * self.rho0[i] / (self.density[i] + self.density[j])
        s = 10.0 * coeff * (self.m[j] / self.density[j]) * (vxy / (self.rn ** 2 + eps2)) * ok
        acc = _sum_rows(self.P.n, i, (K * s)[:, None] * self.g)
        rows = self.dyn & self.fluid
        return np.where(rows[:, None], v + acc * dt, v)

    def weiler18(self, v, coeff, dt):
        """weiler18.py:21-401: block-Jacobi PCG, tol 1e-5 on r.r, cap 1000."""
        i, j = self.P.i, self.P.j
        n = self.P.n
        rows = self.dyn & self.fluid
        eps2 = 0.01 * self.h * self.h
        fac = dt * coeff * self.rho0 / self.density
        cij = np.empty(i.shape[0])

        def fill(ch):
            sl = slice(*ch)
            ok = self.fluid[j[sl]].astype(np.float64)
            cij[sl] = 10.0 * (self.mj[sl] / self.density[j[sl]]) / (self.rn[sl] ** 2 + eps2) * ok

        list(_pool().map(fill, self.P.chunks))

        xij, g = self.xij, self.g

        def A(u):
            def wts(sl):
                uij = u[i[sl]] - u[j[sl]]
                w = cij[sl] * (uij * xij[sl]).sum(1)
                return w[:, None] * g[sl]

            lap = self.P.row_sums(wts, 3)
            return np.where(rows[:, None], u - fac[:, None] * lap, u)

        L = self.P.row_sums(lambda sl: (cij[sl, None, None] * g[sl, :, None] * xij[sl, None, :]).reshape(-1, 9),
                            9).reshape(n, 3, 3)
        M = np.eye(3)[None] - fac[:, None, None] * L
        Pinv = np.linalg.inv(np.where(rows[:, None, None], M, np.eye(3)[None]))

        def prec(r):
            return np.einsum("nab,nb->na", Pinv, r)

        dyn = self.dyn
        r = np.where(dyn[:, None], v - A(v), 0.0)
        z = prec(r)
        p = z.copy()
        dv = np.zeros_like(v)
        rz_old = float((r[dyn] * z[dyn]).sum())
        it = 0
        if rz_old > 1e-12:
            while it < 1000:
                Ap = A(p)
                pAp = float((p[dyn] * Ap[dyn]).sum())
                if pAp <= 0.0:
                    break
                alpha = rz_old / max(pAp, 1e-12)
                dv += alpha * p
                r -= alpha * Ap
                rn = float((r[dyn] * r[dyn]).sum())
                it += 1
                if (rn < 1e-5 and it >= 2) or it >= 1000:
                    break
                z = prec(r)
                rz_new = float((r[dyn] * z[dyn]).sum())
                p = z + (rz_new / max(rz_old, 1e-12)) * p
                rz_old = rz_new
        return np.where(dyn[:, None], v + dv, v), it


def step(x, v, m, m_inv, density0, is_dynamic, material, h, count, ids, C, dt, gravity, viscosity_method,
         viscosity_coeff, domain_size, ip_kw):
    """One Framework.forward in float64 from float32 state and neighbour sets.
    Returns (v after the IP solve, x after advection, v after advection, IP stats)."""
    pairs = Pairs(count, ids, C, is_dynamic)
    s = Step64(x, m, m_inv, density0, is_dynamic, material, h, pairs)
    dyn = s.dyn
    v = np.asarray(v, np.float64).copy()
    v[dyn] += np.asarray(gravity, np.float64) * dt                      # framework.py:26-41
    if viscosity_method == 1:
        v, _ = s.weiler18(v, viscosity_coeff, dt)
    elif viscosity_coeff != 0.0:
        v = s.viscosity_explicit(v, viscosity_coeff, dt)
    v, st = s.ip_solve(v, dt, **ip_kw)
    v_solved = v.copy()
    x = s.x.copy()
    v[~dyn] = 0.0                                                         # framework.py:117-136
    x += dt * v
    L = np.asarray(domain_size, np.float64)
    pad = h
    nrm = np.zeros_like(x)
    for a in range(3):                                                    # boundary.py:27-91
        hi = dyn & (x[:, a] > L[a] - pad)
        nrm[hi, a] += 1.0
        x[hi, a] = L[a] - pad
        lo = dyn & (x[:, a] <= pad)
        nrm[lo, a] -= 1.0
        x[lo, a] = pad
    ln = np.linalg.norm(nrm, axis=1)
    hit = ln > 1e-6
    if hit.any():
        cn = nrm[hit] / ln[hit, None]
        sv = 1.5 * (v[hit] * cn).sum(1)
        v[hit] = v[hit] - sv[:, None] * cn
    return v_solved, x, v, st
