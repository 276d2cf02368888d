"""FP64 numpy restatement of the reference's Broyden solver.

TEST INFRASTRUCTURE ONLY (the checker; the product never imports it).

* BroydenRef  broyden_solver (src/cilia/broyden_solver.cpp:9-129):
              find_update (:26-67) and end_of_iter (:69-129), with J0^{-1}
              (swimmer::jacobian_inv_mult, src/cilia/swimmer.cpp:915-1070) given
              as a block-diagonal list of dense blocks.

The update is Broyden's second ("bad") method in limited-memory form: the
inverse-Jacobian approximation is H_k = -c J0^{-1} - sum_{i<k} C_i D_i^T, and
after a step Delta x = H_k e_k the new columns make H_{k+1} satisfy the secant
condition H_{k+1} (e_{k+1} - e_k) = -Delta x.  tests/test_broyden_oracle.py pins
the restatement by that condition, by the dense rank-one recurrence
H_{k+1} = H_k - H_k e_{k+1} Delta e^T / ||Delta e||^2, and by convergence on
linear and mildly nonlinear problems; dexp by the derivative of the matrix
exponential (d/dt exp(u + t v) exp(u)^T = hat(dexp(u, v))) and dexpinv as its
inverse; compute_errors by closed-form states (a segment moving at its
velocity has zero residual) and the storage order.
"""
from __future__ import annotations

import numpy as np


def block_diag_apply(blocks, v):
    """y = blockdiag(blocks) v."""
    out = np.empty_like(v)
    r = 0
    for b in blocks:
        s = b.shape[0]
        out[r:r + s] = b @ v[r:r + s]
        r += s
    return out


class BroydenRef:
    def __init__(self, n, max_iter, confidence, blocks):
        self.n, self.max_iter, self.c = n, max_iter, confidence
        self.blocks = [np.asarray(b, dtype=np.float64) for b in blocks]
        assert sum(b.shape[0] for b in self.blocks) == n
        self.C = np.zeros((n, max_iter))
        self.D = np.zeros((n, max_iter))
        self.error = np.zeros(n)
        self.iter = 0

    def start(self, error):
        """Residual of the initial guess; iter = 0 (cilia_sim_main.cu:394-411)."""
        self.error = np.array(error, dtype=np.float64)
        self.iter = 0

    def _inverse(self, e):
        u = -self.c * block_diag_apply(self.blocks, e)
        for i in range(self.iter):
            u -= (e @ self.D[:, i]) * self.C[:, i]
        return u

    def find_update(self):
        """broyden_solver.cpp:26-67."""
        return self._inverse(self.error)

    def end_of_iter(self, new_error, error_is_too_large):
        """broyden_solver.cpp:69-129."""
        new_error = np.asarray(new_error, dtype=np.float64)
        if error_is_too_large and self.iter < self.max_iter - 1:
            k = self.iter
            ck = self._inverse(new_error)
            dk = new_error - self.error
            nrm = np.linalg.norm(dk)
            self.C[:, k] = ck / nrm
            self.D[:, k] = dk / nrm
            self.error = new_error.copy()
        self.iter += 1

    def dense_inverse(self):
        """H = -c J0^{-1} - C D^T over the stored columns (for checks)."""
        n = self.n
        H = np.zeros((n, n))
        r = 0
        for b in self.blocks:
            s = b.shape[0]
            H[r:r + s, r:r + s] = -self.c * b
            r += s
        k = min(self.iter, self.max_iter - 1)
        return H - self.C[:, :k] @ self.D[:, :k].T


def solve(F, x0, blocks, confidence=0.1, tol=1e-8, max_iter=400):
    """The reference's Broyden loop (cilia_sim_main.cu:394-502) on a plain
    residual function: returns (x, iterations, residual norms)."""
    x = np.array(x0, dtype=np.float64)
    b = BroydenRef(x.size, max_iter, confidence, blocks)
    e = F(x)
    b.start(e)
    norms = [np.linalg.norm(e)]
    while norms[-1] > tol and b.iter < max_iter:
        x = x + b.find_update()
        e = F(x)
        norms.append(np.linalg.norm(e))
        b.end_of_iter(e, norms[-1] > tol)
    return x, b.iter, norms


SMALL_ANGLE = 1e-5


def dexp(u, v):
    """quaternion.cpp:574-608."""
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    t2 = u @ u
    t = np.sqrt(t2)
    if t < SMALL_ANGLE:
        t4 = t2 * t2
        beta = (1.0 / 3.0 + t2 / 72.0) / (2.0 + t2 / 6.0 + t4 / 72.0)
        alpha = 1.0 / 6.0 + 2.0 * beta * (1.0 - t2 / 12.0)
    else:
        alpha = np.sin(0.5 * t)
        alpha *= 2.0 * alpha / t2
        beta = (t - np.sin(t)) / (t * t2)
    c = np.cross(u, v)
    return v + alpha * c + beta * np.cross(u, c)


def dexpinv(u, v):
    """quaternion.cpp:610-633, in the reference's operation order."""
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    t2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2]
    t = np.sqrt(t2)
    fac = -1.0 / 12.0 if t < SMALL_ANGLE else (0.5 * t * np.cos(0.5 * t) / np.sin(0.5 * t) - 1.0) / t2
    c = np.array([u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]])
    w = np.array([u[1] * c[2] - u[2] * c[1], u[2] * c[0] - u[0] * c[2], u[0] * c[1] - u[1] * c[0]])
    return (v - 0.5 * c) - fac * w


def compute_errors(st, num_swim, num_fil, num_seg, bdf2, dt, dl, tol, blobs_per_swim=0, free_bodies=False):
    """(error, error_is_too_large) of mobility_solver.cpp:2661-2946.

    ``st`` holds seg_x, seg_xm1, seg_xm2, seg_u, seg_um1 (3 per segment),
    v_segs (6 per segment) and body_x, body_xm1, body_xm2, body_u, body_um1
    (3 per body), v_bodies (6 per body).  blobs_per_swim > 0 is the
    Broyden-for-everything form (USE_BROYDEN_FOR_EVERYTHING, :2674-2684,
    2774-2850): per swimmer the 3 NBLOB blob rows V + Omega x (x_blob - x_body)
    - v_blob follow the body's rows; st then also holds x_blobs and v_blobs (3
    per blob).  With prescribed body velocities (free_bodies False) V, Omega
    are v_bodies and the body rows its position / Lie-algebra equations; with
    free bodies the body rows are its force and torque balance
    -swimmer.f + sum (f_blob, (x_blob - x_body) x f_blob) (:2676-2681,
    2837-2847; st then holds f_segs, 6 per segment, and f_blobs, 3 per blob),
    and V, Omega are the velocities the body's positions imply (:2786-2806)."""
    c1, c2, c3 = -4.0 / 3.0, 1.0 / 3.0, -2.0 / 3.0
    per = 6 * (num_fil * num_seg + 1) + 3 * blobs_per_swim
    n = num_swim * per
    err = np.zeros(n)
    too_large = False

    def g3(name, k):
        a = st.get(name)
        return None if a is None else np.asarray(a, dtype=np.float64)[3 * k:3 * k + 3]

    def one(x, xm1, xm2, u, um1, v, om, pid, lid):
        W = dexpinv(u, om)
        if bdf2:
            ep = x + c1 * xm1 + c2 * xm2 + (c3 * dt) * v
            el = u - c2 * um1 + (c3 * dt) * W
        else:
            ep = x - xm1 - dt * v
            el = u - dt * W
        err[pid:pid + 3] = ep
        err[lid:lid + 3] = el
        return bool(np.any(np.abs(ep) > dl * tol) or np.any(np.abs(el) > tol))

    vs = np.asarray(st.get("v_segs", np.zeros(0)), dtype=np.float64)
    vb = np.asarray(st["v_bodies"], dtype=np.float64)
    for s in range(num_swim):
        base = s * per
        for i in range(num_fil):
            for j in range(num_seg):
                g = j + num_seg * (i + s * num_fil)
                pid = base + 6 * i * num_seg + 3 * j
                too_large |= one(g3("seg_x", g), g3("seg_xm1", g), g3("seg_xm2", g), g3("seg_u", g),
                                 g3("seg_um1", g), vs[6 * g:6 * g + 3], vs[6 * g + 3:6 * g + 6], pid,
                                 pid + 3 * num_seg)
        pid = base + 6 * num_fil * num_seg
        X = g3("body_x", s)
        if free_bodies and blobs_per_swim:
            fs = np.asarray(st["f_segs"], dtype=np.float64)
            fb = np.asarray(st["f_blobs"], dtype=np.float64)
            eb = np.zeros(6)          # -swimmer.f: the segments' force and torque about X
            for i in range(num_fil):
                for j in range(num_seg):
                    g = j + num_seg * (i + s * num_fil)
                    Fg, Tg = fs[6 * g:6 * g + 3], fs[6 * g + 3:6 * g + 6]
                    eb[:3] += Fg
                    eb[3:] += Tg + np.cross(g3("seg_x", g) - X, Fg)
            for m in range(blobs_per_swim):
                g = s * blobs_per_swim + m
                fm = fb[3 * g:3 * g + 3]
                eb[:3] += fm
                eb[3:] += np.cross(g3("x_blobs", g) - X, fm)
            err[pid:pid + 6] = eb
            too_large |= bool(np.any(np.abs(eb) > tol))
            if bdf2:
                V = -(X + c1 * g3("body_xm1", s) + c2 * g3("body_xm2", s)) / (c3 * dt)
                w = (c2 * g3("body_um1", s) - g3("body_u", s)) / (c3 * dt)
                Om = dexp(g3("body_u", s), w)
            else:
                V = (X - g3("body_xm1", s)) / dt
                Om = g3("body_u", s) / dt
        else:
            too_large |= one(X, g3("body_xm1", s), g3("body_xm2", s), g3("body_u", s),
                             g3("body_um1", s), vb[6 * s:6 * s + 3], vb[6 * s + 3:6 * s + 6], pid, pid + 3)
            V, Om = vb[6 * s:6 * s + 3], vb[6 * s + 3:6 * s + 6]
        for m in range(blobs_per_swim):
            g = s * blobs_per_swim + m
            diff = g3("x_blobs", g) - X
            w = np.cross(Om, diff)
            e = V + w - g3("v_blobs", g)
            rid = base + 6 * (num_fil * num_seg + 1) + 3 * m
            err[rid:rid + 3] = e
            too_large |= bool(np.any(np.abs(e) > tol))
    return err, too_large
