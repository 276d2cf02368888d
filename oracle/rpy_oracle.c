/*
 * rpy_oracle.c -- CPU FP64 restatement of the reference's RPY mobility hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it.  The product
 * path (timfilsimmod_amd/) never links, loads or calls it.
 *
 * Everything is plain C99 + OpenMP, double precision throughout.  Each function
 * names the reference file:line it restates.  Layouts are the reference's AoS
 * host layouts:
 *   x_segs  [3*Ns]   (x,y,z per segment, body -> filament -> segment order)
 *   f_segs  [6*Ns]   (fx,fy,fz,tx,ty,tz per segment)
 *   v_segs  [6*Ns]   (vx,vy,vz,wx,wy,wz per segment)
 *   x_blobs [3*Nb], f_blobs [3*Nb], v_blobs [3*Nb]
 *
 * Pinning (see tests/test_oracle_*.py): closed-form self mobilities; far-field
 * blocks against finite-difference Faxen/curl operators applied to the Oseen
 * tensor; continuity across r = ai+aj and r = |ai-aj|; symmetry and positive
 * definiteness of the assembled grand mobility; equal-radius overlap blocks
 * against the Rotne-Prager (1969) closed forms; flow kernel against single
 * Stokeslet/rotlet closed forms and the no-slip property of the Blakelet.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846264338327950288

/* Domain of the pair tensor used by the block sums and the dense mobility:
 * 0 unbounded (cuda_functions.cu:207-354), 1 half-space above a no-slip wall
 * at z = 0 (the INFINITE_PLANE_WALL branch, cuda_functions.cu:12-203). */
static int orc_domain = 0;
void orc_set_domain(int d) { orc_domain = d; }
int orc_get_domain(void) { return orc_domain; }

/* ------------------------------------------------------------------------ */
/* Half-space pair tensor: src/cilia/cuda_functions.cu:12-203 (Swan & Brady  */
/* 2007, radius 1 and viscosity 1 as in the reference).  Accumulates into    */
/* V[0..vel_dof) the velocity of particle i due to F[0..force_dof) on j;      */
/* same != 0 when i and j are the same particle (the wall self-mobility).   */
/* The reference evaluates the rotation-translation wall terms with -=, a    */
/* sign fix of the paper's alternating-tensor convention (:17-21); so here.  */
/* ------------------------------------------------------------------------ */
void orc_rpy_pair_wall(int vel_dof, int force_dof, double *V, const double *F, int same,
                       double xi, double yi, double zi, double xj, double yj, double zj)
{
    const double c6 = 1.0 / (6.0 * ORC_PI), c8 = 1.0 / (8.0 * ORC_PI);
    const int tq = force_dof > 3, rot = vel_dof > 3;
    if (same) {
        /* :27-64 */
        const double g1 = 1.0 / zi, g3 = g1 * g1 * g1, g4 = g3 * g1, g5 = g4 * g1;
        const double par = c6 - (9.0 * g1 - 2.0 * g3 + g5) / (96.0 * ORC_PI);
        const double perp = c6 - (9.0 * g1 - 4.0 * g3 + g5) / (48.0 * ORC_PI);
        const double cpl = g4 / (64.0 * ORC_PI);
        V[0] += par * F[0];
        V[1] += par * F[1];
        V[2] += perp * F[2];
        if (tq) {
            V[0] += cpl * F[4];
            V[1] -= cpl * F[3];
        }
        if (rot) {
            V[3] -= cpl * F[1];
            V[4] += cpl * F[0];
            if (tq) {
                const double rpar = c8 - 5.0 * g3 / (128.0 * ORC_PI);
                const double rperp = c8 - g3 / (64.0 * ORC_PI);
                V[3] += rpar * F[3];
                V[4] += rpar * F[4];
                V[5] += rperp * F[5];
            }
        }
        return;
    }
    /* unbounded far-field part, :68-119 */
    {
        const double dx = xi - xj, dy = yi - yj, dz = zi - zj;
        const double q = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
        const double ux = dx * q, uy = dy * q, uz = dz * q;
        const double uf = ux * F[0] + uy * F[1] + uz * F[2];
        const double ca = c8 * q, cb = 1.0 + (2.0 / 3.0) * q * q, cc = 1.0 - 2.0 * q * q;
        V[0] += ca * (cb * F[0] + cc * uf * ux);
        V[1] += ca * (cb * F[1] + cc * uf * uy);
        V[2] += ca * (cb * F[2] + cc * uf * uz);
        const double c2 = c8 * q * q;
        if (tq) {
            V[0] += c2 * (F[4] * uz - F[5] * uy);
            V[1] += c2 * (F[5] * ux - F[3] * uz);
            V[2] += c2 * (F[3] * uy - F[4] * ux);
        }
        if (rot) {
            V[3] += c2 * (F[1] * uz - F[2] * uy);
            V[4] += c2 * (F[2] * ux - F[0] * uz);
            V[5] += c2 * (F[0] * uy - F[1] * ux);
            if (tq) {
                const double c3 = 0.5 * c2 * q, ut = ux * F[3] + uy * F[4] + uz * F[5];
                V[3] += c3 * (3.0 * ut * ux - F[3]);
                V[4] += c3 * (3.0 * ut * uy - F[4]);
                V[5] += c3 * (3.0 * ut * uz - F[5]);
            }
        }
    }
    /* wall corrections from the image at (xj, yj, -zj), :121-199 */
    const double Rx = xi - xj, Ry = yi - yj, Rz = zi + zj;
    const double p1 = 1.0 / sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
    const double p2 = p1 * p1, p3 = p2 * p1, p4 = p3 * p1, p5 = p4 * p1;
    const double ex = Rx * p1, ey = Ry * p1, ez = Rz * p1, ez2 = ez * ez;
    const double h = zj / (zi + zj);
    const double ef = ex * F[0] + ey * F[1] + ez * F[2];
    {
        const double t0 = -(0.75 * p1 * (1.0 + 2.0 * h * (1.0 - h) * ez2) +
                            0.5 * (p3 * (1.0 - 3.0 * ez2) - p5 * (1.0 - 5.0 * ez2)));
        const double t1 = -(0.75 * p1 * (1.0 - 6.0 * h * (1.0 - h) * ez2) - 1.5 * p3 * (1.0 - 5.0 * ez2) +
                            2.5 * p5 * (1.0 - 7.0 * ez2)) * ef;
        const double t2 = ez * F[2] * (1.5 * h * p1 * (1.0 - 6.0 * (1.0 - h) * ez2) - 3.0 * p3 * (1.0 - 5.0 * ez2) +
                                      5.0 * p5 * (2.0 - 7.0 * ez2));
        const double t3 = (1.5 * h * p1 - 5.0 * p5) * ef;
        const double t4 = -((3.0 * h * h * p1 + 3.0 * p3) * ez2 + (2.0 - 15.0 * ez2) * p5) * F[2];
        V[0] += c6 * (t0 * F[0] + (t1 + t2) * ex);
        V[1] += c6 * (t0 * F[1] + (t1 + t2) * ey);
        V[2] += c6 * (t0 * F[2] + (t1 + t2 + t3) * ez + t4);
    }
    if (rot) {
        const double u0 = 0.75 * p2;
        const double u12 = F[2] * (9.0 * h * ez2 * p2 + (1.5 - 15.0 * ez2) * p4) - ez * ef * (4.5 * h * p2 - 7.5 * p4);
        const double u3 = -1.5 * ez * (h * p2 - p4);
        V[3] -= c6 * (u0 * (F[1] * ez - F[2] * ey) - u12 * ey + u3 * F[1]);
        V[4] -= c6 * (u0 * (F[2] * ex - F[0] * ez) + u12 * ex - u3 * F[0]);
        V[5] -= c6 * u0 * (F[0] * ey - F[1] * ex);
        if (tq) {
            const double et = ex * F[3] + ey * F[4] + ez * F[5];
            const double w0 = p3 * (0.375 - 2.25 * ez2), w1 = -1.125 * p3 * et, w2 = 2.25 * ez * p3 * et;
            const double w3 = 2.25 * p3 * (ex * F[4] - ey * F[3]);
            V[3] += c6 * (w0 * F[3] + w1 * ex - w3 * ey);
            V[4] += c6 * (w0 * F[4] + w1 * ey + w3 * ex);
            V[5] += c6 * (w0 * F[5] + w1 * ez + w2);
        }
    }
    if (tq) {
        /* translation from torque: the transpose of the rotation-from-force
         * block of the pair seen from j (image weight 1 - h, e_x and e_y
         * reversed), :181-199 */
        const double hh = 1.0 - h, fx = -ex, fy = -ey;
        const double ext = fx * F[4] - fy * F[3];
        const double q0 = 0.75 * p2;
        const double q1 = (9.0 * hh * ez2 * p2 + (1.5 - 15.0 * ez2) * p4) * ext;
        const double q2 = -ez * (4.5 * hh * p2 - 7.5 * p4) * ext;
        const double q3 = -1.5 * ez * (hh * p2 - p4);
        V[0] -= c6 * (q0 * (F[5] * fy - F[4] * ez) + q2 * fx - q3 * F[4]);
        V[1] -= c6 * (q0 * (F[3] * ez - F[5] * fx) + q2 * fy + q3 * F[3]);
        V[2] -= c6 * (q0 * (F[4] * fx - F[3] * fy) + q2 * ez + q1);
    }
}

/* ------------------------------------------------------------------------ */
/* Pair tensor: src/cilia/cuda_functions.cu:207-354 (unbounded branch).      */
/* Accumulates into V[0..vel_dof) the velocity of particle i (radius ai) due */
/* to force (and torque if force_dof == 6) F on particle j (radius aj).      */
/* ------------------------------------------------------------------------ */
void orc_rpy_pair(int vel_dof, int force_dof, double *V, const double *F,
                  double xi, double yi, double zi, double ai,
                  double xj, double yj, double zj, double aj, double mu)
{
    double dx = xi - xj, dy = yi - yj, dz = zi - zj;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    /* cuda_functions.cu:213 -- r = 0 is not special-cased, it is regularised. */
    const double inv_r = 1.0 / (r + 1e-20);
    const double da = fabs(ai - aj);
    const double sum_a = ai + aj;
    const double ex = dx * inv_r, ey = dy * inv_r, ez = dz * inv_r;

    /* translation-translation: cuda_functions.cu:222-261 */
    double e_dot = ex * F[0] + ey * F[1] + ez * F[2];
    double A, B;
    if (r > sum_a) {
        const double s = inv_r * inv_r * (ai * ai + aj * aj) / 3.0;
        const double c = inv_r / (8.0 * ORC_PI * mu);
        A = (1.0 + s) * c;
        B = (1.0 - 3.0 * s) * c * e_dot;
    } else if (r > da) {
        const double den = 32.0 * r * r * r;
        const double p = da * da + 3.0 * r * r;
        const double q = da * da - r * r;
        const double c = 1.0 / (6.0 * ORC_PI * mu * ai * aj);
        A = (0.5 * (ai + aj) - p * p / den) * c;
        B = (3.0 * q * q / den) * c * e_dot;
    } else {
        A = 1.0 / (6.0 * ORC_PI * mu * (ai > aj ? ai : aj));
        B = 0.0;
    }
    V[0] += A * F[0] + B * ex;
    V[1] += A * F[1] + B * ey;
    V[2] += A * F[2] + B * ez;

    if (force_dof > 3) {
        /* translation-rotation: cuda_functions.cu:263-283 */
        double C;
        if (r > sum_a) {
            C = inv_r * inv_r / (8.0 * ORC_PI * mu);
        } else if (r > da) {
            const double g = aj - ai + r;
            C = inv_r * inv_r * g * g
                * (ai * ai + 2.0 * ai * (r + aj) - 3.0 * (aj - r) * (aj - r))
                / (128.0 * ORC_PI * mu * ai * aj * aj * aj);
        } else {
            C = (aj > ai ? 1.0 : 0.0) * r / (8.0 * ORC_PI * mu * aj * aj * aj);
        }
        /* V_t += C * (tau x e) */
        V[0] += C * (F[4] * ez - F[5] * ey);
        V[1] += C * (F[5] * ex - F[3] * ez);
        V[2] += C * (F[3] * ey - F[4] * ex);
    }

    if (vel_dof > 3) {
        /* rotation-translation: cuda_functions.cu:287-307 (ai <-> aj of tr) */
        double C;
        if (r > sum_a) {
            C = inv_r * inv_r / (8.0 * ORC_PI * mu);
        } else if (r > da) {
            const double g = ai - aj + r;
            C = inv_r * inv_r * g * g
                * (aj * aj + 2.0 * aj * (r + ai) - 3.0 * (ai - r) * (ai - r))
                / (128.0 * ORC_PI * mu * aj * ai * ai * ai);
        } else {
            C = (ai > aj ? 1.0 : 0.0) * r / (8.0 * ORC_PI * mu * ai * ai * ai);
        }
        /* V_r += C * (F x e) */
        V[3] += C * (F[1] * ez - F[2] * ey);
        V[4] += C * (F[2] * ex - F[0] * ez);
        V[5] += C * (F[0] * ey - F[1] * ex);

        if (force_dof > 3) {
            /* rotation-rotation: cuda_functions.cu:309-350 */
            const double t_dot = ex * F[3] + ey * F[4] + ez * F[5];
            double Ar, Br;
            if (r > sum_a) {
                Ar = -inv_r * inv_r * inv_r / (16.0 * ORC_PI * mu);
                Br = -3.0 * t_dot * Ar;
            } else if (r > da) {
                const double r2 = r * r, r3 = r2 * r;
                const double ai2 = ai * ai, aj2 = aj * aj;
                const double d4 = da * da * da * da;
                const double num_a = 5.0 * r3 * r3
                                   - 27.0 * r2 * r2 * (ai2 + aj2)
                                   + 32.0 * r3 * (ai2 * ai + aj2 * aj)
                                   - 9.0 * r2 * (ai2 - aj2) * (ai2 - aj2)
                                   - d4 * (ai2 + aj2 + 4.0 * ai * aj);
                const double w = (ai - aj) * (ai - aj) - r2;
                const double num_b = 3.0 * w * w * (ai2 + aj2 + 4.0 * ai * aj - r2);
                const double c = 1.0 / (512.0 * ORC_PI * mu * r3 * ai2 * ai * aj2 * aj);
                Ar = num_a * c;
                Br = num_b * c * t_dot;
            } else {
                const double am = ai > aj ? ai : aj;
                Ar = 1.0 / (8.0 * ORC_PI * mu * am * am * am);
                Br = 0.0;
            }
            V[3] += Ar * F[3] + Br * ex;
            V[4] += Ar * F[4] + Br * ey;
            V[5] += Ar * F[5] + Br * ez;
        }
    }
}

static void orc_pair_domain(int vel_dof, int force_dof, double *V, const double *F, int same,
                            double xi, double yi, double zi, double ai,
                            double xj, double yj, double zj, double aj, double mu)
{
    if (orc_domain == 1) orc_rpy_pair_wall(vel_dof, force_dof, V, F, same, xi, yi, zi, xj, yj, zj);
    else orc_rpy_pair(vel_dof, force_dof, V, F, xi, yi, zi, ai, xj, yj, zj, aj, mu);
}

/* ------------------------------------------------------------------------ */
/* Block sums: src/cilia/cuda_functions.cu:362-757.  The target range is    */
/* [start, start+count); outputs are indexed from 0 at `start` exactly as   */
/* in the reference kernels (V[6*(i-start)+k]).                              */
/* seg_dof selects the template: 6 -> <6,6>,<6,3>,<3,6>; 3 -> all <3,3>.    */
/* ------------------------------------------------------------------------ */

/* Mss_mult, cuda_functions.cu:362-474: V_s = M_ss F_s (overwrite). */
void orc_Mss(double *V, const double *F, const double *X, int start, int count,
             int num_segs, int seg_dof, double a_seg, double mu)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = start; i < start + count; ++i) {
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < num_segs; ++j) {
            double f[6] = {F[6 * j], F[6 * j + 1], F[6 * j + 2], 0, 0, 0};
            if (seg_dof == 6) { f[3] = F[6 * j + 3]; f[4] = F[6 * j + 4]; f[5] = F[6 * j + 5]; }
            orc_pair_domain(seg_dof, seg_dof, v, f, i == j, X[3 * i], X[3 * i + 1], X[3 * i + 2], a_seg,
                            X[3 * j], X[3 * j + 1], X[3 * j + 2], a_seg, mu);
        }
        for (int k = 0; k < 6; ++k) V[6 * (i - start) + k] = v[k];
    }
}

/* Msb_mult, cuda_functions.cu:551-636: V_s += M_sb F_b (accumulate). */
void orc_Msb(double *V, const double *Fb, const double *Xs, const double *Xb, int start,
             int count, int num_blobs, int seg_dof, double a_seg, double a_blob, double mu)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = start; i < start + count; ++i) {
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < num_blobs; ++j) {
            const double f[3] = {Fb[3 * j], Fb[3 * j + 1], Fb[3 * j + 2]};
            orc_pair_domain(seg_dof, 3, v, f, 0, Xs[3 * i], Xs[3 * i + 1], Xs[3 * i + 2], a_seg,
                            Xb[3 * j], Xb[3 * j + 1], Xb[3 * j + 2], a_blob, mu);
        }
        for (int k = 0; k < 6; ++k) V[6 * (i - start) + k] += v[k];
    }
}

/* Mbb_mult, cuda_functions.cu:476-549: V_b = M_bb F_b (overwrite). */
void orc_Mbb(double *V, const double *F, const double *X, int start, int count,
             int num_blobs, double a_blob, double mu)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = start; i < start + count; ++i) {
        double v[3] = {0, 0, 0};
        for (int j = 0; j < num_blobs; ++j) {
            const double f[3] = {F[3 * j], F[3 * j + 1], F[3 * j + 2]};
            orc_pair_domain(3, 3, v, f, i == j, X[3 * i], X[3 * i + 1], X[3 * i + 2], a_blob,
                            X[3 * j], X[3 * j + 1], X[3 * j + 2], a_blob, mu);
        }
        for (int k = 0; k < 3; ++k) V[3 * (i - start) + k] = v[k];
    }
}

/* Mbs_mult, cuda_functions.cu:638-757: V_b (+= if accumulate, else =) M_bs F_s.
 * accumulate=1 is the USE_BROYDEN_FOR_EVERYTHING || PRESCRIBED_CILIA build
 * (cuda_functions.cu:739-751). */
void orc_Mbs(double *V, const double *Fs, const double *Xb, const double *Xs, int start,
             int count, int num_segs, int seg_dof, int accumulate, double a_seg,
             double a_blob, double mu)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = start; i < start + count; ++i) {
        double v[3] = {0, 0, 0};
        for (int j = 0; j < num_segs; ++j) {
            double f[6] = {Fs[6 * j], Fs[6 * j + 1], Fs[6 * j + 2], 0, 0, 0};
            if (seg_dof == 6) { f[3] = Fs[6 * j + 3]; f[4] = Fs[6 * j + 4]; f[5] = Fs[6 * j + 5]; }
            orc_pair_domain(3, seg_dof, v, f, 0, Xb[3 * i], Xb[3 * i + 1], Xb[3 * i + 2], a_blob,
                            Xs[3 * j], Xs[3 * j + 1], Xs[3 * j + 2], a_seg, mu);
        }
        for (int k = 0; k < 3; ++k) {
            if (accumulate) V[3 * (i - start) + k] += v[k];
            else V[3 * (i - start) + k] = v[k];
        }
    }
}

/* evaluate_full_mobility, rpy_mobility_solver.cu:330-337: Mss(=), Msb(+=),
 * Mbb(=), Mbs(+= or =) in that order, on one device holding every target. */
void orc_full_mobility(double *v_segs, double *v_blobs, const double *f_segs,
                       const double *f_blobs, const double *x_segs, const double *x_blobs,
                       int num_segs, int num_blobs, int seg_dof, int bs_accumulate,
                       double a_seg, double a_blob, double mu)
{
    if (num_segs > 0) {
        orc_Mss(v_segs, f_segs, x_segs, 0, num_segs, num_segs, seg_dof, a_seg, mu);
        orc_Msb(v_segs, f_blobs, x_segs, x_blobs, 0, num_segs, num_blobs, seg_dof, a_seg, a_blob, mu);
    }
    if (num_blobs > 0) {
        orc_Mbb(v_blobs, f_blobs, x_blobs, 0, num_blobs, num_blobs, a_blob, mu);
        orc_Mbs(v_blobs, f_segs, x_blobs, x_segs, 0, num_blobs, num_segs, seg_dof, bs_accumulate,
                a_seg, a_blob, mu);
    }
}

/* Dense grand mobility (tiny N only): column c of M = M e_c, assembled with
 * the pair tensor.  Unknown ordering: seg (seg_dof each) then blob (3 each).
 * Used to pin the block sums against a dense GEMV and to check symmetry and
 * positive definiteness. M is row-major [n x n]. */
void orc_dense_mobility(double *M, const double *x_segs, const double *x_blobs,
                        int num_segs, int num_blobs, int seg_dof, double a_seg,
                        double a_blob, double mu)
{
    const int n = seg_dof * num_segs + 3 * num_blobs;
    memset(M, 0, sizeof(double) * (size_t)n * (size_t)n);
    const int np = num_segs + num_blobs;
    for (int p = 0; p < np; ++p) {
        const int p_is_seg = p < num_segs;
        const int pd = p_is_seg ? seg_dof : 3;
        const double *xp = p_is_seg ? &x_segs[3 * p] : &x_blobs[3 * (p - num_segs)];
        const double ap = p_is_seg ? a_seg : a_blob;
        const int prow = p_is_seg ? seg_dof * p : seg_dof * num_segs + 3 * (p - num_segs);
        for (int q = 0; q < np; ++q) {
            const int q_is_seg = q < num_segs;
            const int qd = q_is_seg ? seg_dof : 3;
            const double *xq = q_is_seg ? &x_segs[3 * q] : &x_blobs[3 * (q - num_segs)];
            const double aq = q_is_seg ? a_seg : a_blob;
            const int qcol = q_is_seg ? seg_dof * q : seg_dof * num_segs + 3 * (q - num_segs);
            for (int c = 0; c < qd; ++c) {
                double f[6] = {0, 0, 0, 0, 0, 0};
                double v[6] = {0, 0, 0, 0, 0, 0};
                f[c] = 1.0;
                orc_pair_domain(pd, qd, v, f, p == q, xp[0], xp[1], xp[2], ap, xq[0], xq[1], xq[2], aq, mu);
                for (int r = 0; r < pd; ++r) M[(size_t)(prow + r) * n + qcol + c] = v[r];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Flow field: src/flow_field/flow_field_evaluator.cu:13-228.                */
/* Stokeslet + rotlet from segments, Stokeslet from blobs, Blakelet image    */
/* system when num_blobs == 0, optional removal of each source's Stokeslet   */
/* re-centred at its body's centre; result scaled by prefac (1/(8 pi mu),   */
/* flow_field_evaluator.cu:549).                                            */
/* Ns and Nb are per body; Nbods bodies are stored contiguously.            */
/* ------------------------------------------------------------------------ */
static void orc_stokeslet_acc(double *v, double dx, double dy, double dz,
                              double fx, double fy, double fz, double sign)
{
    const double inv_r2 = 1.0 / (dx * dx + dy * dy + dz * dz);
    const double inv_r = sqrt(inv_r2);
    const double k = inv_r2 * (dx * fx + dy * fy + dz * fz);
    v[0] += sign * inv_r * (fx + dx * k);
    v[1] += sign * inv_r * (fy + dy * k);
    v[2] += sign * inv_r * (fz + dz * k);
}

void orc_flow_velocities(double *Vf, const double *Xf, const double *Fs, const double *Fb,
                         const double *Xs, const double *Xbody, const double *Xb, int Nf,
                         int Ns, int Nbods, int Nb, double prefac, int remove_net_force)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < Nf; ++i) {
        const double xi = Xf[3 * i], yi = Xf[3 * i + 1], zi = Xf[3 * i + 2];
        double v[3] = {0, 0, 0};
        for (int j = 0; j < Ns * Nbods; ++j) {
            const double dx = xi - Xs[3 * j], dy = yi - Xs[3 * j + 1], dz = zi - Xs[3 * j + 2];
            const double fx = Fs[6 * j], fy = Fs[6 * j + 1], fz = Fs[6 * j + 2];
            const double tx = Fs[6 * j + 3], ty = Fs[6 * j + 4], tz = Fs[6 * j + 5];
            orc_stokeslet_acc(v, dx, dy, dz, fx, fy, fz, 1.0);
            if (Nb == 0) {
                /* Blakelet correction, flow_field_evaluator.cu:83-116 */
                const double zj = Xs[3 * j + 2];
                const double hz = zi + zj; /* image separation, z component */
                const double R2 = dx * dx + dy * dy + hz * hz;
                orc_stokeslet_acc(v, dx, dy, hz, fx, fy, fz, -1.0);
                const double iR = sqrt(1.0 / R2);
                const double g = 2.0 * zj * iR * iR * iR * iR * iR;
                const double Dxx = (3.0 * dx * dx - R2) * zi;
                const double Dxy = 3.0 * dx * dy * zi;
                const double Dxz = (R2 - 3.0 * hz * zi) * dx;
                const double Dyy = (3.0 * dy * dy - R2) * zi;
                const double Dyz = (R2 - 3.0 * hz * zi) * dy;
                const double Dzx = (3.0 * hz * zi + R2) * dx;
                const double Dzy = (3.0 * hz * zi + R2) * dy;
                const double Dzz = (R2 - 3.0 * hz * hz) * zi;
                v[0] += g * (Dxx * fx + Dxy * fy + Dxz * fz);
                v[1] += g * (Dxy * fx + Dyy * fy + Dyz * fz);
                v[2] += g * (Dzx * fx + Dzy * fy + Dzz * fz);
            }
            /* rotlet, flow_field_evaluator.cu:119-124 */
            const double inv_r2 = 1.0 / (dx * dx + dy * dy + dz * dz);
            const double inv_r3 = sqrt(inv_r2) * inv_r2;
            v[0] += inv_r3 * (ty * dz - tz * dy);
            v[1] += inv_r3 * (tz * dx - tx * dz);
            v[2] += inv_r3 * (tx * dy - ty * dx);
            if (remove_net_force) {
                const int b = j / Ns;
                orc_stokeslet_acc(v, xi - Xbody[3 * b], yi - Xbody[3 * b + 1], zi - Xbody[3 * b + 2],
                                  fx, fy, fz, -1.0);
            }
        }
        for (int j = 0; j < Nb * Nbods; ++j) {
            const double fx = Fb[3 * j], fy = Fb[3 * j + 1], fz = Fb[3 * j + 2];
            orc_stokeslet_acc(v, xi - Xb[3 * j], yi - Xb[3 * j + 1], zi - Xb[3 * j + 2], fx, fy, fz, 1.0);
            if (remove_net_force) {
                const int b = j / Nb;
                orc_stokeslet_acc(v, xi - Xbody[3 * b], yi - Xbody[3 * b + 1], zi - Xbody[3 * b + 2],
                                  fx, fy, fz, -1.0);
            }
        }
        Vf[3 * i] = prefac * v[0];
        Vf[3 * i + 1] = prefac * v[1];
        Vf[3 * i + 2] = prefac * v[2];
    }
}

/* RPY flow mode (an extension the reference does not have): velocity of a
 * point tracer (radius 0) using the pair tensor of cuda_functions.cu:207-354
 * with ai = 0: <3,6> from segment sources, <3,3> from blob sources. */
void orc_flow_rpy(double *Vf, const double *Xf, const double *Fs, const double *Fb,
                  const double *Xs, const double *Xb, int Nf, int Ns_total, int Nb_total,
                  double a_seg, double a_blob, double mu)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < Nf; ++i) {
        double v[3] = {0, 0, 0};
        for (int j = 0; j < Ns_total; ++j)
            orc_rpy_pair(3, 6, v, &Fs[6 * j], Xf[3 * i], Xf[3 * i + 1], Xf[3 * i + 2], 0.0,
                         Xs[3 * j], Xs[3 * j + 1], Xs[3 * j + 2], a_seg, mu);
        for (int j = 0; j < Nb_total; ++j)
            orc_rpy_pair(3, 3, v, &Fb[3 * j], Xf[3 * i], Xf[3 * i + 1], Xf[3 * i + 2], 0.0,
                         Xb[3 * j], Xb[3 * j + 1], Xb[3 * j + 2], a_blob, mu);
        Vf[3 * i] = v[0];
        Vf[3 * i + 1] = v[1];
        Vf[3 * i + 2] = v[2];
    }
}

/* ------------------------------------------------------------------------ */
/* Barrier forces: src/cilia/cuda_functions.cu:1240-1436 (barrier_forces),   */
/* all particles of all bodies, sources in the reference's index order      */
/* (segments then blobs).  Segment forces are added to F_s slots 0-2; blob  */
/* forces overwrite F_rep.  strength = REPULSIVE_FORCE_FACTOR *              */
/* END_FORCE_MAGNITUDE; nseg = NSEG; wall selects INFINITE_PLANE_WALL's     */
/* normal force on segments below z = dl/2 (:1271-1278).                    */
static void orc_barrier_sum(double *f, const double *xi, double ai, int i_global, int is_seg, int fili,
                            const double *Xs, const double *Xb, int ns, int nb, int nseg, double a_seg,
                            double a_blob, double strength) {
    for (int j = 0; j < ns + nb; ++j) {
        const int jseg = j < ns;
        const double *xj = jseg ? &Xs[3 * j] : &Xb[3 * (j - ns)];
        const double a_sum = ai + (jseg ? a_seg : a_blob);
        const double chi_fac = 10.0 / a_sum;
        const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        int excluded;
        if (is_seg) {
            const int filj = j / nseg;
            excluded = (fili == filj) && abs(i_global - j) <= 1;
        } else {
            excluded = (i_global == j);
        }
        if (!excluded && dist < 1.1 * a_sum) {
            double fac = fmin(1.0, 1.0 - chi_fac * (dist - a_sum));
            fac *= strength * fac * fac * fac;
            const double dm1 = 1.0 / dist;
            f[0] += fac * dx * dm1;
            f[1] += fac * dy * dm1;
            f[2] += fac * dz * dm1;
        }
    }
}

void orc_barrier_forces(double *F_s, double *F_rep, const double *Xs, const double *Xb, int ns, int nb,
                        int nseg, double a_seg, double a_blob, double strength, int wall, double dl) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ns; ++i) {
        double f[3] = {0.0, 0.0, 0.0};
        const double *xi = &Xs[3 * i];
        if (wall && xi[2] < 0.5 * dl) {
            double fz = fmin(1.0, 1.0 - (xi[2] - a_seg) / (0.5 * dl - a_seg));
            fz *= strength * fz * fz * fz;
            f[2] = fz;
        }
        orc_barrier_sum(f, xi, a_seg, i, 1, i / nseg, Xs, Xb, ns, nb, nseg, a_seg, a_blob, strength);
        F_s[6 * i] += f[0];
        F_s[6 * i + 1] += f[1];
        F_s[6 * i + 2] += f[2];
    }
    if (!F_rep) return;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nb; ++k) {
        double f[3] = {0.0, 0.0, 0.0};
        orc_barrier_sum(f, &Xb[3 * k], a_blob, ns + k, 0, -1, Xs, Xb, ns, nb, nseg, a_seg, a_blob, strength);
        F_rep[3 * k] = f[0];
        F_rep[3 * k + 1] = f[1];
        F_rep[3 * k + 2] = f[2];
    }
}

/* Thread count actually used by the OpenMP loops above (for the CPU
 * baseline report). */
#ifdef _OPENMP
#include <omp.h>
int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }
#else
int orc_num_threads(void) { return 1; }
void orc_set_num_threads(int n) { (void)n; }
#endif
