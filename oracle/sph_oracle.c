/*
 * FP64 CPU oracle for the per-particle SPH step (test infrastructure only).
 *
 * This file is the checker, not the product: only tests/, the smoke check and
 * the bench's cpu_baseline leg load it.  The shipped path is the HIP library
 * under sph_2_amd/csrc and never calls into this code.
 *
 * Every routine restates one function of the reference's FP64 CPU path
 * (Python/Numba) in plain C with OpenMP over rows.  Reference paths are
 * relative to 11_Dec_2022/ of hassanfv/SPH_2:
 *   LIBSX  = GPU_sph/IC/libsx2_2t.py
 *   LIBSX3 = GPU_sph/work_on_smoothing/libsx2_2t.py
 *   SHEAR  = GPU_sph/work_on_smoothing/shear.py
 *   AUGX   = Evrard_collapse_new_h_algorithm/augsphx6.5x.py
 *   LIBV2  = GPU_sph/Important_Test_Simulations/Evrard_collapse/myCppSPHLibs_v2.h
 *   LIBV7  = GPU_sph_Type_cooling_II/hfvCLibs_v7.h
 *   PHOTOPY= GPU_sph/Cooling_v1/photolibs3.py
 *
 * Determinism contract shared with the GPU path (compile with
 * -ffp-contract=off, never -ffast-math):
 *   - pair distance r2 = (dx*dx + dy*dy) + dz*dz in FP64, dx = x_i - x_j;
 *   - neighbour tests compare r2 <= s*s (gather radius s = 2h) or
 *     r2 <= (h_i + h_j)^2 (symmetric pair test, 2*h_ij = h_i + h_j);
 *   - the smoothing-length iteration and the cooling solve use only
 *     +, -, *, / and comparisons, so both sides make identical decisions.
 *
 * Parity pins (see tests/test_oracle_*.py): closed forms and invariants for
 * every function; brute-force kNN for the h iteration; an independent numpy
 * restatement (oracle/npref.py) at small N.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_SIGMA (1.0 / 3.14159265358979323846) /* 3-D M4 normalisation 1/pi */

typedef struct {
    double gamma, alpha, beta, eta, G;
    int32_t nngb, ndown, nup, max_iter;
    int32_t av_form; /* 0: Monaghan-Gingold 1983 (AUGX:103-139); 1: Gadget signal velocity (LIBSX:153-169) */
    int32_t balsara; /* 1: multiply Pi_ij by (f_i+f_j)/2 (LIBV2:400-405) */
    double av_w_eps; /* Gadget form: w = v.r / (r + av_w_eps); 1e-5 LIBSX, 1e-15 augsphx6.5.py */
} orc_params;

typedef struct {
    double unit_u_cgs, unit_rho_cgs, unit_time_s;
    double XH, mu, gamma, heat, tol;
    int32_t max_iter, max_bracket, ntab, pad;
} orc_cool_params;

static inline int64_t row_of(const int64_t *rows, int64_t r) { return rows ? rows[r] : r; }

static inline double pair_r2(const double *pos, int64_t i, int64_t j) {
    double dx = pos[3 * i + 0] - pos[3 * j + 0];
    double dy = pos[3 * i + 1] - pos[3 * j + 1];
    double dz = pos[3 * i + 2] - pos[3 * j + 2];
    double a = dx * dx;
    double b = dy * dy;
    double c = dz * dz;
    return (a + b) + c;
}

/* ------------------------------------------------------------------------ */
/* Cube-root ratio used by the smoothing-length update.                      */
/* cbrt(nngb/k) clamped to [0.5, 2]; Newton iteration with a fixed count so  */
/* the same doubles come out on any IEEE machine.                            */
/* ------------------------------------------------------------------------ */
double orc_h_factor(int32_t nngb, int32_t k) {
    if (8 * (int64_t)k <= (int64_t)nngb) return 2.0;
    if ((int64_t)k >= 8 * (int64_t)nngb) return 0.5;
    double x = (double)nngb / (double)k;
    double y = 1.0;
    for (int it = 0; it < 12; ++it) {
        double y2 = y * y;
        double y3 = y2 * y;
        double num = y3 - x;
        double den = 3.0 * y2;
        y = y - num / den;
    }
    if (y < 0.5) y = 0.5;
    if (y > 2.0) y = 2.0;
    return y;
}

/* Count of j with r2_ij <= s^2, self included (LIBSX:906-915 counting loop). */
static int64_t count_within(const double *pos, int64_t n, int64_t i, double s) {
    double s2 = s * s;
    int64_t k = 0;
    for (int64_t j = 0; j < n; ++j)
        if (pair_r2(pos, i, j) <= s2) ++k;
    return k;
}

void orc_count_within(int64_t n, const double *pos, const double *radius, const int64_t *rows,
                      int64_t nrows, int32_t *k_out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t i = row_of(rows, r);
        k_out[r] = (int32_t)count_within(pos, n, i, radius[r]);
    }
}

/*
 * Smoothing-length iteration (canonical form of LIBSX3:946-1010,
 * LIBV7:255-330 and LIBV2:7-64).
 *
 * Search radius s (= 2h).  Start: s = 2h_prev, or with the Gadget-2 eq. 31
 * predictor s = 2*(0.5*h*(1 + F(k_prev)) + (h*divV)*dt/3) (LIBSX3:960).
 * Each iteration counts k(s) (self included); done when ndown <= k <= nup.
 * Otherwise the step is the log-space Newton update s <- s*F(k) with
 * F(k) = cbrt(nngb/k) clamped to [1/2, 2], kept inside the bracket
 * (lo: k < ndown, hi: k > nup) by bisection once both ends are known.
 * Stops without convergence when the bracket has collapsed, when every
 * particle is already inside s, or after max_iter counts; then the result is
 * the smallest over-full radius seen (else the last radius).  status bit 1 =
 * capped.
 */
static void smooth_one(const double *pos, int64_t n, int64_t i, double h_prev, int32_t k_prev,
                       int use_pred, double divv, double dt, int32_t nngb, int32_t ndown,
                       int32_t nup, int32_t max_iter, double *h_out, int32_t *k_out,
                       int32_t *it_out, int32_t *st_out) {
    double s;
    if (use_pred) {
        double f = orc_h_factor(nngb, k_prev);
        double t1 = 0.5 * h_prev;
        double t2 = 1.0 + f;
        double t3 = t1 * t2;
        double t4 = h_prev * divv;
        double t5 = t4 * dt;
        double t6 = t5 / 3.0;
        s = 2.0 * (t3 + t6);
        if (!(s > 0.0) || !(s < INFINITY)) s = 2.0 * h_prev;
    } else {
        s = 2.0 * h_prev;
    }
    double lo = 0.0, hi = 0.0;
    int have_lo = 0, have_hi = 0;
    int64_t khi = 0, klo = 0;
    int64_t k = 0;
    int32_t it = 0;
    int converged = 0;
    while (it < max_iter) {
        k = count_within(pos, n, i, s);
        ++it;
        if (k >= ndown && k <= nup) {
            converged = 1;
            break;
        }
        if (k < ndown) {
            lo = s; klo = k; have_lo = 1;
            if (k >= n) break; /* every particle inside: growing cannot help */
        } else {
            hi = s; khi = k; have_hi = 1;
        }
        double snew = s * orc_h_factor(nngb, (int32_t)(k > 2147483647 ? 2147483647 : k));
        if (have_lo && have_hi) {
            if (!(snew > lo && snew < hi)) snew = 0.5 * (lo + hi);
            if (!(snew > lo && snew < hi)) break; /* bracket collapsed */
        }
        s = snew;
    }
    int32_t st = 0;
    if (!converged) {
        st = 1;
        if (have_hi) {
            s = hi; k = khi;
        } else {
            s = lo; k = klo;
        }
    }
    *h_out = 0.5 * s;
    *k_out = (int32_t)k;
    *it_out = it;
    *st_out = st;
}

void orc_smoothing(int64_t n, const double *pos, const double *h_prev, const int32_t *k_prev,
                   const double *divv, double dt, int32_t nngb, int32_t ndown, int32_t nup,
                   int32_t max_iter, const int64_t *rows, int64_t nrows, double *h_out,
                   int32_t *k_out, int32_t *it_out, int32_t *st_out) {
    int use_pred = (k_prev != NULL && divv != NULL);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t i = row_of(rows, r);
        smooth_one(pos, n, i, h_prev[i], use_pred ? k_prev[i] : 0, use_pred,
                   use_pred ? divv[i] : 0.0, dt, nngb, ndown, nup, max_iter, &h_out[r], &k_out[r],
                   &it_out[r], &st_out[r]);
    }
}

/*
 * The reference's own smoothing-length rule (LIBSX3:947-1006
 * smoothing_fast_new_mpi3, LIBSX:940-990 smoothing_fast_new_mpi2; the same
 * loop as the CUDA smoothing_h of LIBV2:7-64, there in FP32).  Support
 * radius s (= 2h): from 2h_prev, or with the Gadget-2 eq. 31 start
 * 2 (0.5 h (1 + F) + (1/3) (h divV) dt) (LIBSX3:961, in that order of
 * operations; F = cbrt(nngb/k_prev) by orc_h_factor, where the reference takes
 * (Nngb/Nngb_previous)**(1/3)).  Each pass counts k = #{j : sqrt(r2_ij) <= s}
 * (self included); k < ndown: s += hstep h_prev; k > nup: s -= hstep h_prev
 * (hstep = coeff * 2: 0.002 with LIBSX3's coeff = 0.001, 0.02 with
 * do_smoothingX4's 0.01); the largest s is kept, and after pass
 * max_iter + 1 (N_iter > 100) s becomes that largest value whatever the
 * count.  The loop ends at the first pass whose count lies in [ndown, nup].
 * k_out is the last pass's count (the reference's Nngbx); status 1 = the
 * pass limit was reached.
 *
 * h_mode 4 is the GPU driver variant of LIBV7:255-328 (hfvGpuSphTyp_v7.cu:28,
 * coeff 0.005): the eq. 31 start summed as ((1/3 h) divV) dt; steps of
 * 0.01 h_prev, halved whenever the count jumps across the band from one pass
 * to the next (k < ndown after k > nup or the reverse); no largest-radius
 * rule: after pass 501 (N_iter > 500) s is the last updated radius; its
 * count stops at nup + 1 once over-full (the reference breaks the j loop
 * there), which is what k_out reports then.
 *
 * h_mode 5 is h_smooth_fast (LIBSX:1133-1184, the band 50 +- 10 in
 * augsphx6.5x.py:426-465): from h = h_prev (no predictor), the count is
 * strict, #{j: sqrt(r^2) < 2h}; k > nup: h -= 0.003 h, k < ndown: h +=
 * 0.003 h (a step on the current h, two roundings); the loop ends at the
 * first count in [ndown, nup].  The reference has no pass limit (its counter
 * is never acted on); here the loop stops after SMOOTH_H5_PASSES counts with
 * status 1 and the last h.  With h_min > 0 (h_smooth_fast_h_minimum_set,
 * LIBSX:1192-1245): once eleven steps are taken, a count made at h < h_min
 * sets h = h_min and ends the loop, in the band or not.
 */
#define SMOOTH_H5_PASSES 10000
static int64_t count_within_sqrt(const double *pos, int64_t n, int64_t i, double s, int strict) {
    int64_t k = 0;
    for (int64_t j = 0; j < n; ++j) {
        const double d = sqrt(pair_r2(pos, i, j));
        if (strict ? d < s : d <= s) ++k;
    }
    return k;
}

static void smooth_steps_one(const double *pos, int64_t n, int64_t i, double h_prev,
                             int32_t k_prev, int use_pred, double divv, double dt, int32_t nngb,
                             int32_t ndown, int32_t nup, int32_t max_iter, int32_t h_mode,
                             double h_min, double *h_out, int32_t *k_out, int32_t *it_out,
                             int32_t *st_out) {
    const int halving = h_mode == 4, relative = h_mode == 5;
    double s = 2.0 * h_prev;
    if (use_pred && !relative) {
        double f = orc_h_factor(nngb, k_prev);
        double t1 = 0.5 * h_prev;
        double t2 = 1.0 + f;
        double t3 = t1 * t2;
        double t6;
        if (halving) {
            double t4 = (1.0 / 3.0) * h_prev;
            double t5 = t4 * divv;
            t6 = t5 * dt;
        } else {
            double t4 = h_prev * divv;
            double t5 = (1.0 / 3.0) * t4;
            t6 = t5 * dt;
        }
        double sp = 2.0 * (t3 + t6);
        if (sp > 0.0 && sp < INFINITY) s = sp;
    }
    double coef = h_mode == 3 ? 0.02 : (halving ? 0.01 : 0.002);  /* coeff * 2 */
    const int32_t limit = halving ? 500 : (relative ? SMOOTH_H5_PASSES : max_iter);
    double smax = s;
    int64_t k = 0, kpre = nup;
    int32_t it = 0, st = 0;
    for (;;) {
        k = count_within_sqrt(pos, n, i, s, relative);
        if (halving && k > nup) k = nup + 1;  /* the reference stops counting there */
        const int in_band = k >= ndown && k <= nup;
        /* the steps taken before this count: it (the counter is raised below) */
        const int floor_h = relative && h_min > 0.0 && it > 10 && 0.5 * s < h_min;
        if (halving && ((k < ndown && kpre > nup) || (k > nup && kpre < ndown))) coef = 0.5 * coef;
        if (relative) {
            const double hc = 0.5 * s;
            if (k < ndown) s = 2.0 * (hc + 0.003 * hc);
            if (k > nup) s = 2.0 * (hc - 0.003 * hc);
        } else {
            if (k < ndown) s = s + coef * h_prev;
            if (k > nup) s = s - coef * h_prev;
        }
        if (s > smax) smax = s;
        kpre = k;
        ++it;
        /* a count in the band ends the loop even on the last pass: the
         * reference tests the band before its pass counter */
        if (floor_h) {
            s = 2.0 * h_min;
            break;
        }
        if (in_band) break;
        if (it > limit) {
            if (!halving && !relative) s = smax;
            st = 1;
            break;
        }
    }
    *h_out = 0.5 * s;
    *k_out = (int32_t)k;
    *it_out = it;
    *st_out = st;
}

void orc_smoothing_steps(int64_t n, const double *pos, const double *h_prev, const int32_t *k_prev,
                         const double *divv, double dt, int32_t nngb, int32_t ndown, int32_t nup,
                         int32_t max_iter, int32_t h_mode, double h_min, const int64_t *rows,
                         int64_t nrows,
                         double *h_out, int32_t *k_out, int32_t *it_out, int32_t *st_out) {
    int use_pred = (k_prev != NULL && divv != NULL);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t i = row_of(rows, r);
        smooth_steps_one(pos, n, i, h_prev[i], use_pred ? k_prev[i] : 0, use_pred,
                         use_pred ? divv[i] : 0.0, dt, nngb, ndown, nup, max_iter, h_mode, h_min,
                         &h_out[r], &k_out[r], &it_out[r], &st_out[r]);
    }
}

/* Neighbour lists under the symmetric pair test r2 <= (h_i+h_j)^2 (the gate of
 * LIBV2:89 "rr <= 2*hij").  Two-phase: with idx == NULL only offsets are
 * written (offsets[r+1] = count for row r); then call again to fill. */
void orc_neighbors(int64_t n, const double *pos, const double *h, const int64_t *rows, int64_t nrows,
                   int64_t *offsets, int64_t *idx) {
    if (!idx) {
        offsets[0] = 0;
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t r = 0; r < nrows; ++r) {
            int64_t i = row_of(rows, r);
            int64_t c = 0;
            for (int64_t j = 0; j < n; ++j) {
                double t = h[i] + h[j];
                if (pair_r2(pos, i, j) <= t * t) ++c;
            }
            offsets[r + 1] = c;
        }
        for (int64_t r = 0; r < nrows; ++r) offsets[r + 1] += offsets[r];
        return;
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t i = row_of(rows, r);
        int64_t o = offsets[r];
        for (int64_t j = 0; j < n; ++j) {
            double t = h[i] + h[j];
            if (pair_r2(pos, i, j) <= t * t) idx[o++] = j;
        }
    }
}

/* Cubic-spline kernel W(r, h) (LIBSX:36-45). */
static inline double kernel_w(double q, double hij) {
    double h3 = hij * hij * hij;
    if (q <= 1.0) return ORC_SIGMA / h3 * (1.0 - 1.5 * q * q + 0.75 * q * q * q);
    double w = 2.0 - q;
    if (w < 0.0) w = 0.0;
    return ORC_SIGMA / h3 * 0.25 * w * w * w;
}

/* Gradient factor F with grad_i W = F * (x_i - x_j) (LIBSX:132-150).  The
 * reference's q + 1e-20 guard only matters at q = 0, which this branch never
 * sees (q > 1). */
static inline double kernel_f(double q, double hij) {
    double h5 = hij * hij * hij * hij * hij;
    if (q <= 1.0) return ORC_SIGMA / h5 * (-3.0 + 2.25 * q);
    double w = 2.0 - q;
    if (w < 0.0) w = 0.0;
    return -0.75 * ORC_SIGMA / h5 * w * w / q;
}

/* Density rho_i = sum_j m_j W(r_ij, h_ij), self included (LIBSX:17-51).
 * Also returns the neighbour count under the symmetric pair test. */
void orc_density(int64_t n, const double *pos, const double *m, const double *h,
                 const int64_t *rows, int64_t nrows, double *rho_out, int32_t *nngb_out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t i = row_of(rows, r);
        double s = 0.0;
        int32_t c = 0;
        for (int64_t j = 0; j < n; ++j) {
            double t = h[i] + h[j];
            double r2 = pair_r2(pos, i, j);
            if (r2 <= t * t) {
                double hij = 0.5 * t;
                double q = sqrt(r2) / hij;
                s += m[j] * kernel_w(q, hij);
                ++c;
            }
        }
        rho_out[r] = s;
        if (nngb_out) nngb_out[r] = c;
    }
}

/* Adiabatic EOS: P = (gamma-1) rho u (LIBSX:100-105), c = sqrt(gamma (gamma-1) u) (LIBV2:783-789). */
void orc_eos(int64_t n, const double *rho, const double *u, double gamma, double *P, double *c) {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
        P[i] = (gamma - 1.0) * rho[i] * u[i];
        c[i] = sqrt(gamma * (gamma - 1.0) * u[i]);
    }
}

/* Velocity divergence (signed) and |curl| (SHEAR:8-69). */
void orc_divcurl(int64_t n, const double *pos, const double *v, const double *rho, const double *m,
                 const double *h, const int64_t *rows, int64_t nrows, double *div_out,
                 double *curl_out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t i = row_of(rows, r);
        double sd = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            double t = h[i] + h[j];
            double r2 = pair_r2(pos, i, j);
            if (!(r2 <= t * t)) continue;
            double hij = 0.5 * t;
            double q = sqrt(r2) / hij;
            double F = kernel_f(q, hij);
            double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                   dz = pos[3 * i + 2] - pos[3 * j + 2];
            double gx = F * dx, gy = F * dy, gz = F * dz;
            double vxij = v[3 * i] - v[3 * j], vyij = v[3 * i + 1] - v[3 * j + 1],
                   vzij = v[3 * i + 2] - v[3 * j + 2];
            double w = m[j] / rho[i];
            sd += w * (-vxij * gx - vyij * gy - vzij * gz);
            cx += w * (vyij * gz - vzij * gy);
            cy += w * (vzij * gx - vxij * gz);
            cz += w * (vxij * gy - vyij * gx);
        }
        div_out[r] = sd;
        curl_out[r] = sqrt(cx * cx + cy * cy + cz * cz);
    }
}

/* Artificial viscosity Pi_ij for one pair (returns 0 for receding pairs). */
static inline double av_pi(const orc_params *p, double vr, double r2, double hij, double ci,
                           double cj, double rhoi, double rhoj, double fi, double fj) {
    if (!(vr <= 0.0)) return 0.0;
    double rhob = 0.5 * (rhoi + rhoj);
    double pi;
    if (p->av_form == 0) {
        /* Monaghan & Gingold 1983 (AUGX:103-139) */
        double mu = hij * vr / (r2 + hij * hij * p->eta * p->eta);
        double cb = 0.5 * (ci + cj);
        pi = (-p->alpha * cb * mu + p->beta * mu * mu) / rhob;
    } else {
        /* Gadget signal-velocity form (LIBSX:158-169) */
        double w = vr / (sqrt(r2) + p->av_w_eps);
        double vsig = ci + cj - 3.0 * w;
        pi = -0.5 * p->alpha * vsig * w / rhob;
    }
    if (p->balsara) pi *= 0.5 * (fi + fj);
    return pi;
}

/* Balsara factor f = |divV| / (|divV| + |curlV| + 1e-4 c/h) (LIBV2:400-401). */
static inline double balsara_f(double dv, double cv, double c, double h) {
    double a = fabs(dv);
    return a / (a + cv + 0.0001 * c / h);
}

/* Hydro acceleration (LIBSX:110-186, SHEAR:151-235) and du/dt
 * (LIBSX:363-430, SHEAR:334-406) from the same pair loop and velocity. */
void orc_hydro(int64_t n, const double *pos, const double *v, const double *rho, const double *P,
               const double *c, const double *h, const double *m, const double *divv,
               const double *curlv, const orc_params *p, const int64_t *rows, int64_t nrows,
               double *acc_out, double *du_out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t i = row_of(rows, r);
        double ax = 0.0, ay = 0.0, az = 0.0, du = 0.0;
        double Pi = P[i] / (rho[i] * rho[i]);
        double fi = p->balsara ? balsara_f(divv[i], curlv[i], c[i], h[i]) : 1.0;
        for (int64_t j = 0; j < n; ++j) {
            double t = h[i] + h[j];
            double r2 = pair_r2(pos, i, j);
            if (!(r2 <= t * t)) continue;
            double hij = 0.5 * t;
            double q = sqrt(r2) / hij;
            double F = kernel_f(q, hij);
            double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                   dz = pos[3 * i + 2] - pos[3 * j + 2];
            double vxij = v[3 * i] - v[3 * j], vyij = v[3 * i + 1] - v[3 * j + 1],
                   vzij = v[3 * i + 2] - v[3 * j + 2];
            double vr = vxij * dx + vyij * dy + vzij * dz;
            double fj = p->balsara ? balsara_f(divv[j], curlv[j], c[j], h[j]) : 1.0;
            double pij = av_pi(p, vr, r2, hij, c[i], c[j], rho[i], rho[j], fi, fj);
            double Pj = P[j] / (rho[j] * rho[j]);
            double s = m[j] * (Pi + Pj + pij) * F;
            ax -= s * dx;
            ay -= s * dy;
            az -= s * dz;
            du += m[j] * (Pi + 0.5 * pij) * F * vr;
        }
        acc_out[3 * r + 0] = ax;
        acc_out[3 * r + 1] = ay;
        acc_out[3 * r + 2] = az;
        du_out[r] = du;
    }
}

/* Spline-softened gravity (LIBSX:633-671 / getAcc_g_smthx LIBSX:720-756)
 * and, optionally, the per-particle potential phi_i = -G sum_j m_j phi_ij
 * (the pair term of getPE, LIBSX:593-629). */
void orc_gravity(int64_t n, const double *pos, const double *m, const double *eps, double G,
                 const int64_t *rows, int64_t nrows, double *acc_out, double *pot_out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t i = row_of(rows, r);
        double ax = 0.0, ay = 0.0, az = 0.0, ph = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1],
                   dz = pos[3 * j + 2] - pos[3 * i + 2];
            double rr = sqrt(dx * dx + dy * dy + dz * dz);
            double e = 0.5 * (eps[i] + eps[j]);
            double q = rr / e;
            double fk, pk;
            if (q <= 1.0) {
                fk = (1.0 / (e * e * e)) * ((4.0 / 3.0) - (6.0 / 5.0) * q * q + 0.5 * q * q * q);
                pk = (-2.0 / e) * ((1.0 / 3.0) * q * q - (3.0 / 20.0) * q * q * q * q +
                                   (1.0 / 20.0) * q * q * q * q * q) +
                     7.0 / 5.0 / e;
            } else if (q <= 2.0) {
                double q2 = q * q, q3 = q2 * q, q4 = q3 * q, q5 = q4 * q, q6 = q5 * q;
                fk = (1.0 / (rr * rr * rr)) *
                     ((-1.0 / 15.0) + (8.0 / 3.0) * q3 - 3.0 * q4 + (6.0 / 5.0) * q5 - (1.0 / 6.0) * q6);
                pk = (-1.0 / 15.0) / rr -
                     (1.0 / e) * ((4.0 / 3.0) * q2 - q3 + (3.0 / 10.0) * q4 - (1.0 / 30.0) * q5) +
                     8.0 / 5.0 / e;
            } else {
                fk = 1.0 / (rr * rr * rr);
                pk = 1.0 / rr;
            }
            ax += G * fk * dx * m[j];
            ay += G * fk * dy * m[j];
            az += G * fk * dz * m[j];
            ph -= G * m[j] * pk;
        }
        acc_out[3 * r + 0] = ax;
        acc_out[3 * r + 1] = ay;
        acc_out[3 * r + 2] = az;
        if (pot_out) pot_out[r] = ph;
    }
}

/* ------------------------------------------------------------------------ */
/* Tabulated cooling with an implicit FP64 solve (structure of             */
/* PHOTOPY:489-543 DoCooling_h and PHOTO:613-686).                          */
/* ------------------------------------------------------------------------ */
#define ORC_MH 1.6726e-24 /* g (hfv_sph_gpu_v2_evrard.cu:289) */
#define ORC_KB 1.3807e-16 /* erg/K (hfv_sph_gpu_v2_evrard.cu:290) */

double orc_lambda(double T, const double *Tt, const double *Lt, int32_t nt) {
    if (T <= Tt[0]) return Lt[0];
    if (T >= Tt[nt - 1]) return Lt[nt - 1];
    int32_t lo = 0, hi = nt - 1; /* Tt[lo] <= T < Tt[hi] */
    while (hi - lo > 1) {
        int32_t mid = (lo + hi) >> 1;
        if (Tt[mid] <= T) lo = mid;
        else hi = mid;
    }
    double f = (T - Tt[lo]) / (Tt[hi] - Tt[lo]);
    return Lt[lo] + (Lt[hi] - Lt[lo]) * f;
}

/* g(u) = u - u_old - dt * (nH*heat - nH^2 Lambda(T(u))) / rho, all cgs. */
static inline double cool_g(double u, double u_old, double dt_s, double nH, double rho_cgs,
                            double tfac, const orc_cool_params *cp, const double *Tt,
                            const double *Lt) {
    double T = tfac * u;
    double lam = orc_lambda(T, Tt, Lt, cp->ntab);
    double rate = (nH * cp->heat - nH * nH * lam) / rho_cgs;
    return u - u_old - dt_s * rate;
}

double orc_cool_one(double rho, double u_code, double dt, const orc_cool_params *cp,
                    const double *Tt, const double *Lt, int32_t *iters) {
    double rho_cgs = rho * cp->unit_rho_cgs;
    double dt_s = dt * cp->unit_time_s;
    double nH = cp->XH * rho_cgs / ORC_MH;
    double tfac = (cp->gamma - 1.0) * cp->mu * ORC_MH / ORC_KB;
    /* the gas never cools below the table's first temperature node */
    double u_floor = Tt[0] / tfac;
    double u_old = u_code * cp->unit_u_cgs;
    if (u_old < u_floor) u_old = u_floor;
    int32_t it = 0;
    if (iters) *iters = 0;
    if (!(dt_s > 0.0)) return u_old / cp->unit_u_cgs;
    double g0 = cool_g(u_old, u_old, dt_s, nH, rho_cgs, tfac, cp, Tt, Lt);
    if (g0 == 0.0) return u_old / cp->unit_u_cgs;
    double lo, hi;
    if (g0 > 0.0) { /* net cooling: root below u_old */
        if (cool_g(u_floor, u_old, dt_s, nH, rho_cgs, tfac, cp, Tt, Lt) >= 0.0)
            return u_floor / cp->unit_u_cgs;
        hi = u_old;
        lo = u_old / 1.1;
        if (lo < u_floor) lo = u_floor;
        int32_t nb = 0;
        while (cool_g(lo, u_old, dt_s, nH, rho_cgs, tfac, cp, Tt, Lt) > 0.0 && nb < cp->max_bracket) {
            hi = lo;
            lo = lo / 1.1;
            if (lo < u_floor) lo = u_floor;
            ++nb;
        }
    } else { /* net heating: root above u_old */
        lo = u_old;
        hi = u_old * 1.1;
        int32_t nb = 0;
        while (cool_g(hi, u_old, dt_s, nH, rho_cgs, tfac, cp, Tt, Lt) < 0.0 && nb < cp->max_bracket) {
            lo = hi;
            hi = hi * 1.1;
            ++nb;
        }
    }
    while ((hi - lo) > cp->tol * lo && it < cp->max_iter) {
        double mid = 0.5 * (lo + hi);
        if (cool_g(mid, u_old, dt_s, nH, rho_cgs, tfac, cp, Tt, Lt) > 0.0) hi = mid;
        else lo = mid;
        ++it;
    }
    if (iters) *iters = it;
    return 0.5 * (lo + hi) / cp->unit_u_cgs;
}

void orc_cool(int64_t n, const double *rho, const double *u, double dt, const orc_cool_params *cp,
              const double *Tt, const double *Lt, double *u_out, int32_t *it_out) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        int32_t it = 0;
        u_out[i] = orc_cool_one(rho[i], u[i], dt, cp, Tt, Lt, &it);
        if (it_out) it_out[i] = it;
    }
}

/* Kinetic energy sum 0.5 m v^2 (LIBSX:579-588). */
double orc_ke(int64_t n, const double *v, const double *m) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i)
        s += 0.5 * m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    return s;
}

/* Potential energy sum_{i<j} -G m_i m_j phi_ij (LIBSX:593-629), from the
 * per-particle potential: PE = 0.5 * sum_i m_i phi_i. */
double orc_pe(int64_t n, const double *pos, const double *m, const double *eps, double G) {
    double *phi = (double *)malloc(sizeof(double) * (size_t)n);
    double *acc = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    orc_gravity(n, pos, m, eps, G, NULL, n, acc, phi);
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += 0.5 * m[i] * phi[i];
    free(phi);
    free(acc);
    return s;
}

/* Global adaptive time step of the GPU driver: dt_array (LIBV2:491-522) and
 * the host reduction after it (DRV2:517-561).  The reference's O(N^2) v_sig
 * loop is max_j max(c_i, c_j) = max over all c.  With hmin = min h:
 *   dt_f    = sqrt(hmin / max |a_grav|)
 *   dt_kin  = sqrt(hmin / max |a_total|)
 *   dt_cfl  = c_cfl * hmin / max c
 *   dt_dens = c_cfl * min_i h_i / |h_i divV_i / 3|
 *   dt      = 0.25 * min(dt_f, dt_kin, dt_cfl, dt_dens), clamped to
 *             [dt_min, dt_max] (1e-5 and 1e-3 in the driver).
 * parts (nullable) receives dt_f, dt_kin, dt_cfl, dt_dens. */
double orc_global_dt(int64_t n, const double *h, const double *c, const double *acc,
                     const double *acc_g, const double *divv, double c_cfl, double dt_min,
                     double dt_max, double *parts) {
    double vmax = 0.0, hmin = INFINITY, agmax = 0.0, amax = 0.0, ddmin = INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        const double *a = acc + 3 * i, *g = acc_g + 3 * i;
        double ag = sqrt((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
        double at = sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
        double dd = h[i] / fabs((1.0 / 3.0) * h[i] * divv[i]);
        vmax = c[i] > vmax ? c[i] : vmax;
        hmin = h[i] < hmin ? h[i] : hmin;
        agmax = ag > agmax ? ag : agmax;
        amax = at > amax ? at : amax;
        ddmin = dd < ddmin ? dd : ddmin;
    }
    double dt_f = sqrt(hmin / agmax), dt_kin = sqrt(hmin / amax);
    double dt_cfl = c_cfl * hmin / vmax, dt_dens = c_cfl * ddmin;
    double m = dt_f;
    m = dt_kin < m ? dt_kin : m;
    m = dt_cfl < m ? dt_cfl : m;
    m = dt_dens < m ? dt_dens : m;
    double dt = 0.25 * m;
    if (dt > dt_max) dt = dt_max;
    if (dt < dt_min) dt = dt_min;
    if (parts) {
        parts[0] = dt_f;
        parts[1] = dt_kin;
        parts[2] = dt_cfl;
        parts[3] = dt_dens;
    }
    return dt;
}

/* 2-D (u, rho) cooling table of the reference's v3 driver: applyCooling
 * (GPU_sph/With_cooling/Model_A/myCppSPHLibs_v3.h:794-859; driver
 * hfv_sph_gpu_v3.cu:115-204, 496-508).  u_nodes[nu], rho_nodes[nrho] are the
 * increasing cgs grid nodes, du[j * nrho + i] the cgs energy change of a gas
 * at (u_nodes[j], rho_nodes[i]) over dt_ref_s.  Per particle:
 *   u, rho -> cgs; clamp below the first node to node + 0.001 node and above
 *   the last to node - 0.001 node; j = first bin with g[j] <= x <= g[j+1]
 *   (the reference's linear scans); u_new = u_nodes[j] - (dt_s / dt_ref_s) du
 *   back to code units.  The reference sets no floor; here u_new is floored at
 *   the first u node, so the gas never leaves the table. */
static int64_t orc_grid_bin(const double *g, int64_t ng, double x) {
    for (int64_t j = 0; j < ng - 1; ++j)
        if (x >= g[j] && x <= g[j + 1]) return j;
    return ng - 2;
}

void orc_cool_grid_apply(int64_t n, const double *rho, const double *u, double dt, int64_t nu,
                         const double *u_nodes, int64_t nrho, const double *rho_nodes,
                         const double *du, double dt_ref_s, double unit_u_cgs,
                         double unit_rho_cgs, double unit_time_s, double *u_out) {
    double coeff = (dt * unit_time_s) / dt_ref_s;
    double umin = u_nodes[0], umax = u_nodes[nu - 1];
    double rmin = rho_nodes[0], rmax = rho_nodes[nrho - 1];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double uc = u[i] * unit_u_cgs, rc = rho[i] * unit_rho_cgs;
        if (uc < umin) uc = umin + 0.001 * umin;
        if (uc > umax) uc = umax - 0.001 * umax;
        if (rc < rmin) rc = rmin + 0.001 * rmin;
        if (rc > rmax) rc = rmax - 0.001 * rmax;
        int64_t ju = orc_grid_bin(u_nodes, nu, uc), jr = orc_grid_bin(rho_nodes, nrho, rc);
        double un = u_nodes[ju] - coeff * du[ju * nrho + jr];
        if (un < umin) un = umin;
        u_out[i] = un / unit_u_cgs;
    }
}

/* The table itself (GPU_sph/Cooling_v1/createCoolingGridX2.cu:9-35): the
 * energy change of the implicit solve over dt_ref_s at every node, from this
 * oracle's own solve (orc_cool_one): du = u - u_new, in cgs. */
void orc_cool_grid_build(int64_t nu, const double *u_nodes, int64_t nrho, const double *rho_nodes,
                         double dt_ref_s, const orc_cool_params *cp, const double *Tt,
                         const double *Lt, double *du) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t j = 0; j < nu; ++j)
        for (int64_t i = 0; i < nrho; ++i) {
            double u_code = u_nodes[j] / cp->unit_u_cgs;
            double rho_code = rho_nodes[i] / cp->unit_rho_cgs;
            double dt_code = dt_ref_s / cp->unit_time_s;
            int32_t it = 0;
            double un = orc_cool_one(rho_code, u_code, dt_code, cp, Tt, Lt, &it);
            du[j * nrho + i] = (u_code - un) * cp->unit_u_cgs;
        }
}

/* kNN smoothing length of do_smoothingX((pos, subpos)) (LIBSX:762-783,
 * AUG:305-326): h = 0.5 * (the kth smallest distance, 0-based, self included;
 * the reference takes np.sort(dist)[64]) from each target row to all n
 * particles, distances sqrt((dx^2 + dy^2) + dz^2) with dx = pos_j - pos_i.
 * sqrt is monotonic, so this is 0.5 * sqrt(kth smallest r^2); found with a
 * quickselect on r^2. */
static double orc_kth_smallest(double *v, int64_t n, int64_t k) {
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        double piv = v[lo + (hi - lo) / 2];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (v[i] < piv) ++i;
            while (v[j] > piv) --j;
            if (i <= j) {
                double t = v[i];
                v[i] = v[j];
                v[j] = t;
                ++i;
                --j;
            }
        }
        if (k <= j) hi = j;
        else if (k >= i) lo = i;
        else return v[k];
    }
    return v[k];
}

void orc_knn_h(int64_t n, const double *pos, int64_t kth, const int64_t *rows, int64_t nrows,
               double *h_out) {
#pragma omp parallel
    {
        double *r2 = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(dynamic, 16)
        for (int64_t t = 0; t < nrows; ++t) {
            int64_t i = rows ? rows[t] : t;
            for (int64_t j = 0; j < n; ++j) {
                double dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1],
                       dz = pos[3 * j + 2] - pos[3 * i + 2];
                double a = dx * dx, b = dy * dy, c = dz * dz;
                r2[j] = (a + b) + c;
            }
            h_out[t] = 0.5 * sqrt(orc_kth_smallest(r2, n, kth));
        }
        free(r2);
    }
}

/* Lookup in the 5-D CHIMES table (hcoolingx, LIBV7:1077-1240; its CPU twin is
 * hcooler in GPU_sph_Type_cooling_II/cooling_grids_use_EC2/checker1.py:53-146).
 * The table holds, per (radius in kpc, T node, n_H node, metallicity node,
 * elapsed time), the specific energy reached after that time from the node's
 * initial energy, the mean molecular weight, and n_m species fractions;
 * layouts [kpc][T][nH][Z][time] and [kpc][T][nH][Z][species][time], as the
 * writer GPU_sph_Type_cooling_II/cooling_grids_use_EC2/step_2x.py lays them
 * out.  Per particle, in FP64 on the float32 table values:
 *   - radius r = |x| in kpc, n_H = rho_cgs X_H / m_H: the nearest node, found
 *     as the reference does -- the first node >= the value, moved one down
 *     when the node below is strictly closer; past the last node -> last;
 *   - metallicity node fixed (the reference hard-codes index 1);
 *   - U[t] = table energy at time index 0 of the column (the node energies);
 *     j_u = first t with u_cgs <= U[t], else the last node;
 *   - j_t = first time node >= dt in seconds, else the last;
 *   - j_u > 0: linear interpolation in u between U[j_u-1] and U[j_u] of the
 *     energies at j_t (extrapolating past the last node, as the reference);
 *     j_u == 0: the node value;
 *   - capped at u_max_cgs (T = 1e10 K in the reference), back to code units.
 * ion_out (n x n_m) and mu_out receive the species fractions and mu at node
 * j_u (nullable).  The reference stages U in one global buffer shared by all
 * threads; here every particle reads its own column. */
typedef struct {
    int32_t n_kpc, n_nh, n_z, n_t, n_m, n_time;
    int32_t z_index, pad;
    double xh, m_h, u_max_cgs;
    double unit_length_cm, kpc_cm, unit_rho_cgs, unit_u_cgs, unit_time_s;
} orc_chimes_params;

static int64_t orc_first_ge(const float *g, int64_t ng, int64_t stride, double x) {
    for (int64_t k = 0; k < ng; ++k)
        if (x <= (double)g[k * stride]) return k;
    return -1;
}

static int64_t orc_nearest_node(const float *g, int64_t ng, double x) {
    int64_t j = orc_first_ge(g, ng, 1, x);
    if (j > 0) {
        double d1 = fabs((double)g[j] - x), d2 = fabs((double)g[j - 1] - x);
        if (d2 < d1) j -= 1;
    }
    return j < 0 ? ng - 1 : j;
}

void orc_chimes_apply(int64_t n, const double *pos, const double *rho, const double *u, double dt,
                      const orc_chimes_params *cp, const float *kpc, const float *nh,
                      const float *time_s, const float *u_evol, const float *mu,
                      const float *metalz, double *u_out, float *ion_out, double *mu_out) {
    const int64_t NT = cp->n_t, NH = cp->n_nh, NZ = cp->n_z, NM = cp->n_m, NTIME = cp->n_time;
    const int64_t s_t = NH * NZ * NTIME, s_k = NT * s_t;  /* strides of [kpc][T][nH][Z][time] */
    double dt_s = dt * cp->unit_time_s;
    int64_t jt = orc_first_ge(time_s, NTIME, 1, dt_s);
    if (jt < 0) jt = NTIME - 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        double a = x * x, b = y * y, c = z * z;
        double r = sqrt((a + b) + c);
        double r_kpc = (r * cp->unit_length_cm) / cp->kpc_cm;
        int64_t jk = orc_nearest_node(kpc, cp->n_kpc, r_kpc);
        double nh_i = ((rho[i] * cp->unit_rho_cgs) * cp->xh) / cp->m_h;
        int64_t jn = orc_nearest_node(nh, NH, nh_i);
        int64_t jz = cp->z_index;
        int64_t col = jk * s_k + jn * NZ * NTIME + jz * NTIME;  /* + T * s_t + time */
        double u_i = u[i] * cp->unit_u_cgs;
        int64_t ju = orc_first_ge(u_evol + col, NT, s_t, u_i);
        if (ju < 0) ju = NT - 1;
        double uev;
        if (ju > 0) {
            double U0 = u_evol[col + (ju - 1) * s_t], U1 = u_evol[col + ju * s_t];
            double e0 = u_evol[col + (ju - 1) * s_t + jt], e1 = u_evol[col + ju * s_t + jt];
            double fhi = (u_i - U0) / (U1 - U0);
            double flo = 1.0 - fhi;
            uev = flo * e0 + fhi * e1;
        } else {
            uev = u_evol[col + jt];
        }
        if (uev > cp->u_max_cgs) uev = cp->u_max_cgs;
        u_out[i] = uev / cp->unit_u_cgs;
        if (mu_out) mu_out[i] = mu[col + ju * s_t + jt];
        if (ion_out) {
            int64_t mcol = ((jk * NT + ju) * NH + jn) * NZ * NM * NTIME + jz * NM * NTIME + jt;
            for (int64_t k = 0; k < NM; ++k) ion_out[i * NM + k] = metalz[mcol + k * NTIME];
        }
    }
}

/* ---------------------------------------------------------------------------
 * Analytic H/He cooling network of PHOTOPY (the FP64 v3 functions the
 * reference's DoCooling_h calls), restated function by function:
 *   orc_net_sigma        phCrossSectionX   PHOTOPY:120-134 (Verner et al. 1996)
 *   orc_net_rates        RadiationField    PHOTOPY:230-329 (J_neu_func :215-227)
 *   orc_net_ci           RandCIRates       PHOTOPY:140-166 (Cen 1992)
 *   orc_net_abundance    Abundance_hX      PHOTOPY:169-211
 *   orc_net_heat_cool    coolingHeatingRates PHOTOPY:332-376
 *   orc_net_u_to_temp    convert_u_to_temp_h PHOTOPY:379-422
 *   orc_net_temp_to_u    convert_Temp_to_u PHOTOPY:425-467
 *   orc_net_rate         coolingRateFromU_h PHOTOPY:470-486
 *   orc_net_cool_one     DoCooling_h       PHOTOPY:489-543
 * The radiation field is parameterised: J21 (the reference's Jcoeff: 1e-5 in
 * PHOTOPY, 50 in GPU_sph/With_cooling/Model_A/myPhotoLibsGPU.h:166-296) and
 * the frequency grid (n_grid 0: np.arange(nu_edge, 6e16, 0.02e16) as PHOTOPY;
 * n_grid > 0: n_grid points of (6e16 - nu_edge) / n_grid as the GPU header).
 * The bracket loops of DoCooling_h have no cap in the reference; here at most
 * max_bracket steps (counted as capped).
 * ------------------------------------------------------------------------- */
typedef struct {
    double XH, J21;
    double unit_u_cgs, unit_rho_cgs, unit_time_s;
    double tol;
    int32_t n_grid, max_iter, max_bracket, pad;
} orc_net_params;

typedef struct {
    double gJH0, gJHe0, gJHep, HH0, HHe0, HHep;
} orc_net_rf;

#define ORC_H_EV 4.135667696e-15
#define ORC_CLIGHT 29979245800.0
#define ORC_NET_KB 1.3807e-16
#define ORC_NET_MH 1.6726e-24

static double orc_net_neu_to_eV(double neu) { return ORC_H_EV * neu; }
static double orc_net_eV_to_neu(double eV) { return eV / ORC_H_EV; }
static double orc_net_lamb_to_neu(double lamb) { return ORC_CLIGHT / (lamb * 1e-8); }
static double orc_net_neu_to_lamb(double neu) { return ORC_CLIGHT / neu * 1e8; }

double orc_net_sigma(double neu, double E0, double sig0, double ya, double P, double yw,
                     double y0, double y1) {
    double E = orc_net_neu_to_eV(neu);
    double x = E / E0 - y0;
    double y = sqrt(x * x + y1 * y1);
    double Fy = (((x - 1.0) * (x - 1.0) + yw * yw) * pow(y, 0.5 * P - 5.5)) *
                pow(1.0 + sqrt(y / ya), -P);
    return sig0 * Fy * 1e-18;
}

/* One edge of RadiationField: photoionisation rate g and heating rate H for
 * the species with threshold E (eV) and cross-section parameters c[7]. */
static void orc_net_edge(const orc_net_params *p, double E, const double *c, double *g,
                         double *H) {
    const double hplanck = 6.626196e-27;
    double start = orc_net_eV_to_neu(E) / 1e16;
    int64_t n;
    double step;
    if (p->n_grid > 0) {
        n = p->n_grid;
        step = (6.0 - start) / (double)n;
    } else {
        /* np.arange(start, 6.0, 0.02): ceil((stop - start) / step) values
         * start + i * delta with delta = (start + step) - start */
        n = (int64_t)ceil((6.0 - start) / 0.02);
        step = (start + 0.02) - start;
    }
    double nu_edge = orc_net_lamb_to_neu(orc_net_neu_to_lamb(orc_net_eV_to_neu(E)));
    double nu0 = orc_net_eV_to_neu(E);
    double *nu = (double *)malloc(sizeof(double) * (size_t)n);
    double *fg = (double *)malloc(sizeof(double) * (size_t)n);
    double *fh = (double *)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) nu[i] = (start + (double)i * step) * 1e16;
    for (int64_t i = 0; i < n; ++i) {
        double sig = orc_net_sigma(nu[i], c[0], c[1], c[2], c[3], c[4], c[5], c[6]);
        double J = p->J21 * 1e-21 / (nu[i] / nu_edge);
        fg[i] = 4.0 * M_PI * J * sig / hplanck / nu[i];
        fh[i] = 4.0 * M_PI * J * sig * (nu[i] - nu0) / nu[i];
    }
    double d = n > 1 ? nu[1] - nu[0] : 0.0;
    double sg = 0.0, sh = 0.0;
    for (int64_t i = 0; i + 1 < n; ++i) {
        sg += d * (fg[i] + fg[i + 1]) / 2.0;
        sh += d * (fh[i] + fh[i + 1]) / 2.0;
    }
    free(nu); free(fg); free(fh);
    *g = sg;
    *H = sh;
}

void orc_net_rates(const orc_net_params *p, double *out6) {
    static const double cH0[7] = {4.298e-1, 5.475e4, 3.288e1, 2.963, 0.0, 0.0, 0.0};
    static const double cHe0[7] = {1.361e1, 9.492e2, 1.469e0, 3.188, 2.039, 4.434e-1, 2.136};
    static const double cHep[7] = {1.720e0, 1.369e4, 3.288e1, 2.963, 0.0, 0.0, 0.0};
    orc_net_edge(p, 13.60, cH0, &out6[0], &out6[3]);
    orc_net_edge(p, 24.59, cHe0, &out6[1], &out6[4]);
    orc_net_edge(p, 54.42, cHep, &out6[2], &out6[5]);
}

/* aHp, aHep, aHepp, ad, geH0, geHe0, geHep */
void orc_net_ci(double T, double *r7) {
    double Tfact = 1.0 / (1.0 + sqrt(T / 1e5));
    double aHp = 8.41e-11 * pow(T / 1000.0, -0.2) / (1.0 + pow(T / 1e6, 0.7)) / sqrt(T);
    r7[0] = aHp;
    r7[1] = 1.5e-10 * pow(T, -0.6353);
    r7[2] = 4.0 * aHp;
    r7[3] = 1.9e-3 * pow(T, -1.5) * exp(-470000.0 / T) * (1.0 + 0.3 * exp(-94000.0 / T));
    r7[4] = 5.85e-11 * sqrt(T) * exp(-157809.1 / T) * Tfact;
    r7[5] = 2.38e-11 * sqrt(T) * exp(-285335.4 / T) * Tfact;
    r7[6] = 5.68e-12 * sqrt(T) * exp(-631515.0 / T) * Tfact;
}

/* nH0, nHe0, nHp, ne, nHep, nHepp relative to nH; iters (nullable) */
void orc_net_abundance(double T, double nH, const double *rf, double *ab6, int32_t *iters) {
    double r[7];
    orc_net_ci(T, r);
    double aHp = r[0], aHep = r[1], aHepp = r[2], ad = r[3], geH0 = r[4], geHe0 = r[5],
           geHep = r[6];
    double gJH0 = rf[0], gJHe0 = rf[1], gJHep = rf[2];
    double Y = 0.24;
    double y = Y / (4.0 - 4.0 * Y);
    double ne = 1.0, ne_old;
    double nH0 = 0, nHp = 0, nHep = 0, nHe0 = 0, nHepp = 0;
    int niter = 1;
    double dne = 1.0;
    while (dne > 1e-4 && niter < 100) {
        ne_old = ne;
        nH0 = aHp / (aHp + geH0 + gJH0 / (ne * nH));
        nHp = 1.0 - nH0;
        nHep = y / (1.0 + (aHep + ad) / (geHe0 + gJHe0 / (ne * nH)) +
                    (geHep + gJHep / (ne * nH)) / aHepp);
        nHe0 = nHep * (aHep + ad) / (geHe0 + gJHe0 / (ne * nH));
        nHepp = nHep * (geHep + gJHep / (ne * nH)) / aHepp;
        ne = nHp + nHep + 2.0 * nHepp;
        ne = 0.5 * (ne + ne_old);
        dne = fabs(ne - ne_old);
        ++niter;
    }
    ab6[0] = nH0; ab6[1] = nHe0; ab6[2] = nHp; ab6[3] = ne; ab6[4] = nHep; ab6[5] = nHepp;
    if (iters) *iters = niter;
}

/* heating and cooling per nH^2: out2 = {heat, lambda} */
void orc_net_heat_cool(double T, double nH, const double *rf, double *out2) {
    double r[7], ab[6];
    orc_net_ci(T, r);
    orc_net_abundance(T, nH, rf, ab, NULL);
    double aHp = r[0], aHep = r[1], aHepp = r[2], ad = r[3], geH0 = r[4], geHe0 = r[5],
           geHep = r[6];
    double nH0 = ab[0], nHe0 = ab[1], nHp = ab[2], ne = ab[3], nHep = ab[4], nHepp = ab[5];
    double heat = nH0 * rf[3] / nH + nHe0 * rf[4] / nH + nHep * rf[5] / nH;
    double Tfact = 1.0 / (1.0 + sqrt(T / 1e5));
    double BetaH0 = 7.5e-19 * exp(-118348.0 / T) * Tfact;
    double BetaHep = 5.54e-17 * pow(T, -0.397) * exp(-473638.0 / T) * Tfact;
    double lt = 5.5 - log10(T);
    double Betaff = 1.43e-27 * sqrt(T) * (1.0 + 0.34 * exp(-(lt * lt) / 3.0));
    double exc = BetaH0 * ne * nH0 + BetaHep * ne * nHep;
    double ion = 2.18e-11 * geH0 * ne * nH0 + 3.94e-11 * geHe0 * ne * nHe0 +
                 8.72e-11 * geHep * ne * nHep;
    double rec = 1.036e-16 * T * ne * (aHp * nHp) + 1.036e-16 * T * ne * (aHep * nHep) +
                 1.036e-16 * T * ne * (aHepp * nHepp) + 6.526e-11 * ad * ne * nHep;
    double ff = Betaff * (nHp + nHep + 4.0 * nHepp) * ne;
    out2[0] = heat;
    out2[1] = exc + ion + rec + ff;
}

double orc_net_u_to_temp(double u, double nH, double XH, const double *rf, int32_t *iters) {
    const double gamma = 5.0 / 3.0;
    double yHe = (1.0 - XH) / (4.0 * XH);
    double ne_guess = 1.0;
    double mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + ne_guess);
    double temp = (gamma - 1.0) * ORC_NET_MH / ORC_NET_KB * mu * u;
    double temp_old;
    int niter = 1;
    double dT = temp;
    double ab[6];
    while (dT / temp > 1e-4 && niter < 100) {
        temp_old = temp;
        orc_net_abundance(temp_old, nH, rf, ab, NULL);
        ne_guess = ab[3];
        mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + ne_guess);
        temp = (gamma - 1.0) * ORC_NET_MH / ORC_NET_KB * mu * u;
        temp = 0.5 * (temp + temp_old);
        dT = fabs(temp - temp_old);
        ++niter;
    }
    if (iters) *iters = niter;
    return temp;
}

double orc_net_temp_to_u(double temp, double nH, double XH, const double *rf) {
    const double gamma = 5.0 / 3.0;
    double yHe = (1.0 - XH) / (4.0 * XH);
    double ne_guess = 1.0;
    double mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + ne_guess);
    double u = ORC_NET_KB / ORC_NET_MH / (gamma - 1.0) / mu * temp;
    double u_old;
    int niter = 1;
    double du = u;
    double ab[6];
    while (du / u > 1e-4 && niter < 100) {
        u_old = u;
        double t_old = orc_net_u_to_temp(u_old, nH, XH, rf, NULL);
        orc_net_abundance(t_old, nH, rf, ab, NULL);
        ne_guess = ab[3];
        mu = (1.0 + 4.0 * yHe) / (1.0 + yHe + ne_guess);
        u = ORC_NET_KB / ORC_NET_MH / (gamma - 1.0) / mu * temp;
        u = 0.5 * (u + u_old);
        du = fabs(u - u_old);
        ++niter;
    }
    return u;
}

double orc_net_rate(double u, double nH, double XH, const double *rf) {
    double T = orc_net_u_to_temp(u, nH, XH, rf, NULL);
    if (T < 1e4) T = 1e4;
    if (T > 1e8) T = 1e8;
    double hc[2];
    orc_net_heat_cool(T, nH, rf, hc);
    return hc[0] - hc[1];
}

/* DoCooling_h in cgs (rho g/cm^3, u erg/g, dt s); *status: 0 converged,
 * 1 bisection cap, 2 bracket cap. */
double orc_net_cool_one(double rho, double u_old, double dt, const orc_net_params *p,
                        const double *rf, int32_t *iters, int32_t *status) {
    double XH = p->XH;
    double nH = XH * rho / ORC_NET_MH;
    double ratefact = nH * nH / rho;
    double u = u_old, u_upper = u, u_lower = u;
    int st = 0;
    double G = orc_net_rate(u, nH, XH, rf);
    double g0 = u - u_old - ratefact * G * dt;
    if (iters) *iters = 0;
    if (g0 == 0.0) {
        if (status) *status = 0;
        return u_old;
    }
    int nb = 0;
    if (g0 > 0.0) {
        while (u_upper - u_old - ratefact * orc_net_rate(u_upper, nH, XH, rf) * dt > 0.0) {
            if (nb++ >= p->max_bracket) { st = 2; break; }
            u_upper /= 1.1;
        }
        u_lower = u_upper;
        u_upper = u_lower * 1.1;
    } else {
        while (u_upper - u_old - ratefact * orc_net_rate(u_upper, nH, XH, rf) * dt < 0.0) {
            if (nb++ >= p->max_bracket) { st = 2; break; }
            u_upper *= 1.1;
        }
        u_lower = u_upper / 1.1;
    }
    int niter = 1;
    double du = u;
    while (fabs(du / u) > p->tol && niter < p->max_iter) {
        u = 0.5 * (u_lower + u_upper);
        G = orc_net_rate(u, nH, XH, rf);
        if (u - u_old - ratefact * G * dt > 0.0) u_upper = u;
        else u_lower = u;
        du = fabs(u_upper - u_lower);
        ++niter;
    }
    if (niter >= p->max_iter && st == 0) st = 1;
    if (iters) *iters = niter - 1;
    if (status) *status = st;
    return u;
}

/* Arrays in code units (unit factors of p); rf from orc_net_rates. */
void orc_net_cool(int64_t n, const double *rho, const double *u, double dt,
                  const orc_net_params *p, const double *rf, double *u_out, int32_t *it_out,
                  int32_t *st_out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        int32_t it = 0, st = 0;
        double r = orc_net_cool_one(rho[i] * p->unit_rho_cgs, u[i] * p->unit_u_cgs,
                                    dt * p->unit_time_s, p, rf, &it, &st);
        u_out[i] = r / p->unit_u_cgs;
        if (it_out) it_out[i] = it;
        if (st_out) st_out[i] = st;
    }
}
